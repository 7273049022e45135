// tl_file.hip -- a persistent file format for laid-out (or sorted) key arrays,
// so that a layout built once can be stored and loaded straight back into
// device memory.  The reference has no serialisation.
//
// File (little-endian):
//   [0, 256)        header (LayoutHeader below)
//   [tables_offset) the query table set of a vEB layout: table_size entries of
//                   {u64 L, u64 R, u32 D, u32 pad} (tl_veb_table's 24-byte
//                   stride), then num_tables u64 idx, then num_tables u32
//                   table_idx; empty for the other layouts
//   [data_offset)   the n keys as laid out, w bytes each (4096-aligned offset)
// The header carries a checksum of the keys computed on the device, position
// sensitive (sum over i of mix(key_i + i*phi)), recomputed on load.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "tl_internal.hpp"
#include "tl_plan.hpp"

namespace tl {

constexpr char LAYOUT_MAGIC[8] = {'T', 'L', 'L', 'A', 'Y', 'O', 'U', 'T'};
constexpr uint32_t LAYOUT_VERSION = 1;
constexpr uint64_t FILE_ALIGN = 4096;
constexpr size_t IO_CHUNK = 64u << 20;

struct LayoutHeader {
    char magic[8];
    uint32_t version, header_bytes;
    uint32_t layout, dtype;
    uint64_t n, b;
    uint32_t d, num_tables, table_size, reserved;
    uint64_t tables_offset, tables_bytes;
    uint64_t data_offset, data_bytes;
    uint64_t checksum;
    uint8_t pad[256 - 96];
};
static_assert(sizeof(LayoutHeader) == 256, "header is 256 bytes");

__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename U>
__global__ __launch_bounds__(256) void k_checksum(const U *__restrict__ a, uint64_t n,
                                                  unsigned long long *out) {
    uint64_t s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        s += mix64((uint64_t)a[i] + i * 0x9E3779B97F4A7C15ull);
    __shared__ uint64_t part[256];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) part[threadIdx.x] += part[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)part[0]);
}

static tl_status checksum(Ctx *c, const void *A, uint64_t n, uint32_t W, hipStream_t s,
                          uint64_t *out) {
    unsigned long long *d = nullptr;
    TL_TRY(scratch(c, 3, 64, (void **)&d));
    TL_HIP(hipMemsetAsync(d, 0, 8, s));
    const uint32_t g = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((n + 255) / 256, (uint64_t)c->num_cus * 8));
    if (W == 4)
        hipLaunchKernelGGL(k_checksum<uint32_t>, dim3(g), dim3(256), 0, s, (const uint32_t *)A, n,
                           d);
    else
        hipLaunchKernelGGL(k_checksum<uint64_t>, dim3(g), dim3(256), 0, s, (const uint64_t *)A, n,
                           d);
    TL_LAUNCH_CHECK();
    unsigned long long h = 0;
    TL_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, s));
    TL_HIP(hipStreamSynchronize(s));
    *out = h;
    return TL_OK;
}

// pinned staging buffer for the chunked copies, freed on scope exit
struct Pinned {
    void *p = nullptr;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

static tl_status io_error(const std::string &what, const char *path) {
    set_error(what + ": " + path);
    return TL_EIO;
}

static bool valid_layout(uint32_t layout) { return layout <= 3 || layout == TL_FILE_SORTED; }

static tl_status read_header(FILE *f, const char *path, LayoutHeader &h) {
    if (fread(&h, sizeof(h), 1, f) != 1) return io_error("cannot read the layout header", path);
    uint32_t W;
    if (memcmp(h.magic, LAYOUT_MAGIC, 8) != 0) return io_error("not a layout file", path);
    if (h.version != LAYOUT_VERSION || h.header_bytes != sizeof(LayoutHeader))
        return io_error("unsupported layout file version", path);
    if (!width_of((tl_dtype)h.dtype, W) || !valid_layout(h.layout) || h.data_bytes != h.n * W ||
        h.data_offset % FILE_ALIGN != 0 || h.data_offset < sizeof(LayoutHeader) ||
        h.tables_bytes != (uint64_t)h.table_size * 24 + (uint64_t)h.num_tables * 12 ||
        (h.tables_bytes && h.tables_offset + h.tables_bytes > h.data_offset) ||
        h.num_tables > 64 || h.table_size > 64 * 64)
        return io_error("inconsistent layout header", path);
    return TL_OK;
}

static void fill_info(const LayoutHeader &h, tl_layout_info *info) {
    if (!info) return;
    info->layout = h.layout;
    info->dtype = (tl_dtype)h.dtype;
    info->n = h.n;
    info->b = h.b;
    info->d = h.d;
    info->num_tables = h.num_tables;
    info->table_size = h.table_size;
    info->version = h.version;
    info->checksum = h.checksum;
}

}   // namespace tl

using namespace tl;

extern "C" tl_status tl_layout_save(const char *path, const void *dev_A, uint64_t n,
                                    uint32_t layout, uint64_t b, tl_dtype dtype, void *stream) {
    uint32_t W;
    if (!path || !width_of(dtype, W) || !valid_layout(layout) || (n && !dev_A) ||
        (layout == 1 && b == 0)) {
        set_error("tl_layout_save: bad arguments");
        return TL_EINVAL;
    }
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    LayoutHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, LAYOUT_MAGIC, 8);
    h.version = LAYOUT_VERSION;
    h.header_bytes = sizeof(LayoutHeader);
    h.layout = layout;
    h.dtype = (uint32_t)dtype;
    h.n = n;
    h.b = layout == 1 ? b : (layout == 0 ? 1 : 0);
    h.d = n ? floor_log2_64(n) + 1 : 0;
    // the vEB query tables, built by the library's own device table-set build
    std::vector<tl_veb_table> tables;
    std::vector<uint64_t> idx;
    std::vector<uint32_t> tidx;
    if ((layout == 2 || layout == 3) && n) {
        void *dbuf = nullptr;
        const size_t tb = 64 * 64 * sizeof(tl_veb_table), ib = 64 * 8, tib = 64 * 4, mb = 16;
        TL_HIP(hipMalloc(&dbuf, tb + ib + tib + mb));
        unsigned char *p = (unsigned char *)dbuf;
        tl_status st =
            layout == 2
                ? tl_veb_tableset_device(n, (tl_veb_table *)p, (uint64_t *)(p + tb),
                                         (uint32_t *)(p + tb + ib), (uint32_t *)(p + tb + ib + tib),
                                         stream)
                : tl_veb_classic_tableset_device(n, (tl_veb_table *)p, (uint64_t *)(p + tb),
                                                 (uint32_t *)(p + tb + ib),
                                                 (uint32_t *)(p + tb + ib + tib), stream);
        uint32_t meta[4] = {0, 0, 0, 0};
        if (st == TL_OK && hipStreamSynchronize(s) == hipSuccess &&
            hipMemcpy(meta, p + tb + ib + tib, sizeof(meta), hipMemcpyDeviceToHost) == hipSuccess) {
            tables.resize(meta[2]);
            idx.resize(meta[1]);
            tidx.resize(meta[1]);
            if (hipMemcpy(tables.data(), p, meta[2] * sizeof(tl_veb_table),
                          hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(idx.data(), p + tb, meta[1] * 8, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(tidx.data(), p + tb + ib, meta[1] * 4, hipMemcpyDeviceToHost) !=
                    hipSuccess)
                st = TL_EHIP;
        } else if (st == TL_OK) {
            st = TL_EHIP;
        }
        (void)hipFree(dbuf);
        if (st != TL_OK) {
            set_error("tl_layout_save: table set build failed");
            return st;
        }
        h.d = meta[0];
        h.num_tables = meta[1];
        h.table_size = meta[2];
    }
    h.tables_offset = sizeof(LayoutHeader);
    h.tables_bytes = (uint64_t)h.table_size * 24 + (uint64_t)h.num_tables * 12;
    h.data_offset =
        (sizeof(LayoutHeader) + h.tables_bytes + FILE_ALIGN - 1) / FILE_ALIGN * FILE_ALIGN;
    h.data_bytes = n * W;
    TL_TRY(checksum(c, dev_A, n, W, s, &h.checksum));

    File f;
    f.f = fopen(path, "wb");
    if (!f.f) return io_error("cannot create", path);
    std::vector<unsigned char> head(h.data_offset, 0);
    memcpy(head.data(), &h, sizeof(h));
    unsigned char *tp = head.data() + h.tables_offset;
    for (uint32_t i = 0; i < h.table_size; i++) {
        memcpy(tp, &tables[i].L, 8);
        memcpy(tp + 8, &tables[i].R, 8);
        memcpy(tp + 16, &tables[i].D, 4);
        tp += 24;
    }
    if (h.num_tables) {
        memcpy(tp, idx.data(), h.num_tables * 8);
        memcpy(tp + h.num_tables * 8, tidx.data(), h.num_tables * 4);
    }
    if (fwrite(head.data(), 1, head.size(), f.f) != head.size())
        return io_error("write failed", path);
    Pinned buf;
    if (n) TL_HIP(hipHostMalloc(&buf.p, IO_CHUNK, 0));
    for (uint64_t off = 0; off < h.data_bytes; off += IO_CHUNK) {
        const size_t len = (size_t)std::min<uint64_t>(IO_CHUNK, h.data_bytes - off);
        TL_HIP(hipMemcpyAsync(buf.p, (const unsigned char *)dev_A + off, len,
                              hipMemcpyDeviceToHost, s));
        TL_HIP(hipStreamSynchronize(s));
        if (fwrite(buf.p, 1, len, f.f) != len) return io_error("write failed", path);
    }
    if (fflush(f.f) != 0) return io_error("write failed", path);
    return TL_OK;
}

extern "C" tl_status tl_layout_read_info(const char *path, tl_layout_info *info) {
    if (!path || !info) {
        set_error("tl_layout_read_info: null argument");
        return TL_EINVAL;
    }
    File f;
    f.f = fopen(path, "rb");
    if (!f.f) return io_error("cannot open", path);
    LayoutHeader h;
    TL_TRY(read_header(f.f, path, h));
    fill_info(h, info);
    return TL_OK;
}

extern "C" tl_status tl_layout_load(const char *path, void *dev_A, uint64_t capacity,
                                    tl_veb_table *dev_tables, uint64_t *dev_idx,
                                    uint32_t *dev_table_idx, tl_layout_info *info, void *stream) {
    if (!path) {
        set_error("tl_layout_load: null path");
        return TL_EINVAL;
    }
    File f;
    f.f = fopen(path, "rb");
    if (!f.f) return io_error("cannot open", path);
    LayoutHeader h;
    TL_TRY(read_header(f.f, path, h));
    if (h.n > capacity || (h.n && !dev_A)) {
        set_error("tl_layout_load: the file holds " + std::to_string(h.n) +
                  " keys, the buffer " + std::to_string(capacity));
        return TL_EINVAL;
    }
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    uint32_t W;
    width_of((tl_dtype)h.dtype, W);
    if (h.tables_bytes) {
        std::vector<unsigned char> t(h.tables_bytes);
        if (fseek(f.f, (long)h.tables_offset, SEEK_SET) != 0 ||
            fread(t.data(), 1, t.size(), f.f) != t.size())
            return io_error("cannot read the table set", path);
        std::vector<tl_veb_table> tables(h.table_size);
        for (uint32_t i = 0; i < h.table_size; i++) {
            memcpy(&tables[i].L, &t[24 * (size_t)i], 8);
            memcpy(&tables[i].R, &t[24 * (size_t)i + 8], 8);
            memcpy(&tables[i].D, &t[24 * (size_t)i + 16], 4);
        }
        const unsigned char *ip = t.data() + 24 * (size_t)h.table_size;
        if (dev_tables)
            TL_HIP(hipMemcpy(dev_tables, tables.data(), tables.size() * sizeof(tl_veb_table),
                             hipMemcpyHostToDevice));
        if (dev_idx) TL_HIP(hipMemcpy(dev_idx, ip, h.num_tables * 8, hipMemcpyHostToDevice));
        if (dev_table_idx)
            TL_HIP(hipMemcpy(dev_table_idx, ip + h.num_tables * 8, h.num_tables * 4,
                             hipMemcpyHostToDevice));
    }
    if (fseek(f.f, (long)h.data_offset, SEEK_SET) != 0) return io_error("cannot seek", path);
    Pinned buf;
    if (h.n) TL_HIP(hipHostMalloc(&buf.p, IO_CHUNK, 0));
    for (uint64_t off = 0; off < h.data_bytes; off += IO_CHUNK) {
        const size_t len = (size_t)std::min<uint64_t>(IO_CHUNK, h.data_bytes - off);
        if (fread(buf.p, 1, len, f.f) != len) return io_error("truncated layout file", path);
        TL_HIP(hipMemcpyAsync((unsigned char *)dev_A + off, buf.p, len, hipMemcpyHostToDevice,
                              s));
        TL_HIP(hipStreamSynchronize(s));
    }
    uint64_t sum = 0;
    TL_TRY(checksum(c, dev_A, h.n, W, s, &sum));
    if (sum != h.checksum) {
        set_error(std::string("tl_layout_load: checksum mismatch (corrupt file): ") + path);
        return TL_EIO;
    }
    fill_info(h, info);
    return TL_OK;
}
