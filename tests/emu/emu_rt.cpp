// Runtime of the serial CPU emulation of the batch kernels: a pool of lane
// threads runs one workgroup at a time; __syncthreads is a barrier over the
// workgroup's lanes and a ballot a rendezvous of the wave's lanes.  At the
// end of every workgroup each lane must have passed the same number of
// barriers (and each wave's lanes the same number of ballots), else the run
// aborts naming the sub-task (ticket) that was running.
#include "hip/hip_runtime.h"
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

thread_local dim3 threadIdx, blockIdx;
dim3 blockDim, gridDim;
static thread_local int t_lane = -1;

struct Bar {
    std::mutex m;
    std::condition_variable cv;
    int expected = 0, arrived = 0;
    uint64_t gen = 0, acc = 0, result = 0;
    void reset(int n) { expected = n; arrived = 0; acc = 0; }
    uint64_t arrive(uint64_t bits) {
        std::unique_lock<std::mutex> l(m);
        const uint64_t g = gen;
        acc |= bits;
        if (++arrived == expected) {
            result = acc; acc = 0; arrived = 0; gen++;
            cv.notify_all();
            return result;
        }
        cv.wait(l, [&] { return gen != g; });
        return result;
    }
    void drop() {
        std::unique_lock<std::mutex> l(m);
        expected--;
        if (expected > 0 && arrived == expected) {
            result = acc; acc = 0; arrived = 0; gen++;
            cv.notify_all();
        }
    }
};

static Bar g_block;
static Bar g_wave[16];
static std::vector<uint64_t> g_barriers(1024), g_ballots(1024);
static unsigned char *g_smem = nullptr;
static size_t g_smem_bytes = 0;

void emu_syncthreads() { g_barriers[t_lane]++; g_block.arrive(0); }
uint64_t emu_ballot(bool p) {
    g_ballots[t_lane]++;
    const int w = t_lane / 64;
    return g_wave[w].arrive(p ? (1ull << (t_lane % 64)) : 0ull);
}
void emu_sleep() { std::this_thread::yield(); }
uint64_t emu_clock() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch()).count() / 10;
}
unsigned char *emu_smem() { return g_smem; }

volatile uint64_t emu_ticket = ~0ull;

// a pool of lane threads, reused by every workgroup
static std::mutex g_pm;
static std::condition_variable g_pcv, g_dcv;
static uint64_t g_job = 0;
static int g_nl = 0, g_done = 0;
static const std::function<void()> *g_body = nullptr;
static dim3 g_bdim, g_bid;
static std::vector<std::thread> g_pool;

static void lane_main(int l) {
    uint64_t seen = 0;
    t_lane = l;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(g_pm);
            g_pcv.wait(lk, [&] { return g_job != seen; });
            seen = g_job;
            if (l >= g_nl) continue;
        }
        threadIdx = dim3(l % g_bdim.x, (l / g_bdim.x) % g_bdim.y, l / (g_bdim.x * g_bdim.y));
        blockIdx = g_bid;
        (*g_body)();
        g_wave[l / 64].drop();
        g_block.drop();
        std::unique_lock<std::mutex> lk(g_pm);
        if (++g_done == g_nl) g_dcv.notify_all();
    }
}

void emu_run_block(dim3 grid, dim3 block, dim3 bid, size_t smem, const std::function<void()> &body) {
    const int nl = (int)(block.x * block.y * block.z);
    if (nl > 1024) { fprintf(stderr, "EMU: block of %d lanes\n", nl); abort(); }
    // grow the pool to the widest workgroup seen (the new lanes wait for the
    // next job like the others)
    for (int l = (int)g_pool.size(); l < std::max(nl, 256); l++)
        g_pool.emplace_back(lane_main, l), g_pool.back().detach();
    g_block.reset(nl);
    for (int w = 0; w < (nl + 63) / 64; w++) g_wave[w].reset(std::min(64, nl - 64 * w));
    g_smem_bytes = smem;
    g_smem = (unsigned char *)malloc(smem ? smem : 1);
    memset(g_smem, 0x5A, smem);
    for (int l = 0; l < nl; l++) g_barriers[l] = g_ballots[l] = 0;
    {
        std::unique_lock<std::mutex> lk(g_pm);
        g_nl = nl;
        g_done = 0;
        g_body = &body;
        g_bdim = block;
        g_bid = bid;
        g_job++;
        g_pcv.notify_all();
        g_dcv.wait(lk, [&] { return g_done == g_nl; });
    }
    for (int l = 1; l < nl; l++)
        if (g_barriers[l] != g_barriers[0]) {
            fprintf(stderr, "EMU: barrier count mismatch in block (%u,%u): lane %d %llu vs lane 0 %llu (ticket %llu)\n",
                    bid.x, bid.y, l, (unsigned long long)g_barriers[l], (unsigned long long)g_barriers[0],
                    (unsigned long long)emu_ticket);
            abort();
        }
    for (int l = 0; l < nl; l++)
        if (g_ballots[l] != g_ballots[(l / 64) * 64]) {
            fprintf(stderr, "EMU: ballot count mismatch in wave %d (ticket %llu)\n", l / 64,
                    (unsigned long long)emu_ticket);
            abort();
        }
    free(g_smem);
    g_smem = nullptr;
}
