"""ctypes binding of libtl_amd.so (the gfx950 Tree-Layouts C ABI).

Thin Python surface over include/tree_layouts.h for tests, smoke checks and
the benchmark.  Device buffers are torch tensors on a ROCm device; keys are
stored in torch.int32 / torch.int64 tensors and the key type is named
explicitly ('u32', 'u64', 'i32', 'i64'), since the library compares unsigned
or signed according to that name while layouts only depend on the width.

Nothing here falls back to a CPU path: when the shared library is missing
every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libtl_amd.so")

TL_OK, TL_EINVAL, TL_ENOMEM, TL_EHIP, TL_ETIMEOUT, TL_EUNSUPPORTED, TL_EIO = range(7)
TL_FILE_SORTED = 255
DTYPE = {"u32": 0, "u64": 1, "i32": 2, "i64": 3}
WIDTH = {"u32": 4, "u64": 8, "i32": 4, "i64": 8}
LAYOUT = {"bst": 0, "btree": 1, "mveb": 2, "veb": 3}   # "veb": the classic (CPU) vEB
# limits of tl_permute_btree_noalloc (include/tree_layouts.h)
TL_NOALLOC_SCRATCH_BYTES = 64 << 20
TL_NOALLOC_MAX_KEYS = 1 << 31


class TLError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"tree-layouts status {status}: {msg}")
        self.status = status


class VebTable(ctypes.Structure):
    _fields_ = [("L", ctypes.c_uint64), ("R", ctypes.c_uint64), ("D", ctypes.c_uint32)]


class LayoutInfo(ctypes.Structure):
    _fields_ = [("layout", ctypes.c_uint32), ("dtype", ctypes.c_int), ("n", ctypes.c_uint64),
                ("b", ctypes.c_uint64), ("d", ctypes.c_uint32), ("num_tables", ctypes.c_uint32),
                ("table_size", ctypes.c_uint32), ("version", ctypes.c_uint32),
                ("checksum", ctypes.c_uint64)]


_lib = None

# (name, argtypes) for every exported entry point
_SIGS = {
    "tl_last_error": [],
    "tl_version": [],
    "tl_permute_bst": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                       ctypes.POINTER(ctypes.c_float)],
    "tl_permute_btree": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_permute_mveb": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                        ctypes.POINTER(ctypes.c_float)],
    "tl_permute_btree_noalloc": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                 ctypes.c_void_p],
    "tl_unpermute_bst": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                         ctypes.POINTER(ctypes.c_float)],
    "tl_unpermute_btree": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                           ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_unpermute_mveb": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                          ctypes.POINTER(ctypes.c_float)],
    "tl_permute_veb": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                       ctypes.POINTER(ctypes.c_float)],
    "tl_permute_bst_involutions": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_permute_btree_involutions": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.c_int, ctypes.c_void_p,
                                     ctypes.POINTER(ctypes.c_float)],
    "tl_permute_mveb_involutions": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_shuffle_dk": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                      ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_unshuffle_dk": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                        ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_unshuffle": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                     ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_shift_right": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                       ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_unpermute_veb": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                         ctypes.POINTER(ctypes.c_float)],
    "tl_query_bst": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                     ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                     ctypes.POINTER(ctypes.c_float)],
    "tl_query_binary": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                        ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                        ctypes.POINTER(ctypes.c_float)],
    "tl_query_btree": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                       ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                       ctypes.POINTER(ctypes.c_float)],
    "tl_query_mveb": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                      ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_query_mveb_nonperfect": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.POINTER(ctypes.c_float)],
    "tl_query_veb": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                     ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                     ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
    "tl_query_veb_nonperfect": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.POINTER(ctypes.c_float)],
    "tl_veb_classic_build_table": [ctypes.POINTER(VebTable), ctypes.c_uint64, ctypes.c_uint32,
                                   ctypes.c_uint32],
    "tl_veb_classic_tableset_device": [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "tl_veb_build_table": [ctypes.POINTER(VebTable), ctypes.c_uint64, ctypes.c_uint32,
                           ctypes.c_uint32],
    "tl_veb_tableset_device": [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "tl_plan_export": [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                       ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                       ctypes.POINTER(ctypes.c_uint64)],
    "tl_sync": [ctypes.c_void_p],
    "tl_sort_dedupe": [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                       ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p,
                       ctypes.POINTER(ctypes.c_float)],
    "tl_layout_save": [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                       ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p],
    "tl_layout_read_info": [ctypes.c_char_p, ctypes.POINTER(LayoutInfo)],
    "tl_layout_load": [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                       ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(LayoutInfo),
                       ctypes.c_void_p],
    "tl_selftest_timeout": [ctypes.c_uint32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
}
class BatchItem(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("n", ctypes.c_uint64), ("dtype", ctypes.c_int),
                ("layout", ctypes.c_uint32), ("b", ctypes.c_uint64), ("Q", ctypes.c_void_p),
                ("nq", ctypes.c_uint64), ("ans", ctypes.c_void_p)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("lds_bytes", ctypes.c_uint32), ("wg_per_cu", ctypes.c_uint32),
                ("wait_ms", ctypes.c_uint32), ("tile_k", ctypes.c_uint32),
                ("stream_threads", ctypes.c_uint32), ("btree_qpg", ctypes.c_uint32),
                ("veb_qpl", ctypes.c_uint32), ("query_queue", ctypes.c_uint32)]

BATCH_KNOBS = ("lds_bytes", "tile_k", "stream_threads", "btree_qpg", "veb_qpl", "query_queue")


_OPTIONAL = {
    "tl_batch_create": ([ctypes.POINTER(BatchItem), ctypes.c_uint32, ctypes.POINTER(BatchConfig),
                         ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "tl_batch_run": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                      ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "tl_batch_run_separate": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
                              ctypes.c_int),
    "tl_batch_stats": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                        ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)],
                       ctypes.c_int),
    "tl_batch_autotune": ([ctypes.POINTER(BatchItem), ctypes.c_uint32, ctypes.c_uint32,
                           ctypes.c_void_p, ctypes.POINTER(BatchConfig),
                           ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "tl_batch_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "tl_learned_model_bytes": ([ctypes.c_uint32], ctypes.c_uint64),
    "tl_learned_fit": ([ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32,
                        ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)],
                       ctypes.c_int),
    "tl_batch_export": ([ctypes.POINTER(BatchItem), ctypes.c_uint32, ctypes.POINTER(BatchConfig),
                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64,
                         ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                         ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
    "tl_learned_predict": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "tl_query_btree_learned": ([ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
}


def lib():
    """Load libtl_amd.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TLError(TL_EHIP, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_char_p if name in ("tl_last_error", "tl_version") else ctypes.c_int
    for name, (args, res) in _OPTIONAL.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
    _lib = L
    return L


def _check(st):
    if st != TL_OK:
        raise TLError(st, lib().tl_last_error().decode())


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(stream)


def _ptr(t):
    """Device address of a contiguous ROCm tensor (None -> NULL).  The library
    takes plain pointers and sizes, so a CPU tensor or a strided view would
    be read as the wrong memory: both are rejected here."""
    if t is None:
        return ctypes.c_void_p(0)
    if not t.is_cuda:
        raise ValueError("tree-layouts: expected a tensor on the GPU, got one on " + str(t.device))
    if not t.is_contiguous():
        raise ValueError("tree-layouts: expected a contiguous tensor (got a strided view)")
    return ctypes.c_void_p(t.data_ptr())


def _kind(t, kind):
    """Key type name of t; an explicit kind must match t's element width."""
    if kind is None:
        import torch
        names = {torch.int32: "i32", torch.int64: "i64", torch.uint32: "u32", torch.uint64: "u64"}
        if t.dtype not in names:
            raise ValueError(f"tree-layouts: no key type for tensor dtype {t.dtype}")
        return names[t.dtype]
    if kind not in WIDTH:
        raise ValueError(f"tree-layouts: unknown key type {kind!r}")
    if t.element_size() != WIDTH[kind]:
        raise ValueError(f"tree-layouts: key type {kind} is {WIDTH[kind]} bytes wide but the "
                         f"tensor's elements are {t.element_size()}")
    return kind


def _ans(ans, nq):
    """Answer buffer: a contiguous 8-byte integer tensor of at least nq entries."""
    if ans.element_size() != 8 or ans.is_floating_point():
        raise ValueError("tree-layouts: answers are uint64, pass an int64 tensor")
    if ans.numel() < nq:
        raise ValueError(f"tree-layouts: {nq} queries need {nq} answers, the tensor holds "
                         f"{ans.numel()}")
    return _ptr(ans)


def _queries(Q, kind):
    if Q.element_size() != WIDTH[kind]:
        raise ValueError(f"tree-layouts: queries must have the keys' width ({WIDTH[kind]} bytes)")
    return _ptr(Q)


def sync(stream=None):
    """Wait for the stream and raise TLError(TL_ETIMEOUT) if an in-kernel
    bounded wait of an untimed call gave up since the last report."""
    _check(lib().tl_sync(_stream(stream)))


def selftest_timeout(wait_ms, stream=None):
    """Runs a kernel whose bounded wait can never be satisfied; returns
    (status, elapsed ms).  The status must be TL_ETIMEOUT."""
    out = ctypes.c_float(0.0)
    st = lib().tl_selftest_timeout(wait_ms, _stream(stream), ctypes.byref(out))
    return st, out.value


def _ms(timed):
    return ctypes.byref(ctypes.c_float(0.0)) if timed else None


def _timed_call(fn, *args, timed=True):
    out = ctypes.c_float(0.0)
    _check(fn(*args, ctypes.byref(out) if timed else None))
    return out.value if timed else None


# ---- permutations ----------------------------------------------------------

def permute_bst(A, kind=None, stream=None, timed=True):
    k = _kind(A, kind)
    return _timed_call(lib().tl_permute_bst, _ptr(A), A.numel(), DTYPE[k], _stream(stream),
                       timed=timed)


def permute_btree(A, b, kind=None, stream=None, timed=True):
    k = _kind(A, kind)
    return _timed_call(lib().tl_permute_btree, _ptr(A), A.numel(), b, DTYPE[k],
                       _stream(stream), timed=timed)


def permute_btree_noalloc(A, b, kind=None, stream=None):
    """tl_permute_btree_noalloc: the B-tree layout with all scratch in the
    library's static device array (enqueued only; errors of in-kernel waits
    surface at sync())."""
    k = _kind(A, kind)
    _check(lib().tl_permute_btree_noalloc(_ptr(A), A.numel(), b, DTYPE[k], _stream(stream)))


def permute_mveb(A, kind=None, stream=None, timed=True):
    k = _kind(A, kind)
    return _timed_call(lib().tl_permute_mveb, _ptr(A), A.numel(), DTYPE[k], _stream(stream),
                       timed=timed)


def permute_veb(A, kind=None, stream=None, timed=True):
    """The classic (CPU) vEB layout."""
    k = _kind(A, kind)
    return _timed_call(lib().tl_permute_veb, _ptr(A), A.numel(), DTYPE[k], _stream(stream),
                       timed=timed)


def permute(layout, A, b=16, kind=None, stream=None, timed=True):
    if layout == "bst":
        return permute_bst(A, kind, stream, timed)
    if layout == "btree":
        return permute_btree(A, b, kind, stream, timed)
    if layout == "veb":
        return permute_veb(A, kind, stream, timed)
    return permute_mveb(A, kind, stream, timed)


def permute_involutions(layout, A, b=16, kind=None, stream=None, timed=True):
    """The involution family (tl_permute_{bst,btree,mveb}_involutions): the
    same layouts as permute(), by swap passes with closed-form partners."""
    k = _kind(A, kind)
    L = lib()
    if layout == "bst":
        return _timed_call(L.tl_permute_bst_involutions, _ptr(A), A.numel(), DTYPE[k],
                           _stream(stream), timed=timed)
    if layout == "btree":
        return _timed_call(L.tl_permute_btree_involutions, _ptr(A), A.numel(), b, DTYPE[k],
                           _stream(stream), timed=timed)
    if layout == "mveb":
        return _timed_call(L.tl_permute_mveb_involutions, _ptr(A), A.numel(), DTYPE[k],
                           _stream(stream), timed=timed)
    raise ValueError(f"no involution permutation for layout {layout!r}")


def shuffle_dk(A, k, kind=None, stream=None, timed=True):
    """k-way shuffle of A (numel a multiple of k): key i*d + j -> slot j*k + i."""
    t = _kind(A, kind)
    return _timed_call(lib().tl_shuffle_dk, _ptr(A), A.numel(), k, DTYPE[t], _stream(stream),
                       timed=timed)


def unshuffle_dk(A, k, kind=None, stream=None, timed=True):
    """Inverse of shuffle_dk: every k-th key gathered."""
    t = _kind(A, kind)
    return _timed_call(lib().tl_unshuffle_dk, _ptr(A), A.numel(), k, DTYPE[t], _stream(stream),
                       timed=timed)


def unshuffle(A, k, kind=None, stream=None, timed=True):
    """k-way unshuffle of numel = k^d - 1 keys: stable partition by (i+1) mod k."""
    t = _kind(A, kind)
    return _timed_call(lib().tl_unshuffle, _ptr(A), A.numel(), k, DTYPE[t], _stream(stream),
                       timed=timed)


def shift_right(A, k, kind=None, stream=None, timed=True):
    """Rotation of A right by k by two reversal passes."""
    t = _kind(A, kind)
    return _timed_call(lib().tl_shift_right, _ptr(A), A.numel(), k, DTYPE[t], _stream(stream),
                       timed=timed)


def unpermute(layout, A, b=16, kind=None, stream=None, timed=True):
    """In-place inverse of permute(layout, ...): layout order -> sorted order."""
    k = _kind(A, kind)
    L = lib()
    if layout == "bst":
        return _timed_call(L.tl_unpermute_bst, _ptr(A), A.numel(), DTYPE[k], _stream(stream),
                           timed=timed)
    if layout == "btree":
        return _timed_call(L.tl_unpermute_btree, _ptr(A), A.numel(), b, DTYPE[k],
                           _stream(stream), timed=timed)
    if layout == "veb":
        return _timed_call(L.tl_unpermute_veb, _ptr(A), A.numel(), DTYPE[k], _stream(stream),
                           timed=timed)
    return _timed_call(L.tl_unpermute_mveb, _ptr(A), A.numel(), DTYPE[k], _stream(stream),
                       timed=timed)


# ---- raw keys -> sorted distinct keys; layout files --------------------------

def sort_dedupe(A, kind=None, stream=None):
    """Sorts A in place (the key type's order) and drops repeats; returns
    (number of distinct keys, which fill A[:m], ms)."""
    k = _kind(A, kind)
    m = ctypes.c_uint64(0)
    ms = ctypes.c_float(0.0)
    _check(lib().tl_sort_dedupe(_ptr(A), A.numel(), DTYPE[k], ctypes.byref(m), _stream(stream),
                                ctypes.byref(ms)))
    return m.value, ms.value


def _info_dict(info):
    return {f: getattr(info, f) for f, _ in LayoutInfo._fields_}


def layout_save(path, A, layout, b=0, kind=None, stream=None):
    """Writes A (laid out as `layout`: 'bst', 'btree', 'mveb', 'veb' or
    'sorted') to a layout file."""
    k = _kind(A, kind)
    code = TL_FILE_SORTED if layout == "sorted" else LAYOUT[layout]
    _check(lib().tl_layout_save(os.fsencode(path), _ptr(A), A.numel(), code, b, DTYPE[k],
                                _stream(stream)))


def layout_read_info(path):
    info = LayoutInfo()
    _check(lib().tl_layout_read_info(os.fsencode(path), ctypes.byref(info)))
    return _info_dict(info)


def layout_load(path, A, tables=None, idx=None, tidx=None, stream=None):
    """Loads a layout file into A (capacity A.numel()) and, if given, its vEB
    table set into tables / idx / tidx; returns the header as a dict."""
    info = LayoutInfo()
    _check(lib().tl_layout_load(os.fsencode(path), _ptr(A), A.numel(), _ptr(tables), _ptr(idx),
                                _ptr(tidx), ctypes.byref(info), _stream(stream)))
    return _info_dict(info)


# ---- queries ---------------------------------------------------------------

def query_bst(A, Q, ans, kind=None, stream=None, timed=True):
    k = _kind(A, kind)
    return _timed_call(lib().tl_query_bst, _ptr(A), A.numel(), DTYPE[k], _queries(Q, k), Q.numel(),
                       _ans(ans, Q.numel()), _stream(stream), timed=timed)


def query_binary(A, Q, ans, kind=None, stream=None, timed=True):
    """Binary search of the sorted A: the key's index, its predecessor's
    index when absent, or n when every key is larger."""
    k = _kind(A, kind)
    return _timed_call(lib().tl_query_binary, _ptr(A), A.numel(), DTYPE[k], _queries(Q, k), Q.numel(),
                       _ans(ans, Q.numel()), _stream(stream), timed=timed)


def query_btree(A, b, Q, ans, kind=None, stream=None, timed=True):
    k = _kind(A, kind)
    return _timed_call(lib().tl_query_btree, _ptr(A), A.numel(), b, DTYPE[k], _queries(Q, k),
                       Q.numel(), _ans(ans, Q.numel()), _stream(stream), timed=timed)


def query_mveb(A, table, d, Q, ans, kind=None, stream=None, timed=True):
    k = _kind(A, kind)
    return _timed_call(lib().tl_query_mveb, _ptr(A), _ptr(table), A.numel(), d, DTYPE[k],
                       _queries(Q, k), Q.numel(), _ans(ans, Q.numel()), _stream(stream), timed=timed)


def query_mveb_nonperfect(A, tables, idx, tidx, d, num_tables, Q, ans, kind=None, stream=None,
                          timed=True):
    k = _kind(A, kind)
    return _timed_call(lib().tl_query_mveb_nonperfect, _ptr(A), _ptr(tables), _ptr(idx),
                       _ptr(tidx), A.numel(), d, num_tables, DTYPE[k], _queries(Q, k), Q.numel(),
                       _ans(ans, Q.numel()), _stream(stream), timed=timed)


def veb_tableset_device(n, device="cuda", stream=None, classic=False):
    """Device-built table set of the mvEB (or, classic=True, of the classic
    vEB): (tables int64[64*64*3], idx, tidx, meta) tensors."""
    import torch
    tables = torch.zeros(64 * 64 * 3, dtype=torch.int64, device=device)   # 24-byte entries
    idx = torch.zeros(64, dtype=torch.int64, device=device)
    tidx = torch.zeros(64, dtype=torch.int32, device=device)
    meta = torch.zeros(4, dtype=torch.int32, device=device)
    fn = lib().tl_veb_classic_tableset_device if classic else lib().tl_veb_tableset_device
    _check(fn(n, _ptr(tables), _ptr(idx), _ptr(tidx), _ptr(meta), _stream(stream)))
    return tables, idx, tidx, meta


def veb_build_table(n, d, root_depth=0, classic=False):
    arr = (VebTable * max(d, 1))()
    fn = lib().tl_veb_classic_build_table if classic else lib().tl_veb_build_table
    _check(fn(arr, n, d, root_depth))
    return [(e.L, e.R, e.D) for e in arr]


def query_veb(A, table, d, Q, ans, kind=None, stream=None, timed=True):
    """Classic vEB of a perfect tree, with a classic-rule table."""
    k = _kind(A, kind)
    return _timed_call(lib().tl_query_veb, _ptr(A), _ptr(table), A.numel(), d, DTYPE[k],
                       _queries(Q, k), Q.numel(), _ans(ans, Q.numel()), _stream(stream), timed=timed)


def query_veb_nonperfect(A, tables, idx, tidx, d, num_tables, Q, ans, kind=None, stream=None,
                         timed=True):
    """Classic vEB of any size, with the classic table set."""
    k = _kind(A, kind)
    return _timed_call(lib().tl_query_veb_nonperfect, _ptr(A), _ptr(tables), _ptr(idx),
                       _ptr(tidx), A.numel(), d, num_tables, DTYPE[k], _queries(Q, k), Q.numel(),
                       _ans(ans, Q.numel()), _stream(stream), timed=timed)


def plan_export(layout, n, b=1, kind="u64", lds_bytes=0, inverse=False):
    """Serialized dry-run plan of a permutation, or of its inverse (list of
    uint64 words)."""
    import numpy as np
    used = ctypes.c_uint64(0)
    L = lib()
    lay = LAYOUT[layout] | (0x100 if inverse else 0)
    _check(L.tl_plan_export(lay, n, b, DTYPE[kind], lds_bytes, None, 0, ctypes.byref(used)))
    buf = np.zeros(used.value, dtype=np.uint64)
    _check(L.tl_plan_export(lay, n, b, DTYPE[kind], lds_bytes,
                            buf.ctypes.data_as(ctypes.c_void_p), used.value,
                            ctypes.byref(used)))
    return buf


# ---- learned start node ------------------------------------------------------

def learned_fit(sorted_keys, num_leaves=1 << 16, kind=None, stream=None, timed=True):
    """Fits the rank model on the sorted keys; returns (model tensor, ms)."""
    import torch
    k = _kind(sorted_keys, kind)
    nbytes = lib().tl_learned_model_bytes(num_leaves)
    model = torch.zeros((nbytes + 7) // 8, dtype=torch.int64, device=sorted_keys.device)
    ms = _timed_call(lib().tl_learned_fit, _ptr(sorted_keys), sorted_keys.numel(), DTYPE[k],
                     num_leaves, _ptr(model), _stream(stream), timed=timed)
    return model, ms


def learned_predict(model, keys, kind=None, stream=None):
    """(pred, err) int64 tensors: every key's predicted rank and its leaf's
    error bound, from the device's own arithmetic (-1 for an empty leaf)."""
    import torch
    k = _kind(keys, kind)
    pred = torch.empty(keys.numel(), dtype=torch.int64, device=keys.device)
    err = torch.empty_like(pred)
    _check(lib().tl_learned_predict(_ptr(model), _ptr(keys), keys.numel(), DTYPE[k], _ptr(pred),
                                    _ptr(err), _stream(stream)))
    return pred, err


def learned_max_err(model):
    """largest leaf rank error (uint32 at byte offset 32 of the model)"""
    return int(model[4].item()) & 0xFFFFFFFF


def query_btree_learned(A, b, model, Q, ans, stats=None, kind=None, stream=None, timed=True):
    k = _kind(A, kind)
    return _timed_call(lib().tl_query_btree_learned, _ptr(A), A.numel(), b, DTYPE[k], _ptr(model),
                       _queries(Q, k), Q.numel(), _ans(ans, Q.numel()), _ptr(stats),
                       _stream(stream), timed=timed)


def bench_learned(A, keys, Q, plain_ans, b, steps, warmup, n, num_leaves=1 << 16):
    """Fit on the sorted keys, time the learned queries, compare with plain."""
    import torch
    stream = torch.cuda.current_stream()
    model, fit_ms = learned_fit(keys, num_leaves, kind="u64")
    ans = torch.empty_like(plain_ans)
    stats = torch.zeros(1, dtype=torch.int64, device=A.device)
    query_btree_learned(A, b, model, Q, ans, stats, kind="u64")
    torch.cuda.synchronize()
    same = bool(torch.equal(ans, plain_ans))
    fallbacks = int(stats.item())
    for _ in range(warmup):
        query_btree_learned(A, b, model, Q, ans, None, kind="u64", timed=False)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(steps):
        query_btree_learned(A, b, model, Q, ans, None, kind="u64", timed=False)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    nq = Q.numel()
    return {"ms": round(ms, 4), "Mq_s": round(nq / ms / 1e3, 2), "fit_ms": round(fit_ms, 3),
            "max_rank_err": learned_max_err(model), "fallback_rate": fallbacks / nq,
            "answers_equal_plain": same, "num_leaves": num_leaves}


# ---- uneven batch --------------------------------------------------------------

class Batch:
    """Owns a tl_batch handle over caller tensors (kept alive by this object).
    arrays: list of dicts with A (tensor), kind, layout ('bst'|'btree'|'mveb'),
    b, Q (tensor or None), ans (int64 tensor or None)."""

    def __init__(self, arrays, lds_bytes=0, wg_per_cu=0, wait_ms=0, tile_k=0, stream_threads=0,
                 btree_qpg=0, veb_qpl=0, query_queue=0):
        self.arrays = arrays
        items = (BatchItem * len(arrays))()
        for i, a in enumerate(arrays):
            items[i].A = a["A"].data_ptr()
            items[i].n = a["A"].numel()
            items[i].dtype = DTYPE[a["kind"]]
            items[i].layout = LAYOUT[a["layout"]]
            items[i].b = a.get("b", 1)
            q = a.get("Q")
            items[i].Q = q.data_ptr() if q is not None else None
            items[i].nq = q.numel() if q is not None else 0
            items[i].ans = a["ans"].data_ptr() if q is not None else None
        self._items = items
        cfg = BatchConfig(lds_bytes, wg_per_cu, wait_ms, tile_k, stream_threads, btree_qpg, veb_qpl,
                          query_queue)
        h = ctypes.c_void_p()
        _check(lib().tl_batch_create(items, len(arrays), ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def stats(self):
        o, t, s = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().tl_batch_stats(self.h, ctypes.byref(o), ctypes.byref(t), ctypes.byref(s)))
        return {"ops": o.value, "tasks": t.value, "smem": s.value}

    def run(self, graph=True, stream=None, timed=True):
        return _timed_call(lib().tl_batch_run, self.h, _stream(stream), 1 if graph else 0,
                           timed=timed)

    def run_separate(self, stream=None, timed=True):
        return _timed_call(lib().tl_batch_run_separate, self.h, _stream(stream), timed=timed)

    def close(self):
        if self.h:
            lib().tl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_autotune(arrays, sample=256, stream=None):
    """The planned path's knobs chosen by the library's autotuner on a sample
    of the batch: a dict of BATCH_KNOBS (pass them to Batch) and sample_ms."""
    items = (BatchItem * len(arrays))()
    for i, a in enumerate(arrays):
        items[i].A = a["A"].data_ptr()
        items[i].n = a["A"].numel()
        items[i].dtype = DTYPE[a["kind"]]
        items[i].layout = LAYOUT[a["layout"]]
        items[i].b = a.get("b", 1)
        q = a.get("Q")
        items[i].Q = q.data_ptr() if q is not None else None
        items[i].nq = q.numel() if q is not None else 0
        items[i].ans = a["ans"].data_ptr() if q is not None else None
    best = BatchConfig()
    ms = ctypes.c_float()
    _check(lib().tl_batch_autotune(items, len(arrays), sample, _stream(stream), ctypes.byref(best),
                                   ctypes.byref(ms)))
    out = {k: getattr(best, k) for k in BATCH_KNOBS}
    out["sample_ms"] = ms.value
    return out
