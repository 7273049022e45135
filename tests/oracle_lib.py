"""ctypes access to the CPU oracle (oracle/liboracle.so) and, when it was
built in this container, to the compiled reference CPU code
(oracle/_ref/libref.so).  Test infrastructure: used only as the checker."""
import ctypes
ctypes = ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

C = ctypes.c_uint64
I = ctypes.c_int
VP = ctypes.c_void_p

_o = None
_r = None


class VebTable(ctypes.Structure):
    _fields_ = [("L", ctypes.c_uint64), ("R", ctypes.c_uint64), ("D", ctypes.c_uint32)]


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        o = ctypes.CDLL(ORACLE_SO)
        o.or_ipow.argtypes = [C, ctypes.c_uint32]; o.or_ipow.restype = C
        o.or_ilog.argtypes = [C, C]; o.or_ilog.restype = ctypes.c_uint32
        o.or_inorder_btree.argtypes = [C, C, VP]; o.or_inorder_btree.restype = I
        o.or_btree_pos.argtypes = [C, C, C]; o.or_btree_pos.restype = C
        o.or_mveb_pos.argtypes = [C, C]; o.or_mveb_pos.restype = C
        o.or_mveb_layout.argtypes = [C, VP]; o.or_mveb_layout.restype = I
        o.or_classic_pos.argtypes = [C, C]; o.or_classic_pos.restype = C
        o.or_classic_layout.argtypes = [C, VP]; o.or_classic_layout.restype = I
        o.or_build_table_classic.argtypes = [ctypes.POINTER(VebTable), C, ctypes.c_uint32,
                                             ctypes.c_uint32]
        o.or_classic_tableset.argtypes = [C, ctypes.POINTER(ctypes.c_uint32),
                                          ctypes.POINTER(ctypes.c_uint32),
                                          ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(VebTable),
                                          ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.POINTER(ctypes.c_uint32)]
        o.or_build_table.argtypes = [ctypes.POINTER(VebTable), C, ctypes.c_uint32, ctypes.c_uint32]
        o.or_veb_tableset.argtypes = [C, ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(VebTable),
                                      ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint32)]
        for w in ("u32", "u64"):
            getattr(o, f"or_permute_bst_{w}").argtypes = [VP, C, I]
            getattr(o, f"or_permute_btree_{w}").argtypes = [VP, C, C, I]
            getattr(o, f"or_permute_mveb_{w}").argtypes = [VP, C]
            getattr(o, f"or_permute_classic_{w}").argtypes = [VP, C, I]
        o.or_rev_base_d.argtypes = [C, C, C]; o.or_rev_base_d.restype = C
        o.or_rev_base_b.argtypes = [C, C, C, C]; o.or_rev_base_b.restype = C
        o.or_involution.argtypes = [C, C, C]; o.or_involution.restype = C
        for w in ("u32", "u64"):
            getattr(o, f"or_unshuffle_{w}").argtypes = [VP, C, C, C]
            getattr(o, f"or_shuffle_dk_{w}").argtypes = [VP, C, C]
            getattr(o, f"or_unshuffle_dk_{w}").argtypes = [VP, C, C]
            getattr(o, f"or_rotate_right_{w}").argtypes = [VP, C, C]
            getattr(o, f"or_permute_bst_inv_{w}").argtypes = [VP, C]
            getattr(o, f"or_permute_btree_inv_{w}").argtypes = [VP, C, C]
            getattr(o, f"or_permute_mveb_inv_{w}").argtypes = [VP, C]
        for k in ("u32", "u64", "i32", "i64"):
            getattr(o, f"or_search_bst_{k}").argtypes = [VP, C, VP, C, VP, I]
            getattr(o, f"or_search_btree_{k}").argtypes = [VP, C, C, VP, C, VP, I]
            getattr(o, f"or_search_mveb_{k}").argtypes = [VP, C, VP, C, VP, I]
            getattr(o, f"or_search_binary_{k}").argtypes = [VP, C, VP, C, VP, I]
            getattr(o, f"or_search_classic_{k}").argtypes = [VP, C, VP, C, VP, I]
        _o = o
    return _o


def ref():
    """The compiled reference, or None when it was not built here."""
    global _r
    if _r is None and os.path.exists(REF_SO):
        r = ctypes.CDLL(REF_SO)
        r.ref_permute_bst_u64.argtypes = [VP, C, I]
        r.ref_permute_bst_u32.argtypes = [VP, C, I]
        r.ref_permute_btree_u64.argtypes = [VP, C, C, I]
        r.ref_permute_btree_u32.argtypes = [VP, C, C, I]
        r.ref_permute_veb_u64.argtypes = [VP, C, I]
        r.ref_search_bst_u64.argtypes = [VP, C, VP, C, VP, I]
        r.ref_search_btree_u64.argtypes = [VP, C, C, VP, C, VP, I]
        r.ref_search_binary_u64.argtypes = [VP, C, VP, C, VP, I]
        r.ref_build_table_classic.argtypes = [VP, C, ctypes.c_uint32]
        r.ref_permute_bst_inv_u64.argtypes = [VP, C, I]
        r.ref_permute_btree_inv_u64.argtypes = [VP, C, C, I]
        r.ref_unshuffle_u64.argtypes = [VP, C, C, C]
        r.ref_shuffle_dk_u64.argtypes = [VP, C, C]
        r.ref_unshuffle_dk_u64.argtypes = [VP, C, C]
        _r = r
    return _r


def _p(a):
    return a.ctypes.data_as(VP)


KIND_NP = {"u32": np.uint32, "u64": np.uint64, "i32": np.int32, "i64": np.int64}


def width_suffix(a):
    return "u32" if a.dtype.itemsize == 4 else "u64"


def permute(layout, a, b=16, p=1):
    """Oracle permutation of a numpy array in place (any 4/8-byte dtype)."""
    o = oracle()
    v = a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)
    w = width_suffix(a)
    if layout == "bst":
        rc = getattr(o, f"or_permute_bst_{w}")(_p(v), v.size, p)
    elif layout == "btree":
        rc = getattr(o, f"or_permute_btree_{w}")(_p(v), v.size, b, p)
    elif layout == "veb":   # the classic (CPU) vEB
        rc = getattr(o, f"or_permute_classic_{w}")(_p(v), v.size, p)
    else:
        rc = getattr(o, f"or_permute_mveb_{w}")(_p(v), v.size)
    assert rc == 0
    return a


def permute_involutions(layout, a, b=16):
    """The oracle's restatement of the involution programs, in place."""
    o = oracle()
    v = a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)
    w = width_suffix(v)
    if layout == "bst":
        getattr(o, f"or_permute_bst_inv_{w}")(_p(v), v.size)
    elif layout == "btree":
        getattr(o, f"or_permute_btree_inv_{w}")(_p(v), v.size, b)
    else:
        getattr(o, f"or_permute_mveb_inv_{w}")(_p(v), v.size)
    return a


def layout_ranks(layout, n, b=16, p=1):
    """Rank held by each slot (uint64), from the oracle's restated algorithm
    (layout "bst", "btree", "mveb" or "veb", the classic CPU vEB)."""
    a = np.arange(n, dtype=np.uint64)
    return permute(layout, a, b, p)


def inorder_ranks(n, b):
    """Independent construction: in-order DFS fill of the node tree."""
    pos = np.zeros(n, dtype=np.uint64)
    assert oracle().or_inorder_btree(n, b, _p(pos)) == 0
    out = np.empty(n, dtype=np.uint64)
    out[pos] = np.arange(n, dtype=np.uint64)
    return out


def mveb_struct_ranks(n):
    pos = np.zeros(n, dtype=np.uint64)
    oracle().or_mveb_layout(n, _p(pos))
    out = np.empty(n, dtype=np.uint64)
    out[pos] = np.arange(n, dtype=np.uint64)
    return out


def classic_struct_ranks(n):
    """Independent construction of the classic vEB: the structural recursion"""
    pos = np.zeros(n, dtype=np.uint64)
    oracle().or_classic_layout(n, _p(pos))
    out = np.empty(n, dtype=np.uint64)
    out[pos] = np.arange(n, dtype=np.uint64)
    return out


def search(layout, A, Q, kind, b=16, p=8):
    o = oracle()
    A = np.ascontiguousarray(A, dtype=KIND_NP[kind])
    Q = np.ascontiguousarray(Q, dtype=KIND_NP[kind])
    ans = np.zeros(Q.size, dtype=np.uint64)
    if layout == "bst":
        getattr(o, f"or_search_bst_{kind}")(_p(A), A.size, _p(Q), Q.size, _p(ans), p)
    elif layout == "btree":
        getattr(o, f"or_search_btree_{kind}")(_p(A), A.size, b, _p(Q), Q.size, _p(ans), p)
    elif layout == "veb":
        getattr(o, f"or_search_classic_{kind}")(_p(A), A.size, _p(Q), Q.size, _p(ans), p)
    elif layout == "binary":   # A is the sorted array itself
        getattr(o, f"or_search_binary_{kind}")(_p(A), A.size, _p(Q), Q.size, _p(ans), p)
    else:
        getattr(o, f"or_search_mveb_{kind}")(_p(A), A.size, _p(Q), Q.size, _p(ans), p)
    return ans


def tableset(n, classic=False):
    """(d, num_tables, table_size, tables[list of (L,R,D)], idx, table_idx) of
    the mvEB, or of the classic vEB"""
    o = oracle()
    d = ctypes.c_uint32(); nt = ctypes.c_uint32(); ts = ctypes.c_uint32()
    tables = (VebTable * 4096)()
    idx = (ctypes.c_uint64 * 64)()
    tidx = (ctypes.c_uint32 * 64)()
    fn = o.or_classic_tableset if classic else o.or_veb_tableset
    assert fn(n, ctypes.byref(d), ctypes.byref(nt), ctypes.byref(ts), tables, idx, tidx) == 0
    return (d.value, nt.value, ts.value, [(t.L, t.R, t.D) for t in tables[:ts.value]],
            list(idx[:nt.value]), list(tidx[:nt.value]))


def digest(a):
    """Order-sensitive 64-bit digest of a rank layout: sum over slots of a
    splitmix64 mix of (rank, slot), wrapping mod 2^64."""
    a = np.asarray(a, dtype=np.uint64)
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        for s in range(0, a.size, 1 << 24):
            x = a[s:s + (1 << 24)].astype(np.uint64)
            i = np.arange(s, s + x.size, dtype=np.uint64)
            z = x * np.uint64(0x9E3779B97F4A7C15) + (i + np.uint64(1)) * np.uint64(0xD1B54A32D192ED03)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            total = np.uint64(total + z.sum(dtype=np.uint64))
    return int(total)
