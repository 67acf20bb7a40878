"""Python face of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg;
never by the product.  Wraps oracle/liborc.so (a plain-C, step-by-step
restatement of the reference's default SHCOLL algorithms, see oracle/orc.c)
and adds independent numpy formulations used to pin it:

* reduce_closed_form: the closed form of recursive doubling
  (src/shcoll/src/reduce.c:182-293) -- member k's result is the balanced
  pairwise tree over the folded leaves k^0, k^1, ..., k^(p2s-1) -- written
  without simulating any message, per member; linear_closed_form,
  binomial_closed_form and rabenseifner_closed_form likewise for the
  reference's other reductions (a left fold; a truncated pairwise tree with
  the right subtree first; each element's block owner's tree);
* hash3 / gen: the SplitMix64-keyed input generator in numpy;
* fcollect / alltoall / broadcast / collect closed forms as numpy
  concatenations, alltoalls as one fancy-index gather.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

OPS = {"and": 0, "or": 1, "xor": 2, "min": 3, "max": 4, "sum": 5, "prod": 6}
DTYPES = {"u8": 0, "u16": 1, "u32": 2, "u64": 3, "i8": 4, "i16": 5, "i32": 6,
          "i64": 7, "f32": 8, "f64": 9, "c32": 10, "c64": 11}
NP_DTYPE = {"u8": np.uint8, "u16": np.uint16, "u32": np.uint32,
            "u64": np.uint64, "i8": np.int8, "i16": np.int16,
            "i32": np.int32, "i64": np.int64, "f32": np.float32,
            "f64": np.float64, "c32": np.complex64, "c64": np.complex128}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: make -C oracle")
        L = C.CDLL(LIB_PATH)
        pp = C.POINTER(C.c_void_p)
        L.orc_reduce_rec_dbl.argtypes = [C.c_int, C.c_int, C.c_int,
                                         C.c_size_t, pp, pp]
        for f in ("orc_reduce_linear", "orc_reduce_binomial"):
            getattr(L, f).argtypes = [C.c_int, C.c_int, C.c_int, C.c_size_t,
                                      pp, pp]
        L.orc_reduce_rabenseifner.argtypes = [C.c_int, C.c_int, C.c_int,
                                              C.c_size_t, C.c_int, pp, pp]
        L.orc_reduce_accurate.argtypes = [C.c_int, C.c_int, C.c_int,
                                          C.c_size_t, pp, C.c_void_p,
                                          C.c_void_p]
        L.orc_fsum.restype = C.c_double
        L.orc_fsum.argtypes = [C.c_void_p, C.c_size_t]
        for f in ("orc_fcollect_bruck", "orc_alltoall_shift"):
            getattr(L, f).argtypes = [C.c_int, C.c_size_t, pp, pp]
        L.orc_broadcast_linear.argtypes = [C.c_int, C.c_int, C.c_size_t, pp,
                                           pp]
        L.orc_rotate_inplace.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t]
        L.orc_rotate.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t]
        L.orc_collect_bruck.argtypes = [C.c_int, C.POINTER(C.c_size_t), pp,
                                        pp]
        L.orc_alltoalls_shift.argtypes = [C.c_int, C.c_size_t, C.c_size_t,
                                          C.c_size_t, C.c_size_t, pp, pp]
        L.orc_broadcast_active_set.argtypes = [C.c_int, C.c_int, C.c_size_t,
                                               pp, pp]
        L.orc_strided_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t,
                                       C.c_ssize_t, C.c_ssize_t, C.c_size_t]
        L.orc_strided_copy.restype = None
        L.orc_rotate.restype = None
        L.orc_hash3.restype = C.c_uint64
        L.orc_hash3.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_cpu_bench_procs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_size_t, C.c_int, pp, pp,
                                          C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


# ---- the C restatement ------------------------------------------------------

def rec_dbl(op: str, dtype: str, src: np.ndarray) -> np.ndarray:
    """Reference recursive-doubling reduce of src[P, n]; returns dest[P, n]."""
    src = np.ascontiguousarray(src, dtype=NP_DTYPE[dtype])
    P, n = src.shape
    dest = np.zeros_like(src)
    rows_s = [src[i] for i in range(P)]
    rows_d = [dest[i] for i in range(P)]
    rc = lib().orc_reduce_rec_dbl(OPS[op], DTYPES[dtype], P, n,
                                  _ptrs(rows_s), _ptrs(rows_d))
    if rc != 0:
        raise ValueError(f"oracle rejects {op} on {dtype}")
    return dest


def _reduce_with(fn, op, dtype, src, *extra):
    src = np.ascontiguousarray(src, dtype=NP_DTYPE[dtype])
    P, n = src.shape
    dest = np.zeros_like(src)
    rc = fn(OPS[op], DTYPES[dtype], P, n, *extra,
            _ptrs([src[i] for i in range(P)]),
            _ptrs([dest[i] for i in range(P)]))
    if rc != 0:
        raise ValueError(f"oracle rejects {op} on {dtype}")
    return dest


def linear(op: str, dtype: str, src: np.ndarray) -> np.ndarray:
    """Reference linear reduce (reduce.c:55-91) of src[P, n]."""
    return _reduce_with(lib().orc_reduce_linear, op, dtype, src)


def binomial(op: str, dtype: str, src: np.ndarray) -> np.ndarray:
    """Reference binomial reduce (reduce.c:102-171), children taken in
    ascending order (every child reported before its parent polls)."""
    return _reduce_with(lib().orc_reduce_binomial, op, dtype, src)


def rabenseifner(op: str, dtype: str, src: np.ndarray,
                 ring: bool = False) -> np.ndarray:
    """Reference Rabenseifner reduce (reduce.c:306-511), or its ring-
    allgather variant rabenseifner2 (reduce.c:521-724) with ring=True."""
    return _reduce_with(lib().orc_reduce_rabenseifner, op, dtype, src,
                        int(ring))


def accurate(op: str, dtype: str, src: np.ndarray):
    """(hi, lo) float64 arrays: exact sum / double-double product."""
    src = np.ascontiguousarray(src, dtype=NP_DTYPE[dtype])
    P, n = src.shape
    m = 2 * n if dtype in ("c32", "c64") else n
    hi = np.zeros(m, np.float64)
    lo = np.zeros(m, np.float64)
    rows = [src[i] for i in range(P)]
    rc = lib().orc_reduce_accurate(OPS[op], DTYPES[dtype], P, n, _ptrs(rows),
                                   hi.ctypes.data, lo.ctypes.data)
    if rc != 0:
        raise ValueError("accurate oracle covers float sum/prod only")
    return hi, lo


def fsum(x) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_fsum(x.ctypes.data, x.size)


def fcollect(src: np.ndarray) -> np.ndarray:
    """Bruck in-place fcollect of src[P, nbytes] (uint8); dest[P, P*nbytes]."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    P, nb = src.shape
    dest = np.zeros((P, P * nb), np.uint8)
    lib().orc_fcollect_bruck(P, nb, _ptrs([src[i] for i in range(P)]),
                             _ptrs([dest[i] for i in range(P)]))
    return dest


def alltoall(src: np.ndarray) -> np.ndarray:
    """Shift-exchange alltoall of src[P, P*nbytes] (uint8)."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    P, tot = src.shape
    nb = tot // P
    dest = np.zeros_like(src)
    lib().orc_alltoall_shift(P, nb, _ptrs([src[i] for i in range(P)]),
                             _ptrs([dest[i] for i in range(P)]))
    return dest


def broadcast(src: np.ndarray, root: int) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.uint8)
    P, nb = src.shape
    dest = np.zeros_like(src)
    lib().orc_broadcast_linear(P, root, nb, _ptrs([src[i] for i in range(P)]),
                               _ptrs([dest[i] for i in range(P)]))
    return dest


def rotate_inplace(buf: np.ndarray, dist: int) -> np.ndarray:
    b = np.ascontiguousarray(buf, dtype=np.uint8).copy()
    lib().orc_rotate_inplace(b.ctypes.data, b.size, dist)
    return b


def rotate(buf: np.ndarray, dist: int) -> np.ndarray:
    b = np.ascontiguousarray(buf, dtype=np.uint8).copy()
    lib().orc_rotate(b.ctypes.data, b.size, dist)
    return b


def collect(rows) -> np.ndarray:
    """Bruck collect of per-PE byte rows of any lengths (list of uint8
    arrays); returns dest[P, total]."""
    rows = [np.ascontiguousarray(r, dtype=np.uint8) for r in rows]
    P = len(rows)
    sizes = (C.c_size_t * P)(*[r.size for r in rows])
    total = sum(r.size for r in rows)
    dest = np.zeros((P, max(total, 1)), np.uint8)
    srcs = [r if r.size else np.zeros(1, np.uint8) for r in rows]
    lib().orc_collect_bruck(P, sizes, _ptrs(srcs),
                            _ptrs([dest[i] for i in range(P)]))
    return dest[:, :total]


def alltoalls(src: np.ndarray, es: int, dst: int, sst: int, n: int,
              dest_bytes: int, dest_init=None) -> np.ndarray:
    """Shift-exchange strided alltoall of src[P, bytes] (uint8); returns
    dest[P, dest_bytes] starting from dest_init (zeros if None)."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    P = src.shape[0]
    dest = (np.zeros((P, dest_bytes), np.uint8) if dest_init is None
            else np.ascontiguousarray(dest_init, dtype=np.uint8).copy())
    lib().orc_alltoalls_shift(P, es, dst, sst, n,
                              _ptrs([src[i] for i in range(P)]),
                              _ptrs([dest[i] for i in range(P)]))
    return dest


def broadcast_active_set(src: np.ndarray, root: int,
                         dest_init: np.ndarray) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.uint8)
    P, nb = src.shape
    dest = np.ascontiguousarray(dest_init, dtype=np.uint8).copy()
    lib().orc_broadcast_active_set(P, root, nb,
                                   _ptrs([src[i] for i in range(P)]),
                                   _ptrs([dest[i] for i in range(P)]))
    return dest


def strided_copy(dst: np.ndarray, src: np.ndarray, es: int, tst: int,
                 sst: int, n: int, dst_base: int = 0,
                 src_base: int = 0) -> np.ndarray:
    """Element-wise strided copy between byte buffers; element 0 sits at
    byte dst_base / src_base (so negative strides stay inside)."""
    d = np.ascontiguousarray(dst, dtype=np.uint8).copy()
    s = np.ascontiguousarray(src, dtype=np.uint8)
    lib().orc_strided_copy(d.ctypes.data + dst_base, s.ctypes.data + src_base,
                           es, tst, sst, n)
    return d


def cpu_bench_procs(kind: str, npes: int, n: int, iters: int, op="sum",
                    dtype="u8", src=None, want_dest=False):
    """One forked process per PE over one MAP_SHARED mapping (the reference's
    process-per-PE model): kind "reduce" (rec_dbl, n elements of dtype),
    "fcollect" (Bruck in place, n bytes per PE) or "alltoall" (shift
    exchange, n-byte blocks).  Returns the median seconds per call (the
    slowest PE), and every PE's destination when want_dest."""
    k = {"reduce": 0, "fcollect": 1, "alltoall": 2}[kind]
    es = np.dtype(NP_DTYPE[dtype]).itemsize if k == 0 else 1
    dlen = n * es if k == 0 else n * npes
    dest = np.zeros((npes, dlen), np.uint8) if want_dest else None
    sp = None
    if src is not None:
        src = np.ascontiguousarray(src)
        sp = _ptrs([src[i] for i in range(npes)])
    dp = _ptrs([dest[i] for i in range(npes)]) if want_dest else None
    sec = C.c_double()
    rc = lib().orc_cpu_bench_procs(k, OPS[op], DTYPES[dtype], npes, n, iters,
                                   sp, dp, C.byref(sec))
    if rc != 0:
        raise RuntimeError(f"process baseline failed ({rc})")
    return (sec.value, dest) if want_dest else sec.value


# ---- independent numpy formulations (pins) ----------------------------------

_G = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def hash3(seed: int, pe, idx) -> np.ndarray:
    """SplitMix64-keyed hash h(seed, pe, idx), vectorised over pe/idx."""
    pe = np.asarray(pe, dtype=np.uint64)
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        s = _mix64(np.asarray([seed], dtype=np.uint64))
        a = _mix64(s + _G * (pe + np.uint64(1)))
        return _mix64(a + _G * (idx + np.uint64(1)))


def _umulhi(h: np.ndarray, r: int) -> np.ndarray:
    """high 64 bits of h * r for r < 2**64, exact."""
    h = h.astype(np.uint64)
    lo32 = np.uint64(0xFFFFFFFF)
    r = np.uint64(r)
    r_lo, r_hi = r & lo32, r >> np.uint64(32)
    h_lo, h_hi = h & lo32, h >> np.uint64(32)
    with np.errstate(over="ignore"):
        ll = h_lo * r_lo
        lh = h_lo * r_hi
        hl = h_hi * r_lo
        hh = h_hi * r_hi
        mid = (ll >> np.uint64(32)) + (lh & lo32) + (hl & lo32)
        return hh + (lh >> np.uint64(32)) + (hl >> np.uint64(32)) + \
            (mid >> np.uint64(32))


def gen_int_range(seed, pes, n, lo, hi, dtype, idx=None) -> np.ndarray:
    """Integer uniform [lo, hi] per (seed, pe, idx): bit-identical to the
    device generator (OSC_DIST_INT_RANGE).  idx selects element indices
    (default 0..n-1)."""
    pe = np.asarray(pes, dtype=np.uint64)[:, None]
    idx = (np.arange(n, dtype=np.uint64) if idx is None else
           np.asarray(idx, dtype=np.uint64))[None, :]
    h = hash3(seed, pe, idx)
    rng = (hi - lo + 1) % (1 << 64)
    with np.errstate(over="ignore"):
        v = np.uint64(lo % (1 << 64)) + (_umulhi(h, rng) if rng else h)
    return v.astype(np.uint64).astype(NP_DTYPE[dtype]) if dtype.startswith(
        "u") else v.view(np.int64).astype(NP_DTYPE[dtype])


def _bin(op, a, b):
    """reduce.c:730-736 on numpy arrays of one dtype (elementwise).  A
    complex product is C99's finite formula (ac - bd) + (ad + bc)i in real
    arithmetic: numpy's own complex multiply loop does not round the same
    way in every last bit."""
    if op == "sum":
        return a + b
    if op == "prod":
        if np.iscomplexobj(a):
            ar, ai, br, bi = a.real, a.imag, b.real, b.imag
            out = np.empty(np.broadcast(a, b).shape, a.dtype)
            out.real = ar * br - ai * bi
            out.imag = ar * bi + ai * br
            return out
        return a * b
    if op == "max":
        return np.where(a > b, a, b)
    if op == "min":
        return np.where(a < b, a, b)
    if op == "and":
        return a & b
    if op == "or":
        return a | b
    return a ^ b


def reduce_closed_form(op: str, src: np.ndarray) -> np.ndarray:
    """Per-member closed form of recursive doubling on src[P, n]: member k
    (team index m_k = ceil(k*P/p2s)) gets the balanced pairwise tree over
    leaves k^0..k^(p2s-1), leaf j = op(x[m_j+1], x[m_j]) when m_j+1 is not
    a member, else x[m_j]; the folded neighbour gets its member's value."""
    P = src.shape[0]
    p2s = 1
    while p2s * 2 <= P:
        p2s *= 2
    m = [-(-k * P // p2s) for k in range(p2s + 1)]
    with np.errstate(all="ignore"):
        leaves = []
        for k in range(p2s):
            if m[k + 1] - m[k] == 2:
                leaves.append(_bin(op, src[m[k] + 1], src[m[k]]))
            else:
                leaves.append(src[m[k]])
        out = np.empty_like(src)
        for k in range(p2s):
            level = [leaves[k ^ j] for j in range(p2s)]
            while len(level) > 1:
                level = [_bin(op, level[2 * t], level[2 * t + 1])
                         for t in range(len(level) // 2)]
            out[m[k]] = level[0]
            if m[k + 1] - m[k] == 2:
                out[m[k] + 1] = level[0]
    return out


def linear_closed_form(op: str, src: np.ndarray) -> np.ndarray:
    """Linear reduce (reduce.c:55-91) as a left fold over the team,
    acc = op(acc, x_i), the same value on every member."""
    with np.errstate(all="ignore"):
        acc = src[0].copy()
        for i in range(1, src.shape[0]):
            acc = _bin(op, acc, src[i])
    return np.tile(acc, (src.shape[0], 1))


def binomial_closed_form(op: str, src: np.ndarray) -> np.ndarray:
    """Binomial reduce (reduce.c:102-171) with children taken in ascending
    order, written as a tree: the node of the aligned block [lo, lo + s) is
    op(node of its right half, node of its left half), a right half that
    starts at or past P is absent; the root covers [0, 2^ceil(log2 P))."""
    P = src.shape[0]

    def node(lo, size):
        if size == 1:
            return src[lo]
        half = size // 2
        left = node(lo, half)
        if lo + half >= P:
            return left
        return _bin(op, node(lo + half, half), left)

    top = 1
    while top < P:
        top *= 2
    with np.errstate(all="ignore"):
        r = node(0, top)
    return np.tile(r, (P, 1))


def rabenseifner_closed_form(op: str, src: np.ndarray) -> np.ndarray:
    """Rabenseifner reduce (reduce.c:306-511, :521-724): element e of n
    belongs to block b, the largest b with floor(b*n/p2s) <= e, which the
    reduce-scatter leaves on member k = reverse_bits(b); its value is
    member k's recursive-doubling tree, whose folded leaves are
    op(x[m+1], x[m]) below n//2 and op(x[m], x[m+1]) from n//2 on; every
    member (and folded neighbour) receives it."""
    P, n = src.shape
    p2s, lg = 1, 0
    while p2s * 2 <= P:
        p2s, lg = p2s * 2, lg + 1
    m = [-(-k * P // p2s) for k in range(p2s + 1)]
    upper = np.arange(n) >= n // 2
    with np.errstate(all="ignore"):
        leaves = []
        for k in range(p2s):
            if m[k + 1] - m[k] == 2:
                a, b = src[m[k]], src[m[k] + 1]
                leaves.append(np.where(upper, _bin(op, a, b), _bin(op, b, a)))
            else:
                leaves.append(src[m[k]])
        per_member = []
        for k in range(p2s):
            level = [leaves[k ^ j] for j in range(p2s)]
            while len(level) > 1:
                level = [_bin(op, level[2 * t], level[2 * t + 1])
                         for t in range(len(level) // 2)]
            per_member.append(level[0])
    out = np.empty(n, src.dtype)
    bounds = [(b * n) // p2s for b in range(p2s + 1)]
    for b in range(p2s):
        k = int(format(b, f"0{lg}b")[::-1], 2) if lg else 0
        out[bounds[b]:bounds[b + 1]] = per_member[k][bounds[b]:bounds[b + 1]]
    return np.tile(out, (P, 1))


def fcollect_closed_form(src: np.ndarray) -> np.ndarray:
    P = src.shape[0]
    row = src.reshape(-1)
    return np.tile(row, (P, 1))


def alltoall_closed_form(src: np.ndarray) -> np.ndarray:
    P, tot = src.shape
    nb = tot // P
    blocks = src.reshape(P, P, nb)          # [i, j] = src_i block j
    return blocks.transpose(1, 0, 2).reshape(P, tot)  # dest_j block i


def collect_closed_form(rows) -> np.ndarray:
    """dest_j = concatenation of every PE's row, in PE order."""
    cat = np.concatenate([np.asarray(r, np.uint8) for r in rows])
    return np.tile(cat, (len(rows), 1))


def alltoalls_closed_form(src: np.ndarray, es: int, dst: int, sst: int,
                          n: int, dest_bytes: int) -> np.ndarray:
    """dest_j[(i*n + t)*dst] = src_i[(j*n + t)*sst], written as one fancy-
    index gather over (i, t, byte) per destination j."""
    P = src.shape[0]
    dest = np.zeros((P, dest_bytes), np.uint8)
    i = np.arange(P)[:, None, None]
    t = np.arange(n)[None, :, None]
    b = np.arange(es)[None, None, :]
    for j in range(P):
        d_idx = ((i * n + t) * dst) * es + b                  # [P, n, es]
        s_idx = np.broadcast_to(((j * n + t) * sst) * es + b, d_idx.shape)
        rows = np.broadcast_to(i, d_idx.shape)
        dest[j, d_idx.ravel()] = src[rows.ravel(), s_idx.ravel()]
    return dest
