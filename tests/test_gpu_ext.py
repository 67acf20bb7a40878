"""GPU parity of the collectives and RMA beside the hot path: variable-size
collect, strided alltoalls, the active-set (pre-1.5) forms, typed and
strided put/get and put-with-signal.  Every result is compared bit for bit
with the oracle's step-by-step restatement of the reference algorithm
(Bruck collect with prefix sum and rotation, shift-exchange alltoalls,
recursive-doubling reduce, active-set broadcast, element-wise strided
copy)."""
import ctypes as C
import zlib

import numpy as np
import pytest

import osc
import oracle as O

pytestmark = pytest.mark.gpu

NPES = 16
HEAP = 16 << 20
_H = []


@pytest.fixture
def heap():
    if not _H or not _H[0].alive:
        _H[:] = [osc.Heap(npes=NPES, heap_bytes=HEAP)]
    return _H[0]


def members(start, stride, size):
    return [start + k * stride for k in range(size)]


def put_rows(h, sym, rows, pes):
    for r, pe in zip(rows, pes):
        if len(r):
            h.put(sym, np.ascontiguousarray(r), pe)


def get_rows(h, sym, nbytes, pes):
    return np.stack([h.get(sym, nbytes, np.uint8, pe) for pe in pes])


# ---- collect ------------------------------------------------------------------

@pytest.mark.parametrize("team", [(0, 1, 16), (1, 3, 5), (2, 1, 1),
                                  (0, 2, 7)])
@pytest.mark.parametrize("scale", [1, 37, 5000])
def test_collect_variable_sizes(heap, team, scale):
    start, stride, P = team
    pes = members(start, stride, P)
    t = heap.world if P == NPES else heap.split(start, stride, P)
    rng = np.random.default_rng(scale * 31 + P)
    sizes = rng.integers(0, 3 * scale, P)
    sizes[rng.integers(0, P)] = 0
    rows = [rng.integers(0, 256, int(k), dtype=np.uint8) for k in sizes]
    total = int(sizes.sum())
    s = heap.malloc(int(sizes.max()) + 1)
    d = heap.malloc(total + 64)
    try:
        put_rows(heap, s, rows, pes)
        guard = np.full(total + 64, 0xA5, np.uint8)
        for pe in pes:
            heap.put(d, guard, pe)
        arr = (C.c_size_t * P)(*[int(k) for k in sizes])
        osc.check(heap.L.osc_collect(t, d, s, arr, None), "osc_collect")
        osc.check(heap.L.osc_stream_sync(None), "sync")
        got = get_rows(heap, d, total + 64, pes)
        want = O.collect(rows)
        assert np.array_equal(got[:, :total], want)
        assert np.all(got[:, total:] == 0xA5)      # nothing past the end
    finally:
        heap.free(s)
        heap.free(d)


def test_collect_typed_and_mem(heap):
    rng = np.random.default_rng(2)
    P, n = NPES, 1000
    pes = list(range(P))
    src = rng.integers(-2 ** 31, 2 ** 31, (P, n), dtype=np.int32)
    s, d = heap.malloc(n * 4), heap.malloc(P * n * 4)
    try:
        for k in pes:
            heap.put(s, src[k], k)
        assert osc.typed("shmem_int_collect")(heap.world, d, s, n) == 0
        got = np.stack([heap.get(d, P * n, np.int32, k) for k in pes])
        assert np.array_equal(got, np.tile(src.reshape(-1), (P, 1)))
        assert heap.L.shmem_collectmem(heap.world, d, s, 3 * 4) == 0
        got = get_rows(heap, d, P * 12, pes)
        want = O.collect([src[k].view(np.uint8)[:12] for k in pes])
        assert np.array_equal(got, want)
    finally:
        heap.free(s)
        heap.free(d)


# ---- alltoalls ------------------------------------------------------------------

@pytest.mark.parametrize("tn,es", [("int64", 8), ("int", 4), ("short", 2),
                                   ("char", 1), ("longdouble", 16)])
@pytest.mark.parametrize("dst,sst,n", [(1, 1, 5), (2, 3, 4), (3, 1, 1),
                                       (1, 2, 300)])
def test_alltoalls_typed(heap, tn, es, dst, sst, n):
    start, stride, P = 1, 2, 7
    pes = members(start, stride, P)
    team = heap.split(start, stride, P)
    rng = np.random.default_rng(es * 1000 + dst * 10 + sst + n)
    sbytes = ((P * n - 1) * sst + 1) * es
    dbytes = ((P * n - 1) * dst + 1) * es
    src = rng.integers(0, 256, (P, sbytes), dtype=np.uint8)
    init = rng.integers(0, 256, (P, dbytes), dtype=np.uint8)
    s, d = heap.malloc(sbytes), heap.malloc(dbytes)
    try:
        put_rows(heap, s, src, pes)
        put_rows(heap, d, init, pes)
        rc = osc.typed(f"shmem_{tn}_alltoalls")(team, d, s, dst, sst, n)
        assert rc == 0
        got = get_rows(heap, d, dbytes, pes)
        want = O.alltoalls(src, es, dst, sst, n, dbytes, dest_init=init)
        assert np.array_equal(got, want)
    finally:
        heap.free(s)
        heap.free(d)


def test_alltoallsmem_moves_one_element_of_nelems_bytes(heap):
    P, es, dst, sst = NPES, 3, 2, 5
    pes = list(range(P))
    rng = np.random.default_rng(77)
    sbytes, dbytes = ((P - 1) * sst + 1) * es, ((P - 1) * dst + 1) * es
    src = rng.integers(0, 256, (P, sbytes), dtype=np.uint8)
    init = np.zeros((P, dbytes), np.uint8)
    s, d = heap.malloc(sbytes), heap.malloc(dbytes)
    try:
        put_rows(heap, s, src, pes)
        put_rows(heap, d, init, pes)
        assert heap.L.shmem_alltoallsmem(heap.world, d, s, dst, sst, es) == 0
        got = get_rows(heap, d, dbytes, pes)
        assert np.array_equal(got, O.alltoalls(src, es, dst, sst, 1, dbytes))
    finally:
        heap.free(s)
        heap.free(d)


def test_alltoalls_rejects_bad_strides(heap):
    s, d = heap.malloc(4096), heap.malloc(4096)
    assert osc.typed("shmem_int_alltoalls")(heap.world, d, s, 0, 1, 1) != 0
    assert osc.typed("shmem_int_alltoalls")(heap.world, d, s, 1, -1, 1) != 0
    heap.free(s)
    heap.free(d)


# ---- active-set forms ----------------------------------------------------------

TO_ALL = [("int", "i32", "sum"), ("long", "i64", "prod"),
          ("short", "i16", "and"), ("longlong", "i64", "xor"),
          ("int", "i32", "min"), ("long", "i64", "max"),
          ("double", "f64", "sum"), ("float", "f32", "prod"),
          ("float", "f32", "min"), ("complexd", "c64", "sum")]


@pytest.mark.parametrize("tn,dt,op", TO_ALL)
@pytest.mark.parametrize("aset", [(1, 1, 3), (0, 1, 8), (5, 0, 11)])
def test_to_all_matches_rec_dbl(heap, tn, dt, op, aset):
    """shmem_<T>_<op>_to_all over the active set equals the reference
    recursive doubling bit for bit (floats in reference order)."""
    from gpu_util import inputs
    PE_start, log_stride, PE_size = aset
    pes = members(PE_start, 1 << log_stride, PE_size)
    rng = np.random.default_rng(zlib.crc32(repr((tn, op, aset)).encode()))
    n = 777
    src = inputs(rng, op, dt, PE_size, n)
    es = src.dtype.itemsize
    s, d = heap.malloc(n * es), heap.malloc(n * es)
    try:
        for k, pe in enumerate(pes):
            heap.put(s, src[k], pe)
        osc.check(heap.L.osc_set_fp_mode(osc.FP_REFERENCE), "mode")
        heap.L.osc_clear_error()
        osc.typed(f"shmem_{tn}_{op}_to_all")(d, s, n, PE_start, log_stride,
                                             PE_size, None, None)
        assert heap.L.osc_last_error_code() == 0, \
            heap.L.osc_last_error().decode()
        got = np.stack([heap.get(d, n, src.dtype, pe) for pe in pes])
        want = O.rec_dbl(op, dt, src)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    finally:
        heap.L.osc_set_fp_mode(osc.FP_ACCURATE)
        heap.free(s)
        heap.free(d)


def test_to_all_in_place_and_errors(heap):
    n, pes = 100, members(2, 2, 6)
    rng = np.random.default_rng(5)
    src = rng.integers(-1000, 1000, (6, n), dtype=np.int64)
    s = heap.malloc(n * 8)
    try:
        for k, pe in enumerate(pes):
            heap.put(s, src[k], pe)
        heap.L.osc_clear_error()
        osc.typed("shmem_long_sum_to_all")(s, s, n, 2, 1, 6, None, None)
        assert heap.L.osc_last_error_code() == 0
        got = np.stack([heap.get(s, n, np.int64, pe) for pe in pes])
        assert np.array_equal(got, np.tile(src.sum(0), (6, 1)))
        before = heap.get(s, n, np.int64, 2)
        heap.L.osc_clear_error()
        # active set past the last PE: rejected, target untouched
        osc.typed("shmem_long_sum_to_all")(s, s, n, 2, 2, 6, None, None)
        assert heap.L.osc_last_error_code() == -3
        assert np.array_equal(heap.get(s, n, np.int64, 2), before)
        heap.L.osc_clear_error()
        osc.typed("shmem_longdouble_sum_to_all")(s, s, 1, 0, 0, 2, None, None)
        assert heap.L.osc_last_error_code() == -7
        heap.L.osc_clear_error()
        osc.typed("shmem_int_sum_to_all")(s, s, -1, 0, 0, 2, None, None)
        assert heap.L.osc_last_error_code() == -1
    finally:
        heap.L.osc_clear_error()
        heap.free(s)


@pytest.mark.parametrize("size", [32, 64])
def test_sized_collectives(heap, size):
    L = heap.L
    es = size // 8
    PE_start, log_stride, P = 3, 1, 6
    pes = members(PE_start, 2, P)
    n = 129
    rng = np.random.default_rng(size)
    s, d = heap.malloc(P * n * es * 3), heap.malloc(P * n * es * 3)
    L.osc_clear_error()
    try:
        # fcollect / collect
        src = rng.integers(0, 256, (P, n * es), dtype=np.uint8)
        put_rows(heap, s, src, pes)
        getattr(L, f"shmem_fcollect{size}")(d, s, n, PE_start, log_stride, P,
                                            None)
        assert L.osc_last_error_code() == 0
        assert np.array_equal(get_rows(heap, d, P * n * es, pes),
                              O.fcollect(src))
        getattr(L, f"shmem_collect{size}")(d, s, n, PE_start, log_stride, P,
                                           None)
        assert np.array_equal(get_rows(heap, d, P * n * es, pes),
                              O.collect(list(src)))
        # alltoall
        src = rng.integers(0, 256, (P, P * n * es), dtype=np.uint8)
        put_rows(heap, s, src, pes)
        getattr(L, f"shmem_alltoall{size}")(d, s, n, PE_start, log_stride, P,
                                            None)
        assert np.array_equal(get_rows(heap, d, P * n * es, pes),
                              O.alltoall(src))
        # alltoalls, dst 2, sst 1, 5 elements per pair
        dbytes = ((P * 5 - 1) * 2 + 1) * es
        init = rng.integers(0, 256, (P, dbytes), dtype=np.uint8)
        put_rows(heap, d, init, pes)
        getattr(L, f"shmem_alltoalls{size}")(d, s, 2, 1, 5, PE_start,
                                             log_stride, P, None)
        want = O.alltoalls(src[:, :P * 5 * es], es, 2, 1, 5, dbytes,
                           dest_init=init)
        assert np.array_equal(get_rows(heap, d, dbytes, pes), want)
        # broadcast: the root's target keeps its old contents
        src = rng.integers(0, 256, (P, n * es), dtype=np.uint8)
        init = rng.integers(0, 256, (P, n * es), dtype=np.uint8)
        put_rows(heap, s, src, pes)
        put_rows(heap, d, init, pes)
        getattr(L, f"shmem_broadcast{size}")(d, s, n, 4, PE_start,
                                             log_stride, P, None)
        assert L.osc_last_error_code() == 0
        assert np.array_equal(get_rows(heap, d, n * es, pes),
                              O.broadcast_active_set(src, 4, init))
        L.shmem_barrier(PE_start, log_stride, P, None)
        L.shmem_sync_deprecated(0, 0, NPES, None)
        L.shmem_sync_all()
        assert L.osc_last_error_code() == 0
    finally:
        L.osc_clear_error()
        heap.free(s)
        heap.free(d)


# ---- RMA -------------------------------------------------------------------------

def test_typed_put_get_p_g(heap):
    rng = np.random.default_rng(8)
    a = heap.malloc(8192)
    try:
        x = rng.integers(-2 ** 62, 2 ** 62, 1000, dtype=np.int64)
        osc.typed("shmem_long_put")(a, x.ctypes.data, 1000, 11)
        assert np.array_equal(heap.get(a, 1000, np.int64, 11), x)
        out = np.zeros(1000, np.int64)
        osc.typed("shmem_int64_get")(out.ctypes.data, a, 1000, 11)
        assert np.array_equal(out, x)
        osc.typed("shmem_double_p")(a + 8, 2.5, 6)
        assert osc.typed("shmem_double_g")(a + 8, 6) == 2.5
        osc.typed("shmem_short_p")(a, -7, 6)
        assert osc.typed("shmem_short_g")(a, 6) == -7
        # symmetric -> symmetric, non-blocking, completed by quiet
        b = heap.malloc(8192)
        heap.put(b, x, 0)
        osc.typed("shmem_int64_put_nbi")(a, b, 1000, 3)
        heap.L.shmem_quiet()
        assert np.array_equal(heap.get(a, 1000, np.int64, 3), x)
        heap.L.shmem_get32(out.ctypes.data, a, 2000, 3)
        assert np.array_equal(out, x)
        heap.free(b)
        # out-of-range PE: no write, error recorded
        heap.L.osc_clear_error()
        osc.typed("shmem_long_put")(a, x.ctypes.data, 1, NPES)
        assert heap.L.osc_last_error_code() == -1
    finally:
        heap.L.osc_clear_error()
        heap.free(a)


@pytest.mark.parametrize("es,tst,sst,n", [(4, 1, 1, 100), (8, 3, 2, 50),
                                          (2, 1, 7, 33), (16, 2, 1, 9),
                                          (4, 0, 1, 10)])
def test_iput_iget(heap, es, tst, sst, n):
    L = heap.L
    rng = np.random.default_rng(es + tst * 10 + n)
    span = (max(n - 1, 0) * max(tst, sst, 1) + 1) * es
    sym = heap.malloc(span + 64)
    loc = heap.malloc(span + 64)
    try:
        init = rng.integers(0, 256, span, dtype=np.uint8)
        srcb = rng.integers(0, 256, span, dtype=np.uint8)
        # host source -> PE 5
        heap.put(sym, init, 5)
        getattr(L, f"shmem_iput{es * 8}")(sym, srcb.ctypes.data, tst, sst, n,
                                          5)
        want = O.strided_copy(init, srcb, es, tst, sst, n)
        assert np.array_equal(heap.get(sym, span, np.uint8, 5), want)
        # symmetric source (PE 0's copy) -> PE 9
        heap.put(sym, init, 9)
        heap.put(loc, srcb, 0)
        getattr(L, f"shmem_iput{es * 8}")(sym, loc, tst, sst, n, 9)
        assert np.array_equal(heap.get(sym, span, np.uint8, 9), want)
        # PE 5 -> host, strided both sides
        heap.put(sym, srcb, 5)
        hb = init.copy()
        getattr(L, f"shmem_iget{es * 8}")(hb.ctypes.data, sym, tst, sst, n, 5)
        assert np.array_equal(hb, O.strided_copy(init, srcb, es, tst, sst, n))
    finally:
        heap.free(sym)
        heap.free(loc)


def test_typed_iput_negative_stride(heap):
    a = heap.malloc(4096)
    try:
        x = np.arange(10, dtype=np.int32)
        heap.put(a, np.zeros(10, np.int32), 2)
        # element 0 of the source is its last element in memory
        src_last = x.ctypes.data + 9 * 4
        osc.typed("shmem_int_iput")(a, src_last, 1, -1, 10, 2)
        assert np.array_equal(heap.get(a, 10, np.int32, 2), x[::-1])
    finally:
        heap.free(a)


def test_put_signal(heap):
    L = heap.L
    d = heap.malloc(1 << 16)
    sig = heap.malloc(8)
    try:
        x = np.arange(4096, dtype=np.int64)
        heap.put(sig, np.array([5], np.uint64), 7)
        osc.typed("shmem_long_put_signal")(d, x.ctypes.data, 4096, sig, 3,
                                           1, 7)    # SHMEM_SIGNAL_ADD
        assert np.array_equal(heap.get(d, 4096, np.int64, 7), x)
        sig7 = L.shmem_ptr(sig, 7)
        assert L.shmem_signal_fetch(sig7) == 8
        L.shmem_putmem_signal_nbi(d, x.ctypes.data, 64, sig, 42, 0, 7)
        assert L.shmem_signal_wait_until(sig7, 0, 42) == 42   # CMP_EQ
        L.shmem_put64_signal(d, x.ctypes.data, 10, sig, 1, 1, 7)
        assert L.shmem_signal_wait_until(sig7, 5, 43) == 43   # CMP_GE
        L.osc_clear_error()
        L.shmem_putmem_signal(d, x.ctypes.data, 8, sig, 1, 9, 7)  # bad op
        assert L.osc_last_error_code() == -1
        assert L.shmem_signal_fetch(sig7) == 43
    finally:
        L.osc_clear_error()
        heap.free(d)
        heap.free(sig)


def test_to_all_algorithm_variable(monkeypatch):
    """SHMEM_SUM_TO_ALL_ALGO=rec_dbl gives the active-set sum the reference's
    bits while the team sum keeps its own (accurate) setting."""
    from gpu_util import max_rel_err
    monkeypatch.setenv("SHMEM_SUM_TO_ALL_ALGO", "rec_dbl")
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    P, n = 7, 4096
    with osc.Heap(npes=P, heap_bytes=4 << 20) as h:
        rng = np.random.default_rng(3)
        src = rng.standard_normal((P, n)).astype(np.float32)
        s, d = h.malloc(n * 4), h.malloc(n * 4)
        for pe in range(P):
            h.put(s, src[pe], pe)
        h.L.osc_clear_error()
        osc.typed("shmem_float_sum_to_all")(d, s, n, 0, 0, P, None, None)
        assert h.L.osc_last_error_code() == 0
        got = np.stack([h.get(d, n, np.float32, pe) for pe in range(P)])
        assert np.array_equal(got.view(np.uint32),
                              O.rec_dbl("sum", "f32", src).view(np.uint32))
        assert osc.typed("shmem_float_sum_reduce")(h.world, d, s, n) == 0
        team = np.stack([h.get(d, n, np.float32, pe) for pe in range(P)])
        assert max_rel_err(team, src, "sum", "f32") <= 2.0 ** -23


def test_rma_symmetric_local_side_is_range_checked(heap):
    """A symmetric local buffer must hold the whole transfer inside its
    PE's heap: a get into (or put from) the last bytes of PE 0's heap with
    a byte count that is valid on the remote side is OSC_ERR_PTR, and the
    bytes past the end (PE 1's slice) stay untouched."""
    L = heap.L
    a = heap.malloc(8192)
    hb = L.osc_heap_bytes()
    near_end = L.osc_heap_base() + hb - 100       # PE 0's copy, 100 B left
    guard = np.arange(4096, dtype=np.uint8) ^ 0x5A
    nxt = L.osc_heap_base() + L.osc_slice_bytes()  # PE 1's slice start
    try:
        heap.put(a, np.full(4096, 7, np.uint8), 3)
        L.osc_put_host(C.c_void_p(nxt), guard.ctypes.data, 4096, 0)
        cases = [
            lambda: L.shmem_getmem(C.c_void_p(near_end), a, 4096, 3),
            lambda: L.shmem_putmem(a, C.c_void_p(near_end), 4096, 3),
            lambda: L.shmem_iget64(C.c_void_p(near_end), a, 1, 1, 512, 3),
            lambda: L.shmem_iput64(a, C.c_void_p(near_end), 1, 1, 512, 3),
        ]
        for k, call in enumerate(cases):
            L.osc_clear_error()
            call()
            assert L.osc_last_error_code() == -4, k
            L.shmem_quiet()
            got = np.zeros(4096, np.uint8)
            L.osc_get_host(got.ctypes.data, C.c_void_p(nxt), 4096, 0)
            assert np.array_equal(got, guard), k
        # the remote side is untouched by the rejected puts
        assert np.array_equal(heap.get(a, 4096, np.uint8, 3),
                              np.full(4096, 7, np.uint8))
        # the same calls within the heap still work
        L.osc_clear_error()
        L.shmem_getmem(C.c_void_p(near_end), a, 100, 3)
        assert L.osc_last_error_code() == 0
    finally:
        L.osc_clear_error()
        heap.free(a)
