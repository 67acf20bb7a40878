"""Communication contexts through the C ABI: each context is its own
device stream; on a context made from a team, PE numbers are team
indices.

* create / get_team / destroy, the default context, invalid handles;
* puts, gets, strided and signal puts and atomics on a world context and
  on a team context (team-relative PEs, outsiders rejected and untouched);
* non-blocking puts on two contexts completed by each context's quiet;
* contexts left undestroyed are released when the library finalises;
* SPMD mode: every PE thread puts to its neighbour through a context of
  its own."""
import ctypes as C

import numpy as np
import pytest

import osc

pytestmark = pytest.mark.gpu

VP, I, SZ, L64 = C.c_void_p, C.c_int, C.c_size_t, C.c_long


def fn(name, res, args):
    f = getattr(osc.lib(), name)
    f.restype, f.argtypes = res, args
    return f


def ctx_fns():
    return dict(
        create=fn("shmem_ctx_create", I, [C.c_long, C.POINTER(VP)]),
        team_create=fn("shmem_team_create_ctx", I, [VP, C.c_long, C.POINTER(VP)]),
        destroy=fn("shmem_ctx_destroy", None, [VP]),
        get_team=fn("shmem_ctx_get_team", I, [VP, C.POINTER(VP)]),
        quiet=fn("shmem_ctx_quiet", None, [VP]),
        fence=fn("shmem_ctx_fence", None, [VP]),
        putmem=fn("shmem_ctx_putmem", None, [VP, VP, VP, SZ, I]),
        putmem_nbi=fn("shmem_ctx_putmem_nbi", None, [VP, VP, VP, SZ, I]),
        getmem=fn("shmem_ctx_getmem", None, [VP, VP, VP, SZ, I]),
        long_p=fn("shmem_ctx_long_p", None, [VP, VP, L64, I]),
        long_g=fn("shmem_ctx_long_g", L64, [VP, VP, I]),
        long_iput=fn("shmem_ctx_long_iput", None,
                     [VP, VP, VP, C.c_ssize_t, C.c_ssize_t, SZ, I]),
        long_put_signal=fn("shmem_ctx_long_put_signal", None,
                           [VP, VP, VP, SZ, VP, C.c_uint64, I, I]),
        finc=fn("shmem_ctx_long_atomic_fetch_inc", L64, [VP, VP, I]),
        add=fn("shmem_ctx_long_atomic_add", None, [VP, VP, L64, I]))


def default_ctx():
    return C.c_void_p.in_dll(osc.lib(), "SHMEM_CTX_DEFAULT").value


@pytest.fixture
def heap():
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    with osc.Heap(npes=6, heap_bytes=80 << 20) as h:
        yield h


def test_create_get_team_destroy(heap):
    F, L = ctx_fns(), osc.lib()
    c, t = C.c_void_p(), C.c_void_p()
    assert F["create"](0, C.byref(c)) == 0 and c.value
    assert F["get_team"](c, C.byref(t)) == 0 and t.value == osc.team_world()
    assert F["get_team"](default_ctx(), C.byref(t)) == 0
    assert t.value == osc.team_world()
    odd = C.c_void_p()
    assert L.shmem_team_split_strided(osc.team_world(), 1, 2, 3, None, 0,
                                      C.byref(odd)) == 0
    tc = C.c_void_p()
    assert F["team_create"](odd, 0, C.byref(tc)) == 0 and tc.value
    assert F["get_team"](tc, C.byref(t)) == 0 and t.value == odd.value
    F["fence"](tc)
    F["quiet"](tc)
    F["destroy"](tc)
    F["destroy"](c)
    L.osc_clear_error()
    F["destroy"](default_ctx())                    # refused, still usable
    assert L.osc_last_error_code() != 0
    L.osc_clear_error()
    assert F["get_team"](None, C.byref(t)) != 0 and not t.value
    bad = C.c_void_p(1)
    assert F["team_create"](None, 0, C.byref(bad)) != 0 and not bad.value
    L.shmem_team_destroy(odd)
    L.osc_clear_error()


def test_world_and_team_context_rma_and_atomics(heap):
    h, F, L = heap, ctx_fns(), osc.lib()
    n = 1000
    src = np.arange(n, dtype=np.int64) * 7 - 3
    d = L.shmem_calloc(n + 8, 8)
    sig = L.shmem_calloc(1, 8)
    c = C.c_void_p()
    assert F["create"](0, C.byref(c)) == 0
    F["putmem"](c, d, src.ctypes.data, src.nbytes, 4)
    assert np.array_equal(h.get(d, n, np.int64, 4), src)
    out = np.zeros(n, np.int64)
    F["getmem"](c, out.ctypes.data, d, out.nbytes, 4)
    assert np.array_equal(out, src)
    F["long_p"](c, d + 8 * n, 99, 1)
    assert F["long_g"](c, d + 8 * n, 1) == 99
    # team context: members 1, 3, 5; PE numbers are team indices
    odd = C.c_void_p()
    assert L.shmem_team_split_strided(osc.team_world(), 1, 2, 3, None, 0,
                                      C.byref(odd)) == 0
    tc = C.c_void_p()
    assert F["team_create"](odd, 0, C.byref(tc)) == 0
    F["putmem"](tc, d, src.ctypes.data, 80, 2)            # -> global PE 5
    assert np.array_equal(h.get(d, 10, np.int64, 5), src[:10])
    assert F["long_g"](tc, d, 2) == src[0]
    F["long_iput"](tc, d + 8 * 20, src.ctypes.data, 2, 1, 5, 0)   # -> PE 1
    got = h.get(d + 8 * 20, 9, np.int64, 1)
    assert np.array_equal(got[::2], src[:5])
    F["long_put_signal"](tc, d + 8 * 40, src.ctypes.data, 4, sig, 17, 0, 1)
    assert h.get(sig, 1, np.uint64, 3)[0] == 17           # index 1 -> PE 3
    assert np.array_equal(h.get(d + 8 * 40, 4, np.int64, 3), src[:4])
    assert F["finc"](tc, d + 8 * (n + 1), 2) == 0
    F["add"](tc, d + 8 * (n + 1), 5, 2)
    assert h.get(d + 8 * (n + 1), 1, np.int64, 5)[0] == 6
    # an index outside the team: error, nothing written anywhere
    before = [h.get(d + 8 * n, 1, np.int64, pe)[0] for pe in range(6)]
    L.osc_clear_error()
    F["long_p"](tc, d + 8 * n, -1, 3)
    assert L.osc_last_error_code() != 0
    assert [h.get(d + 8 * n, 1, np.int64, pe)[0] for pe in range(6)] == before
    L.osc_clear_error()
    F["destroy"](tc)
    F["destroy"](c)
    L.shmem_team_destroy(odd)


def test_nbi_on_two_contexts_completed_by_each_quiet(heap):
    h, F, L = heap, ctx_fns(), osc.lib()
    nbytes = 32 << 20                                   # 4 MiB x 8 pieces
    a = L.shmem_malloc(nbytes)
    b = L.shmem_malloc(nbytes)
    osc.check(L.osc_fill(osc.team_world(), a, nbytes // 8,
                         osc.DTYPES["u64"], osc.DIST["bits"], 0.0, 0.0, 7,
                         None), "fill")
    h.sync()
    c1, c2 = C.c_void_p(), C.c_void_p()
    assert F["create"](0, C.byref(c1)) == 0 and F["create"](0, C.byref(c2)) == 0
    half = nbytes // 2
    F["putmem_nbi"](c1, b, a, half, 2)              # device to device
    F["putmem_nbi"](c2, b + half, a + half, half, 2)
    F["quiet"](c1)
    F["quiet"](c2)
    want = h.get(a, nbytes // 8, np.uint64, 0)
    assert np.array_equal(h.get(b, nbytes // 8, np.uint64, 2), want)
    F["destroy"](c1)
    F["destroy"](c2)


def test_undestroyed_contexts_are_released_at_finalize():
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    F = ctx_fns()
    for _ in range(3):
        with osc.Heap(npes=2, heap_bytes=1 << 20):
            cs = [C.c_void_p() for _ in range(4)]
            for c in cs:
                assert F["create"](0, C.byref(c)) == 0
            F["destroy"](cs[0])


PE_FN = C.CFUNCTYPE(C.c_int, C.c_int, C.c_void_p)


def test_spmd_each_pe_puts_through_its_own_context():
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    npes = 5
    F, L = ctx_fns(), osc.lib()
    L.osc_spmd_run.restype = C.c_int
    L.osc_spmd_run.argtypes = [C.c_int, C.c_size_t, C.c_int, PE_FN, VP]
    got, errors = {}, []

    def body(pe, _arg):
        try:
            buf = L.shmem_calloc(64, 8)
            c = C.c_void_p()
            assert F["create"](0, C.byref(c)) == 0
            L.shmem_barrier_all()
            v = (np.arange(64, dtype=np.int64) + 100 * pe)
            F["putmem_nbi"](c, buf, v.ctypes.data, v.nbytes, (pe + 1) % npes)
            F["quiet"](c)
            L.shmem_barrier_all()
            out = np.zeros(64, np.int64)
            F["getmem"](c, out.ctypes.data, buf, out.nbytes, pe)
            got[pe] = out
            F["destroy"](c)
            L.shmem_barrier_all()
            L.shmem_free(buf)
            return 0
        except Exception as e:  # reported by the test, not lost in a thread
            errors.append((pe, repr(e)))
            return 1
    cb = PE_FN(body)
    assert L.osc_spmd_run(npes, 1 << 20, 0, cb, None) == 0
    assert not errors, errors
    for pe in range(npes):
        left = (pe - 1) % npes
        assert np.array_equal(got[pe], np.arange(64) + 100 * left), pe
