"""Blocking calls wait for their device work through a completion slot in
coherent pinned host memory (a one-wave kernel stores a ticket, the host
spins on it for at most OSC_WAIT_SPIN_US microseconds, then falls back to
hipStreamSynchronize).  Every mode must return only after all earlier work
on the library stream is done, and must give the same results:
  * "0": hipStreamSynchronize only;
  * "1": a 1 us spin, so calls longer than that take the fallback;
  * "500": the default, where small calls return through the slot."""
import ctypes as C

import numpy as np
import pytest

import osc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spin", ["0", "1", "500"])
def test_blocking_calls_wait_for_earlier_work(spin, monkeypatch):
    monkeypatch.setenv("OSC_WAIT_SPIN_US", spin)
    P, n = 8, 1 << 20
    rng = np.random.default_rng(int(spin) + 3)
    src = rng.integers(-2 ** 40, 2 ** 40, (P, n), dtype=np.int64)
    L = osc.lib()
    L.oscd_event_query.restype = C.c_int
    L.oscd_event_query.argtypes = [C.c_void_p]
    with osc.Heap(npes=P, heap_bytes=4 * n * 8 + (1 << 20)) as h:
        s, d, a, b = (h.malloc(n * 8) for _ in range(4))
        for pe in range(P):
            h.put(s, src[pe], pe)
        red = osc.typed("shmem_long_sum_reduce")
        ev = C.c_void_p()
        osc.check(L.osc_event_create(C.byref(ev)), "event")
        for rep in range(3):
            # enqueue an asynchronous reduce, mark the stream after it,
            # then make one small blocking call: when it returns, the mark
            # (recorded before its wait) must be complete
            osc.check(L.osc_reduce(h.world, osc.OPS["sum"], osc.DTYPES["i64"],
                                   d, s, n, None), "osc_reduce")
            osc.check(L.osc_event_record(ev, None), "record")
            L.shmem_putmem(b, a, 4096, 0)      # symmetric -> symmetric, blocking
            assert L.oscd_event_query(ev) == 1, (spin, rep)
            # a blocking reduce of its own
            assert red(h.world, a, s, n) == 0
            want = src.sum(axis=0)
            for pe in (0, P - 1):
                assert np.array_equal(h.get(d, n, np.int64, pe), want)
                assert np.array_equal(h.get(a, n, np.int64, pe), want)
        L.osc_event_destroy(ev)


def test_tickets_across_every_slot(monkeypatch):
    """1100 blocking reductions go through all 1024 completion slots and
    reuse some, over the slice walk (8 members), the cross-lane kernel (32
    members, 16 int64) and the gather kernel (32 members, 100 doubles);
    each result is checked.  (Storing the ticket from the reduce kernel's
    last workgroup instead of a kernel of its own was measured slower on
    MI355X, 12.1 vs 10.0 us for 8 x 65536 int32: each workgroup then needs
    an agent-scope release, which writes back its L2, before it counts.)"""
    monkeypatch.setenv("OSC_WAIT_SPIN_US", "500")
    P = 32
    rng = np.random.default_rng(11)
    xi = rng.integers(-1000, 1000, (P, 65536)).astype(np.int32)
    xl = rng.integers(-2 ** 40, 2 ** 40, (P, 16)).astype(np.int64)
    xd = rng.integers(-1000, 1000, (P, 100)).astype(np.float64)  # exact sums
    with osc.Heap(npes=P, heap_bytes=4 << 20) as h:
        si, di = h.malloc(65536 * 4), h.malloc(65536 * 4)
        sl, dl = h.malloc(16 * 8), h.malloc(16 * 8)
        sd, dd = h.malloc(100 * 8), h.malloc(100 * 8)
        for pe in range(P):
            h.put(si, xi[pe], pe)
            h.put(sl, xl[pe], pe)
            h.put(sd, xd[pe], pe)
        team8 = h.split(0, 1, 8)
        ri = osc.typed("shmem_int_sum_reduce")
        rl = osc.typed("shmem_long_sum_reduce")
        rd = osc.typed("shmem_double_sum_reduce")
        wi, wl, wd = xi[:8].sum(axis=0), xl.sum(axis=0), xd.sum(axis=0)
        for k in range(1100):
            which = k % 3
            check = k % 97 < 3 or k >= 1094
            pe = k % 8
            if which == 0:
                assert ri(team8, di, si, 65536) == 0
                if check:
                    assert np.array_equal(h.get(di, 65536, np.int32, pe), wi)
            elif which == 1:
                assert rl(h.world, dl, sl, 16) == 0
                if check:
                    assert np.array_equal(h.get(dl, 16, np.int64, pe), wl)
            else:
                assert rd(h.world, dd, sd, 100) == 0
                if check:
                    assert np.array_equal(h.get(dd, 100, np.float64, pe), wd)
        # a split call (walk + replicate)
        n = (1 << 20) // 8
        s2, d2 = h.malloc(n * 8), h.malloc(n * 8)
        h.fill(None, s2, n, "i64", "int_range", -100, 100, seed=5)
        assert rl(h.world, d2, s2, n) == 0
        src = np.stack([h.get(s2, n, np.int64, pe) for pe in range(P)])
        assert np.array_equal(h.get(d2, n, np.int64, P - 1), src.sum(axis=0))
