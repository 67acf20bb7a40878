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
