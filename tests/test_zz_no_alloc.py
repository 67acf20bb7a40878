"""The allocation-free collect (osc_collect_noalloc) against the existing
entry point (osc_collect) through the C ABI: identical destination bytes
on every member, on two calls in a row (the second smaller, so stale
entries of the static table must not leak into it), bytes past the end
untouched, and a team larger than the static table rejected before any
launch."""
import ctypes as C

import numpy as np
import pytest

import osc
import oracle as O

pytestmark = pytest.mark.gpu


def _sizes(nb):
    return (C.c_size_t * len(nb))(*[int(x) for x in nb])


def _run(h, fn, team, pes, rows, nb, s, d, dlen):
    # the same stale bytes in every member's destination before each run
    h.fill(None, d, (dlen + 7) // 8, "u64", "bits", seed=12)
    for k, pe in enumerate(pes):
        if len(rows[k]):
            h.put(s, rows[k], pe)
    rc = getattr(h.L, fn)(team, d, s, _sizes(nb), None)
    osc.check(rc, fn)
    h.sync()
    return np.stack([h.get(d, dlen, np.uint8, pe) for pe in pes])


def test_noalloc_collect_equals_collect_on_two_calls():
    P = 64
    rng = np.random.default_rng(77)
    with osc.Heap(npes=P, heap_bytes=8 << 20) as h:
        s, d = h.malloc(1 << 20), h.malloc(4 << 20)
        try:
            calls = [(h.world, list(range(P)),
                      rng.integers(0, 40000, P)),
                     # the second call: fewer members, smaller sizes
                     (h.split(3, 2, 20), [3 + 2 * k for k in range(20)],
                      rng.integers(0, 3000, 20))]
            calls[0][2][5] = 0
            for team, pes, nb in calls:
                nb = np.asarray(nb, dtype=np.int64)
                nb[-1] += 7                     # a ragged, odd total
                rows = [rng.integers(0, 256, int(x), dtype=np.uint8)
                        for x in nb]
                dlen = int(nb.sum()) + 4096    # and bytes past the end
                a = _run(h, "osc_collect_noalloc", team, pes, rows, nb, s,
                         d, dlen)
                b = _run(h, "osc_collect", team, pes, rows, nb, s, d, dlen)
                assert np.array_equal(a, b), len(pes)
                want = O.collect(rows)
                assert np.array_equal(a[:, :want.shape[1]], want)
        finally:
            h.free(s)
            h.free(d)


def test_noalloc_collect_rejects_teams_past_its_table():
    with osc.Heap(npes=4097, heap_bytes=64 << 10) as h:
        s, d = h.malloc(256), h.malloc(32 << 10)
        nb = np.ones(4097, np.int64)
        h.L.osc_clear_error()
        rc = h.L.osc_collect_noalloc(h.world, d, s, _sizes(nb), None)
        assert rc == -1 and b"4096" in h.L.osc_last_error()
        small = h.split(0, 1, 4096)
        rc = h.L.osc_collect_noalloc(small, d, s, _sizes(nb[:4096]), None)
        assert rc == 0
        h.sync()
        h.L.osc_clear_error()
