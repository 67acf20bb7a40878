"""The measurement entry points: the XOR read and the 64-bit fill that the
benchmark times as read-only and write-only ceilings, the names of the
kernels a reduce call launched, and the replicate group a split call
chose."""
import ctypes as C

import numpy as np
import pytest

import osc

pytestmark = pytest.mark.gpu

WORLD = 16
_H = []


@pytest.fixture
def heap():
    if not _H or not _H[0].alive:
        _H[:] = [osc.Heap(npes=WORLD, heap_bytes=24 << 20)]
    return _H[0]


def last_kernels(h):
    buf = C.create_string_buffer(4096)
    n = h.L.osc_last_reduce_kernels(buf, len(buf))
    names = [k for k in buf.value.decode().split(";") if k]
    assert len(names) == n
    return names


@pytest.mark.parametrize("nbytes", [16, 4096, 32768 + 16 * 7, 3 << 20])
def test_xor_read_is_the_xor_of_every_unit(heap, nbytes):
    rng = np.random.default_rng(nbytes)
    data = rng.integers(0, 2 ** 63, nbytes // 8, dtype=np.uint64)
    p, sink = heap.malloc(nbytes), heap.malloc(16)
    try:
        heap.put(p, data, 3)
        heap.put(sink, np.array([7, 9], np.uint64), 0)  # cleared by the call
        sl = heap.L.osc_slice_bytes()
        osc.check(heap.L.osc_device_xor_read(p + 3 * sl, nbytes, sink, None),
                  "xor read")
        heap.sync()
        got = heap.get(sink, 2, np.uint64, 0)
        want = np.bitwise_xor.reduce(data.reshape(-1, 2), axis=0)
        assert np.array_equal(got, want)
    finally:
        heap.free(p)
        heap.free(sink)


@pytest.mark.parametrize("nbytes", [16, 1024 * 7 + 16, 1 << 20,
                                    (1 << 20) * 5 + 48])
def test_fill64_fills_exactly_the_range(heap, nbytes):
    guard = 4096
    p = heap.malloc(nbytes + 2 * guard)
    try:
        before = np.random.default_rng(1).integers(
            0, 256, nbytes + 2 * guard, dtype=np.uint8)
        heap.put(p, before, 0)
        osc.check(heap.L.osc_device_fill64(p + guard, nbytes,
                                           0x0123456789ABCDEF, None), "fill")
        heap.sync()
        got = heap.get(p, nbytes + 2 * guard, np.uint8, 0)
        assert np.array_equal(got[:guard], before[:guard])
        assert np.array_equal(got[guard + nbytes:], before[guard + nbytes:])
        words = got[guard:guard + nbytes].view(np.uint64)
        assert np.all(words == np.uint64(0x0123456789ABCDEF))
    finally:
        heap.free(p)


def test_misaligned_ranges_are_rejected(heap):
    p, sink = heap.malloc(4096), heap.malloc(16)
    try:
        assert heap.L.osc_device_fill64(p + 8, 64, 1, None) != 0
        assert heap.L.osc_device_fill64(p, 24, 1, None) != 0
        assert heap.L.osc_device_xor_read(p + 4, 64, sink, None) != 0
    finally:
        heap.free(p)
        heap.free(sink)


def test_kernel_names_of_a_split_call_and_a_small_call(heap):
    n = (4 << 20) // 8  # 4 MiB per member on 16 members: walk + replicate
    s, d = heap.malloc(n * 8), heap.malloc(n * 8)
    try:
        heap.fill(None, s, n, "i64", "int_range", -100, 100, seed=3)
        st = heap.L.osc_default_stream()
        osc.check(heap.L.osc_reduce(heap.world, osc.OPS["sum"],
                                    osc.DTYPES["i64"], d, s, n, st), "reduce")
        names = last_kernels(heap)
        assert len(names) == 2, names
        # int64 sums wrap as unsigned arithmetic
        assert names[0].startswith("k_reduce<unsigned long, osc::PolSeq<"
                                   "unsigned long, 5, false>"), names
        assert names[0].endswith(", true>")  # stores member 0 only
        # below 8 MiB per member: one member per workgroup, width 32
        assert names[1] == "k_replicate<1, 32, false>", names
        width = C.c_int(0)
        assert heap.L.osc_reduce_split_group(WORLD, n * 8,
                                             C.byref(width)) == 1
        assert width.value == 32
        # float MIN/MAX: the replicate skips the spans the walk stored
        osc.check(heap.L.osc_reduce(heap.world, osc.OPS["min"],
                                    osc.DTYPES["f64"], d, s, n, st), "reduce")
        names = last_kernels(heap)
        assert names[0].startswith("k_reduce<double, osc::PolSeq<double, 3, "
                                   "true>"), names
        assert names[1] == "k_replicate<1, 32, true>", names
        # one kernel below 3 MiB per member on fewer than 32 members
        osc.check(heap.L.osc_reduce(heap.world, osc.OPS["sum"],
                                    osc.DTYPES["i64"], d, s, n // 2, st),
                  "reduce")
        names = last_kernels(heap)
        assert len(names) == 1 and names[0].endswith(", false>"), names
        # one kernel when the split is off
        osc.check(heap.L.osc_set_reduce_split(0), "split off")
        osc.check(heap.L.osc_reduce(heap.world, osc.OPS["sum"],
                                    osc.DTYPES["i64"], d, s, n, st), "reduce")
        names = last_kernels(heap)
        assert len(names) == 1 and names[0].endswith(", false>"), names
        osc.check(heap.L.osc_set_reduce_split(-1), "split auto")
        # a small call on 16 members: the cross-lane kernel
        osc.check(heap.L.osc_reduce(heap.world, osc.OPS["sum"],
                                    osc.DTYPES["i64"], d, s, 16, st), "reduce")
        names = last_kernels(heap)
        assert names == ["k_reduce_lanes<unsigned long, osc::PolSeq<"
                         "unsigned long, 5, false>, 16>"], names
        heap.sync()
    finally:
        heap.free(s)
        heap.free(d)


def test_timed_layout_is_decided_without_waiting(heap):
    """From 8 MiB per member the first eight calls of a size class run the
    four replicate layouts (members per workgroup, interleave width) twice,
    each between events; a later call that finds every timed run complete
    fixes the fastest, and later calls keep it."""
    n = (8 << 20) // 8 + 2 * 512  # 8 MiB per member and a bit: timed
    s, d = heap.malloc(n * 8), heap.malloc(n * 8)
    cands = ["k_replicate<1, 16, false>", "k_replicate<1, 32, false>",
             "k_replicate<1, 64, false>", "k_replicate<2, 32, false>"]
    width = C.c_int(0)
    try:
        heap.fill(None, s, n, "u64", "bits", seed=4)
        st = heap.L.osc_default_stream()

        def call():
            osc.check(heap.L.osc_reduce(heap.world, osc.OPS["xor"],
                                        osc.DTYPES["u64"], d, s, n, st),
                      "reduce")
            return last_kernels(heap)[-1]
        assert heap.L.osc_reduce_split_group(WORLD, n * 8, None) == 0
        src = np.stack([heap.get(s, n, np.uint64, k) for k in range(WORLD)])
        want = np.bitwise_xor.reduce(src, axis=0)
        sl = heap.L.osc_slice_bytes()
        ran = []
        for _ in range(8):
            # every layout writes every member in full: poison first
            for k in range(WORLD):
                osc.check(heap.L.osc_device_fill64(d + k * sl, n * 8,
                                                   0xA5A5A5A5A5A5A5A5, st),
                          "poison")
            ran.append(call())
            heap.sync()
            for k in range(WORLD):
                assert np.array_equal(heap.get(d, n, np.uint64, k), want), \
                    (ran[-1], k)
        assert ran == cands * 2
        assert heap.L.osc_reduce_split_group(WORLD, n * 8, None) == 0
        heap.sync()
        last = call()
        g = heap.L.osc_reduce_split_group(WORLD, n * 8, C.byref(width))
        assert "k_replicate<%d, %d, false>" % (g, width.value) == last
        assert last in cands
        assert call() == last
        heap.sync()
        for k in (0, 1, 7, WORLD - 1):
            assert np.array_equal(heap.get(d, n, np.uint64, k), want)
    finally:
        heap.free(s)
        heap.free(d)
