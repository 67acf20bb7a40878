"""Large reductions of 2-, 4-, 8- and 16-byte elements run in two kernels:
the slice walk stores member 0's result only (and a last 16-byte unit that
holds fewer elements to every member), then a replicate kernel copies the
whole units to the other members, each loaded piece to a group of consecutive
members, in 1 KiB pieces interleaved over 16, 32 or 64 workgroups.  Checked against
the oracle just above and below the sizes where the split starts (1 MiB per
member on 32 members or more, 3 MiB on 8 to 31), with whole and partial last units, on power-of-two and
odd teams, a strided team, in place, in both float modes, with the bytes
after each destination untouched; 1-byte types (one kernel) at the same
sizes; sizes that leave the last interleave group of pieces partly empty;
bit-identical to the one-kernel path for every replicate group size and the
size-picked group, on teams that are not a multiple of the group; and float
min/max with NaN and signed-zero extremes planted in some spans (those
spans keep per-member results, the others are replicated)."""
import zlib

import numpy as np
import pytest

import osc
import oracle as O
from gpu_util import TYPENAME, bit_equal, inputs, max_rel_err

pytestmark = pytest.mark.gpu

WORLD = 64
_H = []


@pytest.fixture
def heap():
    if not _H or not _H[0].alive:
        _H[:] = [osc.Heap(npes=WORLD, heap_bytes=24 << 20)]
    return _H[0]


def reduce_checked(h, team, pes, op, dtype, src, mode, inplace=False):
    P, n = src.shape
    es = src.dtype.itemsize
    tail = 4096
    s = h.malloc(n * es + tail)
    d = s if inplace else h.malloc(n * es + tail)
    guard = np.frombuffer(np.random.default_rng(9).bytes(tail), np.uint8)
    try:
        for k, pe in enumerate(pes):
            h.put(s, src[k], pe)
            h.put(d + n * es, guard, pe)
        osc.check(h.L.osc_set_fp_mode(mode), "mode")
        rc = osc.typed(f"shmem_{TYPENAME[dtype]}_{op}_reduce")(team, d, s, n)
        osc.check(rc, "reduce")
        got = np.stack([h.get(d, n, src.dtype, pe) for pe in pes])
        for pe in pes:
            assert np.array_equal(h.get(d + n * es, tail, np.uint8, pe), guard)
        return got
    finally:
        h.L.osc_set_fp_mode(0)
        h.free(s)
        if not inplace:
            h.free(d)


CASES = [  # (dtype, op, n): at, above and off the 1 MiB/member threshold
    ("i64", "sum", 131072), ("i64", "sum", 131072 + 2 * 777),
    ("u64", "xor", 131074), ("i64", "prod", 140000),
    ("f64", "sum", 200002), ("f64", "prod", 150000),
    ("f64", "sum", 131071),  # below 1 MiB per member: one kernel
    ("c64", "prod", 70001), ("c64", "sum", 65535),  # 16-byte elements
    ("c32", "sum", 140000), ("c32", "prod", 131073),
    ("f32", "sum", 300004), ("f32", "prod", 262144),
    ("i32", "max", 262143 * 4), ("i16", "sum", (1 << 19) + 8 * 3),
    ("u16", "and", 1 << 19),
    # float min/max: split with per-span flags
    ("f32", "max", 262148), ("f64", "min", 131072 + 10),
    # one kernel: 1-byte elements
    ("u8", "xor", (1 << 20) + 48),
    # 8 MiB per member (the size-picked group of 4) and a size that leaves
    # the last interleave group of 1 KiB pieces partly empty
    ("f64", "sum", 1 << 20), ("i64", "xor", (1 << 20) + 131072 + 16 * 5),
    # a last 16-byte unit with fewer elements: the walk stores it to every
    # member, the replicate copies the whole units
    ("i64", "sum", (1 << 18) + 1), ("f32", "sum", (1 << 19) + 3),
    ("i16", "xor", (1 << 20) + 5), ("f64", "prod", 131072 + 777),
    ("u32", "min", (1 << 18) + 2), ("f32", "min", 300001),
    ("f64", "max", (1 << 20) + 131072 + 3),
    # 2-8 MiB per member: 1x8 walks below 8 MiB, 8x2 at 8 MiB
    ("i64", "max", 1 << 18), ("f32", "sum", (3 << 20) // 4 + 4),
    ("f64", "sum", (3 << 20) // 8 + 2), ("f64", "prod", (5 << 20) // 8 + 6),
    ("f64", "max", (1 << 20) + 6),
]


@pytest.mark.parametrize("dtype,op,n", CASES)
@pytest.mark.parametrize("team", [(0, 1, 8), (3, 1, 19), (0, 1, 64),
                                  (1, 3, 13)])
def test_split_reduce_matches_oracle(heap, dtype, op, n, team):
    start, stride, P = team
    pes = [start + k * stride for k in range(P)]
    t = heap.world if P == WORLD else heap.split(start, stride, P)
    rng = np.random.default_rng(zlib.crc32(repr((dtype, op, n, team)).encode()))
    src = inputs(rng, op, dtype, P, n)
    modes = [osc.FP_ACCURATE]
    if dtype[0] in "fc":
        modes.append(osc.FP_REFERENCE)
    for mode in modes:
        got = reduce_checked(heap, t, pes, op, dtype, src, mode,
                             inplace=(mode == osc.FP_REFERENCE))
        if dtype[0] in "iu" or mode == osc.FP_REFERENCE or \
                op in ("min", "max"):
            assert bit_equal(got, O.rec_dbl(op, dtype, src), dtype), mode
        else:
            tol = 2.0 ** -24 * 1.001 if dtype in ("f32", "c32") else \
                2.0 ** -53 * 1.01
            assert max_rel_err(got, src, op, dtype) <= tol
            assert bit_equal(got, np.broadcast_to(got[:1], got.shape), dtype)


def test_split_reduce_under_each_algorithm(heap):
    """linear / binomial / rabenseifner orders survive the split."""
    rng = np.random.default_rng(77)
    for mode, ref in [(osc.FP_LINEAR, O.linear), (osc.FP_BINOMIAL, O.binomial),
                      (osc.FP_RABENSEIFNER, O.rabenseifner)]:
        for dtype in ("f64", "c32"):
            src = inputs(rng, "sum", dtype, 19, 131072 + 6)
            got = reduce_checked(heap, heap.split(2, 1, 19), range(2, 21),
                                 "sum", dtype, src, mode)
            assert bit_equal(got, ref("sum", dtype, src), dtype), (mode, dtype)


@pytest.mark.parametrize("dtype,op,n", [("f64", "sum", 400002),
                                        ("i64", "prod", 393216),
                                        ("c64", "prod", 196608 + 2),
                                        ("f32", "sum", 786432 + 4 * 33),
                                        ("i64", "sum", 393216 + 1),
                                        ("f32", "prod", 786432 + 3)])
@pytest.mark.parametrize("team", [(0, 1, 8), (1, 2, 13), (0, 1, 64),
                                  (5, 1, 37)])
def test_split_matches_one_kernel_for_every_group(heap, dtype, op, n, team):
    """Every replicate group size gives the one-kernel path's bits, on teams
    of 8, 13, 37 and 64 members (partial last groups of 1..8 members), at
    3 MiB per member, where teams of fewer than 32 members split too."""
    start, stride, P = team
    pes = [start + k * stride for k in range(P)]
    t = heap.world if P == WORLD else heap.split(start, stride, P)
    rng = np.random.default_rng(zlib.crc32(repr((dtype, op, n, team)).encode()))
    src = inputs(rng, op, dtype, P, n)
    try:
        osc.check(heap.L.osc_set_reduce_split(0), "split off")
        want = reduce_checked(heap, t, pes, op, dtype, src, osc.FP_ACCURATE)
        for group in (1, 2, 4, 8, 16, -1):
            osc.check(heap.L.osc_set_reduce_split(group), "split group")
            got = reduce_checked(heap, t, pes, op, dtype, src, osc.FP_ACCURATE)
            assert bit_equal(got, want, dtype), group
    finally:
        heap.L.osc_set_reduce_split(-1)
    if dtype[0] == "i":
        assert bit_equal(want, O.rec_dbl(op, dtype, src), dtype)


def test_split_setting_rejects_other_groups(heap):
    for bad in (-2, 3, 5, 32):
        assert heap.L.osc_set_reduce_split(bad) == -1  # OSC_ERR_ARG
    assert heap.L.osc_set_reduce_split(8) == 0
    assert heap.L.osc_set_reduce_split(-1) == 0


def planted(rng, dtype, P, n):
    """Normal inputs with special elements planted in a few spans: a NaN in
    one member, a +0/-0 extremum, +-inf, and a NaN in the ragged tail."""
    npt = O.NP_DTYPE[dtype]
    src = rng.standard_normal((P, n)).astype(npt)
    src[3 % P, 5] = np.nan
    zeros = np.zeros(P, npt)
    zeros[1::2] = -0.0  # +0 on even members, -0 on odd ones
    src[:, n // 3] = zeros
    src[P // 2, n // 2 + 1] = np.inf
    src[0, n // 2 + 2] = -np.inf
    src[P - 1, n - 2] = np.nan
    return src


@pytest.mark.parametrize("dtype,op", [("f32", "max"), ("f64", "min"),
                                      ("f64", "max"), ("f32", "min")])
@pytest.mark.parametrize("team", [(0, 1, 8), (1, 3, 13), (0, 1, 64),
                                  (5, 1, 37)])
def test_split_minmax_special_spans(heap, dtype, op, team):
    """Split float MIN/MAX: spans holding a special element get every
    member's own result (the walk stores them, the replicate skips them);
    bit-exact per member against the oracle's recursive doubling and
    Rabenseifner orders, in place, and equal to the one-kernel path."""
    start, stride, P = team
    pes = [start + k * stride for k in range(P)]
    t = heap.world if P == WORLD else heap.split(start, stride, P)
    es = 4 if dtype == "f32" else 8
    n = (1 << 20) // es + 4 * 1000 + 3
    rng = np.random.default_rng(zlib.crc32(repr((dtype, op, team)).encode()))
    src = planted(rng, dtype, P, n)
    want = O.rec_dbl(op, dtype, src)
    for mode, ref, inplace in [(osc.FP_ACCURATE, want, False),
                               (osc.FP_REFERENCE, want, True),
                               (osc.FP_RABENSEIFNER,
                                O.rabenseifner(op, dtype, src), False)]:
        got = reduce_checked(heap, t, pes, op, dtype, src, mode,
                             inplace=inplace)
        assert bit_equal(got, ref, dtype), mode
    try:
        osc.check(heap.L.osc_set_reduce_split(0), "split off")
        one = reduce_checked(heap, t, pes, op, dtype, src, osc.FP_ACCURATE)
    finally:
        heap.L.osc_set_reduce_split(-1)
    assert bit_equal(one, want, dtype)


@pytest.mark.parametrize("mib", [4, 8])
@pytest.mark.parametrize("group", [-1, 2])
@pytest.mark.parametrize("dtype,op", [("f32", "max"), ("f64", "min")])
@pytest.mark.parametrize("team", [(0, 1, 16), (3, 1, 29)])
def test_split_minmax_special_spans_whole_super_chunks(heap, dtype, op, team,
                                                       group, mib):
    """4 and 8 MiB per member and a ragged rest (1x8 and 8x2 walks): the
    replicate's branch-free body for whole super-chunks (2 MiB at the default width 32, 1 MiB at width 16 with 2
    members per workgroup) reads the skip flags as a wave mask; the special
    elements planted inside those super-chunks and in the tail keep their
    per-member results."""
    start, stride, P = team
    pes = [start + k * stride for k in range(P)]
    t = heap.split(start, stride, P)
    es = 4 if dtype == "f32" else 8
    n = (mib << 20) // es + 1004  # a multiple of 16 bytes: split
    rng = np.random.default_rng(zlib.crc32(repr((dtype, op, team)).encode()))
    src = planted(rng, dtype, P, n)
    want = O.rec_dbl(op, dtype, src)
    osc.check(heap.L.osc_set_reduce_split(group), "split group")
    try:
        got = reduce_checked(heap, t, pes, op, dtype, src, osc.FP_ACCURATE)
    finally:
        heap.L.osc_set_reduce_split(-1)
    assert bit_equal(got, want, dtype)
