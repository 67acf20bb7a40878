"""GPU parity of the small-call reduction kernel (lanes over element and
member group, cross-lane butterfly), which the default launch takes for
integer and bitwise operators on teams of 16 or more members and at most
2048 elements.

Every result must be bit-identical to the oracle's step-by-step recursive
doubling (oracle/orc.c, after the reference's
src/shcoll/src/reduce.c:182-293) and to the slice-walk kernel on the same
inputs (forced with osc_set_reduce_shape).  The team sizes cover each
butterfly width (16, 32 and 64 lanes per group) and members left over when
P is not a multiple of the group; the element counts straddle the group
layout and the 2048-element switch to the slice walk.
"""
import numpy as np
import pytest

import osc
import oracle as O
from gpu_util import bit_equal, inputs, run_reduce

pytestmark = pytest.mark.gpu

INT_TYPES = ["u8", "u16", "u32", "u64", "i8", "i16", "i32", "i64"]
OPS = ["and", "or", "xor", "min", "max", "sum", "prod"]
TEAMS = [16, 17, 33, 64, 100, 256]
COUNTS = [1, 3, 64, 65, 2047, 2048, 2049]

_HEAP = []


@pytest.fixture
def heap():
    """256 PEs x 1 MiB, re-created if another test re-initialised the
    library with a different world."""
    if not _HEAP or not _HEAP[0].alive:
        _HEAP[:] = [osc.Heap(npes=256, heap_bytes=1 << 20)]
    return _HEAP[0]


def team_of(h, P):
    return h.world if P == h.npes else h.split(0, 1, P)


@pytest.mark.parametrize("dtype", INT_TYPES)
def test_lanes_matrix(heap, dtype):
    rng = np.random.default_rng(1000 + INT_TYPES.index(dtype))
    for P in TEAMS:
        team = team_of(heap, P)
        for op in OPS:
            for n in COUNTS:
                src = inputs(rng, op, dtype, P, n)
                got = run_reduce(heap, op, dtype, src, team=team)
                want = O.rec_dbl(op, dtype, src)
                assert bit_equal(got, want, dtype), (P, op, n)
        if team != heap.world:
            heap.L.shmem_team_destroy(team)


@pytest.mark.parametrize("dtype,op", [("u64", "sum"), ("i32", "min"),
                                      ("u8", "xor"), ("i16", "prod")])
def test_lanes_equal_slice_walk(heap, dtype, op):
    """The default (lanes) launch and the slice walk write the same bytes."""
    rng = np.random.default_rng(41)
    L = heap.L
    for P in (16, 64, 256):
        team = team_of(heap, P)
        for n in (1, 100, 2048):
            src = inputs(rng, op, dtype, P, n)
            lanes = run_reduce(heap, op, dtype, src, team=team)
            osc.check(L.osc_set_reduce_shape(3), "shape")
            try:
                walk = run_reduce(heap, op, dtype, src, team=team)
            finally:
                L.osc_set_reduce_shape(-1)
            assert np.array_equal(lanes, walk), (P, n)
        if team != heap.world:
            L.shmem_team_destroy(team)


@pytest.mark.parametrize("dtype,op", [("u64", "sum"), ("i8", "max"),
                                      ("u32", "or"), ("i64", "xor")])
def test_lanes_in_place(heap, dtype, op):
    """dest == source: each group reads all members of its element before
    any lane writes, so the in-place result equals the out-of-place one."""
    rng = np.random.default_rng(43)
    for P in (16, 64, 256):
        team = team_of(heap, P)
        for n in (1, 129, 2048):
            src = inputs(rng, op, dtype, P, n)
            got = run_reduce(heap, op, dtype, src, team=team, inplace=True)
            assert bit_equal(got, O.rec_dbl(op, dtype, src), dtype), (P, n)
        if team != heap.world:
            heap.L.shmem_team_destroy(team)


def test_lanes_strided_team(heap):
    """Team of 40 members at global PEs 7, 13, ..., 241 (start 7, stride
    6): only the members' slices are read and written."""
    rng = np.random.default_rng(47)
    P, n = 40, 777
    pes = [7 + 6 * k for k in range(P)]
    team = heap.split(7, 6, P)
    src = inputs(rng, "sum", "i64", P, n)
    s, d = heap.malloc(n * 8), heap.malloc(n * 8)
    sentinel = np.full(n, -12345, np.int64)
    for pe in range(heap.npes):
        heap.put(d, sentinel, pe)
    for k, pe in enumerate(pes):
        heap.put(s, src[k], pe)
    assert osc.typed("shmem_int64_sum_reduce")(team, d, s, n) == 0
    got = np.stack([heap.get(d, n, np.int64, pe) for pe in pes])
    assert np.array_equal(got, O.rec_dbl("sum", "i64", src))
    for pe in set(range(heap.npes)) - set(pes):
        assert np.array_equal(heap.get(d, n, np.int64, pe), sentinel), pe
    heap.free(s)
    heap.free(d)
    heap.L.shmem_team_destroy(team)


@pytest.mark.parametrize("dtype,op", [("u16", "and"), ("i32", "sum")])
def test_lanes_element_offset(heap, dtype, op):
    """Arrays that start one element past a 16-byte boundary."""
    rng = np.random.default_rng(53)
    src = inputs(rng, op, dtype, 64, 301)
    team = heap.split(0, 1, 64)
    got = run_reduce(heap, op, dtype, src, team=team, misalign=1)
    heap.L.shmem_team_destroy(team)
    assert bit_equal(got, O.rec_dbl(op, dtype, src), dtype)


def test_lanes_large_team():
    """P = 1000 (not a power of two; at G = 64 a lane has 15 or 16 members:
    its first load, then one chunk of up to 16) and P = 1100 (lanes 0..11
    have 18 members, so a second, partial chunk); n = 4096 takes the slice
    walk for comparison."""
    rng = np.random.default_rng(59)
    for P in (1000, 1100):
        with osc.Heap(npes=P, heap_bytes=1 << 20) as h:
            for dtype, op in (("i64", "sum"), ("u8", "xor"), ("i32", "min"),
                              ("u16", "prod")):
                for n in (1, 777, 2048, 4096):
                    src = inputs(rng, op, dtype, P, n)
                    got = run_reduce(h, op, dtype, src)
                    want = O.rec_dbl(op, dtype, src)
                    assert bit_equal(got, want, dtype), (P, dtype, op, n)
