"""GPU parity of the small-call float kernel (lanes gather the members into
LDS, then every lane folds them in team order), which the default launch
takes for float and complex sums, products, minima and maxima on teams of
16 to 1024 members with at most 16 KiB per member.

The fold uses the slice walk's own arithmetic in the walk's order, so every
result must be bit-identical to the slice walk on the same inputs (forced
with osc_set_reduce_shape), and within one rounding of the exact value
(the accurate-mode oracle: exact sums, double-double products), as the
reduce matrix tests require.
"""
import numpy as np
import pytest

import osc
import oracle as O
from gpu_util import bit_equal, inputs, max_rel_err, run_reduce

pytestmark = pytest.mark.gpu

FP32_TOL = 1.2e-7
FP64_TOL = 2.3e-16
CASES = [(d, o) for d in ("f32", "f64", "c32", "c64") for o in ("sum", "prod")]
ESIZE = {"f32": 4, "f64": 8, "c32": 8, "c64": 16}

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


def counts(dtype):
    """1 .. the 16 KiB cut, and one element past it (the walk)."""
    cut = 16384 // ESIZE[dtype]
    return [1, 3, 64, 511, cut, cut + 1]


def walk(h, op, dtype, src, team):
    osc.check(h.L.osc_set_reduce_shape(3), "shape")
    try:
        return run_reduce(h, op, dtype, src, team=team)
    finally:
        h.L.osc_set_reduce_shape(-1)


@pytest.mark.parametrize("dtype,op", CASES)
def test_gather_equals_slice_walk(heap, dtype, op):
    rng = np.random.default_rng(2000 + CASES.index((dtype, op)))
    tol = FP32_TOL if dtype in ("f32", "c32") else FP64_TOL
    for P in (16, 17, 64, 100, 256):
        team = team_of(heap, P)
        for n in counts(dtype):
            src = inputs(rng, op, dtype, P, n)
            got = run_reduce(heap, op, dtype, src, team=team)
            assert np.array_equal(got.view(np.uint8), walk(heap, op, dtype, src, team).view(np.uint8)), (P, n)
            err = max_rel_err(got, src, op, dtype)
            assert err <= tol, (P, n, err)
        if team != heap.world:
            heap.L.shmem_team_destroy(team)


@pytest.mark.parametrize("dtype,op", [("f64", "sum"), ("c32", "prod")])
def test_gather_in_place(heap, dtype, op):
    """dest == source: every load of an element precedes the workgroup
    barrier and every store follows it."""
    rng = np.random.default_rng(61)
    for P in (16, 64, 256):
        team = team_of(heap, P)
        for n in (1, 129, 1024):
            src = inputs(rng, op, dtype, P, n)
            want = run_reduce(heap, op, dtype, src, team=team)
            got = run_reduce(heap, op, dtype, src, team=team, inplace=True)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (P, n)
        if team != heap.world:
            heap.L.shmem_team_destroy(team)


def test_gather_strided_team(heap):
    """Team of 50 members at global PEs 3, 8, ..., 248 (start 3, stride 5):
    only the members' slices are read and written."""
    rng = np.random.default_rng(67)
    P, n = 50, 333
    pes = [3 + 5 * k for k in range(P)]
    team = heap.split(3, 5, P)
    src = inputs(rng, "sum", "f32", P, n)
    s, d = heap.malloc(n * 4), heap.malloc(n * 4)
    sentinel = np.full(n, 7.0, np.float32)
    for pe in range(heap.npes):
        heap.put(d, sentinel, pe)
    for k, pe in enumerate(pes):
        heap.put(s, src[k], pe)
    assert osc.typed("shmem_float_sum_reduce")(team, d, s, n) == 0
    got = np.stack([heap.get(d, n, np.float32, pe) for pe in pes])
    assert max_rel_err(got, src, "sum", "f32") <= FP32_TOL
    assert np.array_equal(got, np.broadcast_to(got[0], got.shape))
    for pe in set(range(heap.npes)) - set(pes):
        assert np.array_equal(heap.get(d, n, np.float32, pe), sentinel), pe
    heap.free(s)
    heap.free(d)
    heap.L.shmem_team_destroy(team)


def test_gather_team_limits():
    """1024 members (the most the kernel's LDS holds) and 1025 (the walk
    takes the call): both give the walk's bits."""
    rng = np.random.default_rng(71)
    with osc.Heap(npes=1025, heap_bytes=1 << 20) as h:
        for P in (1024, 1025):
            team = h.world if P == 1025 else h.split(0, 1, P)
            for dtype, op in (("f64", "sum"), ("c64", "prod")):
                for n in (1, 300):
                    src = inputs(rng, op, dtype, P, n)
                    got = run_reduce(h, op, dtype, src, team=team)
                    want = walk(h, op, dtype, src, team)
                    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (P, dtype, n)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("op", ["min", "max"])
def test_gather_min_max_special_values(heap, dtype, op):
    """Float MIN/MAX through the gather kernel: elements holding a NaN or
    whose extremum is a signed zero give every member its own recursive-
    doubling result (bit-exact against the oracle, per member); the others
    give every member the same exact extremum.  Both equal the walk."""
    rng = np.random.default_rng(73 + (op == "max") + 2 * (dtype == "f64"))
    specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.5, -2.0])
    for P in (16, 17, 64, 100, 256):
        team = team_of(heap, P)
        for n in (1, 64, 301):
            src = rng.choice(specials, size=(P, n)).astype(O.NP_DTYPE[dtype])
            src[:, ::3] = rng.standard_normal((P, len(range(0, n, 3))))
            got = run_reduce(heap, op, dtype, src, team=team)
            want = O.rec_dbl(op, dtype, src)
            assert bit_equal(got, want, dtype), (P, n)
            assert bit_equal(got, walk(heap, op, dtype, src, team), dtype), (P, n)
        if team != heap.world:
            heap.L.shmem_team_destroy(team)


def walk_mode(h, op, dtype, src, team, mode):
    osc.check(h.L.osc_set_reduce_shape(3), "shape")
    try:
        return run_reduce(h, op, dtype, src, team=team, mode=mode)
    finally:
        h.L.osc_set_reduce_shape(-1)


@pytest.mark.parametrize("dtype,op", CASES)
def test_gather_reference_order(heap, dtype, op):
    """Reference mode (rec_dbl): the gather kernel evaluates member 0's
    recursive-doubling tree over the LDS copy; every member's result must
    be bit-identical to the oracle's step-by-step recursive doubling and to
    the walk's tree."""
    rng = np.random.default_rng(3000 + CASES.index((dtype, op)))
    for P in (16, 17, 64, 100, 256):
        team = team_of(heap, P)
        for n in (1, 3, 257):
            src = inputs(rng, op, dtype, P, n)
            got = run_reduce(heap, op, dtype, src, team=team, mode=1)
            assert bit_equal(got, O.rec_dbl(op, dtype, src), dtype), (P, n)
            assert bit_equal(got, walk_mode(heap, op, dtype, src, team, 1), dtype), (P, n)
        if team != heap.world:
            heap.L.shmem_team_destroy(team)


def test_gather_reference_order_large_team():
    """P = 1000: p2s = 512 leaves, 488 of them folded pairs."""
    rng = np.random.default_rng(79)
    with osc.Heap(npes=1000, heap_bytes=1 << 20) as h:
        for dtype, op in (("f32", "sum"), ("f64", "prod"), ("c64", "sum")):
            src = inputs(rng, op, dtype, 1000, 33)
            got = run_reduce(h, op, dtype, src, mode=1)
            assert bit_equal(got, O.rec_dbl(op, dtype, src), dtype), (dtype, op)
