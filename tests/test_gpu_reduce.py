"""GPU parity of the team reductions, through the OpenSHMEM C ABI.

Integer, bitwise and MIN/MAX results must be bit-identical to the oracle's
step-by-step recursive doubling (oracle/orc.c, after the reference's
src/shcoll/src/reduce.c:182-293); float SUM/PROD in reference-order mode
("rec_dbl") too; in the default accurate mode they must be within
FP32_TOL / FP64_TOL relative error of the exact / double-double oracle.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import osc
import oracle as O
from gpu_util import bit_equal, inputs, max_rel_err, run_reduce

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
# accurate mode: one rounding of a (near-)exact value -> <= 2^-24 / 2^-53
# relative, plus the double rounding FP64 -> FP32 can add (still < 2^-23).
FP32_TOL = 1.2e-7
FP64_TOL = 2.3e-16
TEAM_SIZES = [1, 2, 3, 5, 8, 13, 64]
N = 1037  # ragged: not a multiple of any vector width

OPS_FOR = {
    "u8": ["and", "or", "xor", "min", "max", "sum", "prod"],
    "u16": ["and", "or", "xor", "min", "max", "sum", "prod"],
    "u32": ["and", "or", "xor", "min", "max", "sum", "prod"],
    "u64": ["and", "or", "xor", "min", "max", "sum", "prod"],
    "i8": ["and", "or", "xor", "min", "max", "sum", "prod"],
    "i16": ["and", "or", "xor", "min", "max", "sum", "prod"],
    "i32": ["and", "or", "xor", "min", "max", "sum", "prod"],
    "i64": ["and", "or", "xor", "min", "max", "sum", "prod"],
    "f32": ["min", "max", "sum", "prod"],
    "f64": ["min", "max", "sum", "prod"],
    "c32": ["sum", "prod"],
    "c64": ["sum", "prod"],
}
CASES = [(d, o) for d, ops in OPS_FOR.items() for o in ops]


_HEAP = []


@pytest.fixture
def heap():
    """64 PEs x 80 MiB, re-created if another test re-initialised the
    library with a different world."""
    if not _HEAP or not _HEAP[0].alive:
        _HEAP[:] = [osc.Heap(npes=64, heap_bytes=80 << 20)]
    return _HEAP[0]


def team_of(h, P):
    return h.world if P == h.npes else h.split(0, 1, P)


def test_int_sum_world_of_8():
    """shmem_int_sum_reduce on WORLD, 8 PEs x 65536 int32 uniform
    in [-1000, 1000] from seed 0 (generated on the device), bit-exact."""
    with open(os.path.join(HERE, "golden", "golden.json")) as f:
        gold = json.load(f)["int_sum_8pe"]
    P, n = 8, 65536
    with osc.Heap(npes=P, heap_bytes=4 << 20) as h:
        s, d = h.malloc(n * 4), h.malloc(n * 4)
        h.fill(None, s, n, "i32", "int_range", -1000, 1000, seed=0)
        h.sync()
        src = np.stack([h.get(s, n, np.int32, pe) for pe in range(P)])
        want_src = O.gen_int_range(0, range(P), n, -1000, 1000, "i32")
        assert np.array_equal(src, want_src)       # generator parity
        assert hashlib.sha256(src.tobytes()).hexdigest() == gold["src_sha256"]
        fn = osc.typed("shmem_int_sum_reduce")
        assert fn(h.world, d, s, n) == 0
        dest = np.stack([h.get(d, n, np.int32, pe) for pe in range(P)])
        assert np.array_equal(dest, O.rec_dbl("sum", "i32", src))
        assert hashlib.sha256(dest.tobytes()).hexdigest() == \
            gold["dest_sha256"]


@pytest.mark.parametrize("dtype,op", CASES)
def test_reduce_matrix(heap, dtype, op):
    rng = np.random.default_rng(abs(hash((dtype, op))) % 2 ** 32)
    for P in TEAM_SIZES:
        src = inputs(rng, op, dtype, P, N)
        team = team_of(heap, P)
        for mode in (0, 1):
            got = run_reduce(heap, op, dtype, src, team=team, mode=mode)
            exact_kind = dtype[0] in "ui" or op in ("min", "max") or mode == 1
            if exact_kind:
                want = O.rec_dbl(op, dtype, src)
                assert bit_equal(got, want, dtype), (P, mode)
            else:
                tol = FP32_TOL if dtype in ("f32", "c32") else FP64_TOL
                err = max_rel_err(got, src, op, dtype)
                assert err <= tol, (P, err)
            if exact_kind and dtype[0] in "ui":
                break  # integer results do not depend on the mode


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("op", ["min", "max", "sum", "prod"])
def test_special_values(heap, dtype, op):
    """NaN, +-inf and signed zeros.  MIN/MAX must give every member its own
    butterfly result (they differ on ties of +0/-0 and on NaN)."""
    rng = np.random.default_rng(17)
    specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.5, -2.0])
    for P in (2, 3, 5, 8, 13, 64):
        src = rng.choice(specials, size=(P, 301)).astype(O.NP_DTYPE[dtype])
        team = team_of(heap, P)
        got = run_reduce(heap, op, dtype, src, team=team, mode=1)
        want = O.rec_dbl(op, dtype, src)
        assert bit_equal(got, want, dtype), P


@pytest.mark.parametrize("dtype,op,mode", [("i64", "sum", 0),
                                           ("f32", "sum", 1),
                                           ("f64", "max", 0),
                                           ("u64", "xor", 0),
                                           ("c64", "prod", 1)])
def test_in_place(heap, dtype, op, mode):
    """dest == source (allowed by OpenSHMEM) gives the out-of-place result;
    for f64 max the inputs hold signed zeros so the exact per-member path
    (which snapshots the inputs before writing) is exercised."""
    rng = np.random.default_rng(23)
    for P in (3, 8, 13):
        src = inputs(rng, op, dtype, P, 777)
        if dtype == "f64":
            src[:, ::5] = 0.0
            src[1::2, ::5] = -0.0
        got = run_reduce(heap, op, dtype, src, team=team_of(heap, P),
                         mode=mode, inplace=True)
        assert bit_equal(got, O.rec_dbl(op, dtype, src), dtype), P


@pytest.mark.parametrize("dtype,op", [("f32", "sum"), ("i16", "max"),
                                      ("f64", "prod"), ("u32", "and")])
def test_misaligned_offsets(heap, dtype, op):
    """Offsets that are not 16-byte aligned take the element-wise path."""
    rng = np.random.default_rng(29)
    src = inputs(rng, op, dtype, 5, 333)
    got = run_reduce(heap, op, dtype, src, team=team_of(heap, 5), mode=1,
                     misalign=1)
    assert bit_equal(got, O.rec_dbl(op, dtype, src), dtype)


def test_strided_team(heap):
    """A team split with start 3, stride 5: members are global PEs
    3, 8, ..., 58; only their slices are read and written."""
    rng = np.random.default_rng(31)
    team = heap.split(3, 5, 12)
    pes = [3 + 5 * k for k in range(12)]
    n = 500
    src = inputs(rng, "sum", "f32", 12, n)
    s, d = heap.malloc(n * 4), heap.malloc(n * 4)
    sentinel = np.full(n, 7.0, np.float32)
    for pe in range(heap.npes):
        heap.put(d, sentinel, pe)
    for k, pe in enumerate(pes):
        heap.put(s, src[k], pe)
    heap.L.osc_set_fp_mode(1)
    assert osc.typed("shmem_float_sum_reduce")(team, d, s, n) == 0
    heap.L.osc_set_fp_mode(0)
    got = np.stack([heap.get(d, n, np.float32, pe) for pe in pes])
    assert bit_equal(got, O.rec_dbl("sum", "f32", src), "f32")
    for pe in set(range(heap.npes)) - set(pes):
        assert np.array_equal(heap.get(d, n, np.float32, pe), sentinel), pe
    heap.free(s)
    heap.free(d)
    assert heap.L.shmem_team_translate_pe(team, 2, heap.world) == 13
    assert heap.L.shmem_team_translate_pe(heap.world, 13, team) == 2
    assert heap.L.shmem_team_translate_pe(heap.world, 4, team) == -1
    assert heap.L.shmem_team_n_pes(team) == 12
    heap.L.shmem_team_destroy(team)


def test_one_member_team_with_any_stride(heap):
    """A one-member split accepts any stride: from a parent with stride 3,
    start 2 and stride 2^30 give global PE 1 + 2*3 = 7, and a reduce over
    it copies that PE's source."""
    parent = heap.split(1, 3, 10)              # global PEs 1, 4, ..., 28
    child = heap.split(2, 1 << 30, 1, parent=parent)
    L = heap.L
    assert L.shmem_team_n_pes(child) == 1
    assert L.shmem_team_translate_pe(child, 0, heap.world) == 7
    assert L.shmem_team_translate_pe(heap.world, 7, child) == 0
    n = 64
    x = np.arange(n, dtype=np.int64) * 3 - 50
    s, d = heap.malloc(n * 8), heap.malloc(n * 8)
    heap.put(s, x, 7)
    assert osc.typed("shmem_long_sum_reduce")(child, d, s, n) == 0
    assert np.array_equal(heap.get(d, n, np.int64, 7), x)
    heap.free(s)
    heap.free(d)
    L.shmem_team_destroy(child)
    L.shmem_team_destroy(parent)


def test_zero_elements_and_errors(heap):
    L = heap.L
    s = heap.malloc(1024)
    f = osc.typed("shmem_float_sum_reduce")
    assert f(heap.world, s, s, 0) == 0
    assert f(heap.world, s, s + 4, 10) != 0            # partial overlap
    assert f(heap.world, s, s, 1 << 40) != 0           # past the heap
    host = np.zeros(4, np.float32)
    assert f(heap.world, host.ctypes.data, s, 4) != 0  # not symmetric
    assert f(None, s, s, 4) != 0                       # invalid team
    g = osc.typed("shmem_longdouble_max_reduce")
    assert g(heap.world, s, s, 4) != 0                 # no 80-bit type
    heap.free(s)
    assert L.osc_last_error()


def test_large_team_tree_path():
    """P = 300 (p2s = 256): the reference-order tree needs 9 levels."""
    rng = np.random.default_rng(37)
    with osc.Heap(npes=300, heap_bytes=1 << 20) as h:
        for dtype, op in (("f32", "sum"), ("f64", "prod"), ("f32", "max")):
            src = inputs(rng, op, dtype, 300, 129)
            if op == "max":
                src[:, :20] = -0.0
                src[::7, :20] = 0.0
            got = run_reduce(h, op, dtype, src, mode=1)
            assert bit_equal(got, O.rec_dbl(op, dtype, src), dtype)


def test_reduce_matrix_full_size_sampled():
    """The reduce matrix shape: 64 PEs x 4,194,304 elements.  u64 xor and i64 sum
    (inputs regenerated on the host: bit-exact on 2048 sampled columns),
    f32 sum in accurate mode (rel. error vs the FP64 oracle <= 1e-6 on the
    sampled columns), and every member's dest identical."""
    P, n = 64, 4194304
    rng = np.random.default_rng(41)
    cols = np.sort(rng.choice(n, 2048, replace=False))
    with osc.Heap(npes=P, heap_bytes=(2 * n * 8) + (4 << 20)) as h:
        s, d = h.malloc(n * 8), h.malloc(n * 8)
        # u64 xor, uniform random bits
        h.fill(None, s, n, "u64", "bits", seed=3)
        assert osc.typed("shmem_uint64_xor_reduce")(h.world, d, s, n) == 0
        srcc = O.hash3(3, np.arange(P)[:, None], cols[None, :])
        want = np.bitwise_xor.reduce(srcc, axis=0)
        for pe in (0, 31, 63):
            got = h.get(d, n, np.uint64, pe)
            assert np.array_equal(got[cols], want)
        # i64 sum, uniform in [-2^40, 2^40]
        h.fill(None, s, n, "i64", "int_range", -2.0 ** 40, 2.0 ** 40, seed=4)
        assert osc.typed("shmem_int64_sum_reduce")(h.world, d, s, n) == 0
        srcc = O.gen_int_range(4, range(P), n, -2 ** 40, 2 ** 40, "i64",
                               idx=cols)
        want = srcc.sum(axis=0)
        rows = [h.get(d, n, np.int64, pe) for pe in (0, 63)]
        assert np.array_equal(rows[0], rows[1])
        assert np.array_equal(rows[0][cols], want)
        # f32 sum, N(0,1), accurate mode
        h.fill(None, s, n, "f32", "normal", 0.0, 1.0, seed=5)
        assert osc.typed("shmem_float_sum_reduce")(h.world, d, s, n) == 0
        src = np.stack([h.get(s, n, np.float32, pe)[cols] for pe in range(P)])
        got = np.stack([h.get(d, n, np.float32, pe)[cols] for pe in (0, 63)])
        assert max_rel_err(got, src, "sum", "f32") <= 1e-6
