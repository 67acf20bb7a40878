"""The reference's other reduction algorithms on the GPU, bit for bit
against the oracle's step-by-step restatements: linear (reduce.c:55-91),
binomial (:102-171, children in ascending order) and rabenseifner /
rabenseifner2 (:306-724).  Float and complex SUM/PROD follow each
algorithm's combine order in the type's own arithmetic; float MIN/MAX
inputs carry NaN and signed zeros, where the operand order shows.  Calls
go through shmem_<T>_<op>_reduce with the algorithm chosen by
osc_set_fp_mode, by SHMEM_<OP>_REDUCE_ALGO, and inside a batch."""
import zlib

import numpy as np
import pytest

import osc
import oracle as O
from gpu_util import TYPENAME, bit_equal, inputs, run_reduce

pytestmark = pytest.mark.gpu

WORLD = 128
MODES = {"linear": (osc.FP_LINEAR, O.linear),
         "binomial": (osc.FP_BINOMIAL, O.binomial),
         "rabenseifner": (osc.FP_RABENSEIFNER, O.rabenseifner)}
CELLS = [("f32", "sum"), ("f64", "prod"), ("c32", "sum"), ("c64", "prod"),
         ("f32", "min"), ("f64", "max"), ("i32", "sum"), ("u64", "xor")]
_H = []


@pytest.fixture
def heap():
    if not _H or not _H[0].alive:
        _H[:] = [osc.Heap(npes=WORLD, heap_bytes=4 << 20)]
    return _H[0]


def special_inputs(rng, op, dtype, P, n):
    x = inputs(rng, op, dtype, P, n)
    if dtype in ("f32", "f64") and op in ("min", "max"):
        x[rng.random((P, n)) < 0.25] = 0.0
        x[rng.random((P, n)) < 0.25] = -0.0
        x[rng.random((P, n)) < 0.03] = np.nan
    return x


@pytest.mark.parametrize("algo", sorted(MODES))
@pytest.mark.parametrize("dtype,op", CELLS)
def test_algorithm_matches_oracle(heap, algo, dtype, op):
    mode, ref = MODES[algo]
    rng = np.random.default_rng(zlib.crc32(f"{algo}{dtype}{op}".encode()))
    for start, stride, P in [(0, 1, 1), (4, 1, 3), (2, 1, 5), (0, 1, 8),
                             (1, 3, 13), (7, 1, 17), (0, 1, 64),
                             (20, 1, 100)]:
        team = heap.world if P == WORLD else heap.split(start, stride, P)
        pes = [start + k * stride for k in range(P)]
        # 1 and 37: small calls (the gather kernel from 16 members on);
        # 5000 and 70001: the slice walk, ragged
        for n in (1, 37, 5000, 70001):
            src = special_inputs(rng, op, dtype, P, n)
            got = run_reduce(heap, op, dtype, src, team=team, pes=pes,
                             mode=mode)
            want = ref(op, dtype, src)
            assert bit_equal(got, want, dtype), (algo, P, n)


@pytest.mark.parametrize("algo", ["binomial", "rabenseifner"])
@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_algorithm_through_every_shape(heap, algo, shape):
    mode, ref = MODES[algo]
    rng = np.random.default_rng(shape * 10 + len(algo))
    osc.check(heap.L.osc_set_reduce_shape(shape), "shape")
    try:
        for dtype, op in [("f32", "sum"), ("f64", "max"), ("c64", "sum")]:
            for P in (19, 64):
                team = heap.split(0, 1, P)
                for n in (257, 4097 * 4 + 3):
                    src = special_inputs(rng, op, dtype, P, n)
                    got = run_reduce(heap, op, dtype, src, team=team,
                                     pes=range(P), mode=mode)
                    assert bit_equal(got, ref(op, dtype, src), dtype), \
                        (dtype, op, P, n)
    finally:
        heap.L.osc_set_reduce_shape(-1)


def test_deep_trees_at_300_members():
    """300 members: nine tree levels (the 13-level kernel instances)."""
    rng = np.random.default_rng(300)
    with osc.Heap(npes=300, heap_bytes=1 << 20) as h:
        for algo in ("binomial", "rabenseifner", "linear"):
            mode, ref = MODES[algo]
            for dtype, op, n in [("f32", "sum", 129), ("f32", "sum", 5000),
                                 ("f64", "min", 700), ("c32", "prod", 33)]:
                src = special_inputs(rng, op, dtype, 300, n)
                got = run_reduce(h, op, dtype, src, mode=mode)
                assert bit_equal(got, ref(op, dtype, src), dtype), \
                    (algo, dtype, op, n)


def test_algorithm_names_from_the_environment(monkeypatch):
    """SHMEM_SUM_REDUCE_ALGO=linear and SHMEM_MAX_REDUCE_ALGO=
    rabenseifner2:double resolve at init (the reference's variable names and
    "algo[:type]" format); the other types keep the default."""
    monkeypatch.setenv("SHMEM_SUM_REDUCE_ALGO", "linear")
    monkeypatch.setenv("SHMEM_MAX_REDUCE_ALGO", "rabenseifner2:double")
    monkeypatch.setenv("SHMEM_PROD_REDUCE_ALGO", "binomial:float")
    rng = np.random.default_rng(11)
    with osc.Heap(npes=13, heap_bytes=1 << 20) as h:
        P, n = 13, 301
        s, d = h.malloc(n * 8), h.malloc(n * 8)
        try:
            for dtype, op, ref in [("f32", "sum", O.linear),
                                   ("f64", "max", O.rabenseifner),
                                   ("f32", "prod", O.binomial),
                                   ("f32", "max", O.rec_dbl)]:
                src = special_inputs(rng, op, dtype, P, n)
                for k in range(P):
                    h.put(s, src[k], k)
                fn = osc.typed(f"shmem_{TYPENAME[dtype]}_{op}_reduce")
                assert fn(h.world, d, s, n) == 0
                got = np.stack([h.get(d, n, src.dtype, k) for k in range(P)])
                assert bit_equal(got, ref(op, dtype, src), dtype), (dtype, op)
        finally:
            h.free(s)
            h.free(d)


@pytest.mark.parametrize("algo", sorted(MODES))
def test_batched_calls_follow_the_algorithm(algo):
    """A batch of small and large calls under the algorithm: digests equal
    the one-by-one path's, and every call equals the oracle."""
    mode, ref = MODES[algo]
    rng = np.random.default_rng(len(algo))
    with osc.Heap(npes=32, heap_bytes=8 << 20) as h:
        h.L.osc_set_fp_mode(mode)
        base0 = h.L.osc_heap_base()
        src_region = h.malloc(2 << 20)
        dst_region = h.malloc(4 << 20)
        try:
            calls, checks = [], []
            off = 0
            for k, (dtype, op) in enumerate(CELLS * 2):
                P = int(rng.integers(2, 33))
                n = int(rng.choice([5, 300, 20000]))
                es = np.dtype(O.NP_DTYPE[dtype]).itemsize
                src = special_inputs(rng, op, dtype, P, n)
                soff = src_region + off
                for p in range(P):
                    h.put(soff, src[p], p)
                c = osc.Call()
                c.kind, c.op, c.dtype = osc.CALL_REDUCE, osc.OPS[op], \
                    osc.DTYPES[dtype]
                c.team_start, c.team_stride, c.team_npes = 0, 1, P
                c.count, c.src_off = n, soff - base0
                calls.append(c)
                checks.append((dtype, op, src))
                off += (n * es + 255) // 256 * 256
            calls, nw = osc.plan(calls, dst_region - base0, 4 << 20)
            for w in range(nw):
                idx = [i for i, c in enumerate(calls) if c.wave == w]
                b = osc.Batch([calls[i] for i in idx])
                b.launch_serial()
                d1 = b.digests()
                b.launch()
                assert np.array_equal(b.digests(), d1), w
                h.sync()
                for i in idx:
                    dtype, op, src = checks[i]
                    c = calls[i]
                    got = np.stack([h.get(base0 + c.dest_off, c.count,
                                          src.dtype, p)
                                    for p in range(c.team_npes)])
                    assert bit_equal(got, ref(op, dtype, src), dtype), \
                        (w, dtype, op, c.team_npes, c.count)
                b.close()
        finally:
            h.L.osc_set_fp_mode(0)
