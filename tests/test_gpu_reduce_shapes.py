"""Every reduce kernel shape at the span boundaries: sizes one unit short of,
at, and past whole spans, ragged rests of several 256-unit pieces, and
calls small enough to take the narrow shape by default.  Each shape is
forced with osc_set_reduce_shape and compared bit for bit with the
oracle's recursive doubling (floats in reference order and, for the
accurate mode, within one rounding of the exact sum)."""
import zlib

import numpy as np
import pytest

import osc
import oracle as O
from gpu_util import bit_equal, inputs, max_rel_err, run_reduce

pytestmark = pytest.mark.gpu

# team sizes: 5 (no slice left over after the groups of 2, 4 or 8), 19
# (leftover slices for the groups of 4 and 8, and an odd tree fold) and 64
# (the reduce matrix's team: one slice left over after the groups of 2)
TEAMS = [5, 19, 64]


@pytest.fixture(scope="module", params=TEAMS)
def heap(request):
    h = osc.Heap(npes=request.param, heap_bytes=8 << 20)
    yield h
    h.close()


# (dtype, op): 8-, 4- and 2-byte elements, a float special-value policy
CASES = [("i64", "sum"), ("u32", "xor"), ("f32", "min"), ("i16", "max"),
         ("f64", "sum"), ("f32", "sum")]
SPANS = {0: 2048, 1: 2048, 2: 512, 3: 256}   # units per span by shape


def sizes_for(shape, vec):
    span = SPANS[shape] * vec
    piece = 256 * vec
    return sorted({1, vec - 1 if vec > 1 else 1, piece + 1, span - 1, span,
                   span + 1, 2 * span + 3 * piece + 5, 3 * span - vec})


@pytest.mark.parametrize("shape", [-1, 0, 1, 2, 3])
@pytest.mark.parametrize("dtype,op", CASES)
def test_shapes_at_span_boundaries(heap, shape, dtype, op):
    vec = 16 // np.dtype(O.NP_DTYPE[dtype]).itemsize
    P = heap.L.osc_npes()
    rng = np.random.default_rng(zlib.crc32(repr((shape, dtype, op,
                                                 P)).encode()))
    osc.check(heap.L.osc_set_reduce_shape(shape), "shape")
    try:
        for n in sizes_for(shape if shape >= 0 else 0, vec):
            src = inputs(rng, op, dtype, P, n)
            for mode in (osc.FP_REFERENCE, osc.FP_ACCURATE):
                if mode == osc.FP_ACCURATE and not (dtype.startswith("f")
                                                    and op == "sum"):
                    continue
                got = run_reduce(heap, op, dtype, src, mode=mode)
                if mode == osc.FP_REFERENCE:
                    want = O.rec_dbl(op, dtype, src)
                    assert bit_equal(got, want, dtype), (P, shape, n)
                else:
                    tol = 2.0 ** -23 if dtype == "f32" else 2.0 ** -52
                    assert max_rel_err(got, src, op, dtype) <= tol, (P, shape,
                                                                     n)
    finally:
        heap.L.osc_set_reduce_shape(-1)
