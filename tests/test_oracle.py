"""Pins for the CPU oracle (oracle/orc.c): every restated function is checked
against something it was not derived from -- numpy's own reductions, the
closed form of recursive doubling written per member, math.fsum, exact
rational arithmetic, numpy.roll and hand-derived known answers.  Runs on
the CPU only.
"""
import math
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TEAM_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 12, 13, 16, 24, 64]


@pytest.fixture(scope="session", autouse=True)
def built_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")],
                   check=True)
    O.lib()


def _rand_int(rng, dtype, shape):
    info = np.iinfo(O.NP_DTYPE[dtype])
    return rng.integers(info.min, info.max, size=shape, endpoint=True,
                        dtype=O.NP_DTYPE[dtype])


INT_TYPES = ["u8", "u16", "u32", "u64", "i8", "i16", "i32", "i64"]


@pytest.mark.parametrize("dtype", INT_TYPES)
@pytest.mark.parametrize("op", ["and", "or", "xor", "min", "max", "sum",
                                "prod"])
def test_integer_rec_dbl_equals_numpy_reduction(dtype, op):
    rng = np.random.default_rng(hash((dtype, op)) & 0xFFFF)
    npf = {"and": np.bitwise_and, "or": np.bitwise_or, "xor": np.bitwise_xor,
           "min": np.minimum, "max": np.maximum, "sum": np.add,
           "prod": np.multiply}[op]
    for P in TEAM_SIZES:
        src = _rand_int(rng, dtype, (P, 37))
        got = O.rec_dbl(op, dtype, src)
        with np.errstate(over="ignore"):
            want = npf.reduce(src, axis=0, dtype=O.NP_DTYPE[dtype])
        assert np.array_equal(got, np.broadcast_to(want, got.shape)), (P, op)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("op", ["sum", "prod", "min", "max"])
def test_float_rec_dbl_equals_closed_form(dtype, op):
    rng = np.random.default_rng(7)
    ut = np.uint32 if dtype == "f32" else np.uint64
    for P in TEAM_SIZES:
        src = rng.standard_normal((P, 257)).astype(O.NP_DTYPE[dtype])
        if op == "prod":
            src = (1 + (src - 0.5) * 0.002).astype(O.NP_DTYPE[dtype])
        got = O.rec_dbl(op, dtype, src)
        want = O.reduce_closed_form(op, src)
        assert np.array_equal(got.view(ut), want.view(ut)), (P, op)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("op", ["sum", "prod", "min", "max"])
def test_float_rec_dbl_special_values(dtype, op):
    """NaN, infinities and signed zeros: per-member results of the butterfly
    differ for min/max; the closed form must follow them exactly (NaN
    payloads excepted)."""
    rng = np.random.default_rng(11)
    ut = np.uint32 if dtype == "f32" else np.uint64
    specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0])
    for P in TEAM_SIZES:
        src = rng.choice(specials, size=(P, 400)).astype(O.NP_DTYPE[dtype])
        got = O.rec_dbl(op, dtype, src)
        want = O.reduce_closed_form(op, src)
        nan = np.isnan(want)
        assert np.array_equal(np.isnan(got), nan), (P, op)
        assert np.array_equal(got.view(ut)[~nan], want.view(ut)[~nan]), (P, op)


def test_max_signed_zero_known_answer():
    """Hand-derived: P=2, MAX_OP(A,B) = A > B ? A : B (reduce.c:731).  PE0
    computes MAX(+0, -0) -> -0, PE1 computes MAX(-0, +0) -> +0."""
    src = np.array([[0.0], [-0.0]], dtype=np.float64)
    got = O.rec_dbl("max", "f64", src)
    assert math.copysign(1, got[0, 0]) == -1.0
    assert math.copysign(1, got[1, 0]) == 1.0


def test_rec_dbl_fold_order_known_answer():
    """Hand-derived for P=3 (p2s=2): PE0 folds PE1 as op(x1, x0); the
    butterfly then computes (x1 + x0) + x2 on both members; PE1 receives
    PE0's value.  Chosen so that another association rounds differently."""
    x = np.array([[1.0], [2.0 ** 53], [-(2.0 ** 53)]], dtype=np.float64)
    got = O.rec_dbl("sum", "f64", x)
    want = (x[1, 0] + x[0, 0]) + x[2, 0]
    assert want != x[0, 0] + (x[1, 0] + x[2, 0])
    assert np.all(got[:, 0] == want)


def test_fsum_equals_math_fsum():
    rng = np.random.default_rng(3)
    for _ in range(200):
        m = int(rng.integers(1, 80))
        x = rng.standard_normal(m) * 10.0 ** rng.integers(-30, 30, size=m)
        assert O.fsum(x) == math.fsum(x)


def test_accurate_sum_is_exact():
    rng = np.random.default_rng(5)
    for dtype in ("f32", "f64"):
        src = (rng.standard_normal((13, 50)) *
               10.0 ** rng.integers(-8, 8, size=(13, 50))).astype(
                   O.NP_DTYPE[dtype])
        hi, lo = O.accurate("sum", dtype, src)
        for i in range(src.shape[1]):
            exact = sum(Fraction(float(v)) for v in src[:, i])
            assert hi[i] == float(exact)  # correctly rounded
            assert Fraction(hi[i]) + Fraction(lo[i]) - exact == 0 or \
                abs(Fraction(hi[i]) + Fraction(lo[i]) - exact) <= \
                abs(exact) * Fraction(1, 2 ** 100)


def test_accurate_prod_vs_exact_rational():
    rng = np.random.default_rng(9)
    for dtype in ("f32", "f64"):
        src = rng.uniform(0.999, 1.001, size=(64, 20)).astype(
            O.NP_DTYPE[dtype])
        hi, lo = O.accurate("prod", dtype, src)
        for i in range(src.shape[1]):
            exact = Fraction(1)
            for v in src[:, i]:
                exact *= Fraction(float(v))
            err = abs(Fraction(hi[i]) + Fraction(lo[i]) - exact) / exact
            assert err < Fraction(1, 2 ** 95)


def test_accurate_complex_vs_exact_rational():
    rng = np.random.default_rng(12)
    src = (rng.uniform(0.9, 1.1, (6, 8)) +
           1j * rng.uniform(-0.1, 0.1, (6, 8))).astype(np.complex128)
    hi, lo = O.accurate("prod", "c64", src)
    for i in range(8):
        re, im = Fraction(1), Fraction(0)
        for v in src[:, i]:
            a, b = Fraction(float(v.real)), Fraction(float(v.imag))
            re, im = re * a - im * b, re * b + im * a
        assert abs(Fraction(hi[2 * i]) + Fraction(lo[2 * i]) - re) < \
            Fraction(1, 2 ** 95)
        assert abs(Fraction(hi[2 * i + 1]) + Fraction(lo[2 * i + 1]) - im) < \
            Fraction(1, 2 ** 95)
    hi, lo = O.accurate("sum", "c64", src)
    assert np.allclose(hi[0::2] + 1j * hi[1::2], src.sum(axis=0), rtol=1e-15)


def test_complex_product_is_c99_finite_formula():
    """The oracle uses C's complex multiply; for finite operands it must be
    (ac - bd) + (ad + bc)i with every product and sum rounded separately --
    the formula the GPU kernel implements."""
    rng = np.random.default_rng(4)
    for dt, ft in (("c32", np.float32), ("c64", np.float64)):
        a = (rng.standard_normal(500) + 1j * rng.standard_normal(500)).astype(
            O.NP_DTYPE[dt])
        b = (rng.standard_normal(500) + 1j * rng.standard_normal(500)).astype(
            O.NP_DTYPE[dt])
        got = O.rec_dbl("prod", dt, np.stack([a, b]))[0]
        ar, ai = a.real.astype(ft), a.imag.astype(ft)
        br, bi = b.real.astype(ft), b.imag.astype(ft)
        re = ar * br - ai * bi
        im = ar * bi + ai * br
        assert np.array_equal(got.real, re) and np.array_equal(got.imag, im)


@pytest.mark.parametrize("P", [1, 2, 3, 5, 8, 13, 16, 33])
@pytest.mark.parametrize("nb", [1, 3, 8, 17, 64])
def test_fcollect_bruck_equals_concatenation(P, nb):
    rng = np.random.default_rng(P * 100 + nb)
    src = rng.integers(0, 256, size=(P, nb), dtype=np.uint8)
    assert np.array_equal(O.fcollect(src), O.fcollect_closed_form(src))


@pytest.mark.parametrize("size,dist", [(10, 3), (12, 8), (4096, 1000),
                                       (1, 0), (7, 6), (64, 16)])
def test_rotate_inplace_is_right_rotation(size, dist):
    buf = np.arange(size, dtype=np.uint8)
    assert np.array_equal(O.rotate_inplace(buf, dist), np.roll(buf, dist))


@pytest.mark.parametrize("P", [1, 2, 3, 5, 8, 13, 64])
@pytest.mark.parametrize("nb", [1, 7, 16, 40])
def test_alltoall_shift_equals_block_transpose(P, nb):
    rng = np.random.default_rng(P + nb)
    src = rng.integers(0, 256, size=(P, P * nb), dtype=np.uint8)
    assert np.array_equal(O.alltoall(src), O.alltoall_closed_form(src))


def test_alltoall_self_describing_pattern():
    """Config pattern int64 src_i[j*blk + idx] = i<<40 | j<<20 | idx: after
    alltoall, dest_j[i*blk + idx] must read back (i, j, idx)."""
    P, blk = 6, 5
    src = np.zeros((P, P * blk), np.int64)
    for i in range(P):
        for j in range(P):
            src[i, j * blk:(j + 1) * blk] = (i << 40) | (j << 20) | np.arange(
                blk)
    dest = O.alltoall(src.view(np.uint8).reshape(P, -1)).view(np.int64)
    for j in range(P):
        for i in range(P):
            v = dest[j, i * blk:(i + 1) * blk]
            assert np.all(v >> 40 == i) and np.all((v >> 20) & 0xFFFFF == j)
            assert np.array_equal(v & 0xFFFFF, np.arange(blk))


@pytest.mark.parametrize("root", [0, 2, 4])
def test_broadcast_linear(root):
    rng = np.random.default_rng(root)
    src = rng.integers(0, 256, size=(5, 33), dtype=np.uint8)
    dest = O.broadcast(src, root)
    assert np.all(dest == src[root])


def test_hash_numpy_equals_c():
    L = O.lib()
    for seed in (0, 1, 12345, 2 ** 63 + 5):
        for pe in (0, 1, 7, 255):
            idx = np.arange(0, 5000, 37)
            want = np.array([L.orc_hash3(seed, pe, int(i)) for i in idx],
                            dtype=np.uint64)
            assert np.array_equal(O.hash3(seed, pe, idx), want)


def test_umulhi_exact():
    rng = np.random.default_rng(1)
    h = rng.integers(0, 2 ** 64, size=1000, dtype=np.uint64)
    for r in (3, 2001, 2 ** 41 + 1, 2 ** 63 + 12345):
        got = O._umulhi(h, r)
        want = np.array([(int(x) * r) >> 64 for x in h], dtype=np.uint64)
        assert np.array_equal(got, want)
