"""Pins of the oracle's restatements of the reference's other reductions
(src/shcoll/src/reduce.c: linear :55-91, binomial :102-171, rabenseifner
:306-511, rabenseifner2 :521-724).

The C oracle simulates each algorithm's messages and local combines for all
PEs; it is checked against (a) numpy's own reductions for integer and
bitwise types, where every order gives the same result, (b) independent
numpy closed forms (a left fold, a truncated pairwise tree with the right
subtree first, each element's block owner's tree) for float and complex
types, with NaN and signed zeros for MIN/MAX, and (c) hand-derived known
answers whose three orders round differently."""
import numpy as np
import pytest

import oracle as O

TEAMS = [1, 2, 3, 5, 6, 7, 8, 12, 13, 17]
NUMPY_RED = {"and": np.bitwise_and, "or": np.bitwise_or,
             "xor": np.bitwise_xor, "min": np.minimum, "max": np.maximum,
             "sum": np.add, "prod": np.multiply}
ALGOS = {"linear": lambda op, dt, x: O.linear(op, dt, x),
         "binomial": lambda op, dt, x: O.binomial(op, dt, x),
         "rabenseifner": lambda op, dt, x: O.rabenseifner(op, dt, x),
         "rabenseifner2": lambda op, dt, x: O.rabenseifner(op, dt, x, True)}
CLOSED = {"linear": O.linear_closed_form, "binomial": O.binomial_closed_form,
          "rabenseifner": O.rabenseifner_closed_form,
          "rabenseifner2": O.rabenseifner_closed_form}


def same_bits(a, b):
    if a.dtype.kind in "fc":
        na, nb = np.isnan(a), np.isnan(b)
        return bool(np.array_equal(na, nb) and np.array_equal(
            np.where(na, 0, a).view(np.uint8), np.where(nb, 0, b).view(np.uint8)))
    return bool(np.array_equal(a, b))


@pytest.mark.parametrize("algo", sorted(ALGOS))
@pytest.mark.parametrize("dt", ["u8", "i16", "i32", "u64", "i64"])
def test_integer_results_equal_numpy_reductions(algo, dt):
    rng = np.random.default_rng(len(algo) * 7 + len(dt))
    info = np.iinfo(O.NP_DTYPE[dt])
    for P in TEAMS:
        for n in (1, 9, 40):
            x = rng.integers(info.min, info.max, (P, n), endpoint=True,
                             dtype=O.NP_DTYPE[dt])
            for op, red in NUMPY_RED.items():
                got = ALGOS[algo](op, dt, x)
                want = red.reduce(x, axis=0, dtype=x.dtype)  # wraps
                assert np.array_equal(got, np.tile(want, (P, 1))), (op, P, n)


@pytest.mark.parametrize("algo", sorted(ALGOS))
@pytest.mark.parametrize("dt", ["f32", "f64", "c32", "c64"])
def test_float_results_equal_closed_forms(algo, dt):
    rng = np.random.default_rng(len(algo) * 13 + len(dt))
    ops = ["sum", "prod"] + (["min", "max"] if dt[0] == "f" else [])
    for P in TEAMS + [16, 31]:
        for n in (1, 2, 3, 7, 33):
            for op in ops:
                x = rng.standard_normal((P, n))
                if dt[0] == "c":
                    x = x + 1j * rng.standard_normal((P, n))
                x = x.astype(O.NP_DTYPE[dt])
                if op in ("min", "max"):
                    x[rng.random((P, n)) < 0.3] = 0.0
                    x[rng.random((P, n)) < 0.2] = -0.0
                    x[rng.random((P, n)) < 0.05] = np.nan
                assert same_bits(ALGOS[algo](op, dt, x),
                                 CLOSED[algo](op, x)), (op, P, n)


def test_rabenseifner_sums_and_products_have_recursive_doublings_bits():
    """Every member's butterfly gives the same bits for + and *, so the
    owner-per-block result equals recursive doubling's."""
    rng = np.random.default_rng(3)
    for P in TEAMS:
        for dt in ("f32", "f64", "c32", "c64"):
            x = rng.standard_normal((P, 11))
            if dt[0] == "c":
                x = x + 1j * rng.standard_normal((P, 11))
            x = x.astype(O.NP_DTYPE[dt])
            for op in ("sum", "prod"):
                assert same_bits(O.rabenseifner(op, dt, x),
                                 O.rec_dbl(op, dt, x)), (P, dt, op)


def test_known_answers_three_orders_round_differently():
    """P = 5, float32, x = [1, u, u, u, u] with u = 2^-24 (half an ulp of
    1; ties round to even):
      linear   (((1 + u) + u) + u) + u           = 1
      rec_dbl  ((x1 + x0) + x2) + (x3 + x4)      = 1 + 2u = 1 + 2^-23
               (members 0, 2, 3, 4; PE 1 folded into member 0)
      binomial x4 + ((x3 + x2) + (x1 + x0))      = u + (2u + 1)
               = 1 + 3u, a tie between 1 + 2^-23 and 1 + 2^-22 -> even
               = 1 + 2^-22
      rabenseifner = rec_dbl's bits."""
    u = np.float32(2.0 ** -24)
    x = np.array([[1.0], [u], [u], [u], [u]], np.float32)
    one = np.float32(1.0)
    assert O.linear("sum", "f32", x)[0, 0] == one
    assert O.rec_dbl("sum", "f32", x)[0, 0] == one + np.float32(2.0 ** -23)
    assert O.binomial("sum", "f32", x)[0, 0] == one + np.float32(2.0 ** -22)
    assert O.rabenseifner("sum", "f32", x)[3, 0] == \
        one + np.float32(2.0 ** -23)


def test_known_answer_rabenseifner_max_signed_zero():
    """P = 3 (members: PE 0 with folded PE 1, and PE 2), n = 2, MAX with
    the ternary a > b ? a : b.  x0 = [+0, +0], x1 = [-0, -0], x2 = [+0, +0].
      fold, element 0 (lower half):  op(x1, x0) = MAX(-0, +0) -> +0
      fold, element 1 (upper half):  op(x0, x1) = MAX(+0, -0) -> -0
      block 0 = element 0, owner member 0: op(v0, v1) = MAX(+0, +0) -> +0
      block 1 = element 1, owner member 1: op(v1, v0) = MAX(+0, -0) -> -0
    so every PE holds [+0, -0]; recursive doubling folds both elements as
    op(x1, x0) and gives [+0, +0] everywhere."""
    x = np.array([[0.0, 0.0], [-0.0, -0.0], [0.0, 0.0]], np.float64)
    for ring in (False, True):
        got = O.rabenseifner("max", "f64", x, ring)
        assert np.array_equal(np.signbit(got),
                              np.tile([False, True], (3, 1))), ring
    assert not np.signbit(O.rec_dbl("max", "f64", x)).any()
