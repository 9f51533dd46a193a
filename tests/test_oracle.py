"""Pins the oracle (the CPU restatement used as the GPU checker) against
golden vectors, the reference's known answers, and mathematical
invariants.  CPU only."""
import math
import os

import numpy as np
import pytest

import oracle
from libfive_amd import Shape, X, Y, Z
from libfive_amd import workloads as W

from helpers import BINARY, UNARY, assert_same_bits, op_shape, ulp_distance

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ieee_ops.npz")


def od(shape):
    return oracle.OracleDeck(shape.deck())


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name", BINARY + UNARY)
def test_ieee_ops_match_golden(golden, name):
    a, b = golden["a"], golden["b"]
    d = od(op_shape(name))
    got = d.values(a, b, np.zeros_like(a))
    assert_same_bits(got, golden[name], name)


def test_sincos_within_one_ulp_of_float64(golden):
    x = golden["trig_x"]
    s, c = oracle.sincos(x)
    assert ulp_distance(s, golden["sin_rn"]).max() <= 1
    assert ulp_distance(c, golden["cos_rn"]).max() <= 1
    # nearly always the correctly rounded result
    assert (ulp_distance(s, golden["sin_rn"]) == 0).mean() > 0.999


def test_sincos_dense_within_one_ulp():
    """Two million float arguments in [-6.5, 6.5] (every quadrant, the
    reduced argument over its whole range): within one ulp of the float64
    value rounded to float, and nearly always equal to it."""
    x = np.linspace(-6.5, 6.5, 2000001).astype(np.float32)
    s, c = oracle.sincos(x)
    rs = np.sin(x.astype(np.float64)).astype(np.float32)
    rc = np.cos(x.astype(np.float64)).astype(np.float32)
    us, uc = ulp_distance(s, rs), ulp_distance(c, rc)
    assert us.max() <= 1 and uc.max() <= 1
    assert (us == 0).mean() > 0.9999 and (uc == 0).mean() > 0.9999


def test_two_over_pi_table():
    """The 2/pi bits used for large-argument reduction, recomputed from
    Machin's formula."""
    def arctan_inv(x, prec):
        one = 1 << prec
        total, term, n, sign, x2 = 0, one // x, 1, 1, x * x
        while term:
            total += sign * (term // n)
            term //= x2
            n += 2
            sign = -sign
        return total
    prec = 400
    pi = 4 * (4 * arctan_inv(5, prec) - arctan_inv(239, prec))
    K = 320
    twoopi = (2 << (prec + K)) // pi
    words = [(twoopi >> (K - 32 * (j + 1))) & 0xFFFFFFFF for j in range(10)]
    src = open(os.path.join(os.path.dirname(__file__), "..", "libfive_amd",
                            "csrc", "fmath.hpp")).read()
    for w in words:
        assert ("0x%08Xu" % w) in src


# ---- ArrayEvaluator known answers (libfive/test/eval_array.cpp:19-121)
def test_array_known_answers():
    p = (np.float32(1), np.float32(2), np.float32(3))
    assert od(X()).values(*[[v] for v in p])[0] == 1.0
    assert od(Y()).values(*[[v] for v in p])[0] == 2.0
    assert od(Shape.constant(3.14)).values([1], [2], [3])[0] == np.float32(3.14)
    assert od(X() * 2 + Y() * 3).values([1], [2], [3])[0] == 8.0
    assert od(X() + 1).values([1], [2], [3])[0] == 2.0
    assert od(X() + Z()).values([1], [2], [3])[0] == 4.0
    v = od(Shape.binary("nth-root", X(), 3)).values([-0.5], [0], [0])[0]
    assert v == pytest.approx(-0.7937, rel=1e-4)


@pytest.mark.parametrize("a,b,expected", [
    (1.1, 1.0, 0.1), (-1.1, 1.0, 0.9), (-1.1, 1.1, 0.0), (2.1, 2.5, 2.1),
    (2.6, 2.5, 0.1), (2.1, -2, -1.9), (2.4, -1, -0.6), (-2.1, -2, -0.1),
    (-2.4, -1, -0.4)])
def test_mod_known_answers(a, b, expected):
    v = od(Shape.binary("mod", X(), Y())).values([a], [b], [0])[0]
    assert v == pytest.approx(expected, abs=1e-5)


def test_every_opcode_runs():
    # eval_array.cpp:100-112, eval_interval.cpp:35-46
    for op in range(7, 32):
        if op in (17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 31):
            t = Shape.binary(op, X(), 5)
        else:
            t = Shape.unary(op, X())
        d = od(t)
        d.values([0.0], [0.0], [0.0])
        d.derivs([0.0], [0.0], [0.0])
        d.intervals([[0, 0, 0]], [[1, 1, 1]])


# ---- DerivArrayEvaluator known answers (eval_deriv_array.cpp:41-91)
def test_deriv_known_answers():
    v, dx, dy, dz = od(X()).derivs([0, 1], [0, 2], [0, 3])
    assert list(zip(dx, dy, dz, v)) == [(1, 0, 0, 0), (1, 0, 0, 1)]
    v, dx, dy, dz = od(X() + Z()).derivs([1, 1], [1, 2], [1, 3])
    assert list(zip(dx, dy, dz, v)) == [(1, 0, 1, 2), (1, 0, 1, 4)]
    v, dx, dy, dz = od(Shape.binary("nth-root", X(), 3)).derivs(
        [0, 1], [0, 2], [0, 3])
    assert math.isinf(dx[0]) and (dy[0], dz[0], v[0]) == (0, 0, 0)
    assert dx[1] == pytest.approx(0.33333, rel=1e-4)
    assert (dy[1], dz[1], v[1]) == (0, 0, 1)


def test_deriv_closed_forms():
    """Partials against analytic derivatives on smooth shapes."""
    rng = np.random.default_rng(1)
    p = rng.uniform(-1, 1, (3, 500)).astype(np.float32)
    s = W.sphere(0.7, (0.1, -0.2, 0.3))
    v, dx, dy, dz = od(s).derivs(*p)
    c = np.array([0.1, -0.2, 0.3])[:, None]
    q = p.astype(np.float64) - c
    r = np.sqrt((q ** 2).sum(0))
    assert np.allclose(v, r - 0.7, atol=1e-6)
    for k, g in enumerate((dx, dy, dz)):
        assert np.allclose(g, q[k] / r, atol=1e-5)
    g = (X() * 3).sin() * Y().cos()
    v, dx, dy, dz = od(g).derivs(*p)
    x, y = p[0].astype(np.float64), p[1].astype(np.float64)
    assert np.allclose(dx, 3 * np.cos(3 * x) * np.cos(y), atol=1e-5)
    assert np.allclose(dy, -np.sin(3 * x) * np.sin(y), atol=1e-5)
    assert np.all(dz == 0)


# ---- IntervalEvaluator known answers (libfive/test/eval_interval.cpp)
def iv(shape, lo, hi):
    l, h, n = od(shape).intervals([lo], [hi])
    return float(l[0]), float(h[0]), bool(n[0])


def test_interval_known_answers():
    assert iv(X() + 1, [1, 1, 1], [2, 2, 2])[:2] == (2.0, 3.0)
    assert iv((X() + Y()) * (X() - Y()), [0, 0, 0], [1, 1, 1])[:2] == (-2.0, 2.0)
    lo, hi, _ = iv(0.1 / Y(), [-0.625, -0.3125, -0.625], [-0.3125, 0, -0.2125])
    assert lo == -math.inf and hi == math.inf


def test_interval_nan_flags():
    # eval_interval.cpp:203-238
    assert not iv(X(), [-1, 1, 0], [1, 2, 0])[2]
    assert not iv(X(), [-math.inf, 1, 0], [1, 2, 0])[2]
    d = X() / Y()
    assert not iv(d, [-1, 1, 0], [1, 2, 0])[2]
    assert iv(d, [-1, 0, 0], [1, 2, 0])[2]
    assert iv(d, [-1, -1, 0], [1, 2, 0])[2]
    assert not iv(d, [1, -1, 0], [2, 2, 0])[2]
    m = Z() * (X() / Y())
    assert not iv(m, [-1, 1, 0], [1, 2, 0])[2]
    assert iv(m, [-1, 0, 0], [1, 2, 0])[2]
    assert iv(m, [-1, -1, 1], [1, 2, 2])[2]
    assert iv(m, [1, -1, -1], [2, 2, 1])[2]
    assert not iv(m, [1, -1, 1], [2, 2, 2])[2]
    assert iv(m, [1, -1, -1], [2, 2, 2])[2]


def test_interval_push_known_answers():
    # eval_interval.cpp:67-109
    d = od((X() + 1).min(Y() + 1))
    n, changed, (lo, hi, _) = d.interval_push([-5, 8, 0], [-4, 9, 0])
    assert (lo, hi) == (-4.0, -3.0)
    assert changed and n < d.n_clauses
    t = (Y().min(X())).min(X().min(Y() + 3))
    d = od(t)
    n, changed, _ = d.interval_push([-5, 0, 0], [-4, 1, 0])
    assert n == 0 and changed


def test_mod_interval_encloses_samples():
    # eval_interval.cpp:153-200: brute-force enclosure of mod
    d = od(Shape.binary("mod", X(), Y()))
    ivs = [(a, b) for a in range(-5, 6) for b in range(-5, 6) if a <= b]
    lo, hi = [], []
    for a in ivs:
        for b in ivs:
            lo.append([a[0], b[0], 0])
            hi.append([a[1], b[1], 0])
    lo = np.array(lo, np.float32)
    hi = np.array(hi, np.float32)
    rl, rh, rn = d.intervals(lo, hi)
    N = 10
    t = (np.arange(N + 1) / np.float32(N)).astype(np.float32)
    for k in range(0, lo.shape[0], 7):
        if rn[k]:
            continue
        xa = lo[k, 0] + t * (hi[k, 0] - lo[k, 0])
        xb = lo[k, 1] + t * (hi[k, 1] - lo[k, 1])
        A, B = np.meshgrid(xa, xb)
        v = d.values(A.ravel(), B.ravel(), np.zeros(A.size))
        assert np.all(v >= rl[k]) and np.all(v <= rh[k]), k


@pytest.mark.parametrize("build", [
    lambda: W.csg_c1(), lambda: W.gyroid_c4(), lambda: W.random_dag(200),
    lambda: (X() * 3).sin() / (Y() + 2) + Z().square().sqrt(),
])
def test_interval_encloses_point_values(build):
    """Every point value inside a box lies within the box's interval
    (unless the interval may be NaN)."""
    s = build()
    d = od(s)
    rng = np.random.default_rng(7)
    c = rng.uniform(-1.2, 1.2, (64, 3)).astype(np.float32)
    w = rng.uniform(0.01, 0.5, (64, 3)).astype(np.float32)
    lo, hi = c - w, c + w
    rl, rh, rn = d.intervals(lo, hi)
    for k in range(lo.shape[0]):
        u = rng.uniform(0, 1, (3, 200)).astype(np.float32)
        pts = lo[k][:, None] + u * (hi[k] - lo[k])[:, None]
        v = d.values(*pts)
        if rn[k]:
            continue
        ok = (v >= rl[k]) & (v <= rh[k]) | np.isnan(v)
        assert ok.all(), (k, rl[k], rh[k], v[~ok][:4])


def test_pushed_tape_is_exact_inside_box():
    """A tape pruned on a box gives bit-identical values inside the box
    (no NaN in the shape): the pruning invariant the mesher relies on."""
    s = W.sphere_union(50, seed=99)
    deck = s.deck()
    d = oracle.OracleDeck(deck)
    rng = np.random.default_rng(3)
    for _ in range(20):
        c = rng.uniform(-1, 1, 3).astype(np.float32)
        lo, hi = c - 0.1, c + 0.1
        n, changed, _ = d.interval_push(lo, hi)
        assert n <= d.n_clauses


def test_f64_twin_close_to_f32():
    s = W.gyroid_c4()
    d = od(s)
    rng = np.random.default_rng(5)
    p = rng.uniform(-1, 1, (3, 1000)).astype(np.float32)
    v32 = d.values(*p)
    v64 = d.values_f64(*p)
    assert np.abs(v32 - v64).max() < 1e-5
