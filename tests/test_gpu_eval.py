"""GPU parity of the evaluation kernels against the oracle (and golden
vectors), through the C ABI."""
import math
import os

import numpy as np
import pytest

import oracle
from libfive_amd import Shape, X, Y, Z, ffi
from libfive_amd import workloads as W

from helpers import (BINARY, UNARY, assert_same_bits, op_shape, ulp_distance,
                     wide_registers)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ieee_ops.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name", BINARY + UNARY)
def test_ieee_ops_golden(gpu, golden, name):
    a, b = golden["a"], golden["b"]
    got = op_shape(name).values(a, b, np.zeros_like(a))
    assert_same_bits(got, golden[name], name)


def test_sincos_golden(gpu, golden):
    x = golden["trig_x"]
    s = X().sin().values(x, x, x)
    c = X().cos().values(x, x, x)
    assert ulp_distance(s, golden["sin_rn"]).max() <= 2
    assert ulp_distance(c, golden["cos_rn"]).max() <= 2
    os_, oc = oracle.sincos(x)
    assert_same_bits(s, os_, "sin vs oracle")
    assert_same_bits(c, oc, "cos vs oracle")


def test_c_api_known_answers(gpu):
    # libfive/test/api.cpp:52-83
    L = ffi.lib()
    a, b = L.libfive_tree_x(), L.libfive_tree_y()
    c = L.libfive_tree_binary(22, a, b)
    assert L.libfive_tree_eval_f(c, ffi.libfive_vec3(1, 2, 3)) == 0.5
    assert L.libfive_tree_eval_f(c, ffi.libfive_vec3(1, 4, 3)) == 0.25
    assert L.libfive_tree_eval_f(c, ffi.libfive_vec3(1, -1, 3)) == -1
    d = L.libfive_tree_binary(21, a, b)
    r = L.libfive_tree_eval_r(d, ffi.libfive_region3(
        ffi.libfive_interval(1, 2), ffi.libfive_interval(2, 3),
        ffi.libfive_interval(0, 0)))
    assert (r.lower, r.upper) == (-2, 0)
    g = L.libfive_tree_eval_d(c, ffi.libfive_vec3(1, 2, 3))
    assert (g.x, g.y, g.z) == (0.5, -0.25, 0.0)
    for t in (a, b, c, d):
        L.libfive_tree_delete(t)


def test_known_answers_gpu(gpu):
    assert X().eval_f(1, 2, 3) == 1.0
    assert (X() * 2 + Y() * 3).eval_f(1, 2, 3) == 8.0
    assert Shape.constant(3.14).eval_f(0, 0, 0) == np.float32(3.14)
    assert Shape.binary("nth-root", X(), 3).eval_f(-0.5, 0, 0) == \
        pytest.approx(-0.7937, rel=1e-4)
    assert (X() + 1).eval_r((1, 1, 1), (2, 2, 2)) == (2.0, 3.0)
    assert ((X() + Y()) * (X() - Y())).eval_r((0, 0, 0), (1, 1, 1)) == (-2.0, 2.0)
    lo, hi = (0.1 / Y()).eval_r((-0.625, -0.3125, -0.625), (-0.3125, 0, -0.2125))
    assert lo == -math.inf and hi == math.inf


SHAPES = {
    "c1_csg": lambda: W.csg_c1(),
    "c4_gyroid": lambda: W.gyroid_c4(),
    "c2_dag": lambda: W.random_dag(),
    "dag_small": lambda: W.random_dag(120, seed=5),
    "spheres50": lambda: W.sphere_union(50, seed=3),
    # common opcodes only, with sin / cos arguments on both sides of 2^17:
    # the derivative interpreter's call-free pass and its redo of the
    # points it cannot evaluate
    "large_sincos": lambda: ((X() * 2e5).sin() * (Y() * 1e5 + 3e4).cos()
                             + (Z() * 0.5).sin() + abs(X()).sqrt()),
    # sin / cos of infinities (x * 3e38 * 2 overflows for |x| > 0.57)
    "inf_sincos": lambda: ((X() * 3e38 * 2).sin() + (Y() * 3e38 * 2).cos()
                           + Z().sin()),
    "all_ops": lambda: (Shape.binary("atan2", X(), Y() + 0.5)
                        + Shape.binary("mod", X() * 7, Y() + 2)
                        + Shape.binary("pow", X(), 3)
                        + Shape.binary("nth-root", Y() * 4, 3)
                        + X().sin() * Z().cos() + abs(Y()).sqrt()
                        + Shape.unary("exp", Z()) + Shape.unary("log", abs(X()) + 1)
                        + Shape.unary("atan", Y()) + Shape.unary("tan", Z() * 0.5)
                        + Shape.unary("asin", X() * 0.5)
                        + Shape.unary("acos", Y() * 0.5)
                        + Shape.binary("compare", X(), Y())
                        + Shape.binary("nanfill", (X() - 0.5).sqrt(), Y())
                        + Shape.unary("recip", Z() + 3)),
    "wide_75_registers": wide_registers,
    # 200 registers: eight banks, part of each file outside the VGPRs
    "wide_200_registers": lambda: wide_registers(195),
}


def points(n, seed, scale=1.2):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, (3, n)).astype(np.float32)
    p[:, :9] = np.array([[0, 0.5, -0.5, 1, -1, 0.25, 0, 0, 1e-20]] * 3,
                        np.float32)
    return p


@pytest.mark.parametrize("name", sorted(SHAPES))
def test_values_match_oracle(gpu, name):
    s = SHAPES[name]()
    p = points(20000, hash(name) & 0xffff)
    got = s.values(*p)
    ref = oracle.OracleDeck(s.deck()).values(*p)
    assert_same_bits(got, ref, name)


@pytest.mark.parametrize("name", sorted(SHAPES))
def test_derivs_match_oracle(gpu, name):
    s = SHAPES[name]()
    p = points(20000, (hash(name) + 1) & 0xffff)
    got = s.derivs(*p)
    ref = oracle.OracleDeck(s.deck()).derivs(*p)
    for k, nm in enumerate(("v", "dx", "dy", "dz")):
        if name == "all_ops" and nm != "v":
            # pow / tan / nth-root partials use double-precision libm
            # values on both sides; compare within 4 ulp
            assert ulp_distance(got[k], ref[k]).max() <= 4, nm
        else:
            assert_same_bits(got[k], ref[k], f"{name}:{nm}")


@pytest.mark.parametrize("name", sorted(SHAPES))
def test_intervals_match_oracle(gpu, name):
    s = SHAPES[name]()
    rng = np.random.default_rng(11)
    c = rng.uniform(-1.2, 1.2, (4000, 3)).astype(np.float32)
    w = (rng.uniform(0, 1, (4000, 3)) ** 3 * 0.6).astype(np.float32)
    lo, hi = c - w, c + w
    gl, gh, gn = s.intervals(lo, hi)
    rl, rh, rn = oracle.OracleDeck(s.deck()).intervals(lo, hi)
    assert (gn == rn).all()
    assert_same_bits(gl, rl, name + ":lower")
    assert_same_bits(gh, rh, name + ":upper")


def lattice_points(n, lo=-1.5, hi=1.5):
    """An n^3 lattice that lands exactly on the CSG shapes' planes, corners
    and tie surfaces (the points where MIN/MAX operands compare equal)."""
    t = np.linspace(lo, hi, n).astype(np.float32)
    x, y, z = np.meshgrid(t, t, t, indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


AMBIG_SHAPES = {
    "min_x_negx": lambda: X().min(-X()),
    "deriv_known_answer": lambda: X().min(Y()).min(X().min(4 + 0.2 * Y())),
    "equal_partials": lambda: X().min(X() * 1.0 + 0.0 * Y()),
    "c1_csg": lambda: W.csg_c1(),
    "box": lambda: W.box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)),
    "spheres50": lambda: W.sphere_union(50, seed=3),
    "c2_dag": lambda: W.random_dag(),
}


@pytest.mark.parametrize("name", sorted(AMBIG_SHAPES))
@pytest.mark.parametrize("derivs", [False, True])
def test_ambiguity_flags_match_oracle(gpu, name, derivs):
    """Device getAmbiguous / getAmbiguousDerivs flags equal the oracle's,
    point by point, on a lattice through the shapes' ties and on random
    points."""
    s = AMBIG_SHAPES[name]()
    lx, ly, lz = lattice_points(33)
    r = points(4000, 17)
    p = [np.concatenate([a, b]) for a, b in zip((lx, ly, lz), r)]
    got = s.ambiguous(*p, derivs=derivs)
    o = oracle.OracleDeck(s.deck())
    ref = o.deriv_ambiguous(*p) if derivs else o.ambiguous(*p)
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, (name, bad[:8], [q[bad[:4]] for q in p])
    if name in ("min_x_negx", "c1_csg", "box", "deriv_known_answer"):
        assert ref.any()        # the lattice does meet ties


def test_ambiguity_known_answers_gpu(gpu):
    """libfive/test/eval_array.cpp:203-221 and
    libfive/test/eval_deriv_array.cpp:94-111 on the device."""
    a = X().min(-X()).ambiguous([0, 1, 2, 0], [0] * 4, [0] * 4)
    assert a.tolist() == [True, False, False, True]
    s = X().min(Y()).min(X().min(4 + 0.2 * Y()))
    d = s.ambiguous([0, 0, 5, 0], [0, 1, 5, 5], [0] * 4, derivs=True)
    assert d.tolist() == [True, False, True, False]


def test_grid_c2_exact_and_enclosed(gpu):
    """C2 DAG on a 64^3 grid (the bench runs 256^3): bit-exact vs the
    oracle, including sin/cos."""
    s = W.random_dag()
    x, y, z = W.grid_points(64)
    got = s.values(x, y, z)
    ref = oracle.OracleDeck(s.deck()).values(x, y, z, threads=8)
    assert_same_bits(got, ref, "c2 grid")


def test_directed_rounding_div_sqrt(gpu):
    """Interval bounds of x / y and sqrt(x) for random point boxes equal the
    oracle's fesetround results bit for bit."""
    rng = np.random.default_rng(13)
    n = 20000
    a = (rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, n)).astype(np.float32)
    b = (rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, n)).astype(np.float32)
    lo = np.stack([a, b, np.zeros(n)], 1).astype(np.float32)
    for s in (X() / Y(), abs(X()).sqrt(), X() * Y(), X() + Y(), X() - Y()):
        gl, gh, _ = s.intervals(lo, lo)
        rl, rh, _ = oracle.OracleDeck(s.deck()).intervals(lo, lo)
        assert_same_bits(gl, rl, str(s))
        assert_same_bits(gh, rh, str(s))


def _device_eval(tape, p, derivs):
    """Runs a compiled tape (its current path) on the points p (3, n)."""
    L = ffi.lib()
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in p)
    n = x.size
    f = lambda a: a.ctypes.data_as(ffi.fp)
    if derivs:
        out = [np.empty(n, np.float32) for _ in range(4)]
        rc = L.libfive_tape_eval_derivs(tape, f(x), f(y), f(z), n,
                                        *[f(o) for o in out])
    else:
        out = [np.empty(n, np.float32)]
        rc = L.libfive_tape_eval_values(tape, f(x), f(y), f(z), n, f(out[0]))
    assert rc == 0, ffi.last_error()
    return out


@pytest.mark.parametrize("name", ["c1_csg", "c4_gyroid", "dag_small",
                                  "spheres50", "all_ops", "large_sincos",
                                  "inf_sincos"])
def test_specialized_kernels_match_interpreter(gpu, name):
    """The hipRTC-specialised kernels give the interpreter's bits (values
    and partials), which are the oracle's."""
    s = SHAPES[name]()
    p = points(20000, (hash(name) + 7) & 0xffff)
    L = ffi.lib()
    tape = L.libfive_tape_new(s.ptr)
    assert tape, ffi.last_error()
    try:
        iv = _device_eval(tape, p, False)
        idr = _device_eval(tape, p, True)
        ms = ctypes_double()
        assert L.libfive_tape_specialize(tape, ms) == 0, ffi.last_error()
        print(f"{name}: specialised in {ms.value:.0f} ms", flush=True)
        sv = _device_eval(tape, p, False)
        sdr = _device_eval(tape, p, True)
    finally:
        L.libfive_tape_delete(tape)
    assert_same_bits(sv[0], iv[0], name + ":values")
    for k, nm in enumerate(("v", "dx", "dy", "dz")):
        assert_same_bits(sdr[k], idr[k], f"{name}:{nm}")
    ref = oracle.OracleDeck(s.deck()).values(*p)
    assert_same_bits(sv[0], ref, name + ":oracle")


def test_specialized_c2_grid_matches_oracle(gpu):
    """The 809-clause random DAG, specialised, on a 32^3 grid: values and
    partials bit-exact against the oracle."""
    s = W.random_dag()
    x, y, z = W.grid_points(32)
    p = np.stack([x, y, z])
    L = ffi.lib()
    tape = L.libfive_tape_new(s.ptr)
    try:
        assert L.libfive_tape_specialize(tape, None) == 0, ffi.last_error()
        got = _device_eval(tape, p, True)
    finally:
        L.libfive_tape_delete(tape)
    ref = oracle.OracleDeck(s.deck()).derivs(x, y, z, threads=8)
    for k, nm in enumerate(("v", "dx", "dy", "dz")):
        assert_same_bits(got[k], ref[k], "c2:" + nm)


@pytest.mark.parametrize("group", ["1", "4"])
def test_specialized_pin_groups_match_oracle(gpu, group, monkeypatch):
    """Other clause groupings of the specialised kernels (read from
    LIBFIVE_AMD_JIT_GROUP when the source is generated): with one clause
    per group the benchmark tape's derivative kernel needs 170 VGPRs and
    is compiled a second time with a request for three waves per SIMD.
    Values and partials stay bit-exact against the oracle."""
    monkeypatch.setenv("LIBFIVE_AMD_JIT_GROUP", group)
    s = W.random_dag()
    x, y, z = W.grid_points(24)
    p = np.stack([x, y, z])
    L = ffi.lib()
    tape = L.libfive_tape_new(s.ptr)
    try:
        assert L.libfive_tape_specialize(tape, None) == 0, ffi.last_error()
        got = _device_eval(tape, p, True)
        gv = _device_eval(tape, p, False)
    finally:
        L.libfive_tape_delete(tape)
    o = oracle.OracleDeck(s.deck())
    ref = o.derivs(x, y, z, threads=8)
    assert_same_bits(gv[0], o.values(x, y, z), f"group {group}: values")
    for k, nm in enumerate(("v", "dx", "dy", "dz")):
        assert_same_bits(got[k], ref[k], f"group {group}: {nm}")


def ctypes_double():
    import ctypes
    return ctypes.c_double()


def test_directed_rounding_special_values(gpu):
    """Interval bounds of x + y, x - y and x * y on point boxes built from
    special values (signed zeros, exact cancellation, overflow to
    infinity, subnormals, widely different exponents, infinities) equal
    the oracle's fesetround results bit for bit."""
    sv = np.array([0.0, -0.0, 1.0, -1.0, 3.0, -3.0, 1e-30, -1e-30, 1e30, -1e30,
                   3e38, -3e38, 3.4028235e38, -3.4028235e38, 1.4e-45, -1.4e-45,
                   1.1754942e-38, 1e-40, 0.1, -0.1, 1.0000001, 16777217.0,
                   np.inf, -np.inf], np.float32)
    a, b = np.meshgrid(sv, sv)
    a, b = a.ravel(), b.ravel()
    lo = np.stack([a, b, np.zeros_like(a)], 1).astype(np.float32)
    for s in (X() + Y(), X() - Y(), X() * Y()):
        gl, gh, gn = s.intervals(lo, lo)
        rl, rh, rn = oracle.OracleDeck(s.deck()).intervals(lo, lo)
        assert (gn == rn).all(), str(s)
        assert_same_bits(gl, rl, str(s) + ":lower")
        assert_same_bits(gh, rh, str(s) + ":upper")


def _division_shapes():
    """Quotients and square roots whose partials divide operands spanning
    2^-70 .. 2^70, with signed zeros and tiny results."""
    out = []
    for e in (-70, -45, -41, -40, -20, 0, 20, 40, 41, 70):
        c = float(2.0 ** e)
        out.append((X() * c + Y()) / (Y() * (1.0 / c) + Z() * 3 + 0.5))
        out.append(abs(X() * X() * c + Y() * (c * 0.5)).sqrt())
        out.append(abs(X() * c).sqrt() * Y() + abs(Z()).sqrt())
    return out


def test_div_sqrt_partials_all_magnitudes(gpu):
    """Partials of x / y and sqrt(x) (three divisions each) on
    the interpreter and the specialised kernels: bit-identical to the
    oracle's division operator, including signed zeros, zero divisors and
    tiny results."""
    rng = np.random.default_rng(8)
    n = 6000
    p = (rng.standard_normal((3, n)) *
         10.0 ** rng.integers(-8, 8, (3, n))).astype(np.float32)
    p[:, :6] = np.array([[0.0, -0.0, 1.0, 0.0, 1e-30, -3.0]] * 3, np.float32)
    p[1, 6:12] = -0.0
    L = ffi.lib()
    for s in _division_shapes():
        ref = oracle.OracleDeck(s.deck()).derivs(*p)
        got = s.derivs(*p)
        for k, nm in enumerate(("v", "dx", "dy", "dz")):
            assert_same_bits(got[k], ref[k], f"{s}: {nm}")
    # the specialised kernels, on two of the shapes
    for s in _division_shapes()[3:5]:
        tape = L.libfive_tape_new(s.ptr)
        assert tape, ffi.last_error()
        try:
            assert L.libfive_tape_specialize(tape, ctypes_double()) == 0, ffi.last_error()
            sd = _device_eval(tape, p, True)
        finally:
            L.libfive_tape_delete(tape)
        ref = oracle.OracleDeck(s.deck()).derivs(*p)
        for k in range(4):
            assert_same_bits(sd[k], ref[k], f"{s}: specialised {k}")
