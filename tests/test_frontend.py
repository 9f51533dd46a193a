"""Front end (tree construction, optimiser, Deck, walk) and C ABI surface,
checked against the reference's own expectations.  CPU only."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from libfive_amd import Shape, X, Y, Z, ffi
from libfive_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# libfive/test/tree.cpp:283-505 (Tree::optimized), printed forms
@pytest.mark.parametrize("build,expected", [
    (lambda x, y, z: (z + 2) / (2 * (z + 3)), "(/ (+ z 2) (+ (* z 2) 6))"),
    (lambda x, y, z: -3 * x.cos() + z + 2, "(- (+ z 2) (* (cos x) 3))"),
    (lambda x, y, z: x * 2 + y * 5 + z.cos() + 5 * z.cos(),
     "(+ (* x 2) (* y 5) (* (cos z) 6))"),
    (lambda x, y, z: (z - 10) / -z, "(/ (- z 10) (- z))"),
    (lambda x, y, z: (2 * x + y) + (2 * x + y), "(+ (* y 2) (* x 4))"),
    (lambda x, y, z: (2 * x + y) * (2 * x + y), "(square (+ y (* x 2)))"),
    (lambda x, y, z: (x + 2 * y) + 3 * (x + 2 * y).cos(),
     "(+ x (* y 2) (* (cos (+ x (* y 2))) 3))"),
    (lambda x, y, z: (x + 2 * y + 7) + 3 * (x + 2 * y + 7).sin().cos(),
     "(+ x (* y 2) (* (cos (sin (+ x (* y 2) 7))) 3) 7)"),
    (lambda x, y, z: Shape.binary("atan2", -z / (z - 10), -z / (z - 100)),
     "(atan2 (/ (- z) (- z 10)) (/ (- z) (- z 100)))"),
    (lambda x, y, z: (-z) * (-z), "(square (- z))"),
    (lambda x, y, z: -0.091 - (z - 2.7) / 0.6, "(- 4.409 (* z 1.66667))"),
])
def test_optimizer_prints(build, expected):
    t = build(X(), Y(), Z())
    assert str(t.optimized()) == expected


def test_print_unoptimized():
    # tree.cpp:404-409
    z = Z()
    assert str(-0.091 - (z - 2.7) / 0.6) == "(- -0.091 (/ (- z 2.7) 0.6))"
    assert str(X() + 5) == "(+ x 5)"


def test_remap_known_answers():
    # libfive/test/tree.cpp:72-156 (each result compared after flatten,
    # which libfive_tree_remap applies at once)
    x, y, z = X(), Y(), Z()
    L = ffi.lib()
    assert L.libfive_tree_id(x.remap(y, x, x).ptr) == L.libfive_tree_id(y.ptr)
    assert str(x.remap(12, x, x)) == "12"
    assert str((x + 5).remap(3, x, x)) == "8"          # folds while remapping
    f = x * 2 + x * 3 + x
    assert str(f.remap(y, y, z)) == "(+ (* y 2) (* y 3) y)"
    y_plus_two = x.remap(y + 2, z, x)
    assert str((y + z).remap(y, y_plus_two, y_plus_two)) == "(+ y 2 y 2)"
    c = (x * x + y * y).sqrt()
    c = c.remap(x / 2, y, z)
    c = c.remap(x - -0.5, y, z - 0.25)
    assert str(c) == "(sqrt (+ (square (/ (- x -0.5) 2)) (square y)))"
    # identity remap and a tree without x, y, z return the tree itself
    assert L.libfive_tree_id(f.remap(x, y, z).ptr) == L.libfive_tree_id(f.ptr)
    k = Shape.constant(3.0) + 0
    assert L.libfive_tree_id(k.remap(y, z, x).ptr) == L.libfive_tree_id(k.ptr)


def test_remap_deep_chain_and_composition():
    import oracle
    # "Deep remapping" (tree.cpp:96-129): a chain 32768 additions deep is
    # substituted, optimised and compiled without exhausting the stack
    x, y, z = X(), Y(), Z()
    t = x
    for i in range(32768):
        t = t + y * i
    r = t.remap(z, x, y)
    # (y * 0 and y * 1 were simplified to 0 and y when t was built)
    assert str(r)[:22] == "(+ z x (* x 2) (* x 3)"
    clauses = r.deck()[0]
    assert len(clauses) == len(t.deck()[0])
    # composition: g = f(X', Y', Z') equals f evaluated at the remapped
    # coordinates (the optimiser may merge constants across the seam, so
    # the comparison allows a few float32 roundings)
    rng = np.random.default_rng(7)
    p = rng.uniform(-1, 1, (3, 4096)).astype(np.float32)
    f = x * 2 + y.sin() * z - x.min(z)
    X2, Y2, Z2 = y * 3, z - x, (x + 1).cos()
    got = oracle.OracleDeck(f.remap(X2, Y2, Z2).deck()).values(*p)
    q = [oracle.OracleDeck(c.deck()).values(*p) for c in (X2, Y2, Z2)]
    want = oracle.OracleDeck(f.deck()).values(*q)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)
    assert not np.allclose(got, oracle.OracleDeck(f.deck()).values(*p))


def n_slots(t):
    return t.deck()[3]


def test_optimizer_sizes():
    # tree.cpp:375-392: min(X, Y, X, Y) collapses to 3 nodes; the version
    # with X+1 / Y+1 to 6.
    x, y = X(), Y()
    t = (x.min(y)).min(x.min(y))
    assert len(t.optimized().deck()[0]) == 1         # one MIN clause
    xa, xb, ya, yb = x + 1, x + 1, y + 1, y + 1
    t = xa.min(ya).min(xb.min(yb))
    cl = t.optimized().deck()[0]
    assert len(cl) == 3                               # 2 ADD + 1 MIN


def test_identities():
    # tree.cpp:640-746 identity and folding rules
    x = X()
    assert str(x + 0) == "x"
    assert str(0 + x) == "x"
    assert str(x - 0) == "x"
    assert str(0 - x) == "(- x)"
    assert str(x * 1) == "x"
    assert str(1 * x) == "x"
    assert str(x * -1) == "(- x)"
    assert str(x * 0) == "0"
    assert str(x * x) == "(square x)"
    assert str(x / 1) == "x"
    assert str(abs(abs(x))) == "(abs x)"
    assert str(abs(x.square())) == "(square x)"
    assert str(-(-x)) == "x"
    assert str(x.min(x)) == "x"
    assert str(x.max(x)) == "x"
    assert str(x + (-Y())) == "(- x y)"
    assert str(x - (-Y())) == "(+ x y)"
    assert str(Shape.constant(2) + 3) == "5"


def test_deck_layout():
    # libfive/test/deck.cpp:19-41
    clauses, consts, xyz, slots = (X() + 1).deck()
    assert slots == 5
    clauses, consts, xyz, slots = (X() + 2).deck()
    assert xyz == (3, 4, 5)
    clauses, consts, xyz, slots = (X() + 5).deck()
    assert consts == {2: 5.0}
    assert clauses.tolist() == [[17, 1, 3, 2]]


def test_walk_order():
    # tree.cpp:219-262: leaves first, root last; X and Y before Z.
    t = X().min(Y()).min(Z().min(1))
    clauses = t.deck()[0]
    # root-first list: the root is MIN with id 1
    assert clauses[0][0] == 19 and clauses[0][1] == 1
    ids = [int(c[1]) for c in clauses]
    assert ids == sorted(ids)


def test_opcode_enum_and_args():
    L = ffi.lib()
    assert L.libfive_opcode_enum(b"min") == 19
    assert L.libfive_opcode_enum(b"max") == 20
    assert L.libfive_opcode_enum(b"VAR-X") == 2
    assert L.libfive_opcode_enum(b"wat") == -1
    assert L.libfive_opcode_enum(b"") == -1
    assert L.libfive_opcode_args(17) == 2
    assert L.libfive_opcode_args(7) == 1
    assert L.libfive_opcode_args(2) == 0
    assert L.libfive_opcode_args(40) == -1


def test_tree_handles():
    # libfive/test/api.cpp:29-40: X is a singleton, Y differs
    L = ffi.lib()
    a, b, c = L.libfive_tree_x(), L.libfive_tree_x(), L.libfive_tree_y()
    assert a == b and b != c
    for t in (a, b, c):
        L.libfive_tree_delete(t)
    # invalid arity / opcode give NULL (libfive.cpp:116-142)
    x = L.libfive_tree_x()
    assert L.libfive_tree_unary(17, x) is None
    assert L.libfive_tree_binary(7, x, x) is None
    assert L.libfive_tree_binary(17, x, None) is None
    assert L.libfive_tree_nullary(99) is None
    L.libfive_tree_delete(x)


def test_workload_tapes():
    info = {}
    for name, f in [("c1", W.csg_c1), ("c2", W.random_dag),
                    ("c3", W.sphere_union), ("c4", W.gyroid_c4)]:
        info[name] = f().tape_info()
    # clause counts match the Deck, registers are small enough for VGPRs
    assert info["c3"][0] == 10999
    assert info["c2"][0] == len(W.random_dag().deck()[0])
    for c, r in info.values():
        assert r <= 256


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "libfive.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(libfive_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_abi_exports_every_declared_symbol():
    so = ffi.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", so],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in _declared_symbols() if n not in exported]
    assert not missing, missing
    L = ffi.lib()
    for n in _declared_symbols():
        assert hasattr(L, n)


def test_no_gpu_fails_loudly():
    """Without a GPU, evaluation entry points report an error instead of
    computing on the CPU."""
    import libfive_amd
    if libfive_amd.gpu_available():
        pytest.skip("GPU present")
    t = X() + 1
    v = t.eval_f(1.0, 2.0, 3.0)
    assert np.isnan(v)
    assert "GPU" in ffi.last_error()
    with pytest.raises(libfive_amd.GpuError):
        t.values(np.zeros(4), np.zeros(4), np.zeros(4))
    c = W.circle(1.0)
    for call in (lambda: c.render_slice((-2, -2), (2, 2), 0.0, 10),
                 lambda: c.render_pixels((-2, -2), (2, 2), 0.0, 10),
                 lambda: c.render_mesh([-2] * 3, [2] * 3, res=8)):
        with pytest.raises(libfive_amd.GpuError):
            call()


def test_specialized_kernels_compile_on_host():
    """The straight-line kernels generated from a tape compile for gfx950
    with hipRTC without a GPU (the GPU tests check their results)."""
    c1 = W.csg_c1()
    assert ffi.lib().libfive_tree_specialized_code_size(c1.ptr) > 0, \
        ffi.last_error()
    s = Shape.binary("atan2", X(), Y()) + Shape.unary("exp", Z()) + \
        Shape.binary("mod", X(), 0.5) + X().sin() * Y().cos()
    assert ffi.lib().libfive_tree_specialized_code_size(s.ptr) > 0, \
        ffi.last_error()


def _vars(pairs):
    n = len(pairs)
    return ffi.libfive_vars((ctypes.c_void_p * n)(*[p for p, _ in pairs]),
                            (ctypes.c_float * n)(*[v for _, v in pairs]), n)


def test_evaluator_update_vars():
    """libfive_evaluator_update_vars (Evaluator::updateVars,
    libfive/include/libfive/eval/evaluator.hpp:41-49): true only when a
    variable the tree uses changes value."""
    L = ffi.lib()
    v, w = Shape.var(), Shape.var()
    t = X() + v * 2.0
    vid, wid = L.libfive_tree_id(v.ptr), L.libfive_tree_id(w.ptr)
    ev = L.libfive_tree_evaluator(t.ptr, _vars([(vid, 1.0)]))
    try:
        assert L.libfive_evaluator_update_vars(ev, _vars([(vid, 2.0)]))
        assert not L.libfive_evaluator_update_vars(ev, _vars([(vid, 2.0)]))
        assert not L.libfive_evaluator_update_vars(ev, _vars([(wid, 5.0)]))
    finally:
        L.libfive_evaluator_delete(ev)


def test_autotune_report_without_decisions():
    """The autotuner's report is valid JSON before anything has been tuned
    (no GPU needed), and enabling, disabling and resetting are plain
    calls."""
    from libfive_amd import autotune_report, ffi
    L = ffi.lib()
    L.libfive_autotune_reset()
    assert autotune_report() == []
    L.libfive_autotune_enable(0)
    L.libfive_autotune_enable(1)
    L.libfive_autotune_force(b"leaf_edges/banks=1/fast=1", 1)
    assert autotune_report() == []          # forcing records no timing
    L.libfive_autotune_reset()
