"""Shared test helpers."""
import numpy as np

from libfive_amd import Shape, X, Y, Z


def same_bits(a, b):
    """Element-wise bit equality of float arrays, with every NaN equal to
    every other NaN (payloads differ between CPUs and GPUs)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    eq = a.view(np.uint32) == b.view(np.uint32)
    return (na & nb) | (~na & ~nb & eq)


def assert_same_bits(a, b, what=""):
    ok = same_bits(a, b)
    if not ok.all():
        i = np.flatnonzero(~ok)[:8]
        raise AssertionError(
            f"{what}: {int((~ok).sum())} of {ok.size} differ; first at "
            f"{i.tolist()}: {np.asarray(a).ravel()[i]} vs "
            f"{np.asarray(b).ravel()[i]}")


def ulp_distance(a, b):
    """|a - b| in units in the last place of float32 (NaN vs NaN = 0)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)

    def key(x):
        i = x.view(np.int32).astype(np.int64)
        return np.where(i < 0, -(i & 0x7fffffff), i)

    d = np.abs(key(a) - key(b))
    both_nan = np.isnan(a) & np.isnan(b)
    return np.where(both_nan, 0, d)


BINARY = ["add", "sub", "mul", "div", "min", "max"]
UNARY = ["neg", "abs", "square", "sqrt"]


def op_shape(name):
    """Single-clause tree over x (and y for binary opcodes)."""
    x, y = Shape.X(), Shape.Y()
    if name in BINARY:
        return Shape.binary(name, x, y)
    return Shape.unary(name, x)


def wide_registers(n=70):
    """A shape whose tape keeps about n + 5 values live at once: n sines,
    summed in one order and max-folded in the other, plus a sphere.  With
    n = 70 it needs 75 registers (four register banks).  The phases differ
    so that no lattice point makes the sines equal (a max over many exact
    ties has exponentially many features)."""
    from libfive_amd import workloads as W
    x, y = X(), Y()
    a = [((x * (0.1 * i + 1.0)) + y * (0.05 * i + 0.3) + (0.37 * i + 0.1)).sin()
         for i in range(n)]
    s1 = a[0]
    for t in a[1:]:
        s1 = s1 + t
    s2 = a[-1]
    for t in reversed(a[:-1]):
        s2 = s2.max(t)
    return s1 * 0.02 + s2 * 0.1 + W.sphere(0.8)
