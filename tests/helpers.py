"""Shared test helpers."""
import numpy as np

from libfive_amd import Shape


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
