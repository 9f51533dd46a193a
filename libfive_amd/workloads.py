"""Seeded synthetic workloads (shapes and point sets) for tests and bench.

Shapes follow the reference's formulas: sphere / box / cylinder / blend
from libfive/test/util/shapes.cpp:71-168 and the CSG operators of
libfive/stdlib/stdlib_impl.cpp:51-65 (union = min, intersection = max,
difference = max(a, -b)).  Random numbers come from SplitMix64 with seed
0x11B5F1 + config index, so every run builds the same trees.
"""
import math

import numpy as np

from .shape import Shape, X, Y, Z

SEED_BASE = 0x11B5F1
MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, n):
        return self.next_u64() % n


def sphere(r, c=(0.0, 0.0, 0.0)):
    x, y, z = X(), Y(), Z()
    return ((x - c[0]).square() + (y - c[1]).square()
            + (z - c[2]).square()).sqrt() - r


def box(lo, hi):
    x, y, z = X(), Y(), Z()
    return (((lo[0] - x).max(x - hi[0])).max((lo[1] - y).max(y - hi[1]))
            .max((lo[2] - z).max(z - hi[2])))


def circle(r, c=(0.0, 0.0)):
    return ((X() - c[0]).square() + (Y() - c[1]).square()).sqrt() - r


def cylinder(r, h, base=(0.0, 0.0, 0.0)):
    z = Z()
    c = circle(r, (base[0], base[1]))
    return c.max((base[2] - z).max(z - (base[2] + h)))


# ---------------------------------------------------------------------------
# Config 1: sphere r=1 UNION box [-.5,.5]x[-.5,.5]x[0,1.5] DIFFERENCE an
# infinite z-axis cylinder of radius 0.3.  Region [-2,2]^3, res 16.
def csg_c1():
    body = sphere(1.0).min(box((-0.5, -0.5, 0.0), (0.5, 0.5, 1.5)))
    hole = circle(0.3)
    return body.max(-hole)


C1_REGION = ((-2.0, -2.0, -2.0), (2.0, 2.0, 2.0))
C1_RES = 16.0

# ---------------------------------------------------------------------------
# Config 2: random expression DAG of 1000 ops over x, y, z and constants.
C2_OPS = ("add", "sub", "mul", "min", "max", "neg", "abs", "square",
          "sqrtabs", "sin", "cos")


def random_dag(n_ops=1000, seed=SEED_BASE + 1):
    """Random expression of n_ops operations drawn uniformly from C2_OPS
    ("sqrtabs" is sqrt(abs(.)) and counts as one draw).  The expression is
    a tree over shared leaves x, y, z and constants U(-2, 2): each binary
    operation splits its remaining operation budget uniformly between its
    two operands.  Shared interior nodes are avoided on purpose: the
    optimiser's affine and commutative flattening re-expands every shared
    node inside a group, which is exponential on heavily shared DAGs (in
    the reference as well)."""
    rng = SplitMix64(seed)
    leaves = [X(), Y(), Z()]

    def leaf():
        u = rng.randint(4)
        if u == 3:
            return Shape.constant(rng.uniform(-2.0, 2.0))
        return leaves[u]

    # Iterative construction (post-order) to avoid Python recursion limits.
    # Each task is (budget, op or None); results are pushed on `out`.
    out = []
    todo = [("build", n_ops)]
    while todo:
        kind, arg = todo.pop()
        if kind == "build":
            n = arg
            if n == 0:
                out.append(leaf())
                continue
            op = C2_OPS[rng.randint(len(C2_OPS))]
            if op in ("add", "sub", "mul", "min", "max"):
                k = rng.randint(n)              # ops in the left operand
                todo.append(("binary", op))
                todo.append(("build", n - 1 - k))
                todo.append(("build", k))
            else:
                todo.append(("unary", op))
                todo.append(("build", n - 1))
        elif kind == "binary":
            b = out.pop()
            a = out.pop()
            out.append(Shape.binary(arg, a, b))
        else:
            a = out.pop()
            if arg == "sqrtabs":
                out.append(abs(a).sqrt())
            else:
                out.append(Shape.unary(arg, a))
    assert len(out) == 1
    return out[0]


def grid_points(n, lo=-1.0, hi=1.0):
    """n^3 regular grid on [lo, hi]^3, x fastest: x_i = lo + (hi-lo) i/(n-1)."""
    t = (lo + (hi - lo) * np.arange(n, dtype=np.float64) / (n - 1)).astype(np.float32)
    zz, yy, xx = np.meshgrid(t, t, t, indexing="ij")
    return xx.ravel(), yy.ravel(), zz.ravel()


# ---------------------------------------------------------------------------
# Config 3: min-union of 1000 spheres.
def sphere_union(n=1000, seed=SEED_BASE + 2):
    rng = SplitMix64(seed)
    out = None
    for _ in range(n):
        c = (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1))
        r = rng.uniform(0.02, 0.1)
        s = sphere(r, c)
        out = s if out is None else out.min(s)
    return out


C3_REGION = ((-1.2, -1.2, -1.2), (1.2, 1.2, 1.2))
C3_DEPTH = 8

# ---------------------------------------------------------------------------
# Config 4: thickened gyroid intersected with the unit sphere.
def gyroid_c4(k=10.0):
    x, y, z = X(), Y(), Z()
    g = ((k * x).cos() * (k * y).sin() + (k * y).cos() * (k * z).sin()
         + (k * z).cos() * (k * x).sin())
    return (abs(g) - 0.1).max(sphere(1.0))


C4_REGION = ((-1.2, -1.2, -1.2), (1.2, 1.2, 1.2))
C4_RES = 512 / 2.4

# ---------------------------------------------------------------------------
# Config 5: random CSG models.
def random_csg(n_ops, rng):
    """A tree of primitives (sphere / box / cylinder, random affine remaps
    written out as expressions) combined with min / max / neg until about
    n_ops tree operations are used."""
    def prim():
        kind = rng.randint(3)
        c = [rng.uniform(-0.6, 0.6) for _ in range(3)]
        s = rng.uniform(0.5, 1.5)
        ang = rng.uniform(0, 2 * math.pi)
        ca, sa = math.cos(ang), math.sin(ang)
        # rotate about z, scale, translate: p' = R (p - c) / s
        x, y, z = X() - c[0], Y() - c[1], Z() - c[2]
        xr = (ca * x + sa * y) * (1.0 / s)
        yr = (ca * y - sa * x) * (1.0 / s)
        zr = z * (1.0 / s)
        if kind == 0:
            r = rng.uniform(0.1, 0.4)
            return (xr.square() + yr.square() + zr.square()).sqrt() - r
        if kind == 1:
            h = [rng.uniform(0.05, 0.3) for _ in range(3)]
            return ((abs(xr) - h[0]).max(abs(yr) - h[1])).max(abs(zr) - h[2])
        r = rng.uniform(0.05, 0.3)
        h = rng.uniform(0.1, 0.5)
        return ((xr.square() + yr.square()).sqrt() - r).max(abs(zr) - h)

    ops_per_prim = 18
    n_prims = max(1, n_ops // ops_per_prim)
    items = [prim() for _ in range(n_prims)]
    while len(items) > 1:
        i = rng.randint(len(items))
        a = items.pop(i)
        j = rng.randint(len(items))
        b = items.pop(j)
        u = rng.randint(4)
        if u < 2:
            t = a.min(b)
        elif u == 2:
            t = a.max(b)
        else:
            t = a.max(-b)
        items.append(t)
    return items[0]


def csg_batch(n_models=256, seed=SEED_BASE + 4):
    rng = SplitMix64(seed)
    out = []
    for _ in range(n_models):
        n_ops = 10 + rng.randint(1991)
        res = (32, 64, 128, 256)[rng.randint(4)]
        out.append((random_csg(n_ops, rng), res))
    return out
