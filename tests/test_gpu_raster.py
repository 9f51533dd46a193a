"""GPU Z-slice occupancy raster (libfive_tree_render_pixels) through the C
ABI: the reference's expected images (libfive/test/heightmap.cpp:52-181),
bit-identical images against the oracle's restatement of the reference
recursion on ragged and CSG cases, and at large sizes the size-independent
property that a pixel is set exactly where the shape's value at the voxel
centre is below zero (GPU point values, centres from the oracle's voxel
grid)."""
import numpy as np
import pytest

import oracle
from libfive_amd import X, Y
from libfive_amd.shape import Shape
from libfive_amd import workloads as W

from test_raster_oracle import CASES, CIRCLE, ORIENT_X, ORIENT_Y

from helpers import wide_registers

pytestmark = pytest.mark.gpu


def od(shape):
    return oracle.OracleDeck(shape.deck())


@pytest.mark.parametrize("z", [0.0, 1.0])
def test_reference_circle_image(z):
    img = W.circle(1.0).render_pixels((-1, -1), (1, 1), z, 5)
    assert img.shape == (10, 10)
    assert (img == CIRCLE).all()


def test_reference_orientation_and_shape():
    c = W.circle(1.0)
    assert (c.max(Y()).render_pixels((-1, -1), (1, 1), 0, 5) == ORIENT_Y).all()
    assert (c.max(X()).render_pixels((-1, -1), (1, 1), 0, 5) == ORIENT_X).all()
    assert c.render_pixels((0, -1), (1, 1), 0, 5).shape == (10, 5)
    assert c.render_pixels((-1, 0), (1, 1), 0, 5).shape == (5, 10)


BIG = CASES + [
    ("gyroid_833", lambda: W.gyroid_c4(), (-1.2, -1.2), (1.2, 1.2), -0.3,
     347.0),
    ("sphere_union_700", lambda: W.sphere_union(300), (-1, -1), (1, 1), 0.2,
     350.0),
    ("random_csg", lambda: W.random_csg(200, np.random.RandomState(7)),
     (-1.3, -1.1), (1.2, 1.4), -0.05, 211.0),
]


@pytest.mark.parametrize("name,mk,lo,hi,z,res", BIG, ids=[c[0] for c in BIG])
def test_matches_oracle(name, mk, lo, hi, z, res):
    s = mk()
    print(f"raster parity: {name} res={res}", flush=True)
    g = s.render_pixels(lo, hi, z, res)
    o = od(s).render_pixels(lo, hi, z, res)
    assert g.shape == o.shape
    assert (g == o).all()


def test_wide_register_file_matches_oracle():
    """37 registers: the raster's kernels with two register banks."""
    s = W.random_dag() - 0.3
    assert s.tape_info()[1] > 32
    g = s.render_pixels((-1, -1), (1, 1), 0.1, 200)
    o = od(s).render_pixels((-1, -1), (1, 1), 0.1, 200)
    assert g.shape == o.shape and (g == o).all()
    assert g.any() and not g.all()


def test_four_register_banks_match_oracle():
    s = wide_registers(70)
    assert s.tape_info()[1] > 64
    g = s.render_pixels((-1, -1), (1, 1), 0.1, 200)
    o = od(s).render_pixels((-1, -1), (1, 1), 0.1, 200)
    assert g.shape == o.shape and (g == o).all()
    assert g.any() and not g.all()


def test_eight_register_banks_match_oracle():
    s = wide_registers(195)
    assert s.tape_info()[1] > 128
    g = s.render_pixels((-1, -1), (1, 1), 0.1, 40.0)
    o = od(s).render_pixels((-1, -1), (1, 1), 0.1, 40.0)
    assert (g == o).all()


def test_empty_full_and_single_pixel():
    far = W.sphere(0.5, (10, 10, 10))
    assert not far.render_pixels((-1, -1), (1, 1), 0, 50).any()
    assert W.sphere(10.0).render_pixels((-1, -1), (1, 1), 0, 50).all()
    one = W.sphere(1.0).render_pixels((0, 0), (0, 0), 0, 10)
    assert one.shape == (1, 1) and one[0, 0]


def _nanfill(a, b):
    return Shape.binary("nanfill", a, b)


# Min / max operands that may be NaN in a view: the reference evaluates a
# view's pixels with the tape its ancestors pushed, whose MIN / MAX keep
# rule ignores the may-be-NaN flag (eval_interval.cpp:127-160), so pixels
# where the dropped operand is NaN take the kept one.  On these shapes that
# differs from the unpruned value's sign at 1024 and 512 pixels.
NAN_CASES = [
    ("min_sqrt", lambda: X().sqrt().min(Y() - 10.0), (-2, -2), (2, 2), 0.0, 16.0),
    ("nanfill_min", lambda: _nanfill(X().sqrt().min(Y() - 0.5), Shape.constant(1.0)),
     (-2, -2), (2, 2), 0.0, 16.0),
    ("nanfill_min_fine", lambda: _nanfill(X().sqrt().min(Y() - 0.5), Shape.constant(1.0)),
     (-2.1, -1.7), (1.9, 2.3), 0.0, 77.0),
]


@pytest.mark.parametrize("name,mk,lo,hi,z,res", NAN_CASES, ids=[c[0] for c in NAN_CASES])
def test_pruned_tapes_with_nan_operands_match_oracle(name, mk, lo, hi, z, res):
    s = mk()
    print(f"raster parity: {name} res={res}", flush=True)
    g = s.render_pixels(lo, hi, z, res)
    o = od(s).render_pixels(lo, hi, z, res)
    assert g.shape == o.shape
    assert (g == o).all()


def test_long_tape_matches_oracle():
    """27 499 clauses: interval work arrays in the global workspace."""
    s = W.sphere_union(2500)
    g = s.render_pixels((-1.2, -1.2), (1.2, 1.2), 0.1, 100.0)
    o = od(s).render_pixels((-1.2, -1.2), (1.2, 1.2), 0.1, 100.0)
    assert g.shape == o.shape
    assert (g == o).all()


def test_large_raster_matches_oracle():
    s = W.gyroid_c4()
    res = 4096 / 2.4
    print("raster parity: gyroid 4096x4096", flush=True)
    g = s.render_pixels((-1.2, -1.2), (1.2, 1.2), 0.05, res)
    o = od(s).render_pixels((-1.2, -1.2), (1.2, 1.2), 0.05, res)
    assert g.shape == o.shape == (4096, 4096)
    assert (g == o).all()


@pytest.mark.parametrize("n", [4096])
def test_large_raster_matches_point_signs(n):
    s = W.gyroid_c4()
    res = n / 2.4
    print(f"raster property: gyroid {n}x{n}", flush=True)
    img = s.render_pixels((-1.2, -1.2), (1.2, 1.2), 0.05, res)
    assert img.shape == (n, n)
    xs, _, _ = oracle.voxel_pts(-1.2, 1.2, res)
    zs, _, _ = oracle.voxel_pts(0.05, 0.05, res)
    gx, gy = np.meshgrid(xs, xs)
    v = s.values(gx.ravel(), gy.ravel(), np.full(gx.size, zs[-1], np.float32))
    assert (img == (v < 0).reshape(n, n)).all()
