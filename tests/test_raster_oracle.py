"""Pins the oracle's Z-slice occupancy raster (the checker for
libfive_tree_render_pixels): the voxel grid against the reference's
known answers (libfive/test/voxels.cpp:16-82), whole images against the
reference's expected depth images (libfive/test/heightmap.cpp:52-181), and
every image against brute force (the sign of the shape at each voxel
centre: interval bounds only ever decide blocks whose points all share
one sign).  CPU only."""
import numpy as np
import pytest

import oracle
from libfive_amd import X, Y, Z
from libfive_amd import workloads as W


def od(shape):
    return oracle.OracleDeck(shape.deck())


def mask(rows):
    """Rows of the reference's expected depth image ('#' = finite depth),
    top row first as written in the test = lowest y."""
    return np.array([[c == "#" for c in r] for r in rows])


# heightmap.cpp:56-67 (circle(1) at 5 voxels per unit over [-1, 1]^2)
CIRCLE = mask([
    "...####...",
    ".########.",
    ".########.",
    "##########",
    "##########",
    "##########",
    "##########",
    ".########.",
    ".########.",
    "...####...",
])

# heightmap.cpp:118-131 (max(circle(1), Y))
ORIENT_Y = mask([
    "...####...",
    ".########.",
    ".########.",
    "##########",
    "##########",
] + ["." * 10] * 5)

# heightmap.cpp:141-154 (max(circle(1), X))
ORIENT_X = mask([
    "...##.....",
    ".####.....",
    ".####.....",
    "#####.....",
    "#####.....",
    "#####.....",
    "#####.....",
    ".####.....",
    ".####.....",
    "...##.....",
])


def brute(shape, xy_lo, xy_hi, z, res):
    xs, _, _ = oracle.voxel_pts(xy_lo[0], xy_hi[0], res)
    ys, _, _ = oracle.voxel_pts(xy_lo[1], xy_hi[1], res)
    zs, _, _ = oracle.voxel_pts(z, z, res)
    gx, gy = np.meshgrid(xs, ys)
    v = od(shape).values(gx.ravel(), gy.ravel(), np.full(gx.size, zs[-1]))
    return (v < 0).reshape(gx.shape)


def test_voxel_grid_known_answers():
    # voxels.cpp:18-28 exact values
    assert oracle.voxel_pts(0, 1, 1)[0][0] == 0.5
    assert oracle.voxel_pts(-0.5, 0.5, 3)[0][1] == 0
    assert oracle.voxel_pts(-1, 1, 0)[0][0] == 0
    # voxels.cpp:30-37 expanding interval
    pts, lower, upper = oracle.voxel_pts(0, np.float32(1.2), 1)
    assert lower == pytest.approx(-0.4, abs=1e-6)
    assert upper == pytest.approx(1.6, abs=1e-6)
    assert pts[0] == pytest.approx(0.1, abs=1e-6)
    assert pts[1] == pytest.approx(1.1, abs=1e-6)
    # voxels.cpp:42-64 sizes
    assert len(oracle.voxel_pts(0, 2, 1)[0]) == 2
    for hi, n in ((10, 100), (5, 50), (2.5, 25)):
        assert len(oracle.voxel_pts(0, hi, 10)[0]) == n
    for hi, n in ((1.1, 2), (2.1, 3), (3.1, 4)):
        assert len(oracle.voxel_pts(0, np.float32(hi), 1)[0]) == n


@pytest.mark.parametrize("z", [0.0, 1.0])
def test_circle_matches_reference_image(z):
    # heightmap.cpp:52-109 (empty Z at z = 0, and the slice at z = 1)
    img = od(W.circle(1.0)).render_pixels((-1, -1), (1, 1), z, 5)
    assert img.shape == (10, 10)
    assert (img == CIRCLE).all()


def test_orientation_matches_reference_images():
    c = W.circle(1.0)
    assert (od(c.max(Y())).render_pixels((-1, -1), (1, 1), 0, 5)
            == ORIENT_Y).all()
    assert (od(c.max(X())).render_pixels((-1, -1), (1, 1), 0, 5)
            == ORIENT_X).all()


def test_image_shape():
    # heightmap.cpp:163-181: rows follow Y, columns follow X
    c = od(W.circle(1.0))
    assert c.render_pixels((0, -1), (1, 1), 0, 5).shape == (10, 5)
    assert c.render_pixels((-1, 0), (1, 1), 0, 5).shape == (5, 10)


CASES = [
    ("sphere_api", lambda: W.sphere(1.0), (-2, -2), (2, 2), 0.0, 10),
    ("circle_ragged", lambda: W.circle(0.7, (0.1, -0.2)), (-1, -0.9),
     (1.3, 0.77), 0.0, 37),
    ("csg_c1", W.csg_c1, (-1.5, -1.5), (1.5, 1.5), 0.3, 64),
    ("gyroid", lambda: W.gyroid_c4(), (-1.2, -1.2), (1.2, 1.2), 0.1, 100),
    ("sphere_union", lambda: W.sphere_union(60), (-1, -1), (1, 1), 0.05, 90),
]


@pytest.mark.parametrize("name,mk,lo,hi,z,res", CASES, ids=[c[0] for c in CASES])
def test_raster_equals_brute_force(name, mk, lo, hi, z, res):
    s = mk()
    img = od(s).render_pixels(lo, hi, z, res)
    ref = brute(s, lo, hi, z, res)
    assert img.shape == ref.shape
    assert (img == ref).all()
    assert img.any() and not img.all()


def test_empty_and_full():
    far = od(W.sphere(0.5, (10, 10, 10)))
    assert not far.render_pixels((-1, -1), (1, 1), 0, 50).any()
    big = od(W.sphere(10.0))
    assert big.render_pixels((-1, -1), (1, 1), 0, 50).all()
