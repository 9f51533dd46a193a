"""Pins the oracle's 2D dual contouring (the checker for
libfive_tree_render_slice): the reference's known answers
(libfive/test/contours.cpp:22-76, libfive/test/api.cpp:75-105), closed
and consistently wound contours, vertices on the zero set of linear
fields and segment structure invariants.  CPU only."""
import numpy as np
import pytest

import oracle
from libfive_amd import X, Y, Z
from libfive_amd import workloads as W


def od(shape):
    return oracle.OracleDeck(shape.deck())


def contours(shape, lo=(-1, -1), hi=(1, 1), z=0.0, mf=0.1):
    return od(shape).render_contours(lo, hi, z, mf)


def area(p):
    return 0.5 * float(np.sum(p[:-1, 0] * p[1:, 1] - p[1:, 0] * p[:-1, 1]))


def rect(x0, x1, y0, y1):
    x, y = X(), Y()
    return ((x0 - x).max(x - x1)).max((y0 - y).max(y - y1))


def test_segment_welding_and_accuracy():
    # contours.cpp:22-51: one contour, radius within 0.01
    c = contours(W.circle(0.5))
    assert len(c.contours) == 1
    r = np.linalg.norm(c.polylines[0], axis=1)
    assert r.max() < 0.51 and r.min() > 0.49


def test_api_circle():
    # api.cpp:75-105: render_slice of r = 1 at res 10 over [-2, 2]^2
    c = contours(W.circle(1.0), (-2, -2), (2, 2), 0.0, 0.1)
    assert len(c.contours) == 1
    r2 = np.sum(c.polylines[0].astype(np.float64) ** 2, axis=1)
    assert r2.min() > 0.99 and r2.max() < 1.01


def test_with_ambiguities():
    # contours.cpp:53-64
    x, y = X(), Y()
    t = (((-x).max(x - 1)).max((-y).max(y - 1))).max(-x)
    c = contours(t, (-2, -2), (2, 2))
    assert len(c.contours) == 1


def test_adjacent_rectangles():
    # contours.cpp:66-76
    rects = rect(-1, 0, -1, 1).min(rect(0, 1, -1, 1))
    assert len(contours(rects, (-2, -2), (2, 2)).contours) == 1
    assert len(contours(-rects, (-2, -2), (2, 2)).contours) == 1


def test_closed_and_wound():
    """Every contour of a shape inside the region is closed; filled
    regions are on the left (positive area), and negating the shape
    reverses every contour."""
    s = W.circle(0.5, (0.1, -0.2)).min(W.circle(0.3, (-0.5, 0.5)))
    c = contours(s, (-1, -1), (1, 1), 0.0, 0.05)
    n = od(-s).render_contours((-1, -1), (1, 1), 0.0, 0.05)
    assert len(c.contours) == 2 and len(n.contours) == 2
    for p in c.polylines:
        assert (p[0] == p[-1]).all() and area(p) > 0
    for p in n.polylines:
        assert (p[0] == p[-1]).all() and area(p) < 0


def test_segments_form_cycles():
    """Each vertex starts exactly one segment and ends exactly one."""
    c = contours(W.gyroid_c4(), (-1.2, -1.2), (1.2, 1.2), 0.3, 2.4 / 128)
    starts = np.bincount(c.segs[:, 0], minlength=len(c.verts))
    ends = np.bincount(c.segs[:, 1], minlength=len(c.verts))
    inner = (starts[1:] > 0) | (ends[1:] > 0)
    assert inner.all()
    # contours that touch the region boundary are open; all others closed
    assert (starts[1:] <= 1).all() and (ends[1:] <= 1).all()


def test_linear_field_vertices_on_line():
    """A half-plane: every vertex lies on the line (closed form)."""
    x, y = X(), Y()
    s = 0.6 * x + 0.8 * y - 0.1
    c = contours(s, (-1, -1), (1, 1), 0.0, 0.05)
    v = c.verts[1:].astype(np.float64)
    assert len(v) > 10
    assert np.abs(0.6 * v[:, 0] + 0.8 * v[:, 1] - 0.1).max() < 1e-6


def test_square_corners_are_exact():
    """A square with grid-aligned sides keeps its four corners (rank-2
    vertices where two features meet)."""
    c = contours(rect(-0.5, 0.5, -0.5, 0.5), (-1, -1), (1, 1), 0.0, 0.1)
    assert len(c.contours) == 1
    pts = {tuple(np.round(p, 6)) for p in c.polylines[0]}
    for corner in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)):
        assert corner in pts
