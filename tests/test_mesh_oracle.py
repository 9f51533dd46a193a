"""Pins the oracle's dual-contouring mesher (the checker for the GPU
mesher): lookup tables against the reference's tables, the vertex solve
against linear algebra, and whole meshes against the reference's known
answers (libfive/test/mesh.cpp, xtree.cpp, dual.cpp, api.cpp,
eval_feature.cpp, region.cpp).  CPU only."""
import os

import numpy as np
import pytest

import oracle
from libfive_amd import Shape, X, Y, Z
from libfive_amd import workloads as W

TABLES = os.path.join(os.path.dirname(__file__), "golden", "dc_tables.npz")


def od(shape):
    return oracle.OracleDeck(shape.deck())


def mesh(shape, lo, hi, min_feature, max_err=1e-8):
    lo = [lo] * 3 if np.isscalar(lo) else lo
    hi = [hi] * 3 if np.isscalar(hi) else hi
    return od(shape).render_mesh(lo, hi, min_feature, max_err)


def check_edge_pairs(tris):
    """Every undirected edge is used once in each direction
    (libfive/test/util/mesh_checks.cpp)."""
    seen = {}
    for t in tris:
        for i in range(3):
            a, b = int(t[i]), int(t[(i + 1) % 3])
            key = (min(a, b), max(a, b))
            bit = 1 if a < b else 2
            assert not (seen.get(key, 0) & bit), ("edge used twice", key)
            seen[key] = seen.get(key, 0) | bit
    assert all(v == 3 for v in seen.values())


def tri_normals(verts, tris):
    a, b, c = (verts[tris[:, i]].astype(np.float64) for i in range(3))
    n = np.cross(b - a, c - a)
    return n / np.linalg.norm(n, axis=1, keepdims=True)


# ---- tables ---------------------------------------------------------------

def test_corner_manifold_table_matches_reference():
    g = np.load(TABLES)["corner_manifold"]
    got = np.array([oracle.corner_manifold(i) for i in range(256)])
    assert (got == g.astype(bool)).all()


def test_marching_table_matches_reference():
    g = np.load(TABLES)
    v, e, p = oracle.marching_table()
    assert (v == g["marching_v"]).all()
    assert (e == g["marching_e"]).all()
    assert (p == g["marching_p"]).all()


def test_marching_table_properties():
    """libfive/test/marching.cpp:29-42, plus the dual marching cubes
    invariant: a mask's patches use every edge with a sign change exactly
    once, always directed filled -> empty."""
    v, e, p = oracle.marching_table()
    assert v[0, 0, 0, 0] == -1 and v[255, 0, 0, 0] == -1
    cube_edges = [(a, a ^ (1 << k)) for a in range(8) for k in range(3)
                  if not a & (1 << k)]
    for mask in range(1, 255):
        assert v[mask, 0, 0, 0] != -1
        used = []
        for patch in range(4):
            for k in range(12):
                a, b = (int(q) for q in v[mask, patch, k])
                if a == -1:
                    break
                assert mask & (1 << a) and not mask & (1 << b)
                used.append((min(a, b), max(a, b)))
                assert p[mask, e[a, b]] == patch
        crossing = [ed for ed in cube_edges
                    if bool(mask & (1 << ed[0])) != bool(mask & (1 << ed[1]))]
        assert sorted(used) == sorted(crossing)
    assert sorted(e[e >= 0].tolist()) == list(range(24))


def test_resolution_level_known_answers():
    """libfive/test/region.cpp:160-167"""
    lo, hi = [-1, -1, -1], [1, 1, 1]
    assert oracle.resolution_level(lo, hi, 5) == 0
    assert oracle.resolution_level(lo, hi, 1) == 1
    assert oracle.resolution_level(lo, hi, 1.2) == 1
    assert oracle.resolution_level(lo, hi, 0.9) == 2


# ---- vertex solve ---------------------------------------------------------

def test_jacobi_eigen_against_numpy():
    rng = np.random.default_rng(5)
    for k in range(300):
        n = rng.standard_normal((rng.integers(1, 6), 3))
        if k % 3 == 0:
            n[:, 2] = 0.0                          # rank deficient
        a = n.T @ n
        w, v = oracle.eigen3(a)
        ref = np.linalg.eigvalsh(a)
        scale = max(1.0, np.abs(ref).max())
        assert np.allclose(np.sort(w), ref, atol=1e-12 * scale)
        assert np.allclose(v.T @ v, np.eye(3), atol=1e-12)
        assert np.allclose(v @ np.diag(w) @ v.T, a, atol=1e-12 * scale)


def test_find_vertex_three_planes_meet_at_point():
    """Three independent planes through p: the QEF minimum is p, error 0."""
    rng = np.random.default_rng(6)
    for _ in range(100):
        p = rng.uniform(-1, 1, 3)
        acc = np.zeros(14)
        while True:
            ns = rng.standard_normal((3, 3))
            ns /= np.linalg.norm(ns, axis=1, keepdims=True)
            if np.linalg.eigvalsh(ns.T @ ns).min() > 0.2:
                break
        for nrm in ns:
            pos = p + rng.uniform(-0.1, 0.1) * np.cross(nrm, [0.3, 0.5, 0.7])
            oracle.push_sample(acc, pos, nrm * 2.5, 0.0)
        ata = acc[[4, 5, 6, 5, 7, 8, 6, 8, 9]].reshape(3, 3)
        v, rank, err = oracle.find_vertex(ata, acc[10:13], acc[13], acc[:4])
        assert rank == 3
        assert np.allclose(v, p, atol=1e-10)
        assert abs(err) < 1e-12


def test_find_vertex_truncated_pseudo_inverse():
    """Against numpy: v = c + U D U^T (A^T b - A^T A c), D = 1/w for
    |w| >= 0.1 else 0, error = v'Av - 2v'b + b'b."""
    rng = np.random.default_rng(7)
    for k in range(200):
        m = rng.integers(1, 5)
        ns = rng.standard_normal((m, 3))
        if k % 2:
            ns[:, 1] = 0.01 * ns[:, 0]
        ns /= np.linalg.norm(ns, axis=1, keepdims=True)
        pts = rng.uniform(-1, 1, (m, 3))
        vals = rng.uniform(-0.01, 0.01, m)
        acc = np.zeros(14)
        for nrm, pt, val in zip(ns, pts, vals):
            oracle.push_sample(acc, pt, nrm, val)
        ata = acc[[4, 5, 6, 5, 7, 8, 6, 8, 9]].reshape(3, 3)
        atb, btb = acc[10:13], acc[13]
        w, u = np.linalg.eigh(ata)
        d = np.where(np.abs(w) < 0.1, 0.0, 1.0 / np.where(w == 0, 1, w))
        c = acc[:3] / acc[3]
        want = u @ np.diag(d) @ u.T @ (atb - ata @ c) + c
        v, rank, err = oracle.find_vertex(ata, atb, btb, acc[:4])
        assert rank == int((d != 0).sum())
        assert np.allclose(v, want, atol=1e-9)
        assert err == pytest.approx(v @ ata @ v - 2 * v @ atb + btb, abs=1e-9)


def _np_find_vertex(ata, atb, btb, mass):
    """findVertex (dc_tree.inl:740-813) written with numpy.linalg.eigh:
    truncated pseudo-inverse with |w| < 0.1 dropped, solved about the mass
    point; returns (vertex, count of kept eigenvalues, error)."""
    w, u = np.linalg.eigh(ata)
    keep = np.abs(w) >= 0.1
    d = np.where(keep, 1.0 / np.where(keep, w, 1.0), 0.0)
    c = mass[:3] / mass[3]
    v = u @ np.diag(d) @ u.T @ (atb - ata @ c) + c
    return v, int(keep.sum()), float(v @ ata @ v - 2 * v @ atb + btb)


# The QEF tolerance: the product's Jacobi solve and LAPACK agree to a few
# ulp of the largest eigenvalue; vertices are compared to 1e-9 of the
# problem's scale (|c| + |A^+ A^T b| stay below 10 here), errors to 1e-9.
QEF_TOL = 1e-9


@pytest.mark.parametrize("kind", ["full", "deficient", "near_cutoff", "tiny", "samples"])
def test_find_vertex_many_random_qefs(kind):
    """3000 QEFs per kind against the numpy restatement: full-rank and
    rank-deficient systems, eigenvalues placed 1e-7 either side of the 0.1
    cutoff (the rank must come out as numpy's), near-zero matrices, and
    systems accumulated by push_sample from random planes."""
    rng = np.random.default_rng({"full": 11, "deficient": 12, "near_cutoff": 13,
                                 "tiny": 14, "samples": 15}[kind])
    for k in range(3000):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if kind == "full":
            w = rng.uniform(0.2, 6.0, 3)
        elif kind == "deficient":
            w = rng.uniform(0.2, 6.0, 3)
            w[rng.integers(0, 3, rng.integers(1, 3))] = 0.0
        elif kind == "near_cutoff":
            w = rng.uniform(0.2, 6.0, 3)
            w[k % 3] = 0.1 * (1 + (1e-7 if k % 2 else -1e-7))
        elif kind == "tiny":
            w = rng.uniform(0.0, 1e-6, 3)
        if kind == "samples":
            m = int(rng.integers(1, 7))
            acc = np.zeros(14)
            for _ in range(m):
                oracle.push_sample(acc, rng.uniform(-1, 1, 3), rng.standard_normal(3) * 3,
                                   rng.uniform(-0.05, 0.05))
            ata = acc[[4, 5, 6, 5, 7, 8, 6, 8, 9]].reshape(3, 3)
            atb, btb, mass = acc[10:13], acc[13], acc[:4]
        else:
            ata = q @ np.diag(w) @ q.T
            ata = (ata + ata.T) / 2
            atb = rng.uniform(-3, 3, 3)
            btb = float(rng.uniform(0, 5))
            mass = np.append(rng.uniform(-2, 2, 3), rng.integers(1, 9)).astype(np.float64)
        v, rank, err = oracle.find_vertex(ata, atb, btb, mass)
        wv, wrank, werr = _np_find_vertex(ata, atb, btb, mass)
        if kind == "samples":
            # sample sums can put an eigenvalue within rounding of the cutoff
            w_np = np.linalg.eigvalsh(ata)
            if np.min(np.abs(np.abs(w_np) - 0.1)) < 1e-12:
                continue
        assert rank == wrank, (kind, k)
        assert np.allclose(v, wv, rtol=0, atol=QEF_TOL), (kind, k, v, wv)
        assert err == pytest.approx(werr, abs=QEF_TOL * max(1.0, abs(werr))), (kind, k)


def test_push_sample_many_random_against_numpy():
    """Intersection::push (intersection.hpp:32-59) restated with numpy on
    random samples, including zero, tiny and non-finite derivatives."""
    rng = np.random.default_rng(16)
    for k in range(2000):
        acc = np.zeros(14)
        want = np.zeros(14)
        for _ in range(int(rng.integers(1, 6))):
            pos = rng.uniform(-2, 2, 3)
            d = rng.standard_normal(3) * 10 ** rng.uniform(-3, 3)
            if k % 7 == 0:
                d = d * 1e-14
            if k % 11 == 0:
                d[int(rng.integers(0, 3))] = np.inf
            val = rng.uniform(-1, 1)
            oracle.push_sample(acc, pos, d, val)
            want[:3] += pos
            want[3] += 1
            nrm = np.sqrt(d @ d)
            with np.errstate(invalid="ignore"):
                dn, vn = d / nrm, val / nrm
            if nrm <= 1e-12 or not np.isfinite(dn).all():
                continue
            want[4:10] += [dn[0] * dn[0], dn[0] * dn[1], dn[0] * dn[2],
                           dn[1] * dn[1], dn[1] * dn[2], dn[2] * dn[2]]
            b = dn @ pos - vn
            want[10:13] += dn * b
            want[13] += b * b
        assert np.allclose(acc, want, rtol=1e-13, atol=1e-13), k


def test_compensated_push_is_exact_to_float_pair_rounding():
    """push_sample_f2 (the precision form's accumulation) against exact
    arithmetic: every sum is carried by its float pair to within a few
    units of 2^-44 of the terms' magnitude, where a plain float sum drifts
    by 2^-24 per term.  The reference values come from fractions (exact
    rational arithmetic) on the float-rounded normals, so a dropped error
    term or a wrong product fails."""
    from fractions import Fraction as Fr
    rng = np.random.default_rng(17)
    for k in range(60):
        st = np.zeros(28, np.float32)
        m = int(rng.integers(1, 40))
        exact = [Fr(0)] * 14
        for _ in range(m):
            pos = rng.uniform(-2, 2, 3).astype(np.float32)
            d = (rng.standard_normal(3) * 10 ** rng.uniform(-2, 2)).astype(np.float32)
            val = np.float32(rng.uniform(-0.1, 0.1))
            oracle.push_sample_f2(st, pos, d, val)
            # the normal and scaled value as the routine rounds them
            nrm = np.sqrt(np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]),
                          dtype=np.float32)
            n = (d / nrm).astype(np.float32)
            v = np.float32(val / nrm)
            for q in range(3):
                exact[q] += Fr(float(pos[q]))
            exact[3] += 1
            nf = [Fr(float(x)) for x in n]
            pf = [Fr(float(x)) for x in pos]
            b = nf[0] * pf[0] + nf[1] * pf[1] + nf[2] * pf[2] - Fr(float(v))
            for q, (i, j) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
                exact[4 + q] += nf[i] * nf[j]
            for j in range(3):
                exact[10 + j] += nf[j] * b
            exact[13] += b * b
        got = oracle.f2_widen(st)
        for q in range(14):
            scale = max(1.0, m * 16.0)
            assert abs(Fr(got[q]) - exact[q]) <= Fr(scale) * Fr(2) ** -42, (k, q)


def test_compensated_sum_of_accumulators():
    """f2_add_acc sums widened accumulators (entries 4..13) as float pairs:
    within 2^-42 relative of the exact sum, mass-point entries untouched."""
    from fractions import Fraction as Fr
    rng = np.random.default_rng(18)
    for k in range(200):
        st = np.zeros(28, np.float32)
        accs = rng.standard_normal((int(rng.integers(1, 13)), 14)) * 10 ** rng.uniform(-3, 3)
        for a in accs:
            oracle.f2_add_acc(st, a)
        got = oracle.f2_widen(st)
        assert (got[:4] == 0).all()
        for q in range(4, 14):
            ex = sum(Fr(float(a[q])) for a in accs)
            mag = sum(abs(Fr(float(a[q]))) for a in accs)
            assert abs(Fr(got[q]) - ex) <= mag * Fr(2) ** -42, (k, q)


def test_push_sample_normalises_and_skips_bad_normals():
    """libfive/include/libfive/render/brep/dc/intersection.hpp:32-59"""
    acc = np.zeros(14)
    oracle.push_sample(acc, [1.0, 2.0, 3.0], [0.0, 0.0, 2.0], 0.5)
    assert acc[:4].tolist() == [1.0, 2.0, 3.0, 1.0]
    assert acc[9] == 1.0 and acc[4:9].tolist() == [0.0] * 5
    b = 3.0 - 0.25
    assert acc[12] == b and acc[13] == b * b
    oracle.push_sample(acc, [1.0, 1.0, 1.0], [0.0, 0.0, 1e-13], 0.0)
    oracle.push_sample(acc, [1.0, 1.0, 1.0], [np.inf, 0.0, 0.0], 0.0)
    assert acc[3] == 3.0                    # mass point still counts them
    assert acc[9] == 1.0 and acc[13] == b * b


# ---- feature evaluator (libfive/test/eval_feature.cpp:19-250) ------------

def test_is_inside_known_answers():
    x, y, z = X(), Y(), Z()
    a = od(x)
    assert a.is_inside([0, 0, 0]) and a.is_inside([-1, 0, 0])
    assert not a.is_inside([1, 0, 0])
    b = od(x.min(-x))
    assert all(b.is_inside(p) for p in ([0, 0, 0], [1, 0, 0], [-1, 0, 0]))
    c = od(x.max(-x))
    assert not any(c.is_inside(p) for p in ([0, 0, 0], [1, 0, 0], [-1, 0, 0]))
    d = od(x.min(-x).min(y.min(-y)))
    assert d.is_inside([0, 0, 0])
    t = od(W.sphere(0.5, (0, 0, 1)).min(W.box((-1, -1, -1), (1, 1, 1))))
    for p in ([0, 0, 0], [0.5, 0, 1], [-0.5, 0, 1], [0, 0.5, 1], [0, -0.5, 1],
              [0, 0, 1.5]):
        assert t.is_inside(p), p
    assert not t.is_inside([0, 0, 2])
    e = od(W.box((-50, -50, 0), (50, 50, 10)).max(
        W.cylinder(10, 100).max(-W.cylinder(5, 100))))
    assert not e.is_inside([0, 0, 0])
    f = od(W.box((-1, 0, -1), (1, 2, -0.2)).max(
        -W.box((0.1, 0, -1), (1, 2, -0.1))))
    assert not f.is_inside([0.625, 0.0, -1.0])
    g = od(x.max(y).max(-(x.max(y))))
    assert not g.is_inside([0.0, 0.0, -0.5])


def rectangle(xmin, xmax, ymin, ymax):
    x = 1.0 * X() + 0.0 * Y() + 0.0 * Z() + 0.0
    y = 0.0 * X() + 1.0 * Y() + 0.0 * Z() + 0.0
    return (xmin - x).max(x - xmax).max((ymin - y).max(y - ymax))


def test_features_known_answers():
    x, y, z = X(), Y(), Z()

    def fs(shape, p):
        f = od(shape).features(p)
        return sorted(tuple(float(c) for c in r) for r in f)

    assert fs(x, [0, 0, 0]) == [(1, 0, 0)]
    assert fs(x.min(-x), [0, 0, 0]) == [(-1, 0, 0), (1, 0, 0)]
    assert fs(x.max(-x), [0, 0, 0]) == [(-1, 0, 0), (1, 0, 0)]
    assert fs(x.min(y.min(z)), [0, 0, 0]) == [(0, 0, 1), (0, 1, 0), (1, 0, 0)]
    assert len(fs(rectangle(-1, 0, -1, 1), [-0.5, -1, 0])) == 1
    nested = (-x).max(x - 1).max((-y).max(y - 1)).max(-x)
    assert len(fs(nested, [0, 0.2, 0])) == 1
    assert len(fs(x.max(x), [0, 0.2, 0])) == 1
    assert len(fs(x.max(x.max(x)), [0, 0.2, 0])) == 1
    assert len(fs(x.max(x).max(x.max(x)), [0, 0.2, 0])) == 1
    assert len(fs((z - 6).max(z + -6), [0, 0, 6])) == 1
    assert len(fs(x.max(x + y * 1e-8), [0, 0, 0])) == 1
    yy = (-(y + 1.5)).max(y - 1.5)
    xx = (-(x + 1.5)).max(x - 1.5)
    assert len(fs(yy, [0, 0, 0])) == 2
    assert len(fs(xx, [0, 0, 0])) == 2
    assert len(fs(xx.max(yy), [0, 0, 0])) == 4


# ---- whole meshes (libfive/test/mesh.cpp, api.cpp, xtree.cpp, dual.cpp) ---

def cube15():
    x, y, z = X(), Y(), Z()
    return ((-(x + 1.5)).max(x - 1.5).max((-(y + 1.5)).max(y - 1.5))
            .max((-(z + 1.5)).max(z - 1.5)))


def test_cube_face_count():
    m = mesh(cube15(), -3.0, 3.0, 0.15)
    assert m.tris.shape[0] == 12
    assert m.verts.shape[0] == 9
    assert np.allclose(np.abs(m.verts[1:]), 1.5, atol=1e-6)


def test_rectangular_prism_face_count():
    x, y, z = X(), Y(), Z()
    t = (-x).max(x - 4).max((-y).max(y - 1)).max((-z).max(z - 0.25))
    m = mesh(t, [-1, -1, -1], [5, 2, 1.25], 0.125)
    assert m.verts.shape[0] == 9
    assert m.tris.shape[0] == 12


def test_sphere_normals_point_outwards():
    m = mesh(W.sphere(0.5), -1.0, 1.0, 0.2)
    n = tri_normals(m.verts, m.tris)
    c = m.verts[m.tris].astype(np.float64).sum(axis=1)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    assert (np.einsum("ij,ij->i", n, c) > 0.9).all()


def _sphere_on_box():
    return W.sphere(0.7, (0, 0, 0.1)).min(W.box((-1, -1, -1), (1, 1, 0.1)))


def test_no_triangles_that_are_lines():
    m = mesh(_sphere_on_box(), -10.0, 10.0, 0.25)
    t = m.tris
    assert ((t[:, 0] != t[:, 1]) & (t[:, 0] != t[:, 2]) &
            (t[:, 1] != t[:, 2])).all()


def test_no_flipped_triangles_on_top_face():
    m = mesh(_sphere_on_box(), -10.0, 10.0, 0.25)
    top = (np.abs(m.verts[m.tris][:, :, 2] - 0.1) < 1e-3).all(axis=1)
    assert top.any()
    n = tri_normals(m.verts, m.tris[top])
    assert np.allclose(n, [0, 0, 1], atol=0.01)


@pytest.mark.parametrize("case", ["sphere_coarse", "sphere_fine", "c1",
                                  "gyroid"])
def test_edge_pairing(case):
    if case == "sphere_coarse":
        m = mesh(W.sphere(0.5), -1.0, 1.0, 1.1)
    elif case == "sphere_fine":
        m = mesh(W.sphere(0.5), -1.0, 1.0, 0.05)
    elif case == "c1":
        m = mesh(W.csg_c1(), -2.0, 2.0, 1 / 16)
    else:
        m = mesh(W.gyroid_c4(), -1.2, 1.2, 2.4 / 32)
    if m.tris.shape[0]:
        check_edge_pairs(m.tris)


def test_c_api_sphere_radius():
    """libfive/test/api.cpp:107-146: x^2+y^2+z^2-1 over [-2,2]^3 at res 10."""
    x, y, z = X(), Y(), Z()
    s = x.square() + y.square() + z.square() - 1.0
    m = mesh(s, -2.0, 2.0, float(np.float32(1) / np.float32(10)))
    r2 = (m.verts[m.tris].astype(np.float64) ** 2).sum(axis=2)
    assert r2.min() > 0.99 and r2.max() < 1.01


def test_dual_walk_sphere_vertex_norms():
    """libfive/test/dual.cpp:130-138 (default settings: min_feature 0.1)."""
    m = mesh(W.sphere(0.5), -1.0, 1.0, 0.1)
    r = np.linalg.norm(m.verts[1:].astype(np.float64), axis=1)
    assert r.min() > 0.49 and r.max() < 0.51


@pytest.mark.parametrize("case,err", [
    ("sliced_box", 0.001), ("sliced_box_2", 0.001),
    ("sphere_cutout", 0.01), ("nested_rings", 0.01)])
def test_leaf_vertices_lie_on_surface(case, err):
    """libfive/test/xtree.cpp:150-238: every vertex of every ambiguous leaf
    evaluates to ~0."""
    x, y, z = X(), Y(), Z()
    if case == "sliced_box":
        s = W.box((0, 0, 0), (1, 1, 1)).max(x + y + z - 1.3)
        lo, hi, mf = -2.0, 2.0, 0.1
    elif case == "sliced_box_2":
        s = W.box((0, 0, 0), (1, 1, 1)).max(x + y + z - 1.2)
        lo, hi, mf = -10.0, 10.0, 0.1
    elif case == "sphere_cutout":
        s = W.sphere(1.0).max(-W.circle(0.5))
        lo, hi, mf = -5.0, 5.0, float(np.float32(1) / np.float32(9))
    else:
        s = (-z).max(W.circle(1.0).max(-W.circle(0.5))).max(z - 1)
        lo, hi, mf = -2.0, 2.0, 0.2
    d = od(s)
    m = d.render_mesh([lo] * 3, [hi] * 3, mf)
    v = m.leaf_verts.astype(np.float32)
    assert v.shape[0] > 0
    vals = d.values(v[:, 0], v[:, 1], v[:, 2])
    assert np.abs(vals).max() <= err


def test_mesh_is_deterministic_and_keys_unique():
    d = od(W.csg_c1())
    a = d.render_mesh([-2] * 3, [2] * 3, 1 / 16)
    b = d.render_mesh([-2] * 3, [2] * 3, 1 / 16)
    assert (a.verts == b.verts).all() and (a.tris == b.tris).all()
    keys = {tuple(k) for k in a.keys[1:].tolist()}
    assert len(keys) == a.verts.shape[0] - 1


def test_fp64_error_oracle_plane_exact():
    """FP64 error oracle on a linear field: every vertex lies on the plane
    to double precision (a dropped term or wrong partial would move it),
    and the float mesher's vertices are within float rounding of it."""
    s = X() + Y() * 0.5 + Z() * 0.25 - 0.3
    d = oracle.OracleDeck(s.deck())
    lo, hi = [-1.0] * 3, [1.0] * 3
    # max_err < 0 keeps every leaf (a plane would collapse to the root)
    r = d.render_mesh(lo, hi, 0.1, max_err=-1.0, fp64=True)
    v = r.verts64[1:]
    assert v.shape[0] > 100
    f = v[:, 0] + v[:, 1] * 0.5 + v[:, 2] * 0.25 - np.float64(np.float32(0.3))
    assert np.abs(f).max() < 1e-12
    o = d.render_mesh(lo, hi, 0.1, max_err=-1.0)
    err, matched, n = oracle.vertex_error(o.verts, o.keys, r)
    assert matched == n and err < 1e-6


def test_fp64_error_oracle_gyroid_within_1e5():
    """Float (plain-form) mesh of the thickened gyroid against the FP64
    error oracle: every vertex matched by key, within 1e-5."""
    s = W.gyroid_c4()
    d = oracle.OracleDeck(s.deck())
    lo = [float(np.float32(-1.2))] * 3
    hi = [float(np.float32(1.2))] * 3
    o = d.render_mesh(lo, hi, 2.4 / 32, threads=4)
    r = d.render_mesh(lo, hi, 2.4 / 32, threads=4, fp64=True)
    err, matched, n = oracle.vertex_error(o.verts, o.keys, r)
    assert matched == n and err < 1e-5
