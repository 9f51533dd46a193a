"""GPU dual-contouring mesher against the oracle mesher, through the C ABI.

The plain form must reproduce the oracle's octree and mesh exactly: the
same cell states per level, the same pruned tape lengths, the same vertex
set (keyed by owning cell and patch) with bit-identical positions, and the
same triangles as ordered key triples.  Vertex numbering is free (the
reference numbers vertices in its walk order), so meshes are compared
through the keys."""
import numpy as np
import pytest

import oracle
from libfive_amd import Shape, X, Y, Z
from libfive_amd import workloads as W

from helpers import wide_registers

pytestmark = pytest.mark.gpu


def canonical(verts, tris, keys):
    kk = [tuple(k) for k in keys.tolist()]
    vmap = {kk[i]: verts[i] for i in range(1, len(kk))}
    tri = sorted(tuple(kk[i] for i in t) for t in tris.tolist())
    return vmap, tri


def compare(shape, lo, hi, mf, max_err=1e-8, threads=1, precision=0):
    # The C API takes the region as floats (libfive_region3), so the oracle
    # gets the same float-rounded bounds.
    lo3 = [float(np.float32(lo))] * 3
    hi3 = [float(np.float32(hi))] * 3
    print(f"mesh parity: lo={lo} hi={hi} min_feature={mf} precision={precision}",
          flush=True)
    g = shape.render_mesh(lo3, hi3, min_feature=mf, max_err=max_err,
                          precision=precision, details=True)
    o = oracle.OracleDeck(shape.deck()).render_mesh(lo3, hi3, mf, max_err,
                                                    threads=threads,
                                                    qef_f32c=precision == 1)
    assert g.level_counts.shape == o.level_counts.shape
    assert (g.level_counts == o.level_counts).all()
    # identical interval / leaf classification, cell by cell
    gc = {tuple(c[:4]): int(c[4]) for c in g.cells.tolist()}
    oc = {tuple(c[:4]): int(c[4]) for c in o.cells.tolist()}
    assert gc == oc
    assert sorted(g.tape_lengths.tolist()) == sorted(o.tape_lengths.tolist())
    assert g.verts.shape == o.verts.shape
    assert g.tris.shape == o.tris.shape
    gv, gt = canonical(g.verts, g.tris, g.keys)
    ov, ot = canonical(o.verts, o.tris, o.keys)
    assert gv.keys() == ov.keys()
    for k, v in gv.items():
        assert v.view(np.uint32).tolist() == ov[k].view(np.uint32).tolist(), k
    assert gt == ot
    return g


def check_edge_pairs(tris):
    seen = {}
    for t in tris.tolist():
        for i in range(3):
            a, b = t[i], t[(i + 1) % 3]
            key = (min(a, b), max(a, b))
            bit = 1 if a < b else 2
            assert not (seen.get(key, 0) & bit)
            seen[key] = seen.get(key, 0) | bit
    assert all(v == 3 for v in seen.values())


@pytest.mark.parametrize("mf", [1.1, 0.2, 0.1, 0.05])
def test_sphere_matches_oracle(gpu, mf):
    g = compare(W.sphere(0.5), -1.0, 1.0, mf)
    if g.tris.shape[0]:
        check_edge_pairs(g.tris)


@pytest.mark.parametrize("res", [32, 64])
def test_gyroid_matches_oracle(gpu, res):
    g = compare(W.gyroid_c4(), -1.2, 1.2, 2.4 / res)
    check_edge_pairs(g.tris)


def test_sphere_union_matches_oracle(gpu):
    compare(W.sphere_union(n=60), -1.2, 1.2, 2.4 / 64)


def test_c3_sphere_union_depth8_matches_oracle(gpu):
    """The full interval-pruning workload: the seeded 1000-sphere min-union
    (about 11k clauses) meshed to depth 8 over [-1.2, 1.2]^3.  Per-cell
    states at every level, the multiset of pruned tape lengths (base tape
    of 10 999 clauses, pruned tapes down to a few clauses), keyed vertices
    and triangles all equal the oracle mesher's."""
    s = W.sphere_union()
    g = compare(s, -1.2, 1.2, 2.4 / 256, threads=16)
    assert g.level_counts.shape[0] == 9
    assert max(g.tape_lengths.tolist()) > 5000
    assert min(g.tape_lengths.tolist()) < 20


@pytest.mark.parametrize("wide", [-1, 0])
def test_long_tape_sphere_union_matches_oracle(gpu, wide):
    """A 2500-sphere union (27 499 clauses), with the wide interval kernel
    (the default for tapes this long) and with the one-wave walk, which
    then keeps its per-clause work arrays in a global workspace instead of
    LDS: the same octree and mesh as the oracle."""
    from libfive_amd import ffi
    s = W.sphere_union(2500)
    assert s.deck()[0].shape[0] > 25000
    ffi.lib().libfive_autotune_force(b"interval_wide", wide)
    try:
        g = compare(s, -1.2, 1.2, 2.4 / 32, threads=16)
    finally:
        ffi.lib().libfive_autotune_force(b"interval_wide", -1)
    assert max(g.tape_lengths.tolist()) > 25000


PRECISION_CASES = [
    ("sphere", lambda: W.sphere(0.5), -1.0, 1.0, 0.1),
    ("gyroid64", W.gyroid_c4, -1.2, 1.2, 2.4 / 64),
    ("csg_c1", W.csg_c1, -2.0, 2.0, 0.25),
    ("sphere_union", lambda: W.sphere_union(n=60), -1.2, 1.2, 2.4 / 64),
    ("sliced_box", lambda: W.box((0, 0, 0), (1, 1, 1)).max(X() + Y() + Z() - 1.3),
     -2.0, 2.0, 0.1),
]


@pytest.mark.parametrize("name,mk,lo,hi,mf", PRECISION_CASES,
                         ids=[c[0] for c in PRECISION_CASES])
def test_precision_form_matches_oracle(gpu, name, mk, lo, hi, mf):
    """precision = 1 (compensated-FP32 QEF sums, FP64 eigensolve): the same
    octree, vertices and triangles as the oracle's precision form."""
    compare(mk(), lo, hi, mf, precision=1)


def test_precision_form_vertex_error(gpu):
    """Both QEF forms against the FP64 error oracle on the gyroid at 128^3:
    vertices within 1e-5 apart from rank flips, which the compensated sums
    may cause where an eigenvalue sits at the 0.1 cutoff."""
    s = W.gyroid_c4()
    lo3, hi3 = [float(np.float32(-1.2))] * 3, [float(np.float32(1.2))] * 3
    mf = float(np.float32(1) / np.float32(128 / 2.4))
    r64 = oracle.OracleDeck(s.deck()).render_mesh(lo3, hi3, mf, threads=16, fp64=True)
    out = {}
    for p in (0, 1):
        g = s.render_mesh(lo3, hi3, min_feature=mf, precision=p, details=True)
        d, matched, n = oracle.vertex_errors(g.verts, g.keys, r64)
        out[p] = d
        print(f"precision {p}: {matched}/{n} matched, max {d.max():.3g}, "
              f"99.99% {np.quantile(d, 0.9999):.3g}, median {np.median(d):.3g}",
              flush=True)
        assert matched > 0.999 * n
        assert np.quantile(d, 0.9999) < 1e-5
    assert np.median(out[1]) < 1e-6


def test_large_trig_arguments_match_oracle(gpu):
    """sin arguments beyond 2^17 inside the region: the mesher's call-free
    leaf kernels hand such parents to the general per-cell kernel."""
    s = W.sphere(0.8) + (X() * 3.0e5).sin() * 0.001
    compare(s, -1.0, 1.0, 2.0 / 32)


def test_uncommon_opcodes_match_oracle(gpu):
    """A tape with opcodes outside the common set runs the leaf kernels
    that call out to the general opcode rules."""
    s = (W.sphere(0.8) + Shape.unary("exp", X()) * 0.05
         + Shape.binary("atan2", Y(), Z() + 2.0) * 0.05)
    compare(s, -1.0, 1.0, 2.0 / 32)


def test_wide_register_file_matches_oracle(gpu):
    """The 809-clause random expression needs 37 registers, so the mesher's
    kernels run with two register banks."""
    s = W.random_dag() - 0.3
    assert s.tape_info()[1] > 32
    compare(s, -1.0, 1.0, 2.0 / 16)


def test_four_register_banks_match_oracle(gpu):
    s = wide_registers(70)
    assert s.tape_info()[1] > 64
    compare(s, -1.0, 1.0, 2.0 / 16)


def test_eight_register_banks_match_oracle(gpu):
    """200 registers: the kernels' register files spill out of the VGPRs
    (eight banks), with the same mesh."""
    s = wide_registers(195)
    assert s.tape_info()[1] > 128
    compare(s, -1.0, 1.0, 2.0 / 16)


def test_more_than_256_registers_is_refused(gpu):
    s = wide_registers(260)
    assert s.tape_info()[1] > 256
    with pytest.raises(Exception, match="256"):
        s.render_mesh([-1.0] * 3, [1.0] * 3, min_feature=0.25)


def test_more_than_20_levels_is_refused(gpu):
    with pytest.raises(Exception, match="20 levels"):
        W.sphere(0.5).render_mesh([-1.0] * 3, [1.0] * 3, min_feature=2.0 / 2 ** 21)


def test_sliced_box_matches_oracle(gpu):
    """Box faces on the grid: exact zeros on tied MIN/MAX go through the
    feature evaluation."""
    s = W.box((0, 0, 0), (1, 1, 1)).max(X() + Y() + Z() - 1.3)
    compare(s, -2.0, 2.0, 0.1)


@pytest.mark.parametrize("res", [8, 16, 32])
def test_c1_csg_matches_oracle(gpu, res):
    """The CPU-runnable reference configuration (sphere union box minus a
    cylinder over [-2, 2]^3) at resolution 16 and neighbours."""
    g = compare(W.csg_c1(), -2.0, 2.0, float(np.float32(1) / np.float32(res)))
    check_edge_pairs(g.tris)


def cube15():
    x, y, z = X(), Y(), Z()
    return ((-(x + 1.5)).max(x - 1.5).max((-(y + 1.5)).max(y - 1.5))
            .max((-(z + 1.5)).max(z - 1.5)))


def test_cube_face_count(gpu):
    """libfive/test/mesh.cpp:70-89 on the GPU."""
    g = compare(cube15(), -3.0, 3.0, 0.15)
    assert g.tris.shape[0] == 12 and g.verts.shape[0] == 9


def test_rectangular_prism_face_count(gpu):
    """libfive/test/mesh.cpp:91-102 (its region is not a cube)."""
    x, y, z = X(), Y(), Z()
    t = (-x).max(x - 4).max((-y).max(y - 1)).max((-z).max(z - 0.25))
    verts, tris = t.render_mesh([-1, -1, -1], [5, 2, 1.25], min_feature=0.125)
    assert verts.shape[0] == 9 and tris.shape[0] == 12


def test_box_corner_features_match_oracle(gpu):
    """Boxes sharing corners and edges on the grid: many-way ties."""
    B = W.box
    s = (B((-1, -1, -1), (0, 0, 0)).min(B((0, 0, 0), (1, 1, 1)))
         .min(B((-.5, -.5, -.5), (.5, .5, .5)).max(-B((0, -1, -1), (1, 0, 1)))))
    compare(s, -2.0, 2.0, 0.125)
    t = B((-1, -1, -1), (1, 1, 1)).max(-(B((-.5, -2, -.5), (.5, 2, .5)).max(-W.sphere(0.25))))
    compare(t, -2.0, 2.0, 0.125)


@pytest.mark.parametrize("i", range(4))
def test_random_csg_matches_oracle(gpu, i):
    """Seeded random CSG models of the many-model workload."""
    models = W.csg_batch(4, seed=W.SEED_BASE + 40)
    shape, _ = models[i]
    compare(shape, -1.0, 1.0, 2.0 / 32)


@pytest.mark.parametrize("case", ["sphere_cutout", "nested_rings", "sphere_on_box"])
def test_feature_shapes_match_oracle(gpu, case):
    x, y, z = X(), Y(), Z()
    if case == "sphere_cutout":
        compare(W.sphere(1.0).max(-W.circle(0.5)), -5.0, 5.0,
                float(np.float32(1) / np.float32(9)))
    elif case == "nested_rings":
        compare((-z).max(W.circle(1.0).max(-W.circle(0.5))).max(z - 1),
                -2.0, 2.0, 0.2)
    else:
        g = compare(W.sphere(0.7, (0, 0, 0.1)).min(W.box((-1, -1, -1), (1, 1, 0.1))),
                    -10.0, 10.0, 0.25)
        t = g.tris
        assert ((t[:, 0] != t[:, 1]) & (t[:, 0] != t[:, 2]) & (t[:, 1] != t[:, 2])).all()


def test_c_api_unit_sphere(gpu):
    """libfive/test/api.cpp:107-146 through libfive_tree_render_mesh."""
    s = X().square() + Y().square() + Z().square() - 1.0
    verts, tris = s.render_mesh([-2] * 3, [2] * 3, res=10)
    r2 = (verts[tris].astype(np.float64) ** 2).sum(axis=2)
    assert r2.min() > 0.99 and r2.max() < 1.01
    check_edge_pairs(tris)


def test_empty_region_gives_dummy_vertex_only(gpu):
    verts, tris = W.sphere(0.5).render_mesh([2] * 3, [3] * 3, res=4)
    assert verts.shape == (1, 3) and tris.shape == (0, 3)


def test_root_leaf(gpu):
    """min_feature larger than the region: the root is evaluated as a leaf."""
    compare(W.sphere(0.5), -1.0, 1.0, 5.0)


def test_batch_matches_single_renders(gpu):
    """libfive_trees_render_mesh_batch gives, model by model, the meshes of
    libfive_tree_render_mesh (same vertex and triangle arrays)."""
    from libfive_amd import render_mesh_batch
    models = W.csg_batch(24, seed=W.SEED_BASE + 41)
    shapes = [m for m, _ in models]
    res = [r / 2.0 for _, r in models]       # cells per unit over [-1, 1]
    regions = [((-1, -1, -1), (1, 1, 1))] * len(shapes)
    print("batch of", len(shapes), "models, res", sorted(set(res)), flush=True)
    got = render_mesh_batch(shapes, regions, res, streams=6)
    for s, r, (v, t) in zip(shapes, res, got):
        v1, t1 = s.render_mesh([-1] * 3, [1] * 3, res=r)
        assert v.shape == v1.shape and t.shape == t1.shape
        assert (v.view(np.uint32) == v1.view(np.uint32)).all()
        assert (t == t1).all()


def _wide_shapes():
    return {
        "gyroid": (W.gyroid_c4(), 1.2, 2.4 / 64),
        "csg_c1": (W.csg_c1(), 2.0, 4.0 / 64),
        "union60": (W.sphere_union(60), 1.2, 2.4 / 64),
        "dag37": (W.random_dag() - 0.3, 1.0, 2.0 / 16),
        "wide75": (wide_registers(70), 1.0, 2.0 / 16),
        "box": (W.box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)), 1.0, 2.0 / 16),
    }


@pytest.mark.parametrize("name", ["gyroid", "csg_c1", "union60", "dag37", "wide75", "box"])
def test_wide_interval_pass_matches_walk(gpu, name):
    """The wide interval kernel (a workgroup of 8 waves per parent working
    on the tape in dependency order) gives the one-wave walk's cell states,
    pruned tape lengths and mesh bit for bit, forced onto shapes whose
    tapes are too short to get it by default (one, two and four register
    banks; a box whose faces meet the grid)."""
    from libfive_amd import ffi
    L = ffi.lib()
    shape, half, mf = _wide_shapes()[name]
    lo3 = [float(np.float32(-half))] * 3
    hi3 = [float(np.float32(half))] * 3
    out = []
    try:
        for choice in (0, 1):
            L.libfive_autotune_force(b"interval_wide", choice)
            print(name, "interval_wide", choice, flush=True)
            out.append(shape.render_mesh(lo3, hi3, min_feature=mf, details=True))
    finally:
        L.libfive_autotune_force(b"interval_wide", -1)
    a, b = out
    assert (a.level_counts == b.level_counts).all()
    assert a.cells.tolist() == b.cells.tolist()
    assert sorted(a.tape_lengths.tolist()) == sorted(b.tape_lengths.tolist())
    assert a.verts.shape == b.verts.shape and a.tris.shape == b.tris.shape
    assert (a.verts.view(np.uint32) == b.verts.view(np.uint32)).all()
    assert (a.tris == b.tris).all() and (a.keys == b.keys).all()


def test_merged_batch_matches_single_renders(gpu):
    """libfive_trees_render_mesh_merged runs all models through one set of
    passes; each model's mesh is still libfive_tree_render_mesh's, bit for
    bit.  The models differ in region, depth (root levels 0 to 8, in no
    particular order), tape length, register count (one, two and four
    banks), features (exact zeros on tied MIN/MAX) and emptiness."""
    from libfive_amd import render_mesh_batch
    models = []
    for s, r in W.csg_batch(12, seed=W.SEED_BASE + 43):
        models.append((s, ((-1.0,) * 3, (1.0,) * 3), r / 2.0))
    models += [
        (W.sphere(0.7), ((-2.0,) * 3, (2.0,) * 3), 4.0),
        (W.gyroid_c4(), ((-1.2,) * 3, (1.2,) * 3), 32 / 2.4),
        (W.box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)), ((-1.0,) * 3, (1.0,) * 3), 8.0),
        (W.random_dag() - 0.3, ((-1.0,) * 3, (1.0,) * 3), 8.0),
        (wide_registers(70), ((-1.0,) * 3, (1.0,) * 3), 8.0),
        (W.sphere(0.2, (3.0, 3.0, 3.0)), ((-1.0,) * 3, (1.0,) * 3), 16.0),   # empty
        (W.sphere(0.5), ((-1.0,) * 3, (1.0,) * 3), 0.25),                    # root leaf
        (W.sphere_union(60), ((-1.2,) * 3, (1.2,) * 3), 64 / 2.4),
        (W.csg_c1(), ((-2.0, -2.0, -2.0), (2.0, 2.0, 2.0)), 8.0),
        (W.sphere(0.6), ((-0.9, -1.1, -0.7), (0.8, 1.2, 0.75)), 20.0),    # not a cube
    ]
    order = np.random.default_rng(7).permutation(len(models))
    models = [models[i] for i in order]
    shapes = [m[0] for m in models]
    regions = [m[1] for m in models]
    res = [m[2] for m in models]
    print("merged batch of", len(shapes), "models", flush=True)
    got = render_mesh_batch(shapes, regions, res, merged=True)
    for i, (s, (lo, hi), r) in enumerate(models):
        v1, t1 = s.render_mesh(list(lo), list(hi), res=r)
        v, t = got[i]
        assert v.shape == v1.shape and t.shape == t1.shape, i
        assert (v.view(np.uint32) == v1.view(np.uint32)).all(), i
        assert (t == t1).all(), i
    # a second call on the same library state gives the same meshes
    again = render_mesh_batch(shapes, regions, res, merged=True)
    for (v, t), (v2, t2) in zip(got, again):
        assert (v.view(np.uint32) == v2.view(np.uint32)).all() and (t == t2).all()


def _region(lo, hi):
    from libfive_amd import ffi
    return ffi.libfive_region3(ffi.libfive_interval(lo, hi),
                               ffi.libfive_interval(lo, hi),
                               ffi.libfive_interval(lo, hi))


def test_render_mesh_coords_matches_render_mesh(gpu):
    """libfive_tree_render_mesh_coords (libfive.cpp:354-395): the same
    vertices, triangles as index triples followed by -1."""
    from libfive_amd import ffi
    s = W.sphere(0.5)
    verts, tris = s.render_mesh([-1] * 3, [1] * 3, res=16)
    L = ffi.lib()
    m = L.libfive_tree_render_mesh_coords(s.ptr, _region(-1, 1), 16.0)
    assert m, ffi.last_error()
    try:
        mc = m.contents
        v = np.ctypeslib.as_array(ctypes_cast_f(mc.verts), (mc.vert_count, 3)).copy()
        ci = np.ctypeslib.as_array(mc.coord_indices, (mc.coord_index_count,)).copy()
    finally:
        L.libfive_mesh_coords_delete(m)
    assert (v.view(np.uint32) == verts.view(np.uint32)).all()
    ci = ci.reshape(-1, 4)
    assert (ci[:, 3] == -1).all()
    assert (ci[:, :3].astype(np.uint32) == tris).all()


def ctypes_cast_f(p):
    import ctypes
    from libfive_amd import ffi
    return ctypes.cast(p, ffi.fp)


def _read_stl(path):
    data = open(path, "rb").read()
    n = int(np.frombuffer(data[80:84], np.uint32)[0])
    rec = np.frombuffer(data[84:], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)),
                                                   ("a", "<u2")]))
    return data[:80], n, rec


def test_save_mesh_writes_binary_stl(gpu, tmp_path):
    """libfive_tree_save_mesh: binary STL (mesh.cpp:139-198) of the mesh
    libfive_tree_render_mesh returns, triangle by triangle."""
    from libfive_amd import ffi
    s = W.csg_c1()
    verts, tris = s.render_mesh([-2] * 3, [2] * 3, res=8)
    path = str(tmp_path / "c1.stl")
    assert ffi.lib().libfive_tree_save_mesh(s.ptr, _region(-2, 2), 8.0, path.encode())
    header, n, rec = _read_stl(path)
    assert header.startswith(b"This is a binary STL exported from libfive.")
    assert header.rstrip(b" ") == b"This is a binary STL exported from libfive."
    assert n == tris.shape[0] == rec.shape[0]
    assert (rec["n"] == 0).all() and (rec["a"] == 0).all()
    assert (rec["v"].view(np.uint32) == verts[tris].view(np.uint32)).all()


def test_save_meshes_concatenates(gpu, tmp_path):
    from libfive_amd import ffi
    import ctypes
    a, b = W.sphere(0.5), W.sphere(0.3, (0.5, 0.5, 0.5))
    arr = (ffi.tree_t * 3)(a.ptr, b.ptr, None)
    path = str(tmp_path / "two.stl")
    assert ffi.lib().libfive_tree_save_meshes(arr, _region(-1, 1), 8.0, 8.0, path.encode())
    _, n, rec = _read_stl(path)
    me = 10.0 ** -8.0          # max_err = 10^-quality, as the entry point computes it
    mf = float(np.float32(1) / np.float32(8))
    na = a.render_mesh([-1] * 3, [1] * 3, min_feature=mf, max_err=me)[1].shape[0]
    nb = b.render_mesh([-1] * 3, [1] * 3, min_feature=mf, max_err=me)[1].shape[0]
    assert n == na + nb == rec.shape[0]


def test_evaluator_save_mesh_uses_variables(gpu, tmp_path):
    """libfive_evaluator_save_mesh renders with the evaluator's current
    free-variable values (libfive.cpp:408-496): a sphere whose radius is a
    variable, meshed at two radii."""
    import ctypes
    from libfive_amd import Shape, ffi
    L = ffi.lib()
    r = Shape.var()
    t = (X().square() + Y().square() + Z().square()).sqrt() - r
    rid = L.libfive_tree_id(r.ptr)

    def vars_(val):
        return ffi.libfive_vars((ctypes.c_void_p * 1)(rid), (ctypes.c_float * 1)(val), 1)

    ev = L.libfive_tree_evaluator(t.ptr, vars_(0.5))
    try:
        for radius in (0.5, 0.8):
            if radius != 0.5:
                assert L.libfive_evaluator_update_vars(ev, vars_(radius))
            path = str(tmp_path / f"r{radius}.stl")
            assert L.libfive_evaluator_save_mesh(ev, _region(-1, 1), path.encode()), \
                ffi.last_error()
            _, n, rec = _read_stl(path)
            assert n > 0
            rr = np.sqrt((rec["v"].astype(np.float64) ** 2).sum(axis=2))
            assert np.abs(rr - radius).max() < 0.01
    finally:
        L.libfive_evaluator_delete(ev)


def test_variables_match_oracle(gpu, tmp_path):
    """Free variables (libfive.cpp:408-496) against the oracle with each
    variable's value loaded into its slot, as the reference's evaluators do
    (evaluator.hpp:41-49): the mesh libfive_evaluator_save_mesh writes has
    the oracle mesh's triangles, vertex coordinates bit for bit, at two sets
    of values (after libfive_evaluator_update_vars)."""
    import ctypes
    from libfive_amd import ffi
    L = ffi.lib()
    r, cx = Shape.var(), Shape.var()
    sph = ((X() - cx).square() + Y().square() + Z().square()).sqrt() - r
    t = Shape.binary("max", sph, Z() - 0.3)          # a flat cut: features
    ids = [L.libfive_tree_id(v.ptr) for v in (r, cx)]

    def vars_(vals):
        return ffi.libfive_vars((ctypes.c_void_p * 2)(*ids), (ctypes.c_float * 2)(*vals), 2)

    lo3 = [float(np.float32(-1.0))] * 3
    hi3 = [float(np.float32(1.0))] * 3
    ev = L.libfive_tree_evaluator(t.ptr, vars_((0.6, 0.05)))
    try:
        for k, vals in enumerate(((0.6, 0.05), (0.45, -0.2))):
            if k:
                assert L.libfive_evaluator_update_vars(ev, vars_(vals))
            path = str(tmp_path / f"v{k}.stl")
            assert L.libfive_evaluator_save_mesh(ev, _region(-1, 1), path.encode()), \
                ffi.last_error()
            _, n, rec = _read_stl(path)
            got = sorted(tuple(tri.view(np.uint32).ravel().tolist()) for tri in rec["v"])
            o = oracle.OracleDeck(t.deck(vars={r: vals[0], cx: vals[1]})).render_mesh(
                lo3, hi3, 0.1, 1e-8)
            ov = np.asarray(o.verts, np.float32)
            want = sorted(tuple(ov[list(tr)].view(np.uint32).ravel().tolist())
                          for tr in o.tris.tolist())
            print("variables", vals, "triangles", n, len(want), flush=True)
            assert n == len(want) and n > 0
            assert got == want
    finally:
        L.libfive_evaluator_delete(ev)


def test_cpp_mesh_api(gpu):
    """The C++ entry (libfive::Mesh::render, both overloads, saveSTL) and
    the reference's cancellation case, in the test program built with the
    library (tests/cpp/mesh_api_test.cpp)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "mesh_api_test")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120,
                       cwd=os.path.dirname(exe))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")


def test_cancel_from_another_thread(gpu):
    """libfive/test/xtree.cpp:240-272 through the C API: a render cancelled
    from another thread returns NULL within 150 ms, right after it starts
    and 30 ms into the octree build."""
    import ctypes
    import threading
    import time
    from libfive_amd import ffi
    L = ffi.lib()
    s = W.gyroid_c4()
    R = ffi.libfive_region3(ffi.libfive_interval(-1.2, 1.2),
                            ffi.libfive_interval(-1.2, 1.2),
                            ffi.libfive_interval(-1.2, 1.2))
    st = ffi.libfive_mesh_settings()
    L.libfive_mesh_settings_default(ctypes.byref(st))
    st.min_feature = 2.4 / 1024
    flag = ctypes.c_int(0)
    st.cancel = ctypes.pointer(flag)
    t0 = time.perf_counter()
    m = L.libfive_tree_render_mesh_settings(s.ptr, R, ctypes.byref(st))
    full_ms = (time.perf_counter() - t0) * 1e3
    assert m, ffi.last_error()
    L.libfive_mesh_delete(m)
    for delay in (0.0, 0.03):
        flag.value = 0
        out = {}

        def run():
            out["m"] = L.libfive_tree_render_mesh_settings(s.ptr, R, ctypes.byref(st))

        th = threading.Thread(target=run)
        th.start()
        time.sleep(delay)
        t1 = time.perf_counter()
        flag.value = 1
        th.join(60)
        ms = (time.perf_counter() - t1) * 1e3
        print(f"cancel after {delay * 1e3:.0f} ms: returned in {ms:.1f} ms "
              f"(full render {full_ms:.1f} ms)", flush=True)
        assert not th.is_alive()
        assert not out["m"]
        assert ms < 150
