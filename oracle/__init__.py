"""Oracle: CPU restatement of the reference algorithms (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this package, always as the checker or the CPU baseline, never as the
measured product.  The shared library is built by oracle/Makefile.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_u8 = ctypes.POINTER(ctypes.c_uint8)

_SIGS = {
    "oracle_deck_new": (ctypes.c_void_p, [_u32, ctypes.c_size_t, _u32, _fp,
                                          ctypes.c_size_t, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint32]),
    "oracle_deck_delete": (None, [ctypes.c_void_p]),
    "oracle_eval_values": (None, [ctypes.c_void_p, _fp, _fp, _fp,
                                  ctypes.c_size_t, _fp, ctypes.c_int]),
    "oracle_eval_derivs": (None, [ctypes.c_void_p, _fp, _fp, _fp,
                                  ctypes.c_size_t, _fp, _fp, _fp, _fp,
                                  ctypes.c_int]),
    "oracle_eval_values_f64": (None, [ctypes.c_void_p, _dp, _dp, _dp,
                                      ctypes.c_size_t, _dp]),
    "oracle_eval_intervals": (None, [ctypes.c_void_p, _fp, _fp, _fp, _fp,
                                     _fp, _fp, ctypes.c_size_t, _fp, _fp,
                                     _u8]),
    "oracle_interval_push": (ctypes.c_size_t, [ctypes.c_void_p, _fp, _fp,
                                               _fp, _fp, _u8,
                                               ctypes.POINTER(ctypes.c_int)]),
    "oracle_sincos": (None, [_fp, ctypes.c_size_t, _fp, _fp]),
    "oracle_eval_ambiguous": (None, [ctypes.c_void_p, _fp, _fp, _fp,
                                     ctypes.c_size_t, _u8]),
    "oracle_eval_deriv_ambiguous": (None, [ctypes.c_void_p, _fp, _fp, _fp,
                                           ctypes.c_size_t, _u8]),
    "oracle_value_and_push": (ctypes.c_size_t, [ctypes.c_void_p, _fp, _fp, _fp,
                                                _fp, ctypes.c_size_t, _fp, _fp]),
    "oracle_interval_push_eval": (ctypes.c_size_t, [ctypes.c_void_p, _fp, _fp,
                                                    _fp, _fp, _fp,
                                                    ctypes.c_size_t, _fp]),
    "oracle_render_mesh": (ctypes.c_void_p, [ctypes.c_void_p, _dp, _dp,
                                             ctypes.c_double, ctypes.c_double]),
    "oracle_render_mesh_mt": (ctypes.c_void_p, [ctypes.c_void_p, _dp, _dp,
                                                ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int]),
    "oracle_render_mesh_ex": (ctypes.c_void_p, [ctypes.c_void_p, _dp, _dp,
                                                ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int, ctypes.c_int]),
    "oracle_mesh_verts64": (ctypes.c_size_t, [ctypes.c_void_p, _dp]),
    "oracle_mesh_sizes": (None, [ctypes.c_void_p] +
                          [ctypes.POINTER(ctypes.c_size_t)] * 4),
    "oracle_mesh_copy": (None, [ctypes.c_void_p, _fp, _u32, _u32,
                                ctypes.POINTER(ctypes.c_uint64), _u32]),
    "oracle_mesh_leaf_verts": (ctypes.c_size_t, [ctypes.c_void_p, _dp]),
    "oracle_mesh_cells": (ctypes.c_size_t, [ctypes.c_void_p, _u32]),
    "oracle_mesh_delete": (None, [ctypes.c_void_p]),
    "oracle_corner_manifold": (ctypes.c_int, [ctypes.c_int]),
    "oracle_marching_table": (None, [ctypes.POINTER(ctypes.c_int8)] * 3),
    "oracle_is_inside": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "oracle_features": (ctypes.c_size_t, [ctypes.c_void_p, _fp, _fp,
                                          ctypes.c_size_t]),
    "oracle_eigen3": (None, [_dp, _dp, _dp]),
    "oracle_find_vertex": (ctypes.c_double, [_dp, _dp, ctypes.c_double, _dp,
                                             _dp, ctypes.POINTER(ctypes.c_int)]),
    "oracle_push_sample": (None, [_dp, _dp, _dp, ctypes.c_double]),
    "oracle_push_sample_f2": (None, [_fp, _fp, _fp, ctypes.c_float]),
    "oracle_f2_add_acc": (None, [_fp, _dp]),
    "oracle_f2_widen": (None, [_fp, _dp]),
    "oracle_resolution_level": (ctypes.c_int, [_dp, _dp, ctypes.c_double]),
    "oracle_render_pixels": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_float] * 6
                             + [_u8, _u32, _u32]),
    "oracle_voxel_pts": (None, [ctypes.c_float] * 3 + [_fp, _u32, _fp, _fp]),
    "oracle_render_contours": (ctypes.c_void_p, [ctypes.c_void_p, _dp, _dp,
                                                 ctypes.c_float, ctypes.c_double,
                                                 ctypes.c_double]),
    "oracle_contours_sizes": (None, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_size_t)] * 5),
    "oracle_contours_copy": (None, [ctypes.c_void_p, _fp] + [_u32] * 5),
    "oracle_contours_delete": (None, [ctypes.c_void_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(_fp)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class OracleDeck:
    """Wraps a reference-format Deck: (clauses[n,4], constants, xyz, slots)
    as returned by libfive_amd.Shape.deck()."""

    def __init__(self, deck):
        clauses, consts, xyz, nslots = deck
        L = lib()
        cl = np.ascontiguousarray(clauses, np.uint32)
        ci = np.array(list(consts.keys()) or [0], np.uint32)
        cv = np.array(list(consts.values()) or [0.0], np.float32)
        self.ptr = L.oracle_deck_new(cl.ctypes.data_as(_u32), cl.shape[0],
                                     ci.ctypes.data_as(_u32), _f(cv),
                                     len(consts), xyz[0], xyz[1], xyz[2],
                                     nslots)
        self.n_clauses = cl.shape[0]

    def __del__(self):
        try:
            lib().oracle_deck_delete(self.ptr)
        except Exception:
            pass

    def values(self, x, y, z, threads=1):
        x, y, z = _f32(x), _f32(y), _f32(z)
        out = np.empty_like(x)
        lib().oracle_eval_values(self.ptr, _f(x), _f(y), _f(z), x.size,
                                 _f(out), threads)
        return out

    def derivs(self, x, y, z, threads=1):
        x, y, z = _f32(x), _f32(y), _f32(z)
        v, dx, dy, dz = (np.empty_like(x) for _ in range(4))
        lib().oracle_eval_derivs(self.ptr, _f(x), _f(y), _f(z), x.size,
                                 _f(v), _f(dx), _f(dy), _f(dz), threads)
        return v, dx, dy, dz

    def values_f64(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        out = np.empty_like(x)
        lib().oracle_eval_values_f64(self.ptr, x.ctypes.data_as(_dp),
                                     y.ctypes.data_as(_dp),
                                     z.ctypes.data_as(_dp), x.size,
                                     out.ctypes.data_as(_dp))
        return out

    def intervals(self, lo, hi):
        lo = np.asarray(lo, np.float32)
        hi = np.asarray(hi, np.float32)
        cols = [_f32(a) for a in (lo[:, 0], hi[:, 0], lo[:, 1], hi[:, 1],
                                  lo[:, 2], hi[:, 2])]
        n = lo.shape[0]
        ol = np.empty(n, np.float32)
        oh = np.empty(n, np.float32)
        nan = np.empty(n, np.uint8)
        lib().oracle_eval_intervals(self.ptr, *[_f(c) for c in cols], n,
                                    _f(ol), _f(oh), nan.ctypes.data_as(_u8))
        return ol, oh, nan.astype(bool)

    def interval_push(self, lo, hi):
        """(pushed tape length, changed, (lower, upper, maybe_nan))."""
        lo = _f32(lo)
        hi = _f32(hi)
        ol = np.zeros(1, np.float32)
        oh = np.zeros(1, np.float32)
        nan = np.zeros(1, np.uint8)
        ch = ctypes.c_int()
        n = lib().oracle_interval_push(self.ptr, _f(lo), _f(hi), _f(ol),
                                       _f(oh), nan.ctypes.data_as(_u8),
                                       ctypes.byref(ch))
        return n, bool(ch.value), (float(ol[0]), float(oh[0]), bool(nan[0]))

    def ambiguous(self, x, y, z):
        """ArrayEvaluator::getAmbiguous per point (bool array)."""
        x, y, z = _f32(x), _f32(y), _f32(z)
        out = np.zeros(x.size, np.uint8)
        lib().oracle_eval_ambiguous(self.ptr, _f(x), _f(y), _f(z), x.size,
                                    out.ctypes.data_as(_u8))
        return out.astype(bool)

    def deriv_ambiguous(self, x, y, z):
        """DerivArrayEvaluator::getAmbiguousDerivs per point (bool array)."""
        x, y, z = _f32(x), _f32(y), _f32(z)
        out = np.zeros(x.size, np.uint8)
        lib().oracle_eval_deriv_ambiguous(self.ptr, _f(x), _f(y), _f(z),
                                          x.size, out.ctypes.data_as(_u8))
        return out.astype(bool)

    def value_and_push(self, p, qx=(), qy=(), qz=()):
        """ArrayEvaluator::valueAndPush at p: (value at p, pushed tape's
        clause count, the pushed tape's values at the points q)."""
        p = _f32(p)
        qx, qy, qz = _f32(qx), _f32(qy), _f32(qz)
        out = np.zeros(qx.size, np.float32)
        val = np.zeros(1, np.float32)
        n = lib().oracle_value_and_push(self.ptr, _f(p), _f(qx), _f(qy),
                                        _f(qz), qx.size, _f(out), _f(val))
        return float(val[0]), n, out

    def interval_push_values(self, lo, hi, qx, qy, qz):
        """intervalAndPush on [lo, hi]: (pushed tape's clause count, its
        values at the points q)."""
        lo, hi = _f32(lo), _f32(hi)
        qx, qy, qz = _f32(qx), _f32(qy), _f32(qz)
        out = np.zeros(qx.size, np.float32)
        n = lib().oracle_interval_push_eval(self.ptr, _f(lo), _f(hi), _f(qx),
                                            _f(qy), _f(qz), qx.size, _f(out))
        return n, out

    def is_inside(self, p):
        p = _f32(p)
        return bool(lib().oracle_is_inside(self.ptr, _f(p)))

    def features(self, p):
        p = _f32(p)
        out = np.zeros((64, 3), np.float32)
        n = lib().oracle_features(self.ptr, _f(p), _f(out), 64)
        return out[:min(n, 64)]

    def render_mesh(self, lo, hi, min_feature, max_err=1e-8, threads=1,
                    fp64=False, qef_f32c=False):
        """Dual-contouring mesh: a MeshResult (see below).  fp64=True gives
        the FP64 error oracle (double evaluation of values, partials and
        the collapse test); its double vertices are in .verts64.
        qef_f32c=True accumulates the QEF sums as compensated float pairs
        (the product's precision form, mesh setting precision = 1)."""
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        L = lib()
        m = L.oracle_render_mesh_ex(self.ptr, lo.ctypes.data_as(_dp),
                                    hi.ctypes.data_as(_dp), min_feature, max_err,
                                    int(threads),
                                    (1 if fp64 else 0) | (2 if qef_f32c else 0))
        try:
            sz = [ctypes.c_size_t() for _ in range(4)]
            L.oracle_mesh_sizes(m, *[ctypes.byref(x) for x in sz])
            nv, nt, ns, nl = (x.value for x in sz)
            verts = np.empty((nv, 3), np.float32)
            tris = np.empty((nt, 3), np.uint32)
            keys = np.empty((nv, 5), np.uint32)
            stats = np.empty(ns, np.uint64)
            lens = np.empty(nl, np.uint32)
            L.oracle_mesh_copy(m, _f(verts), tris.ctypes.data_as(_u32),
                               keys.ctypes.data_as(_u32),
                               stats.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                               lens.ctypes.data_as(_u32))
            nl = L.oracle_mesh_leaf_verts(m, None)
            leaf = np.empty((nl, 3), np.float64)
            L.oracle_mesh_leaf_verts(m, leaf.ctypes.data_as(_dp))
            nc = L.oracle_mesh_cells(m, None)
            cells = np.empty((nc, 5), np.uint32)
            L.oracle_mesh_cells(m, cells.ctypes.data_as(_u32))
            v64 = np.empty((L.oracle_mesh_verts64(m, None), 3), np.float64)
            L.oracle_mesh_verts64(m, v64.ctypes.data_as(_dp))
        finally:
            L.oracle_mesh_delete(m)
        out = MeshResult(verts, tris, keys, stats, lens)
        out.leaf_verts = leaf
        out.cells = cells
        out.verts64 = v64
        return out

    def render_pixels(self, xy_lo, xy_hi, z, res):
        """Occupancy raster of the slice at height z: bool[height, width]
        (row 0 at the lowest y), as libfive_tree_render_pixels."""
        L = lib()
        args = [float(xy_lo[0]), float(xy_hi[0]), float(xy_lo[1]),
                float(xy_hi[1]), float(z), float(res)]
        w, h = ctypes.c_uint32(), ctypes.c_uint32()
        L.oracle_render_pixels(self.ptr, *args, None, ctypes.byref(w),
                               ctypes.byref(h))
        out = np.zeros((h.value, w.value), np.uint8)
        L.oracle_render_pixels(self.ptr, *args, out.ctypes.data_as(_u8),
                               ctypes.byref(w), ctypes.byref(h))
        return out.astype(bool)

    def render_contours(self, lo, hi, z, min_feature, max_err=1e-8):
        """2D dual contouring of the slice at height z: a ContourResult."""
        L = lib()
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        c = L.oracle_render_contours(self.ptr, lo.ctypes.data_as(_dp),
                                     hi.ctypes.data_as(_dp), float(z),
                                     float(min_feature), float(max_err))
        try:
            sz = [ctypes.c_size_t() for _ in range(5)]
            L.oracle_contours_sizes(c, *[ctypes.byref(x) for x in sz])
            nv, ns, nc, npts, ncell = (x.value for x in sz)
            verts = np.zeros((nv, 2), np.float32)
            keys = np.zeros((nv, 4), np.uint32)
            segs = np.zeros((ns, 2), np.uint32)
            counts = np.zeros(nc, np.uint32)
            points = np.zeros(npts, np.uint32)
            cells = np.zeros((ncell, 4), np.uint32)
            L.oracle_contours_copy(c, _f(verts), keys.ctypes.data_as(_u32),
                                   segs.ctypes.data_as(_u32),
                                   counts.ctypes.data_as(_u32),
                                   points.ctypes.data_as(_u32),
                                   cells.ctypes.data_as(_u32))
        finally:
            L.oracle_contours_delete(c)
        return ContourResult(verts, keys, segs, counts, points, cells)


class ContourResult:
    """verts[n,2] (row 0 is the reserved dummy vertex), keys[n,4] (owning
    cell's lower corner x, y in finest-cell units, its level, its patch),
    segs[m,2] vertex pairs, the chained contours as lists of vertex
    indices (.contours) and point arrays (.polylines), and every evaluated
    cell (x, y, level, state)."""

    def __init__(self, verts, keys, segs, counts, points, cells):
        self.verts, self.keys, self.segs, self.cells = verts, keys, segs, cells
        offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.contours = [points[offs[i]:offs[i + 1]] for i in range(len(counts))]
        self.polylines = [verts[c] for c in self.contours]


def voxel_pts(lo, hi, res):
    """(centres, lower, upper) of one axis of the voxel grid."""
    L = lib()
    n = ctypes.c_uint32()
    L.oracle_voxel_pts(lo, hi, res, None, ctypes.byref(n), None, None)
    pts = np.zeros(n.value, np.float32)
    lower = np.zeros(1, np.float32)
    upper = np.zeros(1, np.float32)
    L.oracle_voxel_pts(lo, hi, res, _f(pts), ctypes.byref(n), _f(lower),
                       _f(upper))
    return pts, float(lower[0]), float(upper[0])


class MeshResult:
    """verts[n,3] (row 0 is the reserved dummy vertex), tris[m,3], keys[n,5]
    (owning cell's lower corner x, y, z in finest-cell units, its level, its
    patch), level statistics and the pushed tapes' clause counts."""

    def __init__(self, verts, tris, keys, stats, tape_lengths):
        self.verts, self.tris, self.keys = verts, tris, keys
        self.root_level = int(stats[0])
        self.ambiguous_leaves = int(stats[1])
        self.collapsed = int(stats[2])
        self.level_counts = stats[3:].reshape(-1, 3)   # root level first
        self.tape_lengths = tape_lengths


def corner_manifold(mask):
    return bool(lib().oracle_corner_manifold(int(mask)))


def marching_table():
    v = np.empty((256, 4, 12, 2), np.int8)
    e = np.empty((8, 8), np.int8)
    p = np.empty((256, 24), np.int8)
    i8 = ctypes.POINTER(ctypes.c_int8)
    lib().oracle_marching_table(v.ctypes.data_as(i8), e.ctypes.data_as(i8),
                                p.ctypes.data_as(i8))
    return v, e, p


def eigen3(a):
    """Eigenvalues and eigenvectors (columns) of a symmetric 3x3 matrix by
    the product's Jacobi routine."""
    a = np.ascontiguousarray(a, np.float64)
    w = np.empty(3)
    v = np.empty((3, 3))
    lib().oracle_eigen3(a.ctypes.data_as(_dp), w.ctypes.data_as(_dp),
                        v.ctypes.data_as(_dp))
    return w, v


def find_vertex(ata, atb, btb, mass):
    """(vertex, rank, error) of the dual-contouring vertex solve."""
    ata = np.ascontiguousarray(ata, np.float64)
    atb = np.ascontiguousarray(atb, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    v = np.empty(3)
    r = ctypes.c_int()
    err = lib().oracle_find_vertex(ata.ctypes.data_as(_dp), atb.ctypes.data_as(_dp),
                                   float(btb), mass.ctypes.data_as(_dp),
                                   v.ctypes.data_as(_dp), ctypes.byref(r))
    return v, r.value, err


def push_sample(acc, pos, deriv, value):
    """Adds one sample to a 14-double accumulator (mass x, y, z, count;
    AtA xx, xy, xz, yy, yz, zz; AtB; BtB) in place."""
    pos = np.ascontiguousarray(pos, np.float64)
    deriv = np.ascontiguousarray(deriv, np.float64)
    lib().oracle_push_sample(acc.ctypes.data_as(_dp), pos.ctypes.data_as(_dp),
                             deriv.ctypes.data_as(_dp), float(value))


def push_sample_f2(state, pos, deriv, value):
    """Compensated-FP32 push into state (28 float32: 14 hi, 14 lo), in
    place."""
    pos = np.ascontiguousarray(pos, np.float32)
    deriv = np.ascontiguousarray(deriv, np.float32)
    lib().oracle_push_sample_f2(_f(state), _f(pos), _f(deriv), float(value))


def f2_add_acc(state, acc):
    acc = np.ascontiguousarray(acc, np.float64)
    lib().oracle_f2_add_acc(_f(state), acc.ctypes.data_as(_dp))


def f2_widen(state):
    out = np.empty(14)
    lib().oracle_f2_widen(_f(state), out.ctypes.data_as(_dp))
    return out


def resolution_level(lo, hi, min_feature):
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    return lib().oracle_resolution_level(lo.ctypes.data_as(_dp),
                                         hi.ctypes.data_as(_dp), min_feature)


def sincos(x):
    x = _f32(x)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().oracle_sincos(_f(x), x.size, _f(s), _f(c))
    return s, c


def _pack_keys(keys):
    k = keys.astype(np.uint64)
    return (k[:, 0] | (k[:, 1] << np.uint64(18)) | (k[:, 2] << np.uint64(36))
            | (k[:, 3] << np.uint64(54)) | (k[:, 4] << np.uint64(59)))


def vertex_error(verts, keys, ref):
    """Vertex error of a mesh against the FP64 error oracle's mesh `ref`:
    vertices are matched by key (owning cell and patch); returns
    (max |v - v64| over matched vertices, matched count, vertex count)."""
    d, matched, n = vertex_errors(verts, keys, ref)
    return (float(d.max()) if d.size else 0.0), matched, n


def vertex_errors(verts, keys, ref):
    """Per matched vertex, max over axes of |v - v64| (see vertex_error);
    returns (errors, matched count, vertex count)."""
    a = _pack_keys(np.asarray(keys)[1:])
    b = _pack_keys(np.asarray(ref.keys)[1:])
    order = np.argsort(b)
    bs = b[order]
    pos = np.clip(np.searchsorted(bs, a), 0, max(len(bs) - 1, 0))
    hit = (bs[pos] == a) if len(bs) else np.zeros(len(a), bool)
    j = order[pos[hit]] + 1
    d = np.abs(np.asarray(verts)[1:][hit].astype(np.float64) - ref.verts64[j])
    return (d.max(axis=1) if d.size else np.zeros(0)), int(hit.sum()), len(a)


def _canonical(verts, tris, keys):
    """Vertices sorted by key (owning cell and patch) and triangles as
    sorted rows of key ranks, winding kept within each row."""
    k = _pack_keys(np.asarray(keys)[1:])
    order = np.argsort(k, kind="stable")
    rank = np.empty(len(k) + 1, np.int64)
    rank[0] = -1
    rank[1 + order] = np.arange(len(k))
    t = rank[np.asarray(tris, np.int64)] if len(tris) else np.zeros((0, 3), np.int64)
    t = t[np.lexsort(t.T[::-1])] if len(t) else t
    v = np.ascontiguousarray(np.asarray(verts, np.float32)[1:][order])
    return k[order], v.view(np.uint32), t


def compare_meshes(verts, tris, keys, ref):
    """Exact comparison of a mesh with the oracle's mesh `ref` through the
    vertex keys (vertex numbering is free): same key set, bit-identical
    positions per key, same triangles as ordered key triples.  Returns a
    dict of booleans and counts."""
    ka, va, ta = _canonical(verts, tris, keys)
    kb, vb, tb = _canonical(ref.verts, ref.tris, ref.keys)
    same_keys = ka.shape == kb.shape and bool((ka == kb).all())
    out = {"verts": int(len(ka)), "tris": int(len(ta)), "same_vertex_keys": same_keys,
           "bit_identical_vertices": False, "same_triangles": False}
    if same_keys:
        out["bit_identical_vertices"] = bool((va == vb).all())
        out["same_triangles"] = ta.shape == tb.shape and bool((ta == tb).all())
    return out
