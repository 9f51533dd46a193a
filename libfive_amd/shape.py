"""Python handle for expression trees built through the C ABI.

Follows the reference's Python ``Shape`` wrapper
(libfive/bind/python/libfive/shape.py): operator overloading builds
trees with libfive_tree_unary / libfive_tree_binary, and evaluation and
meshing call the C entry points.  Opcode numbers are the reference's
(libfive/include/libfive/tree/opcode.hpp:23-61).
"""
import ctypes

import numpy as np

from . import ffi

OPS = {
    "square": 7, "sqrt": 8, "neg": 9, "sin": 10, "cos": 11, "tan": 12,
    "asin": 13, "acos": 14, "atan": 15, "exp": 16, "add": 17, "mul": 18,
    "min": 19, "max": 20, "sub": 21, "div": 22, "atan2": 23, "pow": 24,
    "nth-root": 25, "mod": 26, "nanfill": 27, "abs": 28, "recip": 29,
    "log": 30, "compare": 31, "const-var": 6,
}


class GpuError(RuntimeError):
    """An evaluation or meshing call failed (the C API's last error)."""


def _f32p(a):
    return a.ctypes.data_as(ffi.fp)


class Shape:
    __slots__ = ("ptr",)

    def __init__(self, ptr):
        if not ptr:
            raise ValueError("invalid tree (NULL)")
        self.ptr = ptr

    def __del__(self):
        try:
            if self.ptr:
                ffi.lib().libfive_tree_delete(self.ptr)
        except Exception:
            pass

    # ---- construction ---------------------------------------------------
    @staticmethod
    def X():
        return Shape(ffi.lib().libfive_tree_x())

    @staticmethod
    def Y():
        return Shape(ffi.lib().libfive_tree_y())

    @staticmethod
    def Z():
        return Shape(ffi.lib().libfive_tree_z())

    @staticmethod
    def var():
        return Shape(ffi.lib().libfive_tree_var())

    @staticmethod
    def constant(v):
        return Shape(ffi.lib().libfive_tree_const(float(v)))

    @staticmethod
    def wrap(v):
        return v if isinstance(v, Shape) else Shape.constant(v)

    # The wrapped operands are held in locals for the duration of the call:
    # a temporary Shape would release its tree before the C call runs.
    @staticmethod
    def unary(op, a):
        code = OPS[op] if isinstance(op, str) else op
        sa = Shape.wrap(a)
        return Shape(ffi.lib().libfive_tree_unary(code, sa.ptr))

    @staticmethod
    def binary(op, a, b):
        code = OPS[op] if isinstance(op, str) else op
        sa, sb = Shape.wrap(a), Shape.wrap(b)
        return Shape(ffi.lib().libfive_tree_binary(code, sa.ptr, sb.ptr))

    def __add__(self, o): return Shape.binary("add", self, o)
    def __radd__(self, o): return Shape.binary("add", o, self)
    def __sub__(self, o): return Shape.binary("sub", self, o)
    def __rsub__(self, o): return Shape.binary("sub", o, self)
    def __mul__(self, o): return Shape.binary("mul", self, o)
    def __rmul__(self, o): return Shape.binary("mul", o, self)
    def __truediv__(self, o): return Shape.binary("div", self, o)
    def __rtruediv__(self, o): return Shape.binary("div", o, self)
    def __neg__(self): return Shape.unary("neg", self)
    def __abs__(self): return Shape.unary("abs", self)

    def square(self): return Shape.unary("square", self)
    def sqrt(self): return Shape.unary("sqrt", self)
    def sin(self): return Shape.unary("sin", self)
    def cos(self): return Shape.unary("cos", self)
    def min(self, o): return Shape.binary("min", self, o)
    def max(self, o): return Shape.binary("max", self, o)

    # ---- inspection -----------------------------------------------------
    def __str__(self):
        L = ffi.lib()
        p = L.libfive_tree_print(self.ptr)
        s = ctypes.string_at(p).decode()
        L.libfive_free_str(p)
        return s

    def optimized(self):
        return Shape(ffi.lib().libfive_tree_optimized(self.ptr))

    def remap(self, x, y, z):
        """self(x', y', z') (libfive/src/libfive.cpp:186-189)."""
        x, y, z = (Shape.wrap(v) for v in (x, y, z))
        return Shape(ffi.lib().libfive_tree_remap(self.ptr, x.ptr, y.ptr, z.ptr))

    def deck(self, vars=None):
        """Reference-format clause list: (clauses[n,4] root first,
        constants {id: value}, (X, Y, Z) ids, slot count).  vars maps free
        variables (Shape.var() shapes) to values: their slots are then
        listed with the constants, holding those values (the reference's
        evaluators load a variable's value into its slot; a variable not
        given is 0, as in the GPU compile)."""
        L = ffi.lib()
        n = ctypes.c_size_t()
        nc = ctypes.c_size_t()
        if L.libfive_tree_deck(self.ptr, None, ctypes.byref(n), None, None,
                               ctypes.byref(nc), None, None):
            raise GpuError(ffi.last_error())
        cl = np.zeros((max(n.value, 1), 4), np.uint32)
        ci = np.zeros(max(nc.value, 1), np.uint32)
        cv = np.zeros(max(nc.value, 1), np.float32)
        xyz = np.zeros(3, np.uint32)
        ns = ctypes.c_uint32()
        L.libfive_tree_deck(self.ptr, cl.ctypes.data_as(ffi.u32p),
                            ctypes.byref(n), ci.ctypes.data_as(ffi.u32p),
                            _f32p(cv), ctypes.byref(nc),
                            xyz.ctypes.data_as(ffi.u32p), ctypes.byref(ns))
        consts = {int(i): float(v) for i, v in zip(ci[:nc.value], cv[:nc.value])}
        nv = ctypes.c_size_t()
        if L.libfive_tree_deck_vars(self.ptr, None, None, ctypes.byref(nv)):
            raise GpuError(ffi.last_error())
        if nv.value:
            slots = (ctypes.c_uint32 * nv.value)()
            ids = (ctypes.c_void_p * nv.value)()
            L.libfive_tree_deck_vars(self.ptr, slots, ids, ctypes.byref(nv))
            given = {L.libfive_tree_id(v.ptr): float(np.float32(x))
                     for v, x in (vars or {}).items()}
            for k in range(nv.value):
                consts[int(slots[k])] = given.get(ids[k], 0.0)
        return cl[:n.value].copy(), consts, tuple(int(v) for v in xyz), ns.value

    def tape_info(self):
        """(clause count, register count) of the compiled tape (host only)."""
        c = ctypes.c_uint32()
        r = ctypes.c_uint32()
        if ffi.lib().libfive_tree_tape_info(self.ptr, ctypes.byref(c),
                                            ctypes.byref(r)):
            raise GpuError(ffi.last_error())
        return c.value, r.value

    # ---- evaluation (GPU) -----------------------------------------------
    def eval_f(self, x, y, z):
        return ffi.lib().libfive_tree_eval_f(self.ptr, ffi.libfive_vec3(x, y, z))

    def eval_d(self, x, y, z):
        v = ffi.lib().libfive_tree_eval_d(self.ptr, ffi.libfive_vec3(x, y, z))
        return (v.x, v.y, v.z)

    def eval_r(self, lo, hi):
        r = ffi.libfive_region3(ffi.libfive_interval(lo[0], hi[0]),
                                ffi.libfive_interval(lo[1], hi[1]),
                                ffi.libfive_interval(lo[2], hi[2]))
        i = ffi.lib().libfive_tree_eval_r(self.ptr, r)
        return (i.lower, i.upper)

    def values(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        out = np.empty_like(x)
        if ffi.lib().libfive_tree_eval_array(self.ptr, _f32p(x), _f32p(y),
                                             _f32p(z), x.size, _f32p(out)):
            raise GpuError(ffi.last_error())
        return out

    def derivs(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        v, dx, dy, dz = (np.empty_like(x) for _ in range(4))
        if ffi.lib().libfive_tree_eval_deriv_array(
                self.ptr, _f32p(x), _f32p(y), _f32p(z), x.size, _f32p(v),
                _f32p(dx), _f32p(dy), _f32p(dz)):
            raise GpuError(ffi.last_error())
        return v, dx, dy, dz

    def ambiguous(self, x, y, z, derivs=False):
        """Per point: some MIN/MAX clause has equal operands
        (ArrayEvaluator::getAmbiguous); with derivs=True, equal operands
        with different partials (DerivArrayEvaluator::getAmbiguousDerivs)."""
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        out = np.empty(x.size, np.uint8)
        f = (ffi.lib().libfive_tree_eval_deriv_ambiguous_array if derivs
             else ffi.lib().libfive_tree_eval_ambiguous_array)
        if f(self.ptr, _f32p(x), _f32p(y), _f32p(z), x.size,
             out.ctypes.data_as(ffi.u8p)):
            raise GpuError(ffi.last_error())
        return out.astype(bool)

    def intervals(self, lo, hi):
        """lo, hi: arrays of shape (n, 3).  Returns (lower, upper, maybe_nan)."""
        lo = np.asarray(lo, np.float32)
        hi = np.asarray(hi, np.float32)
        cols = [np.ascontiguousarray(a) for a in
                (lo[:, 0], hi[:, 0], lo[:, 1], hi[:, 1], lo[:, 2], hi[:, 2])]
        n = lo.shape[0]
        ol = np.empty(n, np.float32)
        oh = np.empty(n, np.float32)
        nan = np.empty(n, np.uint8)
        if ffi.lib().libfive_tree_eval_interval_array(
                self.ptr, *[_f32p(c) for c in cols], n, _f32p(ol), _f32p(oh),
                nan.ctypes.data_as(ffi.u8p)):
            raise GpuError(ffi.last_error())
        return ol, oh, nan.astype(bool)

    # ---- meshing (GPU) --------------------------------------------------
    def render_mesh(self, lo, hi, res=None, min_feature=None, max_err=1e-8,
                    precision=0, details=False):
        """Dual-contouring mesh over the box [lo, hi].  Returns (verts,
        tris); with details=True a MeshResult that also holds per-vertex
        keys, per-level cell counts, pruned tape lengths and timings."""
        L = ffi.lib()
        R = ffi.libfive_region3(ffi.libfive_interval(lo[0], hi[0]),
                                ffi.libfive_interval(lo[1], hi[1]),
                                ffi.libfive_interval(lo[2], hi[2]))
        if res is not None and min_feature is None and max_err == 1e-8 \
                and precision == 0 and not details:
            m = L.libfive_tree_render_mesh(self.ptr, R, float(res))
        else:
            s = ffi.libfive_mesh_settings()
            L.libfive_mesh_settings_default(ctypes.byref(s))
            if min_feature is not None:
                s.min_feature = min_feature
            elif res is not None:
                s.min_feature = float(np.float32(1) / np.float32(res))
            s.max_err = max_err
            s.precision = precision
            s.collect_stats = 1 if details else 0
            m = L.libfive_tree_render_mesh_settings(self.ptr, R,
                                                    ctypes.byref(s))
        if not m:
            raise GpuError(ffi.last_error())
        try:
            mm = m.contents
            verts = np.ctypeslib.as_array(
                ctypes.cast(mm.verts, ffi.fp), (mm.vert_count, 3)).copy()
            if mm.tri_count:
                tris = np.ctypeslib.as_array(
                    ctypes.cast(mm.tris, ffi.u32p), (mm.tri_count, 3)).copy()
            else:
                tris = np.zeros((0, 3), np.uint32)
            if not details:
                return verts, tris
            kp = L.libfive_mesh_vertex_keys(m)
            keys = np.ctypeslib.as_array(kp, (mm.vert_count, 5)).copy()
        finally:
            L.libfive_mesh_delete(m)
        n = L.libfive_last_mesh_level_counts(None, 0)
        counts = np.zeros(3 * max(n, 1), np.uint64)
        L.libfive_last_mesh_level_counts(
            counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n)
        nt = L.libfive_last_mesh_tape_lengths(None, 0)
        lens = np.zeros(max(nt, 1), np.uint32)
        L.libfive_last_mesh_tape_lengths(lens.ctypes.data_as(ffi.u32p), nt)
        ms = (ctypes.c_double * 4)()
        L.libfive_last_mesh_times(ms)
        nc = L.libfive_last_mesh_cells(None, 0)
        cells = np.zeros((max(nc, 1), 5), np.uint32)
        L.libfive_last_mesh_cells(cells.ctypes.data_as(ffi.u32p), nc)
        out = MeshResult(verts, tris, keys, counts[:3 * n].reshape(-1, 3),
                         lens[:nt], list(ms))
        out.cells = cells[:nc]
        return out

    def render_pixels(self, xy_lo, xy_hi, z, res):
        """Occupancy raster of the slice at height z over [xy_lo, xy_hi]
        (libfive_tree_render_pixels): a bool array of shape (height, width),
        row 0 at the lowest y."""
        L = ffi.lib()
        R = ffi.libfive_region2(ffi.libfive_interval(xy_lo[0], xy_hi[0]),
                                ffi.libfive_interval(xy_lo[1], xy_hi[1]))
        p = L.libfive_tree_render_pixels(self.ptr, R, float(z), float(res))
        if not p:
            raise GpuError(ffi.last_error())
        try:
            pp = p.contents
            return np.ctypeslib.as_array(
                ctypes.cast(pp.pixels, ctypes.POINTER(ctypes.c_uint8)),
                (pp.height, pp.width)).astype(bool)
        finally:
            L.libfive_pixels_delete(p)

    def render_slice(self, xy_lo, xy_hi, z, res):
        """2D contours of the slice at height z (libfive_tree_render_slice):
        a list of float32 arrays of shape (n, 2)."""
        L = ffi.lib()
        R = ffi.libfive_region2(ffi.libfive_interval(xy_lo[0], xy_hi[0]),
                                ffi.libfive_interval(xy_lo[1], xy_hi[1]))
        cs = L.libfive_tree_render_slice(self.ptr, R, float(z), float(res))
        if not cs:
            raise GpuError(ffi.last_error())
        try:
            out = []
            for i in range(cs.contents.count):
                c = cs.contents.cs[i]
                out.append(np.ctypeslib.as_array(
                    ctypes.cast(c.pts, ffi.fp), (c.count, 2)).copy()
                    if c.count else np.zeros((0, 2), np.float32))
            return out
        finally:
            L.libfive_contours_delete(cs)

    def slice_segments(self, xy_lo, xy_hi, z, min_feature, max_err=1e-8,
                       cells=False):
        """(verts[n,2], keys[n,4], segs[m,2]) of the 2D contourer; vertex 0
        is the reserved dummy vertex, keys are the owning cell's lower
        corner in finest-cell units, its level and its patch.  With
        cells=True also every evaluated cell as (x, y, level, state)."""
        L = ffi.lib()
        R = ffi.libfive_region2(ffi.libfive_interval(xy_lo[0], xy_hi[0]),
                                ffi.libfive_interval(xy_lo[1], xy_hi[1]))
        p = L.libfive_tree_render_slice_segments(self.ptr, R, float(z),
                                                 float(min_feature), float(max_err))
        if not p:
            raise GpuError(ffi.last_error())
        try:
            s = p.contents
            verts = np.ctypeslib.as_array(s.verts, (s.vert_count, 2)).copy()
            keys = np.ctypeslib.as_array(s.keys, (s.vert_count, 4)).copy()
            segs = (np.ctypeslib.as_array(s.segs, (s.seg_count, 2)).copy()
                    if s.seg_count else np.zeros((0, 2), np.uint32))
            if not cells:
                return verts, keys, segs
            cl = (np.ctypeslib.as_array(s.cells, (s.cell_count, 4)).copy()
                  if s.cell_count else np.zeros((0, 4), np.uint32))
            return verts, keys, segs, cl
        finally:
            L.libfive_slice_segments_delete(p)


def _mesh_arrays(m):
    mm = m.contents
    verts = np.ctypeslib.as_array(
        ctypes.cast(mm.verts, ffi.fp), (mm.vert_count, 3)).copy()
    if mm.tri_count:
        tris = np.ctypeslib.as_array(
            ctypes.cast(mm.tris, ffi.u32p), (mm.tri_count, 3)).copy()
    else:
        tris = np.zeros((0, 3), np.uint32)
    return verts, tris


def autotune_report():
    """The launch-schedule decisions made so far in this process
    (libfive_autotune_report): a list of {key, choice, candidates_ms}."""
    import json
    L = ffi.lib()
    n = L.libfive_autotune_report(None, 0)
    buf = ctypes.create_string_buffer(n)
    L.libfive_autotune_report(buf, n)
    return json.loads(buf.value.decode())


def render_mesh_batch(shapes, regions, res, streams=8, keep=True, merged=False):
    """Meshes many shapes in one call: libfive_trees_render_mesh_batch
    (`streams` models at a time on their own streams) or, with merged,
    libfive_trees_render_mesh_merged (one set of passes over all models).
    regions[i] = (lo, hi) 3-vectors, res[i] = cells per unit.  Returns a
    list of (verts, tris), or of (vert_count, tri_count) with keep=False."""
    L = ffi.lib()
    n = len(shapes)
    trees = (ffi.tree_t * n)(*[s.ptr for s in shapes])
    R = (ffi.libfive_region3 * n)()
    for i, (lo, hi) in enumerate(regions):
        R[i] = ffi.libfive_region3(ffi.libfive_interval(lo[0], hi[0]),
                                   ffi.libfive_interval(lo[1], hi[1]),
                                   ffi.libfive_interval(lo[2], hi[2]))
    rr = np.ascontiguousarray(res, np.float32)
    out = (ctypes.POINTER(ffi.libfive_mesh) * n)()
    if merged:
        failed = L.libfive_trees_render_mesh_merged(
            trees, R, rr.ctypes.data_as(ffi.fp), n, out)
    else:
        failed = L.libfive_trees_render_mesh_batch(
            trees, R, rr.ctypes.data_as(ffi.fp), n, int(streams), out)
    try:
        if failed:
            raise GpuError(ffi.last_error())
        if keep:
            return [_mesh_arrays(out[i]) for i in range(n)]
        return [(out[i].contents.vert_count, out[i].contents.tri_count)
                for i in range(n)]
    finally:
        for i in range(n):
            if out[i]:
                L.libfive_mesh_delete(out[i])


class MeshResult:
    """A mesh with its vertex keys (owning cell x, y, z in finest-cell
    units, level, patch), per-level cell counts (root level first: empty,
    filled, ambiguous), pruned tape lengths and the render's timings in ms
    (build, collapse, emit, total)."""

    def __init__(self, verts, tris, keys, level_counts, tape_lengths, ms):
        self.verts, self.tris, self.keys = verts, tris, keys
        self.level_counts = level_counts
        self.tape_lengths = tape_lengths
        self.ms = ms


X = Shape.X
Y = Shape.Y
Z = Shape.Z
