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


def sincos(x):
    x = _f32(x)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().oracle_sincos(_f(x), x.size, _f(s), _f(c))
    return s, c
