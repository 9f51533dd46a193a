"""The allocation-free value entry point (libfive_tape_eval_values_noalloc)
against libfive_tape_eval_values, through the C ABI: bit-identical values
on two calls in a row (the second smaller, so stale staging contents
would show), on the interpreter and the specialised kernels, and a clear
error for more points than the staging array holds."""
import numpy as np
import pytest

from libfive_amd import ffi
from libfive_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 20


def _call(fn, tape, x, y, z):
    out = np.full(x.size, np.nan, np.float32)
    rc = fn(tape, x.ctypes.data_as(ffi.fp), y.ctypes.data_as(ffi.fp),
             z.ctypes.data_as(ffi.fp), x.size, out.ctypes.data_as(ffi.fp))
    return rc, out


@pytest.mark.parametrize("specialised", [False, True])
def test_noalloc_matches_tape_eval(gpu, specialised):
    L = ffi.lib()
    shape = W.random_dag(300, seed=21) if specialised else W.gyroid_c4()
    tape = L.libfive_tape_new(shape.ptr)
    assert tape
    try:
        if specialised:
            ms = ctypes_double()
            assert L.libfive_tape_specialize(tape, ms) == 0, ffi.last_error()
            L.libfive_tape_use_specialized(tape, 1)
        rng = np.random.default_rng(5)
        for n in (MAX_POINTS, 1000, 7):
            print(f"noalloc: specialised={specialised} n={n}", flush=True)
            x, y, z = (rng.uniform(-1.2, 1.2, n).astype(np.float32)
                       for _ in range(3))
            rc0, ref = _call(L.libfive_tape_eval_values, tape, x, y, z)
            rc1, got = _call(L.libfive_tape_eval_values_noalloc, tape, x, y, z)
            assert rc0 == 0 and rc1 == 0, ffi.last_error()
            assert (got.view(np.uint32) == ref.view(np.uint32)).all() or \
                np.array_equal(got, ref, equal_nan=True)
    finally:
        L.libfive_tape_delete(tape)


def ctypes_double():
    import ctypes
    return ctypes.byref(ctypes.c_double())


def test_noalloc_rejects_too_many_points(gpu):
    L = ffi.lib()
    shape = W.sphere(1.0)      # held: a temporary would free its tree first
    tape = L.libfive_tape_new(shape.ptr)
    assert tape, ffi.last_error()
    try:
        n = MAX_POINTS + 1
        x = np.zeros(n, np.float32)
        out = np.full(n, 7.0, np.float32)
        rc = L.libfive_tape_eval_values_noalloc(
            tape, x.ctypes.data_as(ffi.fp), x.ctypes.data_as(ffi.fp),
            x.ctypes.data_as(ffi.fp), n, out.ctypes.data_as(ffi.fp))
        assert rc != 0
        assert "at most 1048576 points" in ffi.last_error()
        assert (out == 7.0).all()
        # the library still works after the refusal
        rc, got = _call(L.libfive_tape_eval_values_noalloc, tape, x[:3], x[:3], x[:3])
        assert rc == 0 and got.tolist() == [-1.0] * 3
    finally:
        L.libfive_tape_delete(tape)
