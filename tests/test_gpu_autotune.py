"""Launch-schedule autotuner: every schedule a kernel offers gives the same
bits, so the tuner's choice never changes a result.  Each test runs every
schedule (forced through libfive_autotune_force), then lets the tuner time
them and checks that the decision is reported and the result unchanged."""
import numpy as np
import pytest

import oracle
from libfive_amd import autotune_report, ffi
from libfive_amd import workloads as W

from helpers import assert_same_bits

pytestmark = pytest.mark.gpu

N_TUNED = (1 << 20) + 4099     # above the evaluation calls' tuning threshold


def _eval(tape, p, derivs):
    L = ffi.lib()
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in p)
    n = x.size
    f = lambda a: a.ctypes.data_as(ffi.fp)
    if derivs:
        out = [np.empty(n, np.float32) for _ in range(4)]
        rc = L.libfive_tape_eval_derivs(tape, f(x), f(y), f(z), n, *[f(o) for o in out])
    else:
        out = [np.empty(n, np.float32)]
        rc = L.libfive_tape_eval_values(tape, f(x), f(y), f(z), n, f(out[0]))
    assert rc == 0, ffi.last_error()
    return out


def _keys(prefix):
    return [d["key"] for d in autotune_report() if d["key"].startswith(prefix)]


def test_evaluation_schedules_give_the_same_bits(gpu):
    """Points per lane of the value interpreter (1, 2, 4) and workgroup
    sizes of the specialised value and derivative kernels (256, 128, 64),
    on 1M+ points of the gyroid (one register bank)."""
    L = ffi.lib()
    s = W.gyroid_c4()
    rng = np.random.default_rng(11)
    p = rng.uniform(-1.2, 1.2, (3, N_TUNED)).astype(np.float32)
    tape = L.libfive_tape_new(s.ptr)
    assert tape, ffi.last_error()
    try:
        outs = []
        for c in range(3):
            L.libfive_autotune_reset()
            L.libfive_autotune_force(b"eval_values/banks=1", c)
            print("interpreter schedule", c, flush=True)
            outs.append(_eval(tape, p, False)[0])
        for c in (1, 2):
            assert_same_bits(outs[c], outs[0], f"points-per-lane schedule {c}")
        L.libfive_autotune_reset()
        tuned = _eval(tape, p, False)[0]
        assert _keys("eval_values/banks=1"), autotune_report()
        assert_same_bits(tuned, outs[0], "tuned interpreter")
        ref = oracle.OracleDeck(s.deck()).values(*p[:, :20000])
        assert_same_bits(outs[0][:20000], ref, "interpreter vs oracle")

        assert L.libfive_tape_specialize(tape, None) == 0, ffi.last_error()
        deriv_ref = None
        for derivs, kernel in ((False, "lf_values"), (True, "lf_derivs")):
            L.libfive_autotune_reset()
            _eval(tape, p, derivs)                       # tunes
            keys = _keys(kernel)
            assert len(keys) == 1, autotune_report()
            for c in range(3):
                L.libfive_autotune_force(keys[0].encode(), c)
                print(kernel, "schedule", c, flush=True)
                got = _eval(tape, p, derivs)
                if not derivs:
                    assert_same_bits(got[0], outs[0], f"{kernel} schedule {c}")
                elif deriv_ref is None:
                    deriv_ref = got
                else:
                    for k in range(4):
                        assert_same_bits(got[k], deriv_ref[k], f"{kernel} schedule {c}")
    finally:
        L.libfive_tape_delete(tape)
        L.libfive_autotune_reset()


def test_leaf_schedules_give_the_same_mesh(gpu):
    """Waves per workgroup (1 or 4) of the leaf corner and edge kernels:
    the gyroid at 128^3 and the 60-sphere union give identical meshes
    under each schedule and under the tuner's choice."""
    L = ffi.lib()
    for shape, res in ((W.gyroid_c4(), 128 / 2.4), (W.sphere_union(60), 96 / 2.4)):
        L.libfive_autotune_reset()
        base = shape.render_mesh([-1.2] * 3, [1.2] * 3, res=res)     # tunes
        keys = _keys("leaf_")
        assert len(keys) == 2, autotune_report()
        for c in (0, 1):
            L.libfive_autotune_reset()
            for k in keys:
                L.libfive_autotune_force(k.encode(), c)
            print("leaf schedule", c, keys, flush=True)
            v, t = shape.render_mesh([-1.2] * 3, [1.2] * 3, res=res)
            assert v.shape == base[0].shape and t.shape == base[1].shape
            assert (v.view(np.uint32) == base[0].view(np.uint32)).all()
            assert (t == base[1]).all()
    L.libfive_autotune_reset()
