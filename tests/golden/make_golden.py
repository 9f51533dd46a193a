"""Writes tests/golden/ieee_ops.npz: inputs and expected float32 outputs of
the IEEE-exact opcodes, computed with numpy only (no code of this project
and no reference code): one correctly rounded float32 operation per
clause, min/max as std::min/std::max ((b < a) ? b : a and (a < b) ? b : a,
libfive/include/libfive/eval/interval.hpp:85-93), and sin/cos rounded from
float64.

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(0x11B5F1)
    special = np.array([0.0, -0.0, 1.0, -1.0, 0.5, -2.5, 3.0, 1e-30, -1e-30,
                        1e-45, 3.4e38, -3.4e38, np.inf, -np.inf, np.nan,
                        2.0 ** -126, 1.0 + 2.0 ** -23],
                       np.float32)
    rand = np.concatenate([
        rng.uniform(-4, 4, 2000),
        rng.standard_normal(1000) * 10.0 ** rng.integers(-20, 20, 1000),
    ]).astype(np.float32)
    a = np.concatenate([np.repeat(special, special.size),
                        rand]).astype(np.float32)
    b = np.concatenate([np.tile(special, special.size),
                        rng.permutation(rand)]).astype(np.float32)
    with np.errstate(all="ignore"):
        out = {
            "a": a, "b": b,
            "add": a + b, "sub": a - b, "mul": a * b, "div": a / b,
            "min": np.where(b < a, b, a).astype(np.float32),
            "max": np.where(a < b, b, a).astype(np.float32),
            "neg": -a, "abs": np.abs(a), "square": a * a,
            "sqrt": np.sqrt(a),
        }
        # sin / cos arguments: moderate, large and special values.
        xs = np.concatenate([
            rng.uniform(-10, 10, 4000),
            rng.uniform(-1e5, 1e5, 2000),
            rng.uniform(-1e30, 1e30, 500),
            np.array([0.0, -0.0, np.pi, -np.pi, np.pi / 2, 1e-20, 3.0e38,
                      524287.9, 524288.0, 1e7, 8.0e37]),
        ]).astype(np.float32)
        out["trig_x"] = xs
        out["sin_rn"] = np.sin(xs.astype(np.float64)).astype(np.float32)
        out["cos_rn"] = np.cos(xs.astype(np.float64)).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "ieee_ops.npz"), **out)
    print("wrote", os.path.join(HERE, "ieee_ops.npz"), a.size, "pairs")


if __name__ == "__main__":
    main()
