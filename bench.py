"""Benchmark: tape evaluation throughput on one MI355X.

Workload (the array-evaluation configuration): the seeded random
expression of 1000 operations over {add, sub, mul, min, max, neg, abs,
square, sqrt(abs(.)), sin, cos} with leaves x, y, z and constants
U(-2, 2) (libfive_amd.workloads.random_dag), evaluated with values and
partial derivatives (the DerivArrayEvaluator pass) on the 256^3 regular
grid over [-1, 1]^3.  A step is one pass over the grid; inputs and outputs
are resident in HBM.  The metric is point-operations per second: grid
points x tape clauses / time.

Prints one JSON line.  --gpus above 1 is rejected: this build measures one
GPU (multi-GPU sharding is described in DESIGN.md, not run).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Nominal FLOPs per point for each opcode (value pass / value+gradient
# pass).  sin and cos count W_TRIG = 20: range reduction (multiply, round,
# two fused multiply-adds) plus the two 7-term polynomials evaluated by
# Horner's rule, counting a fused multiply-add as two FLOPs.
W_TRIG = 20
VALUE_FLOPS = {7: 1, 8: 1, 9: 1, 10: W_TRIG, 11: W_TRIG, 17: 1, 18: 1,
               19: 1, 20: 1, 21: 1, 22: 1, 28: 1}
DERIV_FLOPS = {7: 7, 8: 5, 9: 4, 10: W_TRIG + 3, 11: W_TRIG + 4, 17: 4,
               18: 10, 19: 1, 20: 1, 21: 4, 22: 14, 28: 4}
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
GRID = 256


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu", action="store_true",
                   help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--sweep", action="store_true",
                   help="also time the value kernel at 1, 2 and 4 points "
                        "per lane")
    return p.parse_args()


def cpu_baseline(shape, target_seconds):
    """The oracle (CPU restatement of the reference's 256-lane
    DerivArrayEvaluator) timed on this host on a bounded sample of the same
    grid: whole z-slabs of the 256^3 grid, sized to about target_seconds."""
    import oracle
    from libfive_amd import workloads as W
    threads = max(1, min(16, os.cpu_count() or 1))
    deck = shape.deck()
    T = deck[0].shape[0]
    od = oracle.OracleDeck(deck)
    x, y, z = W.grid_points(GRID)
    slab = GRID * GRID
    # calibrate on one slab
    t0 = time.perf_counter()
    od.derivs(x[:slab], y[:slab], z[:slab], threads=threads)
    dt = time.perf_counter() - t0
    n_slabs = int(max(1, min(GRID, target_seconds / max(dt, 1e-6))))
    n = n_slabs * slab
    t0 = time.perf_counter()
    od.derivs(x[:n], y[:n], z[:n], threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": n * T / dt / 1e9,
        "unit": "Gpoint-op/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n_slabs} of {GRID} z-slabs of the {GRID}^3 grid "
                  f"({n} points x {T} clauses, values+gradients, "
                  f"{dt:.1f} s)",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; "
                          "--gpus above 1 is not supported"}))
        sys.exit(2)

    import torch
    import libfive_amd
    from libfive_amd import ffi
    from libfive_amd import workloads as W

    if not libfive_amd.gpu_available():
        raise SystemExit("no usable GPU: " + ffi.last_error())
    L = ffi.lib()
    shape = W.random_dag()
    deck = shape.deck()
    T = int(deck[0].shape[0])
    ops = [int(c[0]) for c in deck[0]]
    vflops = sum(VALUE_FLOPS.get(o, 1) for o in ops)
    dflops = sum(DERIV_FLOPS.get(o, 4) for o in ops)
    tape = L.libfive_tape_new(shape.ptr)
    if not tape:
        raise SystemExit("tape compile failed: " + ffi.last_error())
    regs = L.libfive_tape_registers(tape)

    dev = torch.device("cuda:0")
    x, y, z = W.grid_points(GRID)
    n = x.size
    xd = torch.from_numpy(x).to(dev)
    yd = torch.from_numpy(y).to(dev)
    zd = torch.from_numpy(z).to(dev)
    out = torch.empty((4, n), dtype=torch.float32, device=dev)
    vout = torch.empty(n, dtype=torch.float32, device=dev)
    # A dedicated stream (default priority): the kernels and the timing
    # events must be on the same stream.
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream

    def deriv_step():
        rc = L.libfive_tape_eval_derivs_device(
            tape, xd.data_ptr(), yd.data_ptr(), zd.data_ptr(), n,
            out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
            out[3].data_ptr(), sp)
        if rc:
            raise RuntimeError(ffi.last_error())

    def value_step():
        rc = L.libfive_tape_eval_values_device(
            tape, xd.data_ptr(), yd.data_ptr(), zd.data_ptr(), n,
            vout.data_ptr(), sp)
        if rc:
            raise RuntimeError(ffi.last_error())

    def timed(step, k, w):
        for _ in range(w):
            step()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True),
                torch.cuda.Event(enable_timing=True)) for _ in range(k)]
        t0 = time.perf_counter()
        for a, b in evs:
            a.record(stream)
            step()
            b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern = [a.elapsed_time(b) for a, b in evs]
        return wall, float(np.mean(kern))

    wall, kern_ms = timed(deriv_step, args.steps, args.warmup)
    v_wall, v_kern_ms = timed(value_step, args.steps, args.warmup)
    sweep = {}
    if args.sweep:
        for ppl in (1, 2, 4):
            L.libfive_set_points_per_lane(ppl)
            w_, k_ = timed(value_step, args.steps, args.warmup)
            sweep[str(ppl)] = round(n * T / (k_ * 1e-3) / 1e9, 2)
        L.libfive_set_points_per_lane(0)

    ms_per_step = wall * 1e3 / args.steps
    value = n * T / (wall / args.steps) / 1e9
    achieved = n * dflops / (kern_ms * 1e-3) / 1e12
    result = {
        "metric": "Gpoint-op/s tape eval (values + partial derivatives)",
        "value": round(value, 2),
        "unit": "Gpoint-op/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random 1000-op expression, 256^3 grid)",
        "config": {
            "workload": "random 1000-op DAG, DerivArrayEvaluator pass, "
                        "256^3 grid on [-1,1]^3",
            "points": n, "tape_clauses": T, "registers": int(regs),
            "values_only_gpoint_op_s": round(
                n * T / (v_wall / args.steps) / 1e9, 2),
            "values_only_kernel_ms": round(v_kern_ms, 4),
            "values_only_by_points_per_lane": sweep or None,
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": None,
            "kernel": "k_eval_derivs",
            "kernel_ms": round(kern_ms, 4),
            "flops_per_point": dflops,
        },
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(shape, args.cpu_seconds)
    L.libfive_tape_delete(tape)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
