"""Benchmark of the RUMDEED per-step particle hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--nic NIC]

Workload (the force/field snapshot): N = 65536 electrons uniform in a
200 x 200 x 1000 nm box above a planar cathode, d = 1000 nm, image series
N_ic_max = 10 (41 partners per pair), surface field on the 256 x 256 cell
centres at z = 0.  One step = all-pairs accelerations + surface field, with
the particle state resident in HBM.  The headline value is pair-interactions/s
of the mixed-precision path (N(N-1) + M N interactions per step); the plain
FP64 path is measured beside it.

Prints one JSON line.  The roofline object is for the dominant kernel (the
mixed all-pairs acceleration kernel), timed with HIP events on the stream it
is launched on.  cpu_baseline times the reference's own OpenMP pair loop
(compiled from its sources into oracle/_ref) on a bounded sample on the host
cores, or the oracle restatement when that library is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_FP32_VALU_TFLOPS = 157.3
PEAK_FP64_VALU_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def cpu_baseline(n_sample: int, nic: int, threads: int):
    """Reference OpenMP planar pair loop (Calculate_Acceleration_Particles_Planar,
    src/mod_verlet.F90:763-884) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from rumdeed_amd import workloads as W

    pos, q, m = W.snapshot(n_sample, seed=2)
    if O.Ref.available():
        R = O.Ref.get(cap=n_sample + 16)
        R.threads(threads)
        R.setup_planar(W.SNAPSHOT_D, W.SNAPSHOT_V, image_charge=True, n_ic_max=nic)
        R.put_snapshot(pos, q, m)
        t0 = time.perf_counter()
        R.accel(0, n_sample)
        dt = time.perf_counter() - t0
        kind = "reference"
        what = "Calculate_Acceleration_Particles_Planar (flang -O3 -fopenmp)"
    else:
        g = O.geom_planar(W.SNAPSHOT_D, W.SNAPSHOT_V, True, nic)
        t0 = time.perf_counter()
        O.accel(g, pos, q, m, nthreads=threads)
        dt = time.perf_counter() - t0
        kind = "port"
        what = "oracle FP64 gather restatement (gcc -O2 -fopenmp)"
    pairs = n_sample * (n_sample - 1)
    return {"value": pairs / dt, "unit": "pair-interactions/s", "cores": threads, "kind": kind,
            "sample": f"one force evaluation, N={n_sample}, N_ic_max={nic}, {what}, "
                      f"{dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--nic", type=int, default=10)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=16384)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from rumdeed_amd import workloads as W
    from rumdeed_amd.diode import Diode

    n, nic, G = args.n, args.nic, args.grid
    M = G * G
    pos, q, m = W.snapshot(n, seed=2)
    dd = Diode(n + 64)
    dd.set_planar(W.SNAPSHOT_V, W.SNAPSHOT_D, True, nic)
    dd.put_particles(pos, q, m)
    box = W.SNAPSHOT_BOX
    lib, h = dd._lib, dd._h
    from rumdeed_amd._lib import check

    def step(prec):
        check(lib.rd_accel(h, prec))
        check(lib.rd_field_grid(h, box[0], box[1], box[2], box[3], G, G, 0.0, prec, None))

    def timed(prec, steps, warmup):
        for _ in range(warmup):
            step(prec)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(prec)
        return (time.perf_counter() - t0) / steps

    interactions = n * (n - 1) + M * n
    ms_mixed = timed(1, args.steps, args.warmup) * 1e3
    ms_fp64 = timed(0, args.fp64_steps, 1) * 1e3

    # dominant kernel: mixed all-pairs accelerations; HIP events on its stream
    tot_ms, ker_ms = dd.time_accel("mixed", reps=5)
    flops = n * (n - 1) * W.pair_flops(nic)
    achieved = flops / (ker_ms * 1e-3) / 1e12
    ftot, fker = dd.time_field_grid(*box, G, G, 0.0, "mixed", reps=3)
    t64, k64 = dd.time_accel("fp64", reps=2)

    value = interactions / (ms_mixed * 1e-3)
    out = {
        "metric": "pair-interactions/s (FP32-mixed vs FP64)",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 accumulation",
        "data": "synthetic (Philox seed 2 positions, uniform box)",
        "config": {"workload": "force/field snapshot: all-pairs Coulomb + planar image series "
                               "+ 256x256 surface field",
                   "n_electrons": n, "n_ic_max": nic, "image_partners": 1 + 4 * nic,
                   "surface_points": M, "gap_nm": 1000, "box_nm": [200, 200, 1000],
                   "interactions_per_step": interactions},
        "fp64": {"value": interactions / (ms_fp64 * 1e-3), "ms_per_step": ms_fp64,
                 "accel_kernel_ms": k64,
                 "achieved_tflops": flops / (k64 * 1e-3) / 1e12},
        "mixed_over_fp64": ms_fp64 / ms_mixed,
        "kernels_ms": {"accel_mixed": ker_ms, "accel_mixed_total": tot_ms,
                       "field_grid_mixed": fker, "field_grid_mixed_total": ftot,
                       "accel_fp64": k64},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VALU_TFLOPS,
                     "traffic": None,
                     "kernel": "k_accel_planar32",
                     "flops_per_launch": flops,
                     "flops_per_interaction": W.pair_flops(nic)},
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_sample, nic, threads)
        except Exception as e:  # reported, never substituted for the GPU number
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    dd.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
