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
cores, or the oracle restatement when that library is absent.  `serving`
covers the metric's batched half on a short cut of the serving stream (64
requests of 200 steps by default): sims/hr and completion latency with
Poisson arrivals, and the mixed / FP64 throughput ratio.
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
PEAK_FP32_VALU_TFLOPS = 157.3     # packed (dual-issue) FP32 vector peak
PEAK_FP32_UNPACKED_TFLOPS = 78.6  # one FP32 FMA per lane per cycle (this build: no packed FP32)
PEAK_FP64_VALU_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
COPY_BYTES = 1 << 30  # calibration copy: 1 GiB read + 1 GiB written per launch

# HBM bytes per dispatch of the dominant kernel from the FETCH_SIZE and
# WRITE_SIZE counter passes over `bench.py --profile-pass` (one rocprofv3
# --pmc run per counter), each scaled by what the counter reads for the 1 GiB
# calibration copy (FETCH_SIZE x 2, WRITE_SIZE x 1), warm dispatch.  Counters
# cannot be read from inside this process, so the figures are those of the
# committed summary profiles/snapshot65536_nic10_hbm_counters.txt, for that
# workload and kernel form only.  The writes are the kernel's FP64 partial
# sums, 3 x 64 blocks x N doubles = 100.66 MB; the reads are the packed
# particles (32 B each) as the block pairs fetch them, mostly from L2.
MEASURED_TRAFFIC = {("k_accel_sym32", 65536, 10): {"read": 27.530e6, "write": 100.663e6}}
# VALU issue of the same dispatch (profiles/snapshot65536_nic10_valu_counters.txt):
# SQ_INSTS_VALU over 1024 SIMDs x 2.4 GHz x duration / 4 cycles per wave64
# instruction.
MEASURED_VALU_ISSUE = {("k_accel_sym32", 65536, 10): 0.976}


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
    host = host_cpu()
    return {"value": pairs / dt, "unit": "pair-interactions/s", "cores": threads, "kind": kind,
            "host": host,
            "sample": f"one force evaluation, N={n_sample}, N_ic_max={nic}, {what}, "
                      f"{dt:.2f} s, {threads} OpenMP threads on {host['model']}"}


def host_cpu():
    """The host CPU: model, sockets x cores x threads (lscpu's figures from
    /proc/cpuinfo) and the CPUs this process may run on."""
    model, phys, cores, siblings = "unknown CPU", set(), 0, 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    phys.add(v)
                elif k == "cpu cores":
                    cores = int(v)
                elif k == "siblings":
                    siblings = int(v)
    except OSError:
        pass
    sockets = max(1, len(phys))
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    return {"model": model, "sockets": sockets, "cores_per_socket": cores,
            "threads_per_core": (siblings // cores) if cores else None,
            "physical_cores": sockets * cores, "cpus_available_to_process": share}


def default_cpu_threads():
    """OpenMP threads of the CPU baseline: OMP_NUM_THREADS when set (the
    host's share for this process), else every CPU the process may use."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def serve_bench(n, precision, realtime, steps, rate, graphs=False):
    """The batched serving stream (Poisson arrivals, continuous batching)."""
    from rumdeed_amd.serve import Server, stream_requests

    reqs = stream_requests(n=n, rate=rate, steps=steps, regime="intended")
    srv = Server(total_particles=1 << 23, max_slots=512, hbm_budget=96 << 30,
                 precision=precision, graphs=graphs)
    done = []

    def progress(res, k):
        if k % 16 == 0 or k == n:
            print(f"[serve] {k}/{n} done, last latency {res.latency:.2f} s", file=sys.stderr,
                  flush=True)
        done.append(res)

    def on_step(k, n_active):
        if k % 200 == 0:
            print(f"[serve] pool step {k}, {n_active} active", file=sys.stderr, flush=True)

    results, stats = srv.run(reqs, realtime=realtime, progress=progress, on_step=on_step)
    srv.close()
    emitted = np.array([r.emitted for r in results])
    nfin = np.array([r.n_final for r in results])
    stats.update({"precision": precision, "graphs": bool(graphs), "realtime_arrivals": realtime,
                  "steps_per_request": steps,
                  "arrival_rate_per_s": rate,
                  "emitted_total_median": float(np.median(emitted)),
                  "emitted_total_max": int(emitted.max()),
                  "electrons_final_median": float(np.median(nfin)),
                  "electrons_final_max": int(nfin.max()),
                  "dropped_total": int(sum(r.dropped for r in results))})
    return stats


E2E = {
    # intended-regime variants of the end-to-end configurations
    "planar_fe": dict(kind="planar", V=8000.0, d=1000e-9, side=100e-9, phi=4.7, dt=0.25e-15,
                      steps=4000, seed=1, nic=0, mode=10),
    "tip_fe": dict(kind="tip", V=1000.0, d_tip=1000e-9, R_base=250e-9, h_tip=500e-9,
                   box_top=1500e-9, phi=4.5, dt=0.25e-15, steps=4000, seed=1, mode=3),
    "checker_gtf": dict(kind="planar", V=6000.0, d=1000e-9, side=1000e-9, T=1500.0, dt=0.25e-15,
                        steps=2000, seed=7, nic=0, mode=9),
}


def e2e_bench(name, steps, precision, graphs=False, quad_levels=0):
    from rumdeed_amd.diode import Diode

    cfg = dict(E2E[name])
    if steps:
        cfg["steps"] = steps
    dd = Diode(1 << 18)
    dd.set_graphs(graphs)
    if cfg["kind"] == "planar":
        dd.set_planar(cfg["V"], cfg["d"], True, cfg["nic"])
    else:
        dd.set_tip(cfg["V"], cfg["d_tip"], cfg["R_base"], cfg["h_tip"], True,
                   box_top=cfg["box_top"])
    dd.set_time_step(cfg["dt"])
    dd.set_seed(cfg["seed"])
    if cfg["mode"] == 10:
        dd.set_emitter(10, 0, -cfg["side"] / 2, -cfg["side"] / 2, cfg["side"], cfg["side"],
                       wf=np.array([[cfg["phi"]]]), quad_levels=quad_levels)
    elif cfg["mode"] == 9:
        wf = np.where((np.arange(10)[:, None] + np.arange(10)[None, :]) % 2 == 0, 4.0, 4.5)
        dd.set_emitter(9, 0, -cfg["side"] / 2, -cfg["side"] / 2, cfg["side"], cfg["side"],
                       T=cfg["T"], wf=wf, quad_levels=quad_levels)
    else:
        dd.set_emitter(3, 0, w_tip=cfg["phi"])
    t0 = time.perf_counter()
    recs = []
    chunk = 100
    done = 0
    pairs = 0
    while done < cfg["steps"]:
        k = min(chunk, cfg["steps"] - done)
        rr = dd.run(k, precision)
        recs += rr
        pairs += sum(r.n * (r.n - 1) for r in rr)
        done += k
        print(f"[e2e {name}] step {done}: n={rr[-1].n}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    half = recs[len(recs) // 2:]
    force_err = None
    if dd.n > 1:
        # the final snapshot: mixed-precision accelerations against plain FP64
        a64 = dd.calculate_acceleration_particles("fp64")
        a32 = dd.calculate_acceleration_particles("mixed")
        rel = np.linalg.norm(a32 - a64, axis=1) / np.linalg.norm(a64, axis=1)
        force_err = {"n": int(dd.n), "max_rel": float(rel.max()),
                     "rms_rel": float(np.sqrt(np.mean(rel ** 2)))}
    out = {"workload": name, "precision": precision, "graphs": bool(graphs),
           "supply_quad_levels": quad_levels,
           "steps": cfg["steps"],
           "wall_s": wall, "ms_per_step": wall / cfg["steps"] * 1e3,
           "force_pair_interactions_per_s": pairs / wall,
           "electrons_final": recs[-1].n, "electrons_max": max(r.n for r in recs),
           "emitted_total": sum(r.emitted for r in recs),
           "supply_levels": {str(k): int(v) for k, v in
                             zip(*np.unique([r.supply_level for r in recs], return_counts=True))},
           "absorbed_total": sum(r.absorbed for r in recs),
           "ramo_current_second_half_A": float(np.mean([r.ramo_current for r in half])),
           "final_accel_mixed_vs_fp64": force_err,
           "config": {k: v for k, v in cfg.items()}}
    dd.close()
    return out


def ensemble_bench(name, seeds, steps):
    """Ensemble-mean current of an end-to-end configuration (the statistical
    parity criterion: current within 1%): `seeds` simulations of one pool
    per precision, `steps` steps each; the mean Ramo current over the second
    half of each run, its mean over seeds and the standard error."""
    from rumdeed_amd.ensemble import DiodeSpec, run_ensemble, steady_current

    cfg = dict(E2E[name])
    if steps:
        cfg["steps"] = steps
    if cfg["kind"] == "planar":
        wf = [[cfg.get("phi", 4.7)]]
        if cfg["mode"] == 9:
            wf = np.where((np.arange(10)[:, None] + np.arange(10)[None, :]) % 2 == 0,
                          4.0, 4.5).tolist()
        spec = DiodeSpec(kind="planar", mode=cfg["mode"], V=cfg["V"], d=cfg["d"],
                         side=cfg["side"], wf=wf, T=cfg.get("T", 300.0), dt=cfg["dt"],
                         nic=cfg["nic"])
    else:
        spec = DiodeSpec(kind="tip", mode=cfg["mode"], V=cfg["V"], d_tip=cfg["d_tip"],
                         R_base=cfg["R_base"], h_tip=cfg["h_tip"], phi=cfg["phi"],
                         box_top=cfg["box_top"], dt=cfg["dt"])
    out = {"workload": name, "seeds": seeds, "steps": cfg["steps"]}
    cur = {}
    for prec in ("mixed", "fp64"):
        t0 = time.perf_counter()

        def progress(k, prec=prec):
            print(f"[ensemble {name} {prec}] step {k}", file=sys.stderr, flush=True)

        recs, _ = run_ensemble(spec, range(1, seeds + 1), cfg["steps"], prec,
                               capacity=1 << 16, progress=progress)
        wall = time.perf_counter() - t0
        cur[prec] = np.array([steady_current(r) for r in recs])
        out[prec] = {"wall_s": wall, "I_mean_A": float(cur[prec].mean()),
                     "I_se_A": float(cur[prec].std(ddof=1) / np.sqrt(seeds)),
                     "I_per_seed_A": cur[prec].tolist(),
                     "electrons_final_mean": float(np.mean([r[-1]["n"] for r in recs])),
                     "emitted_per_seed": [int(sum(x["emitted"] for x in r)) for r in recs]}
    d = cur["mixed"] - cur["fp64"]
    out["rel_diff_mixed_vs_fp64"] = float(d.mean() / cur["fp64"].mean())
    out["rel_se_of_diff"] = float(d.std(ddof=1) / np.sqrt(seeds) / cur["fp64"].mean())
    out["within_1pct"] = bool(abs(out["rel_diff_mixed_vs_fp64"]) <= 0.01)
    out["mixed_over_fp64_wall"] = out["fp64"]["wall_s"] / out["mixed"]["wall_s"]
    return out


def serving_summary(n, steps):
    """The metric's batched half on a short cut of the serving stream: n
    requests of `steps` steps.  Mixed precision with Poisson arrivals at 4/s
    (sims/hr and completion latency); then every request admitted at once in
    both precisions, for the throughput ratio."""
    rt = serve_bench(n, "mixed", True, steps, 4.0)
    tm = serve_bench(n, "mixed", False, steps, 4.0)
    t64 = serve_bench(n, "fp64", False, steps, 4.0)
    return {"requests": n, "steps_per_request": steps, "arrival_rate_per_s": 4.0,
            "mixed_poisson": {k: rt[k] for k in ("sims_per_hour", "latency_p50_s",
                                                  "latency_p99_s", "makespan_s", "max_active")},
            "sims_per_hour_all_admitted": {"mixed": tm["sims_per_hour"],
                                           "fp64": t64["sims_per_hour"]},
            "mixed_over_fp64": tm["sims_per_hour"] / t64["sims_per_hour"]}


def profile_pass(n, nic, G, reps):
    """A short run for hardware-counter passes: `reps` launches of each
    snapshot kernel, then a plain 1 GiB device-to-device copy (three times)
    whose FETCH_SIZE / WRITE_SIZE readings calibrate the counters."""
    from rumdeed_amd import workloads as W
    from rumdeed_amd._lib import check
    from rumdeed_amd.diode import Diode

    pos, q, m = W.snapshot(n, seed=2)
    dd = Diode(n + 64)
    dd.set_planar(W.SNAPSHOT_V, W.SNAPSHOT_D, True, nic)
    dd.put_particles(pos, q, m)
    box = W.SNAPSHOT_BOX
    for prec in (1, 0):
        for _ in range(reps if prec == 1 else 1):
            check(dd._lib.rd_accel(dd._h, prec))
            check(dd._lib.rd_field_grid(dd._h, box[0], box[1], box[2], box[3], G, G, 0.0, prec,
                                        None))
    print(f"[profile] snapshot kernels done: n={n} nic={nic} grid={G} reps={reps}", flush=True)
    ms = dd.time_copy(COPY_BYTES, reps=2)
    dd.close()  # synchronises the context's stream before freeing
    print(f"[profile] calibration copy k_copy16 done: {COPY_BYTES} bytes read and written per "
          f"launch, {ms:.3f} ms, {2 * COPY_BYTES / ms / 1e6:.0f} GB/s", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--profile-pass", action="store_true",
                    help="short run of the snapshot kernels plus a 1 GiB calibration copy")
    ap.add_argument("--serve", type=int, default=0, help="serve N requests of the stream")
    ap.add_argument("--serve-steps", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-precision", default="mixed")
    ap.add_argument("--no-realtime", action="store_true")
    ap.add_argument("--graphs", action="store_true",
                    help="serving and end-to-end runs: each step's jump loop as one HIP graph")
    ap.add_argument("--e2e", default="", help="end-to-end workload: " + ", ".join(E2E))
    ap.add_argument("--e2e-steps", type=int, default=0)
    ap.add_argument("--e2e-precision", default="mixed")
    ap.add_argument("--e2e-quad-levels", type=int, default=0,
                    help="refinement doublings of the planar supply rule (-1: the fixed rule)")
    ap.add_argument("--ensemble", default="",
                    help="ensemble-mean current, mixed vs FP64: " + ", ".join(E2E))
    ap.add_argument("--ensemble-seeds", type=int, default=8)
    ap.add_argument("--ensemble-steps", type=int, default=0)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--nic", type=int, default=10)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="particles of the CPU baseline's force evaluation (0: the workload's)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-summary", type=int, default=64,
                    help="requests of the short serving cut in the default line (0: none)")
    ap.add_argument("--serve-summary-steps", type=int, default=2000,
                    help="steps per request of the serving cut (the serving config's 2000)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.profile_pass:
        profile_pass(args.n, args.nic, args.grid, args.steps)
        return
    if args.serve:
        print(json.dumps({"serving": serve_bench(args.serve, args.serve_precision,
                                                 not args.no_realtime, args.serve_steps,
                                                 args.serve_rate, args.graphs)}), flush=True)
        return
    if args.e2e:
        print(json.dumps({"e2e": e2e_bench(args.e2e, args.e2e_steps, args.e2e_precision,
                                           args.graphs, args.e2e_quad_levels)}), flush=True)
        return
    if args.ensemble:
        print(json.dumps({"ensemble": ensemble_bench(args.ensemble, args.ensemble_seeds,
                                                     args.ensemble_steps)}), flush=True)
        return

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
    far = W.far_expansion_used(pos, W.SNAPSHOT_D, nic)
    tot_ms, ker_ms = dd.time_accel("mixed", reps=5)
    flops = W.force_flops(n, nic, "mixed", far_expansion=far)
    kernel = "k_accel_sym32" if n >= W.SYM_MIN else "k_accel_planar32"
    nic_pair = min(nic, 1) if far else nic
    achieved = flops / (ker_ms * 1e-3) / 1e12
    # like for like: the same accelerations by the reference's algorithm in
    # FP64 and in mixed precision, then the pair-symmetric form, then the far
    # image partners as an expansion (the default)
    forms = {}
    for name, opts in (("mixed_gather_pairwise", dict(no_symmetric=True, no_far_expansion=True)),
                       ("mixed_symmetric_pairwise", dict(no_far_expansion=True)),
                       ("mixed_symmetric_far_expansion", {})):
        dd.set_force_options(**opts)
        forms[name] = dd.time_accel("mixed", reps=3)[0]
    dd.set_force_options()
    ftot, fker = dd.time_field_grid(*box, G, G, 0.0, "mixed", reps=3)
    t64, k64 = dd.time_accel("fp64", reps=2)
    copy_ms = dd.time_copy(COPY_BYTES, reps=3)

    value = interactions / (ms_mixed * 1e-3)
    traffic = MEASURED_TRAFFIC.get((kernel, n, nic))
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
                 "achieved_tflops": W.force_flops(n, nic, "fp64") / (k64 * 1e-3) / 1e12},
        "mixed_over_fp64": ms_fp64 / ms_mixed,
        "accel_forms_ms": dict(forms, fp64_gather_pairwise=t64),
        "accel_speedup_factors": {
            "precision (fp64 -> mixed, same algorithm)": t64 / forms["mixed_gather_pairwise"],
            "pair-symmetric form": forms["mixed_gather_pairwise"] /
                                   forms["mixed_symmetric_pairwise"],
            "far partners as an FP64 expansion": forms["mixed_symmetric_pairwise"] /
                                                 forms["mixed_symmetric_far_expansion"]},
        "kernels_ms": {"accel_mixed": ker_ms, "accel_mixed_total": tot_ms,
                       "field_grid_mixed": fker, "field_grid_mixed_total": ftot,
                       "accel_fp64": k64},
        "hbm_copy_gbs": 2 * COPY_BYTES / (copy_ms * 1e-3) / 1e9,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_VALU_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VALU_TFLOPS,
                     "frac_of_unpacked_fp32_peak": achieved / PEAK_FP32_UNPACKED_TFLOPS,
                     "traffic": traffic and traffic["read"] + traffic["write"],
                     "traffic_read_write": traffic,
                     "traffic_unit": "bytes per launch (HBM read + write, counter passes)",
                     "valu_issue_frac": MEASURED_VALU_ISSUE.get((kernel, n, nic)),
                     "kernel": kernel,
                     "flops_per_launch": flops,
                     "image_partners_in_kernel": 1 + 4 * nic_pair,
                     "far_partners_as_expansion": bool(far),
                     "flops_per_interaction": W.pair_flops(nic_pair),
                     "pairs_per_launch": n * (n - 1) // (2 if kernel == "k_accel_sym32" else 1)},
    }
    dd.close()
    if args.serve_summary > 0:
        out["serving"] = serving_summary(args.serve_summary, args.serve_summary_steps)
    if not args.no_cpu:
        threads = args.cpu_threads or default_cpu_threads()
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_sample or n, nic, threads)
        except Exception as e:  # reported, never substituted for the GPU number
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
