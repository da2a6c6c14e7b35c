"""Benchmark: object-steps/s of the full-force 70x70 RKF78 propagation.

Workload: 16,384 LEO/MEO objects (seed 2 population), 70x70 synthetic field,
NRLMSISE-00 drag, SRP with Earth shadow, Sun/Moon/8 planetary barycenters,
solid + ocean tides, relativity; 60 s fixed step over one day. One "step"
is one RKF78 integration step of the whole batch (13 force evaluations per
object). Inputs are device-resident when the timed region starts.

Prints one JSON line (rank 0, one GPU). --gpus N > 1 exits: this build runs on
one GPU and the path shards by object slots.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "object-steps/sec full-force 70x70 FP64 vs mixed; p99 request latency; mm error"
FP64_PEAK_TFLOPS = 78.6
# Algorithmic FLOPs per object-step of the reference algorithm at degree L
# (FMA = 2, div/sqrt = 1): F = 13 (F_SH(L) + F_other) + F_RK with
# F_SH(L) = 29 (L+1)(L+2)/2 + 15 (L+1) + 300, F_other = 1465 (central, 10 third
# bodies, SRP with shadow, relativity, degree-4 and degree-6 tide gradients;
# NRLMSISE-00 not counted), F_RK = 800.
F_OTHER = 1465.0
F_RK = 800.0


def flops_per_object_step(L):
    f_sh = 29.0 * (L + 1) * (L + 2) / 2 + 15.0 * (L + 1) + 300.0
    return 13.0 * (f_sh + F_OTHER) + F_RK


def cpu_baseline(cfg, t, sample_objects, sample_seconds, threads):
    import oracle
    idx = np.linspace(0, t["n"] - 1, sample_objects).astype(int)
    sub = {k: (v[idx] if k not in ("state", "n") else v) for k, v in t.items()}
    sub["state"] = np.ascontiguousarray(t["state"][:, idx])
    sub["n"] = len(idx)
    o = oracle.Oracle(cfg)
    e, s, st = o.propagate(sub, sample_seconds, threads=threads)
    return sub, s, st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--objects", type=int, default=16384)
    ap.add_argument("--degree", type=int, default=70)
    ap.add_argument("--waves", type=int, default=0, help="waves per 64-object tile (0: library default)")
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-objects", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=3600.0)
    ap.add_argument("--workdir", default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the propagation shards by object slots with no exchange; "
              "multi-GPU runs are not part of this build (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import parahpop_amd
    from parahpop_amd import synth

    work = args.workdir or tempfile.mkdtemp(prefix="phx_bench_")
    duration = 86400.0
    cfg, tg = synth.make_workload(work, 2, n=args.objects, duration=duration, degree=args.degree)
    t = synth.load_targets(tg)
    ctx = parahpop_amd.Context(cfg)
    if args.waves:
        ctx.set_option("waves", args.waves)
    if args.steps_per_launch:
        ctx.set_option("steps_per_launch", args.steps_per_launch)
    b = ctx.batch(t, duration)
    day_steps = int(round(duration / 60.0))

    pos = [0]

    def run(k):
        stats = []
        while k > 0:
            if pos[0] >= day_steps:
                b.reset()
                pos[0] = 0
            c = min(k, day_steps - pos[0])
            stats.append(b.run(c))
            pos[0] += c
            k -= c
        return stats

    if args.warmup:
        run(args.warmup)
    b.reset()
    pos[0] = 0
    t0 = time.perf_counter()
    stats = run(args.steps)
    wall = time.perf_counter() - t0
    obj_steps = sum(s["object_steps"] for s in stats)
    kern = sum(s["kernel_seconds"] for s in stats)
    value = obj_steps / wall
    fps = flops_per_object_step(args.degree)
    launches = sum(s["launches"] for s in stats)
    achieved = fps * obj_steps / kern / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "object-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded population, synthetic 70x70 field, constant space weather/EOP, analytic ephemeris)",
        "config": {"workload": "full-force 70x70, 16384 LEO/MEO objects, 60 s step, 1 day",
                   "objects": args.objects, "degree": args.degree, "waves": stats[0]["waves"],
                   "steps_per_launch": stats[0]["steps_per_launch"]},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel_ms_avg": 1e3 * kern / max(launches, 1), "flops_per_object_step": fps},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        sub, s_cpu, st = cpu_baseline(cfg, t, args.cpu_objects, args.cpu_seconds, threads)
        e_g, s_g, _ = ctx.propagate(sub, args.cpu_seconds)
        err_mm = float(np.max(np.linalg.norm(s_g[:3] - s_cpu[:3], axis=0)) * 1e6)
        line["cpu_baseline"] = {"value": st["object_steps"] / st["seconds"], "unit": "object-steps/s", "cores": threads,
                                "kind": "port",
                                "sample": "%d evenly spaced objects of the same population, %.0f s (%d object-steps)"
                                          % (sub["n"], args.cpu_seconds, int(st["object_steps"]))}
        line["error_mm_vs_cpu"] = err_mm
    print(json.dumps(line))


if __name__ == "__main__":
    main()
