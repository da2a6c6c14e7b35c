"""Benchmark: object-steps/s of the full-force 70x70 RKF78 propagation, FP64
vs mixed precision, uneven batches, and the serving loop's latency.

Headline (`value`): 16,384 LEO/MEO objects (seed-2 population), 70x70
synthetic field, NRLMSISE-00 drag, SRP with Earth shadow, Sun/Moon/8 planetary
barycenters, solid + ocean tides, relativity; 60 s fixed step over one day in
FP64. A "step" is one RKF78 integration step of the whole batch (13 force
evaluations per object); the timed region starts with the batch resident in
HBM.

Further legs (one JSON object each in the same line):
  mixed   16,384 objects, 120x120 Kaula-rule field, 30 s step: FP64 vs mixed
          precision throughput and the mixed result's position error against
          the FP64 run (all objects) and the CPU oracle (64-object subset)
  uneven  131,072 objects, epochs over 30 days, durations log-uniform
          10 min - 7 days (shortened final steps), autotuned tile shape,
          multi-step launches until every object has ended
  serve   Poisson arrivals of requests (1 - 50,000 objects log-uniform, 1 h -
          3 d log-uniform, uneven-batch distributions) on a mixed-precision
          70x70 model: p50/p99 latency, object-steps/s, error vs the FP64 path
  ephemeris  the headline workload with its body positions from a .brie
          Chebyshev ephemeris (synthetic circular orbits in the reference's
          barycenter chains), then also its Earth orientation from a .brot2
          IERS2000 unit, beside the analytic models: object-steps/s over 128
          steps and the epoch-record time of each
  alternate  J2-only Earth gravity and drag in a 28-band exponential
          atmosphere over the geodetic altitude (Sun/Moon, SRP with shadow,
          relativity): one day at 60 s for 16,384 and 262,144 objects, the
          error against the CPU oracle (64 objects) and the oracle's own rate

--gpus N > 1 exits: this build runs on one GPU (the propagation shards by
object slots with no exchange between GPUs).
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "object-steps/sec full-force 70x70 FP64 vs mixed; p99 request latency; mm error"
FP64_PEAK_TFLOPS = 78.6
# Algorithmic FLOPs per object-step of the per-object force model and the RKF78
# step, counted by the FLOP-counting build of the oracle (the oracle's sources
# with a counting scalar; `make -C oracle flops && python oracle/count_flops.py`):
# + - * / and sqrt count 1, so an FMA counts 2. Per SH evaluation 29 FLOPs per
# (n, m) term and 4 per degree, 13 evaluations per step, plus 26,757.4 per
# object-step for the rest (central and third bodies, SRP with the shadow cones,
# relativity, the degree-4 and degree-6 tide gradients, the drag force, the
# step) over the 64-object sample of the headline workload: F(70) = 994,061,
# F(120) = 2,815,686, of which 2,700,750 are the FP32 degrees 21-120 in mixed
# precision. Not in the figure: NRLMSISE-00 (6,589 FLOPs and 422
# transcendental calls per evaluation below the cutoff) and the epoch-only
# inputs (evaluated once per epoch here).
F_SH_TERM = 29.0
F_SH_DEGREE = 4.0
F_REST = 26757.3875
# The alternate model (J2, Sun/Moon, SRP with shadow, relativity, exponential
# drag over the geodetic altitude), counted the same way.
F_ALT = 5698.1265625


def flops_per_object_step(L):
    return 13.0 * (F_SH_TERM * (L + 1) * (L + 2) / 2 + F_SH_DEGREE * (L + 1)) + F_REST


def subset(t, idx):
    idx = np.asarray(idx)
    out = {k: (v[idx] if hasattr(v, "shape") and v.ndim == 1 and k != "state" else v) for k, v in t.items()}
    out["state"] = np.ascontiguousarray(t["state"][:, idx])
    out["n"] = len(idx)
    return out


def pos_err_mm(a, b):
    d = np.linalg.norm(a[:3] - b[:3], axis=0) * 1e6
    return float(np.sqrt(np.mean(d ** 2))), float(np.max(d))


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on,
    capped at OMP_NUM_THREADS when it is set (one GPU's share of the host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(omp))) if omp.isdigit() else max(1, n)


def host_info():
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    flags = ""
    try:
        for ln in open(os.path.join(ROOT, "oracle", "Makefile")):
            if ln.startswith("CXXFLAGS"):
                flags = ln.split("=", 1)[1].strip()
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model, "build": "g++ " + flags}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_steps(b, total, chunk):
    """Runs `total` steps in calls of at most `chunk`; returns summed stats."""
    acc = {"object_steps": 0.0, "kernel_seconds": 0.0, "launches": 0.0}
    done = 0
    while done < total:
        k = min(chunk, total - done)
        st = b.run(k)
        for key in acc:
            acc[key] += st[key]
        for key in ("waves", "steps_per_launch", "kernels_per_step"):
            acc[key] = st[key]
        for key in ("graph_replays", "graph_captures"):
            acc[key] = acc.get(key, 0.0) + st.get(key, 0.0)
        done += k
    return acc


def leg_main(args, work):
    import parahpop_amd
    from parahpop_amd import synth
    duration = 86400.0
    cfg, tg = synth.make_workload(os.path.join(work, "w2"), 2, n=args.objects, duration=duration, degree=args.degree)
    t = synth.load_targets(tg)
    ctx = parahpop_amd.Context(cfg)
    if args.waves:
        ctx.set_option("waves", args.waves)
    if args.split >= 0:
        ctx.set_option("split", args.split)
    if args.graph_steps >= 0:
        ctx.set_option("graph_steps", args.graph_steps)
    spl = args.steps_per_launch or int(ctx.option("steps_per_launch"))
    ctx.set_option("steps_per_launch", spl)
    b = ctx.batch(t, duration)
    day_steps = int(round(duration / 60.0))
    if args.warmup:
        run_steps(b, min(args.warmup, day_steps), spl)
    b.reset()
    steps = min(args.steps, day_steps)
    t0 = time.perf_counter()
    st = run_steps(b, steps, steps)
    wall = time.perf_counter() - t0
    obj_steps = st["object_steps"]
    kern = st["kernel_seconds"]
    fps = flops_per_object_step(args.degree)
    achieved = fps * obj_steps / kern / 1e12
    line = {
        "metric": METRIC,
        "value": obj_steps / wall,
        "unit": "object-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded population, synthetic 70x70 field, constant space weather/EOP, analytic ephemeris)",
        "config": {"workload": "full-force 70x70, 16384 LEO/MEO objects, 60 s step, 1 day",
                   "objects": args.objects, "degree": args.degree, "waves": st["waves"],
                   "steps_per_launch": st["steps_per_launch"], "graph_steps": int(ctx.option("graph_steps")),
                   "graph_replays": st.get("graph_replays", 0.0), "graph_captures": st.get("graph_captures", 0.0),
                   "path": "split-stage" if st["kernels_per_step"] else "fused multi-step"},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "flops_per_object_step": fps},
    }
    kps = st["kernels_per_step"]
    if kps:
        # split-stage path: the unit of work the roofline is priced on is one step
        # (13 stages of SH, species and finish kernels, then the combine kernel)
        sw, ss = int(ctx.option("split_waves")), int(ctx.option("split_species"))
        drag = {0: "splitDragKernel", 8: "splitDragTileKernel"}.get(ss, "splitSpeciesKernel<%d>, splitDragFinishKernel" % ss)
        kern_desc = "step of %d kernels: 13 x (splitShKernel<%d>, %s) + splitCombineKernel" % (kps, sw, drag)
        per, shape = steps, {"path": "split", "split_waves": sw, "split_species": ss}
    else:
        kern_desc = "launch of rkf78Kernel<%d> (%d steps)" % (st["waves"], st["steps_per_launch"])
        per, shape = st["launches"], {"path": "fused", "steps_per_launch": st["steps_per_launch"]}
    line["roofline"].update({"kernel": kern_desc, "kernel_ms_avg": 1e3 * kern / max(per, 1), "launches": per,
                             "object_steps_per_launch": obj_steps / max(per, 1)})
    traffic = os.path.join(ROOT, "profiles", "step_full_force_70x70_traffic.json")
    if os.path.exists(traffic):
        # HBM bytes per launch unit from the FETCH_SIZE / WRITE_SIZE counter runs (scaled
        # by a device copy of known size), valid for the same path and launch shape
        tr = json.load(open(traffic))
        if all(tr.get(k) == v for k, v in shape.items()) and args.objects == 16384 and args.degree == 70:
            line["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
            line["roofline"]["traffic_source"] = os.path.relpath(traffic, ROOT)
    if not args.no_cpu:
        import oracle
        threads = cpu_threads()
        # the timed batch itself: 64 of its objects after the timed steps
        e_t, s_t = b.state()
        idx64 = np.linspace(0, t["n"] - 1, 64).astype(int)
        e_o, s_o, _ = oracle.Oracle(cfg).propagate(subset(t, idx64), steps * 60.0, threads=threads)
        rms, mx = pos_err_mm(s_t[:, idx64], s_o)
        line["timed_batch_vs_oracle"] = {
            "objects": 64, "steps": steps, "pos_mm_rms": rms, "pos_mm_max": mx,
            "vel_um_s_max": float(np.max(np.linalg.norm(s_t[3:, idx64] - s_o[3:], axis=0)) * 1e9),
            "epochs_identical": bool(np.array_equal(e_t[idx64], e_o))}
        idx = np.linspace(0, t["n"] - 1, args.cpu_objects).astype(int)
        sub = subset(t, idx)
        e_o, s_o, so = oracle.Oracle(cfg).propagate(sub, args.cpu_seconds, threads=threads)
        e_g, s_g, _ = ctx.propagate(sub, args.cpu_seconds)
        line["cpu_baseline"] = {"value": so["object_steps"] / so["seconds"], "unit": "object-steps/s", "cores": threads,
                                "kind": "port",
                                "sample": "%d evenly spaced objects of the same population, %.0f s (%d object-steps)"
                                          % (sub["n"], args.cpu_seconds, int(so["object_steps"]))}
        line["cpu_baseline"].update(host_info())
        rms, mx = pos_err_mm(s_g, s_o)
        dv = float(np.max(np.linalg.norm(s_g[3:] - s_o[3:], axis=0)) * 1e9)
        line["error_vs_cpu"] = {"pos_mm_rms": rms, "pos_mm_max": mx, "vel_um_s_max": dv,
                                "epochs_identical": bool(np.array_equal(e_g, e_o))}
    return line


def leg_mixed(args, work):
    import parahpop_amd
    from parahpop_amd import synth
    duration = args.mixed_seconds
    d = os.path.join(work, "w3")
    cfg, tg = synth.make_workload(d, 3, n=args.objects, duration=duration)
    t = synth.load_targets(tg)
    doc = json.load(open(cfg))
    doc["execution"]["precision"] = {"mode": "fp64"}
    cfg64 = os.path.join(d, "CONFIG_fp64.json")
    json.dump(doc, open(cfg64, "w"))
    out = {"workload": "full-force 120x120 Kaula field, 16384 objects, 30 s step, %.0f s" % duration}
    states = {}
    for mode, path in (("fp64", cfg64), ("mixed", cfg)):
        ctx = parahpop_amd.Context(path)
        b = ctx.batch(t, duration)
        b.run(2)
        b.reset()
        t0 = time.perf_counter()
        st = b.run()
        wall = time.perf_counter() - t0
        states[mode] = b.download()[1]
        out[mode] = {"value": st["object_steps"] / wall, "kernel_seconds": st["kernel_seconds"], "waves": st["waves"]}
        log("mixed leg: %s %.3g object-steps/s" % (mode, out[mode]["value"]))
    out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    rms, mx = pos_err_mm(states["mixed"], states["fp64"])
    out["mixed_vs_fp64_pos_mm"] = {"rms": rms, "max": mx}
    if not args.no_cpu:
        import oracle
        idx = np.linspace(0, t["n"] - 1, 64).astype(int)
        e_o, s_o, _ = oracle.Oracle(cfg64).propagate(subset(t, idx), duration, threads=cpu_threads())
        rms, mx = pos_err_mm(states["mixed"][:, idx], s_o)
        out["mixed_vs_oracle_pos_mm_64obj"] = {"rms": rms, "max": mx}
        rms, mx = pos_err_mm(states["fp64"][:, idx], s_o)
        out["fp64_vs_oracle_pos_mm_64obj"] = {"rms": rms, "max": mx}
    return out


def leg_uneven(args, work):
    import parahpop_amd
    from parahpop_amd import synth
    n = args.uneven_objects
    cfg, tg = synth.make_workload(os.path.join(work, "w4"), 4, n=n, duration=604800.0)
    t = synth.load_targets(tg)
    rng = np.random.default_rng(44)
    dur = np.exp(rng.uniform(math.log(600.0), math.log(604800.0), n))
    ctx = parahpop_amd.Context(cfg, steps_per_launch=args.uneven_steps_per_launch)
    if args.graph_steps >= 0:
        ctx.set_option("graph_steps", args.graph_steps)
    b = ctx.batch(t, dur)
    tune = b.autotune(4)
    # run to completion in ten chunks of steps, to see the tail's step time
    total = int(np.ceil(dur.max() / 60.0))
    chunk = (total + 9) // 10
    chunks = []
    t0 = time.perf_counter()
    while not b.done():
        st = b.run(chunk)
        chunks.append(st)
    wall = time.perf_counter() - t0
    e, s = b.download()
    end_err = float(np.max(np.abs((e - t["epoch"]) * 86400.0 - dur)))
    # objects that end non-finite: re-entries (NRLMSISE-00 undefined below the
    # ground); the reference driver rejects such a run (driver.cpp:153-154), so
    # the headline rate of the leg counts only the object-steps of finite objects
    finite = np.all(np.isfinite(s), axis=0) & np.isfinite(e)
    steps_each = np.ceil(dur / 60.0)
    obj_steps = sum(c["object_steps"] for c in chunks)
    kern = sum(c["kernel_seconds"] for c in chunks)
    finite_steps = float(steps_each[finite].sum())
    def chunk_steps(c):
        return c["launches"] / c["kernels_per_step"] if c["kernels_per_step"] else chunk
    chunks = [c for c in chunks if chunk_steps(c) > 0]
    tail = chunks[-1] if chunk_steps(chunks[-1]) >= chunk // 2 else chunks[-2]
    return {"workload": "uneven: %d objects, epochs U[4035,4065], durations log-U[600 s, 7 d], 70x70 full force" % n,
            "value": finite_steps / wall, "unit": "object-steps/s of the objects that end finite",
            "objects_nonfinite": int((~finite).sum()), "object_steps_finite": finite_steps,
            "value_all_lanes": obj_steps / wall, "object_steps": obj_steps,
            "expected_object_steps": float(steps_each.sum()), "wall_seconds": wall,
            "kernel_seconds": kern, "launches": sum(c["launches"] for c in chunks), "epoch_records": chunks[0]["epoch_records"],
            "graph_replays": sum(c["graph_replays"] for c in chunks), "graph_captures": sum(c["graph_captures"] for c in chunks),
            "chunk_steps": [chunk_steps(c) for c in chunks],
            "chunk_ms_per_step": [1e3 * c["kernel_seconds"] / chunk_steps(c) for c in chunks],
            "tail_ms_per_step": 1e3 * tail["kernel_seconds"] / chunk_steps(tail),
            "tuned": {"waves": int(ctx.option("waves")), "share0": ctx.option("share0"),
                      "split_waves": int(ctx.option("split_waves")), "split_species": int(ctx.option("split_species")),
                      "path": "split-stage" if tail["kernels_per_step"] else "fused multi-step",
                      "trial_ms_per_step": 1e3 * tune["kernel_seconds"] / 4},
            "max_end_epoch_error_s": end_err}


def leg_alternate(args, work):
    import parahpop_amd
    from parahpop_amd import synth
    duration = 86400.0
    f = F_ALT
    out = {"workload": "J2-only gravity, 28-band exponential drag over the geodetic altitude, Sun/Moon, SRP with "
                       "Earth shadow, relativity; seed-2 population, 60 s step, 1 d",
           "flops_per_object_step": f}
    states = {}
    for n in (args.objects, 262144):
        cfg, tg = synth.make_alternate_workload(os.path.join(work, "walt%d" % n), n=n, duration=duration)
        t = synth.load_targets(tg)
        ctx = parahpop_amd.Context(cfg)
        b = ctx.batch(t, duration)
        b.run(2)
        b.reset()
        t0 = time.perf_counter()
        st = b.run()
        wall = time.perf_counter() - t0
        states[n] = (cfg, t, b.download()[1])
        rate = st["object_steps"] / wall
        out["objects_%d" % n] = {"value": rate, "kernel_seconds": st["kernel_seconds"], "waves": st["waves"],
                                 "fp64_tflops": rate * f * 1e-12, "fp64_frac": rate * f * 1e-12 / FP64_PEAK_TFLOPS}
        log("alternate leg: %d objects %.4g object-steps/s" % (n, rate))
        ctx.close()
    out["value"] = out["objects_%d" % args.objects]["value"]
    if not args.no_cpu:
        import oracle
        cfg, t, s_g = states[args.objects]
        idx = np.linspace(0, t["n"] - 1, 64).astype(int)
        threads = cpu_threads()
        orc = oracle.Oracle(cfg)
        _, s_o, ost = orc.propagate(subset(t, idx), duration, threads=threads)
        rms, mx = pos_err_mm(s_g[:, idx], s_o)
        out["vs_oracle_pos_mm_64obj"] = {"rms": rms, "max": mx}
        cidx = np.arange(min(1024, t["n"]))
        _, _, cst = orc.propagate(subset(t, cidx), 3600.0, threads=threads)
        out["cpu_baseline"] = {"value": cst["object_steps"] / cst["seconds"], "unit": "object-steps/s", "cores": threads,
                               "kind": "port", "sample": "%d objects x 1 h on the oracle" % len(cidx)}
    return out


def leg_ephemeris(args, work):
    import parahpop_amd
    from parahpop_amd import brie_io, synth
    day = 86400.0
    d = os.path.join(work, "w2eph")
    cfg, tg = synth.make_workload(d, 2, n=args.objects, duration=day)
    t = synth.load_targets(tg)
    et0 = 4035.0 * day - 43200.0
    brie_io.write_brie(os.path.join(d, "eph.brie"), brie_io.ssb_units(brie_io.circular_geocentric, et0 - 2 * day,
                                                                          et0 + 3 * day))
    brie_io.write_brot2(os.path.join(d, "itrf.brot2"), 4035.0 - 4.0, 4035.0 + 6.0,
                        brie_io.iau2006_channels(xp_arcsec=0.3, yp_arcsec=0.3, wobble=True))
    doc = json.load(open(cfg))
    doc["model"]["environment"]["ephemeris"] = ["eph.brie"]
    cfg_brie = os.path.join(d, "CONFIG_brie.json")
    json.dump(doc, open(cfg_brie, "w"))
    doc["model"]["environment"]["orientations"] = ["itrf.brot2"]
    cfg_files = os.path.join(d, "CONFIG_brie_brot2.json")
    json.dump(doc, open(cfg_files, "w"))
    out = {"workload": "full-force 70x70, %d objects, 60 s step; body positions from a .brie file (12 body units, "
                       "degree 10-13) and Earth orientation from a .brot2 IERS2000 unit vs the analytic models" % t["n"], "steps": args.eph_steps}
    for label, path in (("analytic", cfg), ("brie", cfg_brie), ("brie_brot2", cfg_files)):
        ctx = parahpop_amd.Context(path)
        b = ctx.batch(t, day)
        b.run(4)
        st = b.run(args.eph_steps)
        out[label] = {"value": st["object_steps"] / st["kernel_seconds"], "ms_per_step": 1e3 * st["kernel_seconds"] / args.eph_steps,
                      "epoch_records": st["epoch_records"], "epoch_seconds": st["epoch_seconds"]}
        log("ephemeris leg: %s %.4g object-steps/s" % (label, out[label]["value"]))
        ctx.close()
    out["value"] = out["brie_brot2"]["value"]
    out["brie_vs_analytic"] = out["brie"]["value"] / out["analytic"]["value"]
    out["brie_brot2_vs_analytic"] = out["brie_brot2"]["value"] / out["analytic"]["value"]
    return out


def serve_run(ctx, reqs, gaps):
    """Submit reqs with the given inter-arrival gaps, polling meanwhile; returns
    (results by request index, latencies, wall seconds, object-steps)."""
    steps0 = ctx.option("serve_object_steps")
    pending, results = {}, {}
    t0 = time.perf_counter()
    next_t = t0
    for i, (r, du) in enumerate(reqs):
        next_t += gaps[i]
        while time.perf_counter() < next_t:
            for j, k in list(pending.items()):
                res = ctx.poll(k)
                if res is not None:
                    results[j] = res
                    del pending[j]
            time.sleep(0.001)
        pending[i] = ctx.submit(r, du)
        if (i + 1) % 50 == 0:
            log("serve: submitted %d/%d" % (i + 1, len(reqs)))
    for j, k in pending.items():
        results[j] = ctx.wait(k, timeout=900.0)
    wall = time.perf_counter() - t0
    lat = np.array([results[j][3] - results[j][2] for j in range(len(reqs))])
    return results, lat, wall, ctx.option("serve_object_steps") - steps0


def leg_serve(args, work):
    import parahpop_amd
    from parahpop_amd import synth
    d = os.path.join(work, "w5")
    pool_n = args.serve_pool
    cfg, tg = synth.make_workload(d, 4, n=pool_n, duration=3 * 86400.0)
    t = synth.load_targets(tg)
    doc = json.load(open(cfg))
    doc["execution"]["precision"] = {"mode": "mixed", "fp64Degree": 20}
    cfg_mixed = os.path.join(d, "CONFIG_mixed.json")
    json.dump(doc, open(cfg_mixed, "w"))
    rng = np.random.default_rng(5)
    nreq = args.serve_requests
    sizes = np.exp(rng.uniform(0.0, math.log(args.serve_max_objects), nreq)).astype(int).clip(1, pool_n)
    durs = np.exp(rng.uniform(math.log(3600.0), math.log(3 * 86400.0), nreq))
    starts = rng.integers(0, pool_n - sizes + 1)
    reqs = [(subset(t, np.arange(s, s + k)), float(du)) for s, k, du in zip(starts, sizes, durs)]
    ctx = parahpop_amd.Context(cfg_mixed, serve_steps_per_tick=args.serve_steps_per_tick)
    # warm the server (pool allocation, graph capture) with a tiny request
    ctx.wait(ctx.submit(subset(t, [0]), 600.0))
    rates = [float(x) for x in args.serve_rates.split(",") if x]
    out = {"workload": "serving: %d Poisson requests at each of %s /s, 1-%d objects log-U, 1 h-3 d log-U, mixed 70x70"
                       % (nreq, "/".join("%g" % r for r in rates), args.serve_max_objects),
           "requests": nreq, "objects_requested": int(sizes.sum()), "rates": []}
    first = None
    for rate in rates:
        gaps = np.random.default_rng(17).exponential(1.0 / rate, nreq)
        captures0 = ctx.option("serve_graph_captures")
        results, lat, wall, obj_steps = serve_run(ctx, reqs, gaps)
        entry = {"rate_per_s": rate, "object_steps_per_s": obj_steps / wall, "object_steps": obj_steps, "wall_seconds": wall,
                 "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
                 "latency_max_s": float(lat.max()), "graph_captures": ctx.option("serve_graph_captures") - captures0}
        if first is None:
            first = results
        else:
            # the same requests in another mix: the same bits
            entry["results_equal_first_rate"] = all(np.array_equal(results[j][1], first[j][1], equal_nan=True)
                                                    for j in range(nreq))
        out["rates"].append(entry)
        log("serve: rate %g/s p50 %.2f s p99 %.2f s max %.2f s, %.3g object-steps/s"
            % (rate, entry["latency_p50_s"], entry["latency_p99_s"], entry["latency_max_s"], entry["object_steps_per_s"]))
    for k in ("object_steps_per_s", "latency_p50_s", "latency_p99_s", "latency_max_s"):
        out[k] = out["rates"][-1][k]
    # error of every served object against the FP64 drop-in path: all requests
    # as one uneven FP64 batch (per-object durations)
    idx = np.concatenate([np.arange(s, s + k) for s, k in zip(starts, sizes)])
    dur = np.concatenate([np.full(k, du) for k, du in zip(sizes, durs)])
    ctx64 = parahpop_amd.Context(cfg)
    every = subset(t, idx)
    every["id"] = np.arange(len(idx), dtype=np.int64)   # requests overlap in the pool: one id per served object
    b = ctx64.batch(every, dur)
    b.run()
    s64 = b.download()[1]
    served = np.concatenate([first[j][1] for j in range(nreq)], axis=1)
    # objects that reenter end non-finite (NRLMSISE-00 is undefined below the
    # ground) or, having passed through the Earth's field, at a meaningless state
    # that differs between any two paths; the error is stated over the objects
    # that end in orbit (80 km above the equatorial radius to 100,000 km) on both
    fin = np.all(np.isfinite(served), axis=0) & np.all(np.isfinite(s64), axis=0)
    r_s = np.linalg.norm(np.where(fin, served[:3], 0.0), axis=0)
    r_64 = np.linalg.norm(np.where(fin, s64[:3], 0.0), axis=0)
    orbit = fin & (np.minimum(r_s, r_64) > 6378.1363 + 80.0) & (np.maximum(r_s, r_64) < 1e5)
    rms, mx = pos_err_mm(served[:, orbit], s64[:, orbit])
    err = np.linalg.norm(served[:3, orbit] - s64[:3, orbit], axis=0) * 1e6
    out["mixed_vs_fp64_pos_mm"] = {"objects": int(orbit.sum()), "objects_reentered": int((~orbit).sum()),
                                   "objects_nonfinite": int((~fin).sum()), "rms": rms, "max": mx,
                                   "p99": float(np.percentile(err, 99)) if err.size else None,
                                   "nonfinite_on_both_paths": bool(np.array_equal(np.isfinite(served).all(axis=0),
                                                                                  np.isfinite(s64).all(axis=0)))}
    log("serve: mixed vs FP64 over %d served objects in orbit (%d reentered, %d of them non-finite): rms %.3g mm, max %.3g mm"
        % (orbit.sum(), (~orbit).sum(), (~fin).sum(), rms, mx))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1440, help="timed steps of the headline leg (1440 = one day)")
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--objects", type=int, default=16384)
    ap.add_argument("--degree", type=int, default=70)
    ap.add_argument("--waves", type=int, default=0, help="waves per 64-object tile (0: library default)")
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--split", type=int, default=-1, help="1: split-stage path, 0: fused multi-step kernel (-1: library default)")
    ap.add_argument("--graph-steps", type=int, default=-1,
                    help="split stages: steps per replayed graph, 0: launch every kernel on the stream (-1: library default)")
    ap.add_argument("--legs", default="main,mixed,uneven,serve,alternate,ephemeris")
    ap.add_argument("--eph-steps", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-objects", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=86400.0)
    ap.add_argument("--mixed-seconds", type=float, default=86400.0)
    ap.add_argument("--uneven-objects", type=int, default=131072)
    ap.add_argument("--uneven-steps-per-launch", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", default="2,8",
                    help="Poisson arrival rates (requests/s), one run of all requests each (about half and twice the capacity)")
    ap.add_argument("--serve-max-objects", type=int, default=50000)
    ap.add_argument("--serve-pool", type=int, default=65536)
    ap.add_argument("--serve-steps-per-tick", type=int, default=16)
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (scales the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the propagation shards by object slots with no exchange; "
              "multi-GPU runs are not part of this build (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 5}))
        return

    work = args.workdir or tempfile.mkdtemp(prefix="phx_bench_")
    legs = [x for x in args.legs.split(",") if x]
    line = leg_main(args, work) if "main" in legs else {"metric": METRIC}
    log("main: %.4g object-steps/s" % line.get("value", float("nan")))
    if "mixed" in legs:
        line["mixed"] = leg_mixed(args, work)
    if "uneven" in legs:
        line["uneven"] = leg_uneven(args, work)
        log("uneven: %.4g object-steps/s" % line["uneven"]["value"])
    if "serve" in legs:
        line["serve"] = leg_serve(args, work)
        log("serve: p50 %.2f s p99 %.2f s" % (line["serve"]["latency_p50_s"], line["serve"]["latency_p99_s"]))
    if "alternate" in legs:
        line["alternate"] = leg_alternate(args, work)
    if "ephemeris" in legs:
        line["ephemeris"] = leg_ephemeris(args, work)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
