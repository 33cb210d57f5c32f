"""Benchmark of the MI355X MPM step: particle-updates/s on the 8M-particle workload.

Workload (default): 8000x8000-node synthetic DEM at 1 m, 32-vertex star polygon
release (mean radius 800 m, h0 = 4 m, ~8M UNIFORM particles), mu(I) friction
(mu1 0.38, mu2 0.64, I0 0.279, d 0.5 m, phi 0.6), rho 2000, 120 s.  A "step"
is one pass of the step loop over all particles (sort, P2G + grid update,
G2P, dt).  Inputs are resident in device memory before the timed region
starts; the default timed region is 10,000 consecutive steps (most of the
run: the cost per step grows as the flow piles up), per-kernel event times
come from the 100 steps after it.

Prints ONE JSON line.  roofline: dominant kernel's algorithmic bytes per launch
/ its mean HIP-event duration vs the 8 TB/s HBM peak; roofline.traffic: that
kernel's HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
WRITE_SIZE summary of the same workload (profiles/step_<form>_<workload>_*.md),
null when there is none.  cpu_baseline: the FP64 CPU oracle (single-threaded
restatement of the reference step) timed on this host on a bounded number of
steps of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "slab-pro_amd"))

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes (compulsory state traffic, each persistent field read once /
# written once), per particle and per active node (m > 1e-6), for each kernel.
# FP64 form: "p2g" = per-tile reduction reading x y vx vy vz h A m (two field
#            groups), "grid" = partial merge + grid update (terrain
#            Gx Gy cos, h_max v_max2 read+write, write vx vy flipx flipy).
ALG_BYTES_FP64 = {
    "p2g": {"particle": 64, "node": 0},
    "grid": {"particle": 0, "node": 24 + 32 + 32},
    # read x y vx vy h m, write x y vx vy vz h A; per node: read vx vy flipx flipy + terrain z
    "g2p": {"particle": 48 + 56, "node": 32 + 8},
    "tile_sort": {"particle": 8, "node": 0},
    "scatter": {"particle": 12, "node": 0},
}
STEP_BYTES_FP64 = {"particle": 104, "node": 120}
# FP32 form: particle = cell int32 + offsets + vx vy vz h m (32 B read, 28 B
# written by G2P); "p2g" = per-tile LDS reduction of the particle state (reads
# offsets vx vy h m: the cell follows from the tile and the offsets, vz is read
# on output steps only), "grid" = partial merge + grid update (terrain 12 B,
# max maps 16 B, vf 16 B).
ALG_BYTES_FP32 = {
    "p2g": {"particle": 24, "node": 0},
    "grid": {"particle": 0, "node": 12 + 16 + 16},
    "g2p": {"particle": 28 + 28, "node": 16 + 4},
    "tile_sort": {"particle": 8, "node": 0},
    "scatter": {"particle": 12, "node": 0},
}
STEP_BYTES_FP32 = {"particle": 60, "node": 60}

# kernel names of the event timers -> rocprof kernel names, per form
ROCPROF_NAMES = {
    "fp32": {"g2p": "k_eg2p", "p2g": "k_ep2g", "grid": "k_egrid", "tile_sort": "k_esort", "scatter": "k_escatter"},
    "fp64": {"g2p": "k_g2p", "p2g": "k_p2g64", "grid": "k_grid64", "tile_sort": "k_tile_sort",
             "scatter": "k_scatter"},
}


def measured_traffic(form, workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprof summary
    (column 'HBM MB/launch (scaled)'), or (None, None)."""
    import glob
    want = "`" + ROCPROF_NAMES.get(form, {}).get(kernel, "?") + "`"
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"step_{form}_{workload}_*.md"))):
        with open(path) as f:
            for line in f:
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cols) >= 8 and cols[0] == want:
                    try:
                        return float(cols[6].replace(",", "")) * 1e6, os.path.relpath(path, ROOT)
                    except ValueError:
                        pass
    return None, None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(cfg, steps, warm):
    """Time the single-threaded FP64 oracle on `steps` steps of the workload."""
    from oracle import oracle as orc
    t0 = time.time()
    pr = dict(cfg["params"])
    fr = dict(cfg["friction"])
    typ = fr.pop("type")
    s = orc.OracleSim(cfg["z"], cfg["dx"], orc.params(fric_type=typ, **pr, **fr), cfg["x"], cfg["y"], cfg["h"],
                      cfg["A"], cfg["m"])
    s.prime()
    for _ in range(warm):
        s.step()
    t1 = time.time()
    for _ in range(steps):
        s.step()
    el = time.time() - t1
    np_ = len(cfg["x"])
    log(f"cpu oracle: setup {t1 - t0:.1f}s, {steps} steps {el:.2f}s; phases {s.scalars()['phases']}")
    s.close()
    return np_ * steps / el, el


def copy_probe():
    """Streaming kernels of known size written for gfx950 (slab_stream_probe):
    the practical HBM ceiling for a read+write copy, a read, a write, and
    G2P's particle streams at the 8M-particle workload's count.  Their
    FETCH_SIZE / WRITE_SIZE per dispatch also give the factors that turn those
    counters into HBM bytes.  Prints one JSON line."""
    from slabpro import abi
    gib = 1 << 30
    out = {}
    for name, kind, n, nbytes in (("copy_1GiB", abi.PROBE_COPY, gib, 2 * gib),
                                  ("read_1GiB", abi.PROBE_READ, gib, gib),
                                  ("write_1GiB", abi.PROBE_WRITE, gib, gib),
                                  ("particles_8147024", abi.PROBE_PARTICLES, 8147024, 60 * 8147024)):
        ms = abi.stream_probe(kind, n, 20)
        out[name] = {"ms": ms, "bytes": nbytes, "GB_s": nbytes / ms / 1e6, "frac_of_8TBs": nbytes / ms / 1e6 / 8000.0}
        log(f"{name}: {ms:.4f} ms, {nbytes / ms / 1e6:.0f} GB/s")
    print(json.dumps({"stream_probe": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000,
                    help="timed steps (default: 10,000 of the C2 run's ~12,000, i.e. most of its 120 s)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2"])
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner (FP32 form)")
    ap.add_argument("--form", default="fp32", choices=["fp32", "fp64"],
                    help="fp32: the FP32 form (int cell + FP32 offsets, FP32 bulk, FP64 time); fp64: plain form")
    ap.add_argument("--ensemble", action="store_true",
                    help="256-scenario ensemble: FP32 batched vs FP64 one-at-a-time, scenarios/h")
    ap.add_argument("--scenarios", type=int, default=256)
    ap.add_argument("--fp64-sample", type=int, default=16)
    ap.add_argument("--serve", action="store_true", help="continuous-batching server, p50/p95/p99 latency")
    ap.add_argument("--accuracy", action="store_true",
                    help="full run of --workload in both forms: FP32 error vs FP64 (max maps, runout, deposit)")
    ap.add_argument("--rates", default="2,5,10")
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--slots", type=int, default=128, help="serving: batch slots")
    ap.add_argument("--capacity", type=int, default=24_000_000, help="serving: device particle capacity")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 5 plain 1 GiB device copies (calibration of the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the step runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        return copy_probe()

    import numpy as np
    import slabpro
    from slabpro import abi, configs

    if args.ensemble:
        return ensemble_main(args)
    if args.serve:
        return serve_main(args)
    if args.accuracy:
        return accuracy_main(args)

    t0 = time.time()
    cfg = getattr(configs, args.workload)()
    npart = len(cfg["x"])
    nx, ny = cfg["z"].shape
    log(f"workload {args.workload}: {npart} particles, grid {nx}x{ny}, generated in {time.time() - t0:.1f}s")

    t0 = time.time()
    run = slabpro.make_run(cfg, precision=abi.FP32 if args.form == "fp32" else abi.FP64)
    run.prime()
    log(f"device setup + prime {time.time() - t0:.1f}s")
    for _ in range(args.warmup):
        run.step(1)
    tuned = None
    if args.form == "fp32" and not args.no_tune:
        # launch-shape autotuner: its trial steps are ordinary, untimed steps
        t0 = time.time()
        tuned = run.autotune(8)
        log(f"autotune: {tuned} in {time.time() - t0:.1f}s")

    # timed region: K steps, inputs resident, no event instrumentation
    done = 0
    t0 = time.perf_counter()
    while done < args.steps:
        info = run.step(args.steps - done)  # returns after a stream synchronize
        done += info["steps_done"]
        if info["finished"]:
            break
    el = time.perf_counter() - t0
    # per-kernel HIP-event durations on the run's stream, separate short region
    run.timing(True)
    ksteps = 0
    kwant = min(args.steps, 100)
    while ksteps < kwant:
        info = run.step(kwant - ksteps)
        ksteps += info["steps_done"]
        if info["finished"]:
            break
    tim = run.timing(False)
    run.keep_grid(True)
    run.step(1)
    d = run.diag()
    n_act = int(d["n_active"])
    value = npart * done / el
    ms_step = 1e3 * el / max(done, 1)

    # dominant kernel by total event time
    dom = max(tim, key=lambda k: tim[k][0])
    dom_ms, dom_n = tim[dom]
    avg_s = dom_ms / max(dom_n, 1) * 1e-3
    alg_k = ALG_BYTES_FP32 if args.form == "fp32" else ALG_BYTES_FP64
    step_b = STEP_BYTES_FP32 if args.form == "fp32" else STEP_BYTES_FP64
    ab = alg_k.get(dom, {"particle": 0, "node": 0})
    alg = ab["particle"] * npart + ab["node"] * n_act
    achieved = alg / avg_s / 1e9 if avg_s > 0 else 0.0
    step_alg = step_b["particle"] * npart + step_b["node"] * n_act
    traffic, traffic_src = measured_traffic(args.form, args.workload, dom)
    log("kernel times (ms total, launches): " + ", ".join(f"{k} {v[0]:.2f}/{v[1]}" for k, v in tim.items()))

    cpu = None
    if not args.no_cpu:
        cpu_rate, cpu_el = cpu_baseline(cfg, args.cpu_steps, 1)
        cpu = {"value": cpu_rate, "unit": "particle-updates/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps (after prime + 1 warm-up step) of the same {args.workload} "
                         f"workload, {npart} particles, FP64 oracle, single thread, {cpu_el:.1f} s"}

    out = {
        "metric": "particle-updates/s at 8M particles",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": done,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.form == "fp32" else "f64",
        "data": "synthetic (seeded DEM + release generators; no external data)",
        "config": {"workload": f"{args.workload}: " + {
            "c0": "601x151 nodes at 2 m, analytical 35deg plane+arc+flat, 4,800 particles, Coulomb",
            "c1": "2000x2000 DEM at 5 m, raster release of 40 ellipses, Voellmy",
            "c2": "8000x8000 DEM at 1 m, 32-vertex star polygon r=800 m h0=4 m, mu(I)"}[args.workload],
            "particles": npart, "grid": f"{nx}x{ny}", "form": args.form, "parallelism": "single GPU",
            "autotuned": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                     "alg_bytes_per_launch": alg, "avg_launch_ms": avg_s * 1e3,
                     "traffic_source": traffic_src},
        "step_roofline": {"alg_bytes_per_step": step_alg, "achieved_GBs": step_alg / (ms_step * 1e-3) / 1e9,
                          "frac": step_alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, "n_active_nodes": n_act},
        "kernels_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in tim.items()},
        "cpu_baseline": cpu,
    }
    run.close()
    print(json.dumps(out), flush=True)


def ensemble_main(args):
    """C3: 256 scenarios on the 2000x2000 DEM at 5 m; FP32 batched (all) vs the
    plain FP64 one-at-a-time GPU path (a size-stratified sample, stated)."""
    import numpy as np
    from slabpro import abi, configs, ensemble
    t0 = time.time()
    dem = configs.c1_dem()
    scen = [configs.c3_scenario(i) for i in range(args.scenarios)]
    sizes = np.array([len(s["x"]) for s in scen])
    log(f"ensemble: {len(scen)} scenarios, {sizes.sum()} particles (min {sizes.min()}, max {sizes.max()}), "
        f"generated in {time.time() - t0:.1f}s")
    fr = [s["friction"]["type"] for s in scen]
    bat = ensemble.run_batched(dem, 5.0, scen, max_slots=len(scen), progress=log)
    log(f"fp32 batched: {bat['wall_s']:.1f}s, {bat['scenarios_per_hour']:.0f} scen/h, "
        f"{bat['particle_updates_per_s']:.3g} pu/s, failed {bat['failed']}")
    k = max(1, min(args.fp64_sample, len(scen)))
    rank = np.argsort(sizes, kind="stable")
    sample = [int(rank[int((q + 0.5) * len(scen) / k)]) for q in range(k)]
    one64 = ensemble.run_one_at_a_time(dem, 5.0, [scen[i] for i in sample], abi.FP64, progress=log)
    log(f"fp64 one-at-a-time on {k} scenarios: {one64['wall_s']:.1f}s, {one64['scenarios_per_hour']:.0f} scen/h")
    one32 = ensemble.run_one_at_a_time(dem, 5.0, [scen[i] for i in sample], abi.FP32, progress=log)
    log(f"fp32 one-at-a-time on the same {k}: {one32['wall_s']:.1f}s, {one32['scenarios_per_hour']:.0f} scen/h")
    out = {
        "metric": "scenarios/h on the 256-scenario ensemble",
        "value": bat["scenarios_per_hour"], "unit": "scenarios/h", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic (seeded DEM + ellipse releases)",
        "config": {"workload": f"c3: {len(scen)} scenarios on the 2000x2000 DEM at 5 m, 1-10 ellipses each, "
                               f"log-uniform 2k-1M particles, 50% Voellmy / 30% Coulomb / 20% mu(I), 120 s each",
                   "particles_total": int(sizes.sum()),
                   "friction_counts": {"coulomb": fr.count(0), "voellmy": fr.count(1), "mu_I": fr.count(2)}},
        "fp32_batched": {k2: v for k2, v in bat.items() if k2 != "results"},
        "fp64_one_at_a_time": {"sample": f"{k} scenarios, one per size stratum (ranks {sample})",
                               **{k2: v for k2, v in one64.items() if k2 != "per_scenario_s"}},
        "fp32_one_at_a_time": {k2: v for k2, v in one32.items() if k2 != "per_scenario_s"},
        "speedup_vs_fp64_one_at_a_time": bat["scenarios_per_hour"] / one64["scenarios_per_hour"],
    }
    print(json.dumps(out), flush=True)


def accuracy_main(args):
    """Full run of one workload in the FP64 and the FP32 form; FP32 error."""
    from slabpro import accuracy, configs
    t0 = time.time()
    cfg = getattr(configs, args.workload)()
    log(f"accuracy {args.workload}: {len(cfg['x'])} particles, generated in {time.time() - t0:.1f}s, "
        f"total_time {cfg['params']['total_time']} s")
    rep, tm = accuracy.compare_forms(cfg)
    log(f"fp64 {tm['fp64_s']:.1f}s, fp32 {tm['fp32_s']:.1f}s: {rep}")
    out = {"metric": "FP32-form error vs the FP64 form (full run)", "value": rep["hmax_L1"],
           "unit": "relative L1 of the max-height map", "n_gpus": 1, "higher_is_better": False,
           "dtype": "f32 vs f64", "data": "synthetic", "config": {"workload": args.workload,
                                                                  "particles": len(cfg["x"])},
           "report": rep, "wall_s": tm}
    print(json.dumps(out), flush=True)


def serve_main(args):
    """C4: open-loop Poisson arrivals drawn from the ensemble distribution on the
    2000x2000 DEM; continuous batching with admission by device footprint."""
    import numpy as np
    from slabpro import configs, ensemble
    dem = configs.c1_dem()
    res = {}
    n = args.requests
    t0 = time.time()
    reqs = [configs.c3_scenario(1000 + k) for k in range(n)]
    log(f"serve: {n} requests generated in {time.time() - t0:.1f}s, {sum(len(r['x']) for r in reqs)} particles")
    for lam in [float(v) for v in args.rates.split(",")]:
        arr, _ = configs.c4_requests(n, lam, seed_index=int(lam * 10))
        r = ensemble.serve(dem, 5.0, reqs, arr, max_slots=args.slots, capacity=args.capacity, progress=log)
        res[f"lambda_{lam:g}"] = r
        log(f"lambda {lam:g}/s: p50 {r['p50_s']:.2f}s p95 {r['p95_s']:.2f}s p99 {r['p99_s']:.2f}s, "
            f"{r['scenarios_per_hour']:.0f} scen/h over {r['wall_s']:.1f}s")
    first = res[next(iter(res))]
    out = {"metric": "p99 completion latency (serving)", "value": first["p99_s"], "unit": "s", "n_gpus": 1,
           "higher_is_better": False, "dtype": "f32", "data": "synthetic (seeded requests, Poisson arrivals)",
           "config": {"workload": f"c4: {n} of the 2,000 seeded requests per rate, rates {args.rates}/s, "
                                  f"c3 scenario distribution on the 2000x2000 DEM, {args.slots} slots, "
                                  f"{args.capacity / 1e6:g}M-particle capacity"},
           "per_rate": res}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
