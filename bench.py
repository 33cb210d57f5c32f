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
null when there is none.  stream_ceiling: GB/s of hand-written streaming
kernels on the same GPU (the practical ceiling; frac_of_copy_ceiling reads the
dominant kernel against it).  cpu_baseline: the FP64 CPU oracle
(single-threaded restatement of the reference step) timed on this host, 10
untimed + 50 timed steps from the t = 0 prime and from the run's state at
t = 60 s, with the host's core count and CPU model.  ensemble: the
256-scenario ensemble (FP32 batched scenarios/h, the FP64 one-at-a-time rate
on a stratified sample and each sampled scenario's max-height L1 error).
serving: a bounded sample of the request server (300 requests at 5 per
second: p50/p95/p99 completion latency and scenarios/h).
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
# FP32 form: particle = cell int32 + offsets + vx vy vz h m + elevation z;
# "p2g" = the tile sort (bucket read + sorted index write, 8 B) and the
# per-tile LDS reduction of the particle state (reads offsets vx vy h m, 24 B:
# the cell follows from the tile and the offsets, vz is read on output steps
# only) in one launch, "grid" = partial merge + grid update (terrain 12 B, max
# maps 16 B, vf 16 B), "g2p" = reads cell offsets vx vy h and writes them back
# (24 + 24 B; a single run's vz and z only around output steps: not counted),
# per node vf 16 B.
ALG_BYTES_FP32 = {
    "p2g": {"particle": 8 + 24, "node": 0},
    "grid": {"particle": 0, "node": 12 + 16 + 16},
    "g2p": {"particle": 24 + 24, "node": 16},
    "scatter": {"particle": 12, "node": 0},
}
STEP_BYTES_FP32 = {"particle": 52, "node": 60}

# kernel names of the event timers -> rocprof kernel names, per form
ROCPROF_NAMES = {
    "fp32": {"g2p": "k_eg2p", "p2g": "k_esp", "grid": "k_egrid", "scatter": "k_escatter"},
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


def host_facts():
    """nproc, usable cores and the CPU model of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cores": usable, "cpu_model": model}


def cpu_baseline(cfg, steps, warm, state=None, time0=0.0, frame0=0):
    """Time the single-threaded FP64 oracle on `steps` steps (after `warm`
    untimed ones) of the workload: from the t = 0 prime, or from `state`
    (particle arrays x, y, vx, vy, h, A, m of a run at time0)."""
    from oracle import oracle as orc
    t0 = time.time()
    pr = dict(cfg["params"])
    fr = dict(cfg["friction"])
    typ = fr.pop("type")
    src = state if state is not None else cfg
    s = orc.OracleSim(cfg["z"], cfg["dx"], orc.params(fric_type=typ, **pr, **fr), src["x"], src["y"], src["h"],
                      src["A"], src["m"])
    if state is None:
        s.prime()
    else:
        s.set_velocity(state["vx"], state["vy"])
        s.set_time(time0, frame0)
    for _ in range(warm):
        s.step()
    t1 = time.time()
    for _ in range(steps):
        s.step()
    el = time.time() - t1
    np_ = len(src["x"])
    log(f"cpu oracle ({'t=0' if state is None else f't={time0:g} s'}): setup {t1 - t0:.1f}s, {warm}+{steps} steps, "
        f"timed {el:.2f}s; phases {s.scalars()['phases']}")
    s.close()
    return np_ * steps / el, el


def mid_run_state(cfg, t_mid):
    """The workload's particle state at simulated time t_mid, from the FP64
    form on the GPU (the state the CPU baseline's mid-run sample starts from)."""
    import slabpro
    from slabpro import abi
    r = slabpro.make_run(cfg, precision=abi.FP64)
    r.prime()
    info = None
    while True:
        info = r.step(1 << 20)
        if info["time"] >= t_mid - 1e-9 or info["finished"]:
            break
    p = r.particles()
    r.close()
    return p, info["time"], info["frame"]


def stream_ceiling():
    """Hand-written streaming kernels (slab_stream_probe): GB/s of a 1 GiB
    read+write copy and of G2P's particle streams for 8,147,024 particles."""
    from slabpro import abi
    gib = 1 << 30
    cp = min(abi.stream_probe(abi.PROBE_COPY, gib, 10), abi.stream_probe(abi.PROBE_COPY_NT, gib, 10))
    pt = abi.stream_probe(abi.PROBE_PARTICLES, 8147024, 10)
    return {"copy_1GiB_GBs": 2 * gib / cp / 1e6, "g2p_streams_8147024_GBs": 60 * 8147024 / pt / 1e6,
            "source": "slab_stream_probe (slab-pro_amd/csrc/slab_probe.hip), faster of the two copy kernels"}


def ensemble_numbers(scenarios=256, fp64_sample=16):
    """The 256-scenario ensemble on the 2000x2000 DEM at 5 m: FP32 batched
    scenarios/h, the plain FP64 one-at-a-time rate on a size-stratified sample
    (one scenario per size stratum), and each sampled scenario's max-height
    L1 error of the batched FP32 result against its FP64 run."""
    import numpy as np
    from slabpro import abi, configs, ensemble
    t0 = time.time()
    dem = configs.c1_dem()
    scen = [configs.c3_scenario(i) for i in range(scenarios)]
    sizes = np.array([len(x["x"]) for x in scen])
    fr = [x["friction"]["type"] for x in scen]
    log(f"ensemble: {len(scen)} scenarios, {sizes.sum()} particles (min {sizes.min()}, max {sizes.max()}), "
        f"generated in {time.time() - t0:.1f}s")
    k = max(1, min(fp64_sample, len(scen)))
    rank = np.argsort(sizes, kind="stable")
    sample = [int(rank[int((q + 0.5) * len(scen) / k)]) for q in range(k)]
    bat = ensemble.run_batched(dem, 5.0, scen, max_slots=len(scen), progress=log, results=set(sample))
    log(f"fp32 batched: {bat['wall_s']:.1f}s, {bat['scenarios_per_hour']:.0f} scen/h, "
        f"{bat['particle_updates_per_s']:.3g} pu/s, failed {bat['failed']}")
    one64 = ensemble.run_one_at_a_time(dem, 5.0, [scen[i] for i in sample], abi.FP64, progress=log, results=True)
    log(f"fp64 one-at-a-time on {k} scenarios: {one64['wall_s']:.1f}s, {one64['scenarios_per_hour']:.0f} scen/h")
    errs = []
    for q, i in enumerate(sample):
        h32 = bat["results"][i][1][0]
        h64 = one64["results"][q][0]
        errs.append(float(np.abs(h32 - h64).sum() / np.abs(h64).sum()))
    errs = np.array(errs)
    return {
        "workload": f"c3: {len(scen)} scenarios on the 2000x2000 DEM at 5 m, 1-10 ellipses each, log-uniform "
                    f"2k-1M particles, 50% Voellmy / 30% Coulomb / 20% mu(I), 120 s each",
        "particles_total": int(sizes.sum()),
        "friction_counts": {"coulomb": fr.count(0), "voellmy": fr.count(1), "mu_I": fr.count(2)},
        "fp32_batched_scenarios_per_hour": bat["scenarios_per_hour"],
        "fp32_batched_wall_s": bat["wall_s"], "fp32_batched_particle_updates_per_s": bat["particle_updates_per_s"],
        "failed": bat["failed"], "autotuned": bat["autotuned"],
        "fp64_one_at_a_time_scenarios_per_hour": one64["scenarios_per_hour"],
        "fp64_sample": f"{k} scenarios, one per size stratum (indices {sample})",
        "speedup_vs_fp64_one_at_a_time": bat["scenarios_per_hour"] / one64["scenarios_per_hour"],
        "hmax_L1_vs_fp64": {"max": float(errs.max()), "p95": float(np.percentile(errs, 95)),
                            "mean": float(errs.mean()), "per_sampled_scenario": [float(e) for e in errs]},
    }


def serving_sample(n=300, lam=5.0, slots=128, capacity=24_000_000):
    """A bounded sample of the serving workload for the default line: n
    seeded requests from the ensemble distribution (configs.c3_scenario(1000
    + k)) on the 2000x2000 DEM, open-loop Poisson arrivals at lam per second,
    continuous batching (ensemble.serve) with admission by device footprint;
    p50/p95/p99 completion latency and scenarios/h.  The full 2,000 requests
    per rate at 2, 5 and 10 per second run with --serve."""
    from slabpro import configs, ensemble
    t0 = time.time()
    dem = configs.c1_dem()
    reqs = [configs.c3_scenario(1000 + k) for k in range(n)]
    arr, _ = configs.c4_requests(n, lam, seed_index=int(lam * 10))
    log(f"serving sample: {n} requests generated in {time.time() - t0:.1f}s")
    r = ensemble.serve(dem, 5.0, reqs, arr, max_slots=slots, capacity=capacity, progress=log)
    log(f"serving sample: p50 {r['p50_s']:.2f}s p95 {r['p95_s']:.2f}s p99 {r['p99_s']:.2f}s, "
        f"{r['scenarios_per_hour']:.0f} scen/h over {r['wall_s']:.1f}s")
    return {"workload": f"c4 sample: {n} seeded requests (c3 distribution, 2000x2000 DEM), Poisson arrivals "
                        f"at {lam:g}/s, {slots} slots, {capacity / 1e6:g}M-particle capacity",
            "served": r["served"], "rejected": len(r["rejected"]), "p50_s": r["p50_s"], "p95_s": r["p95_s"],
            "p99_s": r["p99_s"], "mean_s": r["mean_s"], "scenarios_per_hour": r["scenarios_per_hour"],
            "wall_s": r["wall_s"], "mean_active_particles": r["mean_active_particles"]}


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
                                  ("copy_nt_1GiB", abi.PROBE_COPY_NT, gib, 2 * gib),
                                  ("read_1GiB", abi.PROBE_READ, gib, gib),
                                  ("write_1GiB", abi.PROBE_WRITE, gib, gib),
                                  ("write_nt_1GiB", abi.PROBE_WRITE_NT, gib, gib),
                                  ("particles_8147024", abi.PROBE_PARTICLES, 8147024, 60 * 8147024)):
        ms = abi.stream_probe(kind, n, 20)
        out[name] = {"ms": ms, "bytes": nbytes, "GB_s": nbytes / ms / 1e6, "frac_of_8TBs": nbytes / ms / 1e6 / 8000.0}
        log(f"{name}: {ms:.4f} ms, {nbytes / ms / 1e6:.0f} GB/s")
    # kernel boundary against grid barrier: eight dependent copy phases of n
    # bytes, as eight launches and as one cooperative launch with seven grid
    # barriers (device-scope release/acquire), same workgroups
    phases = {}
    for per_cu, kl, kb in ((4, abi.PROBE_PHASES_LAUNCHES, abi.PROBE_PHASES_BARRIERS),
                           (1, abi.PROBE_PHASES_LAUNCHES_1, abi.PROBE_PHASES_BARRIERS_1)):
        for n in (1 << 20, 64 << 20, 256 << 20):
            t_l = abi.stream_probe(kl, n, 20)
            t_b = abi.stream_probe(kb, n, 20)
            phases[f"{per_cu}_per_CU_{n >> 20}MiB"] = {
                "launches_ms": t_l, "barriers_ms": t_b, "launch_minus_barrier_us_per_boundary": (t_l - t_b) / 7 * 1e3}
            log(f"{per_cu} workgroup(s) per CU, 8 phases of {n >> 20} MiB: 8 launches {t_l:.4f} ms, "
                f"1 launch + 7 grid barriers {t_b:.4f} ms")
    print(json.dumps({"stream_probe": out, "phase_boundaries": phases}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000,
                    help="timed steps (default: 10,000 of the C2 run's ~12,000, i.e. most of its 120 s)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2"])
    ap.add_argument("--cpu-steps", type=int, default=50, help="timed oracle steps per CPU sample")
    ap.add_argument("--cpu-warmup", type=int, default=10, help="untimed oracle steps before each CPU sample")
    ap.add_argument("--cpu-mid", type=float, default=60.0,
                    help="simulated time (s) of the CPU baseline's mid-run sample (0: t = 0 sample only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the bounded serving sample of the default line")
    ap.add_argument("--batch-window", action="store_true",
                    help="only a window of the 256-scenario batch with eager launches (kernel traces, counters)")
    ap.add_argument("--batch-start", type=int, default=0, help="batch window: untimed steps before it")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 256-scenario ensemble numbers in the default line")
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
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--slots", type=int, default=128, help="serving: batch slots")
    ap.add_argument("--capacity", type=int, default=24_000_000, help="serving: device particle capacity")
    ap.add_argument("--conditioning", default="",
                    help="comma-separated ensemble scenario indices: FP32 error vs FP64 sensitivity of h_max")
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
    if args.batch_window:
        return batch_window_main(args)
    if args.serve:
        return serve_main(args)
    if args.accuracy:
        return accuracy_main(args)
    if args.conditioning:
        return conditioning_main(args)

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

    run.close()
    ceiling = stream_ceiling()
    log(f"stream ceiling: {ceiling}")

    cpu = None
    if not args.no_cpu:
        rate0, el0 = cpu_baseline(cfg, args.cpu_steps, args.cpu_warmup)
        samples = {"t0": {"value": rate0, "timed_s": el0}}
        if args.cpu_mid > 0:
            st, tmid, fmid = mid_run_state(cfg, args.cpu_mid)
            rate1, el1 = cpu_baseline(cfg, args.cpu_steps, args.cpu_warmup, st, tmid, fmid)
            samples[f"t{tmid:g}s"] = {"value": rate1, "timed_s": el1}
        rates = [v["value"] for v in samples.values()]
        cpu = {"value": sum(rates) / len(rates), "unit": "particle-updates/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_warmup} untimed + {args.cpu_steps} timed steps of the same {args.workload} "
                         f"workload ({npart} particles), FP64 oracle, single thread, from the t = 0 prime and "
                         f"from the state at t = {args.cpu_mid:g} s (FP64 form on the GPU); value = mean of "
                         f"the samples",
               "samples": samples, "host": host_facts()}

    ens = None
    if not args.no_ensemble and args.workload == "c2" and args.form == "fp32":
        ens = ensemble_numbers()
    srv = None
    if not args.no_serve and args.workload == "c2" and args.form == "fp32":
        srv = serving_sample()
        if ens:
            srv["vs_fp64_one_at_a_time"] = srv["scenarios_per_hour"] / ens["fp64_one_at_a_time_scenarios_per_hour"]

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
        # every kernel with an algorithmic byte figure: its fraction of 8 TB/s
        # by the same event times (the roofline object is the dominant one)
        "kernels_hbm_frac": {k: (alg_k[k]["particle"] * npart + alg_k[k]["node"] * n_act) /
                                (v[0] / v[1] * 1e-3) / 1e9 / HBM_PEAK_GBS
                             for k, v in tim.items() if k in alg_k and v[1] > 0 and v[0] > 0},
        "stream_ceiling": ceiling,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "serving": srv,
    }
    out["roofline"]["frac_of_copy_ceiling"] = achieved / ceiling["copy_1GiB_GBs"]
    print(json.dumps(out), flush=True)


def ensemble_main(args):
    """C3: 256 scenarios on the 2000x2000 DEM at 5 m; FP32 batched (all) vs the
    plain FP64 one-at-a-time GPU path (a size-stratified sample, stated), with
    the sampled scenarios' max-height L1 error beside the throughput."""
    e = ensemble_numbers(args.scenarios, args.fp64_sample)
    out = {"metric": "scenarios/h on the 256-scenario ensemble", "value": e["fp32_batched_scenarios_per_hour"],
           "unit": "scenarios/h", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
           "data": "synthetic (seeded DEM + ellipse releases)", "config": {"workload": e["workload"]},
           "ensemble": e}
    print(json.dumps(out), flush=True)


def batch_window_main(args):
    """The 256-scenario batch for a window of steps, every launch eager (no
    hipGraph), for kernel traces and counters of the batched kernels: all
    scenarios admitted at once, `--batch-start` untimed steps, then `--steps`
    steps with HIP-event kernel times."""
    from slabpro import abi, configs, ensemble
    dem = configs.c1_dem()
    scen = [configs.c3_scenario(i) for i in range(args.scenarios)]
    cap = sum(len(x["x"]) for x in scen)
    b = abi.Batch(dem, 5.0, len(scen), cap)
    b.use_graph(False)
    for x in scen:
        if b.admit(ensemble._params(x, abi.FP32), ensemble._fric(x), x["x"], x["y"], x["h"], x["A"], x["m"]) is None:
            raise RuntimeError("a scenario does not fit the batch")
    done = 0
    while done < args.batch_start:  # Batch.step returns the number of active slots
        k = min(64, args.batch_start - done)
        b.step(k)
        done += k
    b.timing(True)
    t0 = time.perf_counter()
    b.step(args.steps)
    el = time.perf_counter() - t0
    n = args.steps
    tim = b.timing(False)
    b.close()
    out = {"metric": "batched step kernels (256-scenario batch window)", "value": 1e3 * el / max(n, 1),
           "unit": "ms/step (eager launches, event-timed)", "n_gpus": 1, "higher_is_better": False,
           "dtype": "f32", "data": "synthetic (seeded DEM + ellipse releases)",
           "config": {"workload": f"c3: {len(scen)} scenarios, {cap} particles, steps {args.batch_start}.."
                                  f"{args.batch_start + n}"},
           "kernels_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in tim.items() if v[1]}}
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


def conditioning_main(args):
    """Conditioning of the max-height map for chosen ensemble scenarios: the
    FP32 form's h_max L1 error against the FP64 form beside the change of the
    FP64 result, and of the FP32 result, when every initial position moves by
    one FP32 offset ulp of a cell (dx * 2^-24, random sign per particle and
    axis), and the change of the FP64 result when the friction coefficient mu
    is rounded to FP32 or scaled by 1 +- 1e-7."""
    import numpy as np
    import slabpro
    from slabpro import abi, configs
    dem = configs.c1_dem()
    rows = []
    for i in [int(v) for v in args.conditioning.split(",")]:
        s = configs.c3_scenario(i)
        cfg = dict(s)
        cfg["z"] = dem
        maps = {}
        for tag, prec, pert in (("fp64", abi.FP64, False), ("fp64_moved", abi.FP64, True), ("fp32", abi.FP32, False),
                                ("fp32_moved", abi.FP32, True)):
            c = dict(cfg)
            if pert:
                g = np.random.default_rng(7)
                u = cfg["dx"] * 2.0 ** -24
                c["x"] = np.asarray(cfg["x"]) + g.choice([-1.0, 1.0], len(cfg["x"])) * u
                c["y"] = np.asarray(cfg["y"]) + g.choice([-1.0, 1.0], len(cfg["y"])) * u
            r = slabpro.make_run(c, precision=prec)
            r.prime()
            r.run()
            maps[tag] = r.max_maps()[0]
            r.close()
        mu = cfg["friction"].get("mu", 0.0)
        if mu:
            for tag, m in (("fp64_mu_f32", float(np.float32(mu))), ("fp64_mu_up", mu * (1 + 1e-7)),
                           ("fp64_mu_dn", mu * (1 - 1e-7))):
                c = dict(cfg)
                c["friction"] = dict(cfg["friction"], mu=m)
                r = slabpro.make_run(c, precision=abi.FP64)
                r.prime()
                r.run()
                maps[tag] = r.max_maps()[0]
                r.close()
        l1 = lambda a, b: float(np.abs(a - b).sum() / np.abs(b).sum())
        row = {"scenario": i, "particles": len(s["x"]), "friction": s["friction"],
               "fp32_vs_fp64": l1(maps["fp32"], maps["fp64"]), "fp64_moved_vs_fp64": l1(maps["fp64_moved"], maps["fp64"]),
               "fp32_moved_vs_fp32": l1(maps["fp32_moved"], maps["fp32"])}
        for tag in ("fp64_mu_f32", "fp64_mu_up", "fp64_mu_dn"):
            if tag in maps:
                row[tag + "_vs_fp64"] = l1(maps[tag], maps["fp64"])
        log(row)
        rows.append(row)
    print(json.dumps({"metric": "h_max L1: FP32 error; FP64 / FP32 sensitivity to one FP32 offset ulp and FP64 "
                                "sensitivity to mu rounded to FP32 or scaled by 1 +- 1e-7", "rows": rows}),
          flush=True)


def serve_main(args):
    """C4: open-loop Poisson arrivals drawn from the ensemble distribution on the
    2000x2000 DEM; continuous batching with admission by device footprint
    (free slots, particle capacity, reserved active tiles), beside the plain
    FP64 one-at-a-time rate on a size-stratified sample of the same requests."""
    import numpy as np
    from slabpro import abi, configs, ensemble
    dem = configs.c1_dem()
    res = {}
    n = args.requests
    t0 = time.time()
    reqs = [configs.c3_scenario(1000 + k) for k in range(n)]
    sizes = np.array([len(r["x"]) for r in reqs])
    log(f"serve: {n} requests generated in {time.time() - t0:.1f}s, {sizes.sum()} particles")
    for lam in [float(v) for v in args.rates.split(",")]:
        arr, _ = configs.c4_requests(n, lam, seed_index=int(lam * 10))
        r = ensemble.serve(dem, 5.0, reqs, arr, max_slots=args.slots, capacity=args.capacity, progress=log)
        res[f"lambda_{lam:g}"] = {k: v for k, v in r.items()
                                  if k not in ("latency_s", "co_running", "admit_call", "finish_call", "results")}
        log(f"lambda {lam:g}/s: p50 {r['p50_s']:.2f}s p95 {r['p95_s']:.2f}s p99 {r['p99_s']:.2f}s, "
            f"{r['scenarios_per_hour']:.0f} scen/h over {r['wall_s']:.1f}s, rejected {len(r['rejected'])}")
    k = max(1, min(args.fp64_sample, n))
    rank = np.argsort(sizes, kind="stable")
    sample = [int(rank[int((q + 0.5) * n / k)]) for q in range(k)]
    one64 = ensemble.run_one_at_a_time(dem, 5.0, [reqs[i] for i in sample], abi.FP64, progress=log)
    log(f"fp64 one-at-a-time on {k} requests: {one64['scenarios_per_hour']:.0f} scen/h")
    for v in res.values():
        v["vs_fp64_one_at_a_time"] = v["scenarios_per_hour"] / one64["scenarios_per_hour"]
    first = res[next(iter(res))]
    out = {"metric": "p99 completion latency (serving)", "value": first["p99_s"], "unit": "s", "n_gpus": 1,
           "higher_is_better": False, "dtype": "f32", "data": "synthetic (seeded requests, Poisson arrivals)",
           "config": {"workload": f"c4: {n} seeded requests per rate, rates {args.rates}/s, c3 scenario "
                                  f"distribution on the 2000x2000 DEM, {args.slots} slots, "
                                  f"{args.capacity / 1e6:g}M-particle capacity"},
           "per_rate": res,
           "fp64_one_at_a_time": {"scenarios_per_hour": one64["scenarios_per_hour"],
                                  "sample": f"{k} requests, one per size stratum (indices {sample})"}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
