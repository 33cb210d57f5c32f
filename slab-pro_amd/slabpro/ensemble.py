"""Ensemble and serving drivers over the C ABI.

run_batched       -- all scenarios of an ensemble in one FP32 batch (admitted
                     as slots free up), advanced together
run_one_at_a_time -- the plain FP64 path: one scenario per run, sequentially
serve             -- continuous batching: open-loop arrivals, admission by
                     device footprint (free slots + particle capacity), batch
                     reformed every step, p50/p95/p99 completion latency
"""
import collections
import time

import numpy as np

from . import abi


def _params(s, precision):
    return abi.params(precision=precision, **s["params"])


def _fric(s):
    return abi.friction(**s["friction"])


def _footprint(s):
    return len(s["x"])


def run_batched(dem, dx, scenarios, max_slots=256, capacity=None, steps_per_call=32, graph=True, results=False,
                progress=None, every_s=15.0, autotune=True):
    """Advance every scenario in one FP32 batch.  Scenarios are admitted in
    order whenever a slot and particle capacity are free; finished slots are
    collected and released.  Returns wall time and counts; "results"[i] =
    (slot info, max maps or None): the max maps (h_max, v_max2) are downloaded
    for every scenario when results is True, or for the indices in results
    when it is a set.  progress(msg) is called about every every_s seconds
    with the state of the batch (finished scenarios, active slots,
    simulated-time range)."""
    n = len(scenarios)
    cap = capacity or sum(_footprint(s) for s in scenarios)
    t0 = time.perf_counter()
    b = abi.Batch(dem, dx, min(max_slots, n), cap)
    b.use_graph(graph)
    t_setup = time.perf_counter() - t0
    todo = collections.deque(range(n))
    slot_of = {}
    done = {}
    pupd = 0
    calls = 0
    tuned = None
    t0 = time.perf_counter()
    last = t0
    while todo or slot_of:
        while todo:
            s = scenarios[todo[0]]
            slot = b.admit(_params(s, abi.FP32), _fric(s), s["x"], s["y"], s["h"], s["A"], s["m"])
            if slot is None:
                break
            slot_of[slot] = todo.popleft()
        if not slot_of:
            raise RuntimeError(f"scenario {todo[0]} does not fit an empty batch")
        if autotune and tuned is None and calls > 0:
            # launch shapes timed on the live batch (its trial steps are real steps)
            tuned = b.autotune(4)
            if progress:
                progress(f"batched: autotuned {tuned}")
        b.step(steps_per_call)
        calls += 1
        st, tm = b.statuses()
        if progress and time.perf_counter() - last > every_s:
            last = time.perf_counter()
            act = [tm[k] for k in slot_of]
            progress(f"batched: {len(done)}/{n} done, {len(slot_of)} active, {calls * steps_per_call} steps, "
                     f"t in [{min(act):.2f}, {max(act):.2f}] s, {last - t0:.0f}s wall")
        for slot, idx in list(slot_of.items()):
            if st[slot] in (abi.SLOT_DONE, abi.SLOT_FAILED):
                info = b.slot_info(slot)
                pupd += info["np"] * info["step"]
                want = (idx in results) if isinstance(results, (set, frozenset)) else bool(results)
                done[idx] = (info, b.max_maps(slot) if want else None)
                b.release(slot)
                del slot_of[slot]
    wall = time.perf_counter() - t0
    b.close()
    failed = sum(1 for v in done.values() if v[0]["status"] == abi.SLOT_FAILED)
    return {"wall_s": wall, "setup_s": t_setup, "scenarios": n, "failed": failed, "autotuned": tuned,
            "scenarios_per_hour": n / wall * 3600.0, "particle_updates": pupd,
            "particle_updates_per_s": pupd / wall, "results": done}


def run_one_at_a_time(dem, dx, scenarios, precision=abi.FP64, progress=None, results=False):
    """The plain path: each scenario its own run (create, upload, prime, run to
    total_time, download max maps), one after another.  progress(msg) is
    called after each scenario; with results, the max maps (h_max, v_max2) of
    each scenario are returned in "results"."""
    t0 = time.perf_counter()
    per = []
    maps = []
    pupd = 0
    for s in scenarios:
        t1 = time.perf_counter()
        run = abi.Run(dem, dx, _params(s, precision), _fric(s))
        run.upload(s["x"], s["y"], s["h"], s["A"], s["m"])
        run.prime()
        _, info = run.run()
        mm = run.max_maps()
        if results:
            maps.append(mm)
        run.close()
        pupd += len(s["x"]) * info["step"]
        per.append(time.perf_counter() - t1)
        if progress:
            progress(f"one-at-a-time: {len(per)}/{len(scenarios)} ({len(s['x'])} particles, {info['step']} steps) "
                     f"in {per[-1]:.1f}s")
    wall = time.perf_counter() - t0
    return {"wall_s": wall, "scenarios": len(scenarios), "per_scenario_s": per,
            "scenarios_per_hour": len(scenarios) / wall * 3600.0, "particle_updates": pupd,
            "particle_updates_per_s": pupd / wall, "results": maps}


def serve(dem, dx, requests, arrivals, max_slots=64, capacity=8_000_000, steps_per_call=8, graph=True,
          progress=None, collect=None):
    """Continuous batching.  requests[k] (a scenario dict) arrives at wall
    time arrivals[k] (s, open loop).  Each loop: admit arrived requests in FIFO
    order while a slot and particle capacity are free; run a few batch steps;
    collect and release finished slots.  Latency = completion - arrival.

    A request that an empty batch cannot hold (more particles than the
    capacity) is rejected: it is counted in "rejected", gets no latency, and
    the queue moves on.  collect(k, batch, slot), if given, is called for each
    finished request before its slot is released; its return value is kept in
    "results"[k].  "co_running"[k] is the number of requests that were on the
    device when request k was admitted (-1 if it never was); "admit_call"[k]
    and "finish_call"[k] count the batch step calls made before request k was
    admitted and before it was seen finished."""
    b = abi.Batch(dem, dx, max_slots, capacity)
    b.use_graph(graph)
    n = len(requests)
    order = np.argsort(arrivals, kind="stable")
    nxt = 0
    queue = collections.deque()
    slot_of = {}
    lat = np.full(n, np.nan)
    co_running = np.full(n, -1)
    admit_call = np.full(n, -1)
    finish_call = np.full(n, -1)
    calls = 0
    results = {}
    rejected = []
    busy_particles = []
    t0 = time.perf_counter()
    last_report = t0
    while nxt < n or queue or slot_of:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[order[nxt]] <= now:
            queue.append(order[nxt])
            nxt += 1
        while queue:
            k = queue[0]
            s = requests[k]
            slot = b.admit(_params(s, abi.FP32), _fric(s), s["x"], s["y"], s["h"], s["A"], s["m"])
            if slot is None:
                if not slot_of:  # refused by an empty batch: it can never fit
                    rejected.append(int(k))
                    queue.popleft()
                    continue
                break
            co_running[k] = len(slot_of)
            admit_call[k] = calls
            slot_of[slot] = k
            queue.popleft()
        if slot_of:
            b.step(steps_per_call)
            calls += 1
            st, _ = b.statuses()
            fin = time.perf_counter() - t0
            for slot, k in list(slot_of.items()):
                if st[slot] in (abi.SLOT_DONE, abi.SLOT_FAILED):
                    lat[k] = fin - arrivals[k]
                    finish_call[k] = calls
                    if collect is not None:
                        results[int(k)] = collect(k, b, slot)
                    b.release(slot)
                    del slot_of[slot]
            busy_particles.append(sum(len(requests[k]["x"]) for k in slot_of.values()))
        elif nxt < n:
            time.sleep(max(0.0, min(0.05, arrivals[order[nxt]] - (time.perf_counter() - t0))))
        if progress and time.perf_counter() - last_report > 30:
            last_report = time.perf_counter()
            progress(f"serve: {int(np.sum(np.isfinite(lat)))}/{n} done, queue {len(queue)}, active {len(slot_of)}")
    wall = time.perf_counter() - t0
    b.close()
    served = int(np.sum(np.isfinite(lat)))
    pct = (lambda q: float(np.nanpercentile(lat, q))) if served else (lambda q: float("nan"))
    return {"wall_s": wall, "requests": n, "served": served, "rejected": rejected,
            "p50_s": pct(50), "p95_s": pct(95), "p99_s": pct(99),
            "mean_s": float(np.nanmean(lat)) if served else float("nan"),
            "scenarios_per_hour": served / wall * 3600.0, "latency_s": lat, "co_running": co_running,
            "admit_call": admit_call, "finish_call": finish_call,
            "results": results,
            "mean_active_particles": float(np.mean(busy_particles)) if busy_particles else 0.0}
