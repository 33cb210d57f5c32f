"""The single-run, ensemble and serving workloads on the GPU -- needs an MI355X.

  - C1 (2000 x 2000 DEM at 5 m, 40 ellipse releases, ~576k particles,
    Voellmy, 180 s): the FP32 form against the FP64 form, max-height map L1
    <= 1 %, max-velocity map L1 <= 2 %, runout within 2 cells.
  - C3 (the 256-scenario ensemble on the C1 DEM): a size-stratified sample
    of nine scenarios -- the smallest, a middle and the largest of each
    friction law (2k to 1M particles) -- advanced together in one FP32 batch;
    each slot's max-height map within L1 <= 1 % of the same scenario run
    alone in the FP64 form, and every slot finishes at its own total_time.
  - C4 (serving): requests from the C3 distribution arriving while others
    are on the device; each finished request's max maps are bitwise equal to
    the same request run alone (tolerance: exact), and requests were admitted
    while others were mid-run.
  - Cancellation: a running slot is released mid-run, its capacity reused by
    a new admission; every other scenario is bitwise equal to its alone run.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_DEM = {}


def _dem():
    if "c1" not in _DEM:
        from slabpro import configs
        _DEM["c1"] = configs.c1_dem()
    return _DEM["c1"]


def _hmax_l1(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def test_c1_fp32_error_against_fp64():
    from slabpro import accuracy, configs
    cfg = configs.c1()
    print("c1:", len(cfg["x"]), "particles", flush=True)
    rep, tm = accuracy.compare_forms(cfg)
    print("c1 fp32-vs-fp64 error report", rep, tm, flush=True)
    assert rep["steps_fp32"] == rep["steps_fp64"]
    assert rep["hmax_L1"] <= 0.01
    assert rep["vmax_L1"] <= 0.02
    assert rep["runout_diff_cells"] <= 2.0
    assert abs(rep["mass_fp32"] - rep["mass_fp64"]) <= 1e-6 * rep["mass_fp64"]


# smallest, a middle and the largest scenario of each friction law among the
# 256 (indices into configs.c3_scenario; sizes 2,069 .. 981,804 target particles)
C3_SAMPLE = [147, 134, 229,   # Coulomb: 2.3k, 39k, 982k
             125, 139, 3,     # Voellmy: 2.2k, 26k, 818k
             110, 46, 152]    # mu(I):   2.1k, 86k, 919k


def test_c3_stratified_batch_vs_fp64_alone():
    from slabpro import abi, configs, ensemble
    dem = _dem()
    scen = [configs.c3_scenario(i) for i in C3_SAMPLE]
    assert {s["friction"]["type"] for s in scen} == {0, 1, 2}
    print("c3 sample sizes", [len(s["x"]) for s in scen], flush=True)
    res = ensemble.run_batched(dem, 5.0, scen, max_slots=len(scen), results=True, autotune=True)
    assert res["failed"] == 0
    one = ensemble.run_one_at_a_time(dem, 5.0, scen, results=True)
    errs = []
    for k, s in enumerate(scen):
        info, (h32, v32) = res["results"][k]
        h64, v64 = one["results"][k]
        assert info["status"] == abi.SLOT_DONE
        assert abs(info["time"] - s["params"]["total_time"]) < 1e-9
        errs.append(_hmax_l1(h32, h64))
    print("c3 per-scenario hmax L1 (fp32 batch vs fp64 alone)", [f"{e:.4%}" for e in errs], flush=True)
    print("c3 batch", res["wall_s"], "s; fp64 one at a time", one["wall_s"], "s", flush=True)
    assert max(errs) <= 0.01


def _tiny_dem():
    nx, ny, dx = 128, 112, 5.0
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    z = 800.0 - 0.5 * x[:, None] + 0.0015 * (y[None, :] - 280.0) ** 2
    return np.ascontiguousarray(z), dx


def _small_request(seed, t_range=(4.0, 10.0)):
    """A small scenario on _tiny_dem: 1-3 ellipses, 1-6k particles, any law."""
    from slabpro import configs
    g = np.random.default_rng(seed)
    dx = 5.0
    px, py, h = [], [], []
    for _ in range(int(g.integers(1, 4))):
        cx, cy = g.uniform(80.0, 160.0), g.uniform(180.0, 380.0)
        x, y, _ = configs.ellipse_release(dx, cx, cy, g.uniform(25.0, 60.0), g.uniform(25.0, 60.0), g.uniform(0, 180))
        px.append(x)
        py.append(y)
        h.append(np.full(len(x), g.uniform(0.5, 3.0)))
    px, py, h = np.concatenate(px), np.concatenate(py), np.concatenate(h)
    A = np.full(len(px), dx * dx / 4.0)
    u = g.uniform()
    if u < 0.5:
        fr, rho = {"type": 1, "mu": float(g.uniform(0.1, 0.3)), "xi": float(g.uniform(500.0, 4000.0))}, 300.0
    elif u < 0.8:
        fr, rho = {"type": 0, "mu": float(g.uniform(0.3, 0.6))}, 300.0
    else:
        fr, rho = {"type": 2, "mu1": 0.38, "mu2": 0.64, "I0": 0.279, "d": 0.5, "phi": 0.6}, 2000.0
    params = dict(gravity=9.81, rho=rho, C_cfl=0.5, dt_max=0.01, flip_ratio=0.95, output_dt=1.0,
                  total_time=float(g.uniform(*t_range)))
    return {"x": px, "y": py, "h": h, "A": A, "m": rho * h * A, "params": params, "friction": fr}


def _alone(z, dx, s):
    """One scenario alone in its own batch: (h_max, v_max2, final particles)."""
    from slabpro import abi, ensemble
    b = abi.Batch(z, dx, 1, len(s["x"]))
    slot = b.admit(ensemble._params(s, abi.FP32), ensemble._fric(s), s["x"], s["y"], s["h"], s["A"], s["m"])
    assert slot is not None
    while b.step(64) > 0:
        pass
    assert b.slot_info(slot)["status"] == abi.SLOT_DONE
    out = b.max_maps(slot) + (b.particles(slot),)
    b.close()
    return out


def test_c4_serve_results_equal_alone():
    from slabpro import ensemble
    z, dx = _tiny_dem()
    # 20-40 s of simulated time (2,000-4,000 steps) each, arriving 0.1 s apart:
    # later requests arrive while earlier ones run
    reqs = [_small_request(500 + k, (20.0, 40.0)) for k in range(10)]
    arrivals = np.arange(10) * 0.1
    # room for four at a time: the rest queue and are admitted as slots free up
    cap = sum(sorted(len(r["x"]) for r in reqs)[-4:])

    def collect(k, b, slot):
        return b.max_maps(slot) + (b.particles(slot),)

    res = ensemble.serve(z, dx, reqs, arrivals, max_slots=4, capacity=cap, steps_per_call=8, collect=collect)
    a, f = res["admit_call"], res["finish_call"]
    print("serve:", {k: res[k] for k in ("served", "p50_s", "p99_s", "rejected")},
          "admit_call", a.tolist(), "finish_call", f.tolist(), flush=True)
    assert res["served"] == len(reqs) and not res["rejected"]
    # some request was admitted after another had stepped and before it finished
    mid = [(j, k) for j in range(10) for k in range(10) if a[j] < a[k] < f[j]]
    assert mid, "no request was admitted while another was mid-run"
    for k, r in enumerate(reqs):
        hm, vm, p = res["results"][k]
        ha, va, pa = _alone(z, dx, r)
        assert np.array_equal(hm, ha) and np.array_equal(vm, va), k
        for f in ("x", "y", "vx", "vy", "h"):
            assert np.array_equal(p[f], pa[f]), (k, f)


def test_serve_rejects_a_request_larger_than_the_capacity():
    from slabpro import ensemble
    z, dx = _tiny_dem()
    reqs = [_small_request(600 + k) for k in range(3)]
    cap = max(len(reqs[0]["x"]), len(reqs[2]["x"]))
    big = dict(reqs[1])
    big["x"] = np.concatenate([reqs[1]["x"]] * 3)  # cannot fit an empty batch
    for f in ("y", "h", "A", "m"):
        big[f] = np.concatenate([reqs[1][f]] * 3)
    assert len(big["x"]) > cap
    res = ensemble.serve(z, dx, [reqs[0], big, reqs[2]], np.array([0.0, 0.0, 0.0]), max_slots=2, capacity=cap)
    assert res["rejected"] == [1]
    assert res["served"] == 2 and np.isnan(res["latency_s"][1])


def test_cancel_running_slot_and_reuse():
    from slabpro import abi, ensemble
    z, dx = _tiny_dem()
    reqs = [_small_request(700 + k) for k in range(4)]
    cap = len(reqs[0]["x"]) + max(len(reqs[1]["x"]), len(reqs[3]["x"])) + len(reqs[2]["x"])
    b = abi.Batch(z, dx, 3, cap)
    slots = []
    for r in reqs[:3]:
        s = b.admit(ensemble._params(r, abi.FP32), ensemble._fric(r), r["x"], r["y"], r["h"], r["A"], r["m"])
        assert s is not None
        slots.append(s)
    b.step(40)
    assert b.slot_info(slots[1])["status"] == abi.SLOT_RUNNING
    b.release(slots[1])  # cancelled while running
    st, _ = b.statuses()
    assert st[slots[1]] == abi.SLOT_FREE
    r = reqs[3]
    s3 = b.admit(ensemble._params(r, abi.FP32), ensemble._fric(r), r["x"], r["y"], r["h"], r["A"], r["m"])
    assert s3 == slots[1]  # the freed slot and its capacity are reused
    while b.step(64) > 0:
        pass
    for k, s in ((0, slots[0]), (2, slots[2]), (3, s3)):
        assert b.slot_info(s)["status"] == abi.SLOT_DONE
        hm, vm = b.max_maps(s)
        p = b.particles(s)
        ha, va, pa = _alone(z, dx, reqs[k])
        assert np.array_equal(hm, ha) and np.array_equal(vm, va), k
        assert np.array_equal(p["x"], pa["x"]) and np.array_equal(p["h"], pa["h"]), k
    b.close()
