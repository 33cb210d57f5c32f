"""GPU FP64 form vs the CPU FP64 oracle through the C ABI (needs an MI355X).

Tolerance: the FP64 form differs from the oracle only in the order in which
P2G sums particle contributions into a node, so grid and particle fields after
one step agree to 1e-12 relative (max |a-b| / max |b| per field); after a few
tens of steps to 1e-10.  Everything else (windows, dt, time, frame counters)
is compared exactly.
"""
import math

import numpy as np
import pytest

from conftest import oracle_sim

pytestmark = pytest.mark.gpu


def _cfgs():
    from slabpro import configs
    return configs


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b))) / scale if b.size else 0.0


def _make(cfg):
    import slabpro
    return slabpro.make_run(cfg, keep_grid=True)


def _cmp_grid(run, orc, tol, label=""):
    g = run.grid()
    o = orc.grid()
    assert g["window"] == o["window"], (label, g["window"], o["window"])
    for k_gpu, k_orc in (("mass", "mass"), ("vx", "vx"), ("vy", "vy"), ("h", "h"), ("vz", "vz")):
        e = _rel(g[k_gpu], o[k_orc])
        assert e <= tol, f"{label} grid {k_gpu}: rel err {e:.3e} > {tol:g}"


def _cmp_particles(run, orc, tol, label=""):
    p = run.particles()
    o = orc.particles()
    for k in ("x", "y", "vx", "vy", "vz", "h", "A", "m", "z"):
        e = _rel(p[k], o[k])
        assert e <= tol, f"{label} particle {k}: rel err {e:.3e} > {tol:g}"


def test_device_is_gfx950():
    from slabpro import abi
    info = abi.device_info()
    print("device", info, flush=True)
    assert info["gfx950"]


def test_prime_c0_matches_oracle():
    cfg = _cfgs().c0()
    run = _make(cfg)
    orc = oracle_sim(cfg)
    run.prime()
    orc.prime()
    # t = 0 identity: far-corner weight 2^-10 is cut, so sum m_i = (1 - 2^-10) sum m_p
    g = run.grid()
    assert g["window"] == (34, 54, 33, 43)
    assert abs(g["mass"].sum() - (1 - 2.0 ** -10) * cfg["m"].sum()) <= 1e-9 * cfg["m"].sum()
    _cmp_grid(run, orc, 1e-12, "prime")
    hm, _ = run.max_maps()
    ohm, _ = orc.max_maps()
    assert _rel(hm, ohm) <= 1e-12


def test_one_step_c0_matches_oracle():
    cfg = _cfgs().c0()
    run = _make(cfg)
    orc = oracle_sim(cfg)
    run.prime()
    orc.prime()
    info = run.step(1)
    orc.step()
    sc = orc.scalars()
    assert info["steps_done"] == 1
    assert info["dt"] == sc["dt"] and info["time"] == sc["time"]
    _cmp_grid(run, orc, 1e-12, "step1")
    _cmp_particles(run, orc, 1e-12, "step1")


@pytest.mark.parametrize("fric", [
    {"type": 0, "mu": 0.3},
    {"type": 1, "mu": 0.155, "xi": 2000.0},
    {"type": 2, "mu1": 0.38, "mu2": 0.64, "I0": 0.279, "d": 0.5, "phi": 0.6},
], ids=["coulomb", "voellmy", "mu_I"])
def test_tiny_steps_all_friction_laws(fric):
    cfg = _cfgs().tiny(fric=fric)
    run = _make(cfg)
    orc = oracle_sim(cfg)
    run.prime()
    orc.prime()
    for n in range(1, 31):
        info = run.step(1)
        orc.step()
        sc = orc.scalars()
        assert info["frame"] == sc["frame"] and info["step"] == sc["step"]
        assert abs(info["time"] - sc["time"]) <= 1e-12 * max(1.0, sc["time"])
        if n in (1, 2, 10, 30):
            tol = 1e-12 if n <= 2 else 1e-9
            _cmp_grid(run, orc, tol, f"step{n}")
            _cmp_particles(run, orc, tol, f"step{n}")
    hm, vm = run.max_maps()
    ohm, ovm = orc.max_maps()
    assert _rel(hm, ohm) <= 1e-9 and _rel(vm, ovm) <= 1e-9


def _dense_cfg(n_per_axis, cells, corner):
    """`cells` x `cells` terrain cells (dx = 2 m) from (corner, corner) holding
    n_per_axis^2 particles each on a regular sub-lattice, on a 96 x 80-node
    inclined plane with a valley across it."""
    from slabpro import configs
    nx, ny, dx = 96, 80, 2.0
    xs = np.arange(nx) * dx
    ys = np.arange(ny) * dx
    z = np.ascontiguousarray(500.0 - 0.45 * xs[:, None] + 0.002 * (ys[None, :] - 80.0) ** 2)
    o = (np.arange(n_per_axis) + 0.5) / n_per_axis
    ci = np.arange(cells)
    px = (corner / dx + ci[:, None, None, None] + o[None, None, :, None]) * dx
    py = (corner / dx + ci[None, :, None, None] + o[None, None, None, :]) * dx
    px, py = np.broadcast_arrays(px, py)
    px, py = px.ravel().copy(), py.ravel().copy()
    part = configs.finish_particles(px, py, np.ones(len(px)), (cells * dx) ** 2, 300.0)
    params = dict(gravity=9.81, rho=300.0, C_cfl=0.5, dt_max=0.01, flip_ratio=0.95, output_dt=0.5,
                  total_time=0.3)
    return {"name": "dense", "z": z, "dx": dx, "params": params,
            "friction": {"type": 1, "mu": 0.155, "xi": 2000.0}, **part}


@pytest.mark.parametrize("n_per_axis,cells,corner", [(6, 6, 68.0), (6, 14, 60.0), (20, 4, 66.0)],
                         ids=["1296_in_lds", "5184_beyond_lds", "6400_cells_of_400"])
def test_dense_tiles_match_oracle(n_per_axis, cells, corner):
    """Tiles of 1,296 particles (sorted in LDS), 5,184 and 6,400 particles
    (beyond the LDS capacity of 2,048: sorted in global memory; the last with
    400 particles per cell) follow the oracle: 1e-12 after one step, 1e-9 after
    twenty."""
    cfg = _dense_cfg(n_per_axis, cells, corner)
    run = _make(cfg)
    orc = oracle_sim(cfg)
    run.prime()
    orc.prime()
    _cmp_grid(run, orc, 1e-12, "prime")
    for n in range(1, 21):
        info = run.step(1)
        orc.step()
        sc = orc.scalars()
        assert info["step"] == sc["step"]
        assert abs(info["dt"] - sc["dt"]) <= 1e-12 * sc["dt"]
        if n in (1, 20):
            tol = 1e-12 if n == 1 else 1e-9
            _cmp_grid(run, orc, tol, f"step{n}")
            _cmp_particles(run, orc, tol, f"step{n}")


def test_run_twice_bitwise_identical():
    cfg = _cfgs().tiny(fric={"type": 1, "mu": 0.2, "xi": 1000.0})
    outs = []
    for _ in range(2):
        run = _make(cfg)
        run.prime()
        run.step(25)
        outs.append((run.particles(), run.grid(), run.max_maps()))
        run.close()
    (p1, g1, m1), (p2, g2, m2) = outs
    for k in p1:
        assert np.array_equal(p1[k], p2[k]), k
    for k in ("mass", "vx", "vy", "h"):
        assert np.array_equal(g1[k], g2[k]), k
    assert np.array_equal(m1[0], m2[0]) and np.array_equal(m1[1], m2[1])


def test_separate_tile_scan_launch_matches_merged(monkeypatch):
    """FP64 block offsets of the tile scan from a separate scan launch (the
    form used beyond 2,048 scan blocks, forced with
    SLAB_SCAN_MERGE_MAX_BLOCKS=0) give bitwise the results of the offsets
    formed inside the offsets kernel."""
    cfg = _cfgs().tiny(fric={"type": 0, "mu": 0.3})
    outs = []
    for split in (False, True):
        if split:
            monkeypatch.setenv("SLAB_SCAN_MERGE_MAX_BLOCKS", "0")
        run = _make(cfg)
        run.prime()
        info = run.step(40)
        outs.append((run.particles(), run.grid(), run.max_maps(), info))
        run.close()
    (p1, g1, m1, i1), (p2, g2, m2, i2) = outs
    assert i1["step"] == i2["step"] and i1["time"] == i2["time"]
    for k in p1:
        assert np.array_equal(p1[k], p2[k]), k
    for k in ("mass", "vx", "vy", "h"):
        assert np.array_equal(g1[k], g2[k]), k
    assert np.array_equal(m1[0], m2[0]) and np.array_equal(m1[1], m2[1])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_frames_and_stop_semantics(precision):
    """The run returns at every output frame with the oracle's frame and step
    counters and time, and stops at total_time (both forms: dt_max binds on
    this case, so the FP32 form takes the same steps)."""
    import slabpro
    from slabpro import abi
    cfg = _cfgs().tiny()
    cfg["params"]["total_time"] = 1.2
    cfg["params"]["output_dt"] = 0.25
    run = slabpro.make_run(cfg, precision=abi.FP64 if precision == "fp64" else abi.FP32, keep_grid=True)
    orc = oracle_sim(cfg)
    run.prime()
    orc.prime()
    frames = []
    while True:
        info = run.step(1000)
        if info["frame_hit"]:
            frames.append((info["frame"], info["step"], info["time"]))
        if info["finished"]:
            break
    ofr = []
    while True:
        r = orc.step()
        if r == 1:
            s = orc.scalars()
            ofr.append((s["frame"], s["step"], s["time"]))
        if r == -1 or orc.scalars()["time"] >= 1.2:
            break
    assert len(ofr) == 4 and [f[:2] for f in frames] == [f[:2] for f in ofr]
    for a, b in zip(frames, ofr):
        assert abs(a[2] - b[2]) <= 1e-12


def test_c0_full_run_max_maps():
    """60 s of C0 (6,000 steps) on both sides; max maps agree within 1e-6
    relative (L-inf over the map, normalised by the map maximum)."""
    cfg = _cfgs().c0()
    run = _make(cfg)
    run.keep_grid(False)
    orc = oracle_sim(cfg)
    run.prime()
    orc.prime()
    nsteps, info = run.run()
    orc.run()
    sc = orc.scalars()
    print("c0 steps", nsteps, info, sc["step"], flush=True)
    assert info["step"] == sc["step"]
    hm, vm = run.max_maps()
    ohm, ovm = orc.max_maps()
    eh, ev = _rel(hm, ohm), _rel(vm, ovm)
    print(f"c0 max-map rel err h {eh:.3e} v2 {ev:.3e}", flush=True)
    assert eh <= 1e-6 and ev <= 1e-6
    p = run.particles()
    o = orc.particles()
    assert _rel(p["x"], o["x"]) <= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dt_maxima_follow_the_particles(precision):
    """The per-step max |v_p|^2 partials (integer atomics into shards, consumed
    and reset by the next step's dt) track the particles: diag's max |v_p| is
    the downloaded particles' after priming and after each step, and step 1's
    dt is the CFL dt of the velocities set before priming
    (src/simulation.cpp:565-582)."""
    from slabpro import abi
    import slabpro
    cfg = _cfgs().tiny()
    run = slabpro.make_run(cfg, precision=abi.FP64 if precision == "fp64" else abi.FP32)
    n = len(cfg["x"])
    g = np.random.default_rng(3)
    vx, vy = g.uniform(-150.0, 150.0, n), g.uniform(-150.0, 150.0, n)  # CFL-limited dt
    run.set_velocity(vx, vy)
    run.prime()
    tol = 1e-14 if precision == "fp64" else 1e-6
    v0 = math.sqrt(float(np.max(vx * vx + vy * vy)))
    if precision == "fp64":
        assert run.diag()["vmax"] == pytest.approx(v0, rel=tol)
    info = run.step(1)
    prm = cfg["params"]
    dt_cfl = prm["C_cfl"] * cfg["dx"] / v0
    assert dt_cfl < prm["dt_max"]
    assert info["dt"] == pytest.approx(dt_cfl, rel=tol)
    for _ in range(6):
        info = run.step(1)
        p = run.particles()
        vm = math.sqrt(float(np.max(p["vx"] ** 2 + p["vy"] ** 2)))
        if precision == "fp64":
            assert run.diag()["vmax"] == pytest.approx(vm, rel=tol)
    run.close()
