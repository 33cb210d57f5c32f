"""Batched FP32 engine (many scenarios on one DEM) -- needs an MI355X.

The engine orders particles by (slot, tile, cell, creation index) and every
per-slot quantity (dt, friction, grid, max maps) depends only on that slot's
own particles, so a scenario's results are bitwise the same alone and in any
mix (tolerance: exact).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenario(seed, fric, n_extra=0):
    from slabpro import configs
    g = np.random.default_rng(seed)
    dx = 2.0
    cx = 60.0 + 20.0 * g.uniform()
    cy = 60.0 + 20.0 * g.uniform()
    px, py, area = configs.ellipse_release(dx, cx, cy, 14.0 + n_extra, 10.0 + g.uniform() * 4, g.uniform(0, 180))
    h = np.full(len(px), 1.0 + g.uniform())
    part = configs.finish_particles(px, py, h, area * h[0], 300.0)
    return part, fric


def _terrain():
    nx, ny, dx = 96, 80, 2.0
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    z = 500.0 - 0.45 * x[:, None] + 0.002 * (y[None, :] - 80.0) ** 2
    return np.ascontiguousarray(z), dx


FRICS = [{"type": 0, "mu": 0.3}, {"type": 1, "mu": 0.155, "xi": 2000.0},
         {"type": 2, "mu1": 0.38, "mu2": 0.64, "I0": 0.279, "d": 0.5, "phi": 0.6}]


def _params(total=3.0):
    from slabpro import abi
    return abi.params(rho=300.0, total_time=total, output_dt=0.5, precision=abi.FP32)


def _run_batch(scens, max_slots=8, cap=200000, steps=100000, total=3.0, resort=None, sort_cap=None):
    from slabpro import abi
    z, dx = _terrain()
    b = abi.Batch(z, dx, max_slots, cap)
    if resort is not None:
        b.set_resort(resort)
    if sort_cap is not None:
        b.set_sort_capacity(sort_cap)
        assert b.launch_shape()["sort_capacity"] == sort_cap
    slots = []
    for part, fr in scens:
        s = b.admit(_params(total), abi.friction(**fr), part["x"], part["y"], part["h"], part["A"], part["m"])
        assert s is not None
        slots.append(s)
    n = 0
    while b.step(64) > 0:
        n += 1
        assert n < steps
    out = [(b.max_maps(s), b.particles(s), b.slot_info(s)) for s in slots]
    b.close()
    return out


def test_alone_and_in_mix_are_bitwise_equal():
    target = _scenario(11, FRICS[1])
    others = [_scenario(20 + k, FRICS[k % 3], n_extra=3 * k) for k in range(5)]
    (alone,) = _run_batch([target])
    mix = _run_batch(others[:2] + [target] + others[2:])
    (hm_a, vm_a), p_a, i_a = alone
    (hm_m, vm_m), p_m, i_m = mix[2]
    assert i_a["status"] == i_m["status"] and i_a["step"] == i_m["step"] and i_a["time"] == i_m["time"]
    assert np.array_equal(hm_a, hm_m) and np.array_equal(vm_a, vm_m)
    for k in p_a:
        assert np.array_equal(p_a[k], p_m[k]), k


def test_batch_slot_equals_single_fp32_run():
    import slabpro
    from slabpro import abi
    part, fr = _scenario(5, FRICS[2])
    (res,) = _run_batch([(part, fr)])
    z, dx = _terrain()
    cfg = {"z": z, "dx": dx, "params": dict(rho=300.0, total_time=3.0, output_dt=0.5), "friction": fr, **part}
    run = slabpro.make_run(cfg, precision=abi.FP32)
    run.prime()
    run.run()
    hm, vm = run.max_maps()
    p = run.particles()
    assert np.array_equal(hm, res[0][0]) and np.array_equal(vm, res[0][1])
    for k in ("x", "y", "vx", "vy", "h"):
        assert np.array_equal(p[k], res[1][k]), k


def test_non_finite_request_rejected_and_device_failure_isolated():
    from slabpro import abi
    z, dx = _terrain()
    good = [_scenario(31, FRICS[0]), _scenario(32, FRICS[1])]
    bad_part, bad_fr = _scenario(33, FRICS[2])
    # a NaN input is refused at admission; nothing is admitted
    b = abi.Batch(z, dx, 4, 100000)
    h = bad_part["h"].copy()
    h[7] = np.nan
    with pytest.raises(abi.SlabError) as ei:
        b.admit(_params(), abi.friction(**bad_fr), bad_part["x"], bad_part["y"], h, bad_part["A"])
    assert ei.value.status == abi.EINVAL
    assert b.footprint()["free_slots"] == 4
    b.close()
    # a finite input that overflows on the device fails only its own slot
    ref = _run_batch(good)
    b = abi.Batch(z, dx, 8, 200000)
    s0 = b.admit(_params(), abi.friction(**good[0][1]), *[good[0][0][k] for k in ("x", "y", "h", "A", "m")])
    vx = np.zeros(len(bad_part["x"]))
    vx[7] = 1e30  # |v|^2 overflows FP32 on the device
    s1 = b.admit(_params(), abi.friction(**bad_fr), bad_part["x"], bad_part["y"], bad_part["h"], bad_part["A"],
                 vx=vx)
    s2 = b.admit(_params(), abi.friction(**good[1][1]), *[good[1][0][k] for k in ("x", "y", "h", "A", "m")])
    while b.step(64) > 0:
        pass
    assert b.slot_info(s1)["status"] == abi.SLOT_FAILED
    for r, s in ((ref[0], s0), (ref[1], s2)):
        assert b.slot_info(s)["status"] == abi.SLOT_DONE
        assert np.array_equal(r[0][0], b.max_maps(s)[0])
        assert np.array_equal(r[1]["x"], b.particles(s)["x"])
    b.close()


def test_admission_capacity_release_and_reuse():
    from slabpro import abi
    z, dx = _terrain()
    parts = [_scenario(40 + k, FRICS[k % 3]) for k in range(4)]
    npart = [len(p["x"]) for p, _ in parts]
    cap = npart[0] + npart[1] + 10
    b = abi.Batch(z, dx, 3, cap)
    fp = b.footprint()
    assert fp["free_slots"] == 3 and fp["free_particles"] == cap and fp["bytes_per_particle"] > 0
    s0 = b.admit(_params(1.0), abi.friction(**parts[0][1]), **{k: parts[0][0][k] for k in ("x", "y", "h", "A", "m")})
    s1 = b.admit(_params(1.0), abi.friction(**parts[1][1]), **{k: parts[1][0][k] for k in ("x", "y", "h", "A", "m")})
    assert (s0, s1) == (0, 1)
    # particle capacity exhausted -> refused, nothing changes
    assert b.admit(_params(1.0), abi.friction(**parts[2][1]),
                   **{k: parts[2][0][k] for k in ("x", "y", "h", "A", "m")}) is None
    while b.step(64) > 0:
        pass
    st, tm = b.statuses()
    assert st[0] == abi.SLOT_DONE and st[1] == abi.SLOT_DONE and tm[0] >= 1.0
    p0 = b.particles(0)
    assert len(p0["x"]) == npart[0] and np.all(np.isfinite(p0["x"]))
    b.release(0)
    b.release(1)
    b.step(1)  # released particles drop out of the state
    assert b.footprint()["free_particles"] == cap
    s2 = b.admit(_params(1.0), abi.friction(**parts[2][1]), **{k: parts[2][0][k] for k in ("x", "y", "h", "A", "m")})
    assert s2 == 0
    while b.step(64) > 0:
        pass
    assert b.slot_info(s2)["status"] == abi.SLOT_DONE
    b.close()


def test_graph_replay_matches_eager():
    from slabpro import abi
    z, dx = _terrain()
    scens = [_scenario(50 + k, FRICS[k % 3]) for k in range(3)]
    outs = []
    for graph in (False, True):
        b = abi.Batch(z, dx, 4, 100000)
        b.use_graph(graph)
        slots = [b.admit(_params(2.0), abi.friction(**fr), p["x"], p["y"], p["h"], p["A"], p["m"]) for p, fr in scens]
        for _ in range(5):  # several replays with host work in between
            b.step(17)
            b.statuses()
        while b.step(64) > 0:
            pass
        outs.append([(b.max_maps(s)[0], b.particles(s)["x"]) for s in slots])
        b.close()
    for (ha, xa), (hb, xb) in zip(*outs):
        assert np.array_equal(ha, hb) and np.array_equal(xa, xb)


def test_outflow_particles_are_parked_and_isolated():
    """Particles that flow off the terrain (cell outside [1, N-3]) are frozen
    in place with zero velocity and leave P2G/G2P; a neighbour slot's results
    are unchanged by them (bitwise, alone vs in the mix)."""
    from slabpro import abi, configs
    dx = 2.0
    px, py, area = configs.ellipse_release(dx, 172.0, 80.0, 10.0, 8.0, 0.0)   # 18 m from the +x edge
    h = np.full(len(px), 1.5)
    edge = (configs.finish_particles(px, py, h, area * 1.5, 300.0), {"type": 0, "mu": 0.05})
    inner = _scenario(11, FRICS[1])
    (hm_e, pe, ie), (hm_i, pi_, ii) = _run_batch([edge, inner], total=6.0)
    (hm_a, pa, ia), = _run_batch([inner], total=6.0)
    assert ie["status"] == abi.SLOT_DONE and ii["status"] == abi.SLOT_DONE
    nx, ny = _terrain()[0].shape
    x_out = (nx - 2) * dx
    out = pe["x"] >= x_out
    print(f"outflow: {int(out.sum())} of {len(out)} particles parked", flush=True)
    assert out.sum() > 0.9 * len(out)
    assert np.all(pe["vx"][out] == 0) and np.all(pe["vy"][out] == 0)
    assert np.all(pe["x"] <= (nx - 1) * dx) and np.all(np.isfinite(pe["h"]))
    assert np.allclose(pe["m"].sum(), edge[0]["m"].sum(), rtol=1e-6)
    for k in ("x", "y", "vx", "vy", "h"):
        assert np.array_equal(pi_[k], pa[k]), k
    assert np.array_equal(hm_i[0], hm_a[0])


def test_separate_tile_scan_launch_matches_merged(monkeypatch):
    """Block offsets of the tile scan from a separate scan launch (the form
    used beyond 2,048 scan blocks, forced here with
    SLAB_SCAN_MERGE_MAX_BLOCKS=0) give bitwise the results of the offsets
    formed inside the offsets kernel."""
    scens = [_scenario(41, FRICS[0]), _scenario(42, FRICS[2])]
    a = _run_batch(scens, total=1.0)
    monkeypatch.setenv("SLAB_SCAN_MERGE_MAX_BLOCKS", "0")
    b = _run_batch(scens, total=1.0)
    for (ma, pa, ia), (mb, pb, ib) in zip(a, b):
        assert ia["step"] == ib["step"] and ia["status"] == ib["status"]
        for k in ("x", "y", "vx", "vy", "h", "z"):
            assert np.array_equal(pa[k], pb[k]), k
        assert np.array_equal(ma[0], mb[0]) and np.array_equal(ma[1], mb[1])


@pytest.mark.parametrize("period", [1, 5])
def test_resort_period_does_not_change_batch_results(period):
    """Every scenario of a mix is bitwise the same with storage re-sorted every
    step and every `period` steps (default 16)."""
    scens = [_scenario(3, FRICS[0]), _scenario(4, FRICS[1]), _scenario(5, FRICS[2])]
    a = _run_batch(scens, resort=period)
    b = _run_batch(scens)
    for (ma, pa, ia), (mb, pb, ib) in zip(a, b):
        assert ia["step"] == ib["step"]
        for k in ("x", "y", "vx", "vy", "h", "z"):
            assert np.array_equal(pa[k], pb[k]), k
        assert np.array_equal(ma[0], mb[0]) and np.array_equal(ma[1], mb[1])


def _dense(n_per_axis, cells, x0, y0, h=1.0):
    """A block of `cells` x `cells` terrain cells (dx = 2 m) holding n_per_axis^2
    particles each on a regular sub-lattice: dense tiles and cells."""
    from slabpro import configs
    dx = 2.0
    o = (np.arange(n_per_axis) + 0.5) / n_per_axis
    ci = np.arange(cells)
    px = (x0 / dx + ci[:, None, None, None] + o[None, None, :, None]) * dx
    py = (y0 / dx + ci[None, :, None, None] + o[None, None, None, :]) * dx
    px, py = np.broadcast_arrays(px, py)
    px, py = px.ravel().copy(), py.ravel().copy()
    hh = np.full(len(px), h)
    area = (cells * dx) ** 2
    return configs.finish_particles(px, py, hh, area * h, 300.0)


@pytest.mark.parametrize("n_per_axis,cells,corner",
                         [(6, 6, 68.0), (17, 3, 66.0), (6, 14, 60.0), (20, 4, 66.0), (5, 16, 64.0)])
def test_dense_tiles_sort_canonically(n_per_axis, cells, corner):
    """Tiles within the in-LDS sort capacity whose cells hold 36 particles
    (each placed by its rank in its cell) and 289 particles (bitonic sort of
    the whole tile), and tiles beyond it (4,761 particles in cells of 36;
    6,400 in cells of ~100; ~6,000 in cells of at most 25: LDS rounds of whole
    cells) are ordered canonically: results are bitwise the same alone, in a
    mix and with another storage re-sort period, and stay close to the FP64
    form."""
    from slabpro import abi
    dense = (_dense(n_per_axis, cells, corner, corner), FRICS[1])
    other = _scenario(21, FRICS[0])
    a = _run_batch([dense], cap=400000, total=0.3, resort=1)
    b = _run_batch([other, dense], cap=400000, total=0.3)
    (ma, pa, ia), = a
    (_, _, _), (mb, pb, ib) = b
    assert ia["step"] == ib["step"] and ia["status"] == abi.SLOT_DONE
    for k in ("x", "y", "vx", "vy", "h"):
        assert np.array_equal(pa[k], pb[k]), k
    assert np.array_equal(ma[0], mb[0])
    # FP64 form of the same run (its own global-memory sort)
    z, dx = _terrain()
    part = dense[0]
    r = abi.Run(z, dx, abi.params(rho=300.0, total_time=0.3, output_dt=0.5, precision=abi.FP64),
                abi.friction(**dense[1]))
    r.upload(part["x"], part["y"], part["h"], part["A"], part["m"])
    r.prime()
    r.run()
    q = r.particles()
    r.close()
    assert np.max(np.abs(pa["x"] - q["x"])) < 1e-3
    assert np.max(np.abs(pa["vx"] - q["vx"])) < 1e-2 * max(1.0, np.abs(q["vx"]).max())


@pytest.mark.parametrize("n_per_axis,cells,corner", [(6, 6, 68.0), (17, 3, 66.0), (20, 4, 66.0), (5, 16, 64.0)])
def test_sort_capacity_does_not_change_results(n_per_axis, cells, corner):
    """The in-LDS capacity of the tile sort (1,024, 2,048 or 4,096 particles) is
    a launch-shape choice: tiles of 1,296 (cells of 36), 2,601 (cells of 289),
    6,400 (cells of 400) and ~6,000 particles, sorted in LDS or in rounds of
    whole cells of at most half the capacity, give bitwise the same results.
    The 1,296-particle tiles also take every P2G form: the sort in global
    memory (capacity 1,024), the in-LDS sort with each cell's particles read
    from global memory as its thread goes (2,048: the fields of 720 or more
    particles do not fit the LDS), and with the fields staged in LDS (4,096)."""
    dense = (_dense(n_per_axis, cells, corner, corner), FRICS[1])
    other = _scenario(21, FRICS[0])
    ref = _run_batch([other, dense], cap=400000, total=0.3)
    for sc in (1024, 2048, 4096):
        got = _run_batch([other, dense], cap=400000, total=0.3, sort_cap=sc)
        for (ma, pa, ia), (mb, pb, ib) in zip(ref, got):
            assert ia["step"] == ib["step"] and ia["status"] == ib["status"], sc
            for k in ("x", "y", "vx", "vy", "h", "z"):
                assert np.array_equal(pa[k], pb[k]), (sc, k)
            assert np.array_equal(ma[0], mb[0]) and np.array_equal(ma[1], mb[1]), sc


def test_active_tile_overflow_stops_the_batch(monkeypatch):
    # An active-tile capacity of one tile (SLAB_TILE_CAP, read when the batch
    # is created) against a release that spans several tiles: the step is
    # refused with SLAB_EDOMAIN, after which the half-started step must not be
    # resumed -- step, admit and release are refused with SLAB_ESTATE -- while
    # downloads still return the last completed state.
    from slabpro import abi
    z, dx = _terrain()
    part, fr = _scenario(51, FRICS[0], n_extra=10)
    cells = {(int(x // dx) // 16, int(y // dx) // 16) for x, y in zip(part["x"], part["y"])}
    assert len(cells) >= 2
    monkeypatch.setenv("SLAB_TILE_CAP", "1")
    b = abi.Batch(z, dx, 2, 100000)
    monkeypatch.delenv("SLAB_TILE_CAP")
    b.set_tile_growth(0)  # no tile check at admission: the step meets the overflow
    s = b.admit(_params(), abi.friction(**fr), part["x"], part["y"], part["h"], part["A"], part["m"])
    assert s is not None
    with pytest.raises(abi.SlabError) as ei:
        b.step(4)
    assert ei.value.status == abi.EDOMAIN
    with pytest.raises(abi.SlabError) as ei:
        b.step(1)
    assert ei.value.status == abi.ESTATE
    with pytest.raises(abi.SlabError) as ei:
        b.admit(_params(), abi.friction(**fr), part["x"], part["y"], part["h"], part["A"], part["m"])
    assert ei.value.status == abi.ESTATE
    with pytest.raises(abi.SlabError) as ei:
        b.release(s)
    assert ei.value.status == abi.ESTATE
    p = b.particles(s)
    assert np.max(np.abs(p["x"] - part["x"])) < 1e-3 * dx  # the admitted state (FP32 offsets)
    hm, _ = b.max_maps(s)
    assert np.all(hm == 0.0)  # no step completed
    b.close()


def test_admission_by_active_tiles(monkeypatch):
    # A tile capacity that holds one release's reserved tiles but not two:
    # the second scenario is refused (admit -> None, SLAB_ENOMEM) although a
    # slot and particle capacity are free, and admitted once the first is
    # released.  The grid velocity lives in the batch's grid-tile pool; the
    # scenario's results equal its run alone in a batch with room to spare.
    from slabpro import abi
    z, dx = _terrain()
    part, fr = _scenario(61, FRICS[1])
    ib = np.floor(np.asarray(part["x"]) / dx - 0.5).astype(int)
    jb = np.floor(np.asarray(part["y"]) / dx - 0.5).astype(int)
    pt = {(i >> 4, j >> 4) for i, j in zip(ib, jb)}
    gt = {(a + u, b + v) for a, b in pt for u in (0, 1) for v in (0, 1)}
    growth = 4
    cap = growth * max(len(pt), (len(gt) + 1) // 2)
    monkeypatch.setenv("SLAB_TILE_CAP", str(cap))
    b = abi.Batch(z, dx, 4, 200000)
    monkeypatch.delenv("SLAB_TILE_CAP")
    b.set_tile_growth(growth)
    t = b.tiles()
    print("tiles", t, "release tiles", len(pt), len(gt), flush=True)
    assert t["particle_tile_capacity"] == cap and t["grid_tile_capacity"] == 2 * cap
    assert t["bytes_per_grid_tile"] == 256 * 16
    args = (_params(), abi.friction(**fr), part["x"], part["y"], part["h"], part["A"], part["m"])
    s0 = b.admit(*args)
    assert s0 is not None
    assert b.admit(*args) is None  # tiles, not slots or particles, are exhausted
    fp = b.footprint()
    assert fp["free_slots"] == 3 and fp["free_particles"] >= len(part["x"])
    while b.step(64) > 0:
        pass
    first = (b.max_maps(s0), b.particles(s0))
    b.release(s0)
    s1 = b.admit(*args)
    assert s1 is not None
    while b.step(64) > 0:
        pass
    (ref,) = _run_batch([(part, fr)])
    for (hm, vm), p in (first, (b.max_maps(s1), b.particles(s1))):
        assert np.array_equal(hm, ref[0][0]) and np.array_equal(vm, ref[0][1])
        assert np.array_equal(p["x"], ref[1]["x"]) and np.array_equal(p["h"], ref[1]["h"])
    b.close()


def test_max_map_tiles_return_to_the_pool():
    # A batch keeps each slot's max maps only on the tiles its flow touched.
    # Releasing a slot (running or done) returns its tiles, zeroed; a scenario
    # admitted afterwards onto reused tiles gets the same maps as alone.
    from slabpro import abi
    z, dx = _terrain()
    scens = [_scenario(70 + k, FRICS[k % 3]) for k in range(3)]
    b = abi.Batch(z, dx, 4, 200000)
    slots = []
    for part, fr in scens:
        s = b.admit(_params(), abi.friction(**fr), part["x"], part["y"], part["h"], part["A"], part["m"])
        assert s is not None
        slots.append(s)
    b.step(20)
    t = b.tiles()
    print("tiles", t, flush=True)
    assert t["bytes_per_max_map_tile"] == 256 * 8 and 0 < t["max_map_tiles"] <= t["max_map_tile_capacity"]
    used = t["max_map_tiles"]
    b.release(slots[0])  # cancelled while running
    assert b.tiles()["max_map_tiles"] < used
    while b.step(64) > 0:
        pass
    for s in slots[1:]:
        b.release(s)
    assert b.tiles()["max_map_tiles"] == 0
    part, fr = scens[0]
    s = b.admit(_params(), abi.friction(**fr), part["x"], part["y"], part["h"], part["A"], part["m"])
    while b.step(64) > 0:
        pass
    hm, vm = b.max_maps(s)
    (ref,) = _run_batch([scens[0]])
    assert np.array_equal(hm, ref[0][0]) and np.array_equal(vm, ref[0][1])
    assert np.array_equal(b.particles(s)["x"], ref[1]["x"])
    b.close()


def test_max_map_pool_overflow_stops_the_batch(monkeypatch):
    # A max-map pool of one tile (SLAB_MAXMAP_TILE_CAP) and no tile check at
    # admission: the first step's tile listing runs out of pool tiles, the
    # step returns SLAB_EDOMAIN and the batch then refuses further steps.
    from slabpro import abi
    z, dx = _terrain()
    part, fr = _scenario(81, FRICS[1], n_extra=10)
    monkeypatch.setenv("SLAB_MAXMAP_TILE_CAP", "1")
    b = abi.Batch(z, dx, 2, 100000)
    monkeypatch.delenv("SLAB_MAXMAP_TILE_CAP")
    assert b.tiles()["max_map_tile_capacity"] == 1
    b.set_tile_growth(0)
    s = b.admit(_params(), abi.friction(**fr), part["x"], part["y"], part["h"], part["A"], part["m"])
    assert s is not None
    with pytest.raises(abi.SlabError) as ei:
        b.step(4)
    assert ei.value.status == abi.EDOMAIN and "max-map" in str(ei.value)
    with pytest.raises(abi.SlabError) as ei:
        b.step(1)
    assert ei.value.status == abi.ESTATE
    b.close()
