"""The slab-pro command and the Python front end on the GPU, end to end on the
Brienz example inputs (DEM 279x228 at 5 m, polygon and raster releases,
Voellmy friction, 60 s, POISSON sampling as shipped): output files, their
layout, and their contents against the FP64 oracle run from the same
initial particles."""
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import frontend_oracle as fo
from oracle import oracle
from slabpro import abi, frontend

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "slab-pro_amd", "bin", "slab-pro")
FIX = fo.fixture_dir()


def _write_case(tmp_path, name, **edits):
    cfg = json.load(open(os.path.join(FIX, name + ".json")))
    cfg["project_dir"] = str(tmp_path / "output")
    cfg["terrain"]["dem"]["filepath"] = os.path.join(FIX, "brienz_terrain_5m.asc")
    cfg["release"]["polygon"]["filepath"] = os.path.join(FIX, "brienz_release_polygon.shp")
    cfg["release"]["raster"]["filepath"] = os.path.join(FIX, "brienz_release_raster_5m.asc")
    for path, v in edits.items():
        d = cfg
        keys = path.split(".")
        for k in keys[:-1]:
            d = d[k]
        d[keys[-1]] = v
    p = tmp_path / (name + ".json")
    json.dump(cfg, open(p, "w"))
    return cfg, p


def _oracle_run(cfg, p):
    """Initial state from the front end on the CPU, then the FP64 oracle."""
    s = frontend.Simulation.load_from_json(p)
    s.set_verbose(False)
    s.initialize_terrain()
    s.initialize_particles()
    t, q = s.terrain(), s.particles()
    fr = cfg["friction"]
    kind = {"COULOMB": oracle.COULOMB, "VOELLMY": oracle.VOELLMY, "MU_I": oracle.MU_I}[fr["type"]]
    kw = {}
    if kind == oracle.COULOMB:
        kw = {"mu": fr["coulomb"]["mu"]}
    elif kind == oracle.VOELLMY:
        kw = {"mu": fr["voellmy"]["mu"], "xi": fr["voellmy"]["xi"]}
    else:
        m = fr["mu_I"]
        kw = {"mu1": m["mu_1"], "mu2": m["mu_2"], "I0": m["I0"], "d": m["d"], "phi": m["phi"]}
    ph, tp, nm = cfg["physical_params"], cfg["time_params"], cfg["numerical_params"]
    prm = oracle.params(gravity=ph["gravity"], rho=ph["rho"], C_cfl=nm["C_cfl"], dt_max=nm["dt_max"],
                        flip_ratio=nm["flip_ratio"], output_dt=tp["output_dt"], total_time=tp["total_time"],
                        fric_type=kind, **kw)
    sim = oracle.OracleSim(t["z"], t["dx"], prm, q["x"], q["y"], q["h"], q["A"], q["m"])
    sim.prime()
    sim.run()
    return t, q, sim


def _run_cli(p, *opts):
    print(f"slab-pro {p} {' '.join(opts)}", flush=True)
    r = subprocess.run([CLI, str(p), *opts], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    print(r.stdout[-3000:], flush=True)
    assert r.returncode == 0, r.stdout
    return r.stdout


def _max_tif(sim_dir):
    img, tags = frontend.read_geotiff(os.path.join(sim_dir, "max_values.tif"))
    # north-up rows -> grid order [i, j]
    return img[0][::-1, :].T, img[1][::-1, :].T, tags


def test_cli_brienz_polygon_fp64_matches_oracle(tmp_path):
    cfg, p = _write_case(tmp_path, "brienz_polygon")
    out = _run_cli(p)
    for line in ("Initialization complete.", "Terrain initialized.", "Particles initialized.",
                 "Number of particles: 7700", "Simulation took"):
        assert line in out
    d = os.path.join(cfg["project_dir"], cfg["sim_name"])
    assert os.path.exists(os.path.join(d, "brienz_polygon.json"))
    for k in range(7):
        assert os.path.exists(os.path.join(d, "particles", f"particles_frame_{k}.ply")), k
        assert os.path.exists(os.path.join(d, "rasters", f"rasters_frame_{k}.tif")), k
    t, q, sim = _oracle_run(cfg, p)
    sc = sim.scalars()
    assert sc["frame"] == 6 and abs(sc["time"] - 60.0) < 1e-9

    # max maps: FP64 GPU vs FP64 oracle, stored as float32
    hm, vm, tags = _max_tif(d)
    ohm, ovm = sim.max_maps()
    assert tags[33922] == [0.0, 0.0, 0.0, t["xllcorner"], t["yllcorner"] + 5.0 * (t["z"].shape[1] - 1), 0.0]
    wet = ohm.astype(np.float32) != 0
    assert np.array_equal(hm != -9999.0, wet)
    scale_h = ohm.max()
    assert np.max(np.abs(hm[wet] - ohm[wet])) <= 2e-6 * scale_h
    assert np.max(np.abs(vm[wet] - np.sqrt(ovm[wet]))) <= 1e-5 * np.sqrt(ovm.max())

    # last particle frame
    data, _ = frontend.read_ply(os.path.join(d, "particles", "particles_frame_6.ply"))
    op = sim.particles()
    assert len(data["x"]) == 7700
    for k in ("x", "y", "h"):
        assert np.max(np.abs(data[k] - op[k])) <= 1e-5 * max(1.0, np.abs(op[k]).max()), k

    # last grid frame over the remesh window, north-up, nodata where h == 0
    img, tags = frontend.read_geotiff(os.path.join(d, "rasters", "rasters_frame_6.tif"))
    g = sim.grid()
    iL, jB, gnx, gny = g["window"]
    assert img.shape == (6, gny, gnx)
    assert tags[33922][3] == t["xllcorner"] + iL * 5.0
    assert tags[33922][4] == t["yllcorner"] + (jB + gny - 1) * 5.0
    h = img[0][::-1, :].T
    gw = g["h"].astype(np.float32) != 0
    assert np.array_equal(h != -9999.0, gw)
    assert np.max(np.abs(h[gw] - g["h"][gw])) <= 1e-5 * g["h"].max()


def test_api_brienz_raster_fp32_within_one_percent(tmp_path):
    cfg, p = _write_case(tmp_path, "brienz_raster", sim_name="brienz_raster_fp32")
    s = frontend.Simulation.load_from_json(p)
    s.set_verbose(False)
    s.set_precision(abi.FP32)
    s.initialize()
    s.initialize_terrain()
    s.initialize_particles()
    print("brienz raster fp32: particles", len(s.particles()["x"]), flush=True)
    steps = s.simulate()
    d = s.output()[0]
    hm, vm, _ = _max_tif(d)
    _, _, sim = _oracle_run(cfg, p)
    ohm, _ = sim.max_maps()
    assert steps > 0
    hm = np.where(hm == -9999.0, 0.0, hm)
    l1 = np.abs(hm - ohm).sum() / np.abs(ohm).sum()
    print(f"FP32 h_max L1 relative error {l1:.2e} after {steps} steps", flush=True)
    assert l1 <= 0.01

