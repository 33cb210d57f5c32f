"""Front end on the CPU: JSON configuration, folders, terrain, release
shapes, samplers and the output writers, against the oracle's restatement
(oracle/frontend_oracle.py) and the Brienz example inputs shipped with the
reference (tests/golden/brienz: DEM and release raster at 5 m, release
polygon shapefile, the two example JSON files)."""
import copy
import json
import math
import os

import numpy as np
import pytest

from oracle import frontend_oracle as fo
from oracle import oracle
from slabpro import abi, frontend

FIX = fo.fixture_dir()


def _example(name, tmp_path, **edits):
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


def _sim(p):
    s = frontend.Simulation.load_from_json(p)
    s.set_verbose(False)
    return s


# --- pins of the oracle's own pieces -----------------------------------------------


def test_libc_rand_stream_is_the_documented_one():
    assert fo.libc_rand(3) == [1804289383, 846930886, 1681692777]


def test_oracle_polygon_area_closed_forms():
    assert fo.polygon_area([0, 4, 4, 0], [0, 0, 3, 3]) == 12.0
    assert fo.polygon_area([0, 4, 4, 0, 0], [0, 0, 3, 3, 0]) == 12.0      # closed ring
    assert fo.polygon_area([0, 3, 0], [0, 0, 2]) == 3.0
    assert fo.polygon_area([0, 0, 4, 4], [0, 3, 3, 0]) == 12.0            # clockwise
    g = np.random.default_rng(5)
    th = np.sort(g.uniform(0, 2 * np.pi, 40))
    r = g.uniform(5, 10, 40)
    x, y = 1e3 + r * np.cos(th), 2e3 + r * np.sin(th)
    xc, yc = np.append(x, x[0]), np.append(y, y[0])
    ref = 0.5 * abs(np.sum(xc[:-1] * yc[1:] - xc[1:] * yc[:-1]))
    assert abs(fo.polygon_area(x, y) - ref) <= 1e-12 * ref


def test_oracle_asc_reader_row_order(tmp_path):
    p = tmp_path / "t.asc"
    p.write_text("ncols 3\nnrows 2\nxllcorner 10\nyllcorner 20\ncellsize 5\nNODATA_value -9999\n"
                 "1 2 3\n4 5 6\n")
    r = fo.read_asc(str(p))
    # first file row is the northern (top) row: y index nrows-1
    assert r["z"].shape == (3, 2)
    assert list(r["z"][:, 1]) == [1, 2, 3] and list(r["z"][:, 0]) == [4, 5, 6]


def test_oracle_poisson_invariants():
    dx, Lx, Ly = 1.0, 12.0, 9.0
    x, y = fo.poisson_disk(dx, Lx, Ly, 4)
    r = math.sqrt(15.0 / 24.0 * dx * dx / 4)
    d2 = (x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2
    np.fill_diagonal(d2, np.inf)
    assert d2.min() >= r * r
    assert x.min() >= 0 and x.max() <= Lx and y.min() >= 0 and y.max() <= Ly
    # about one sample per dx^2/4 (the 15/24 factor is chosen for that)
    assert 0.8 * 4 * Lx * Ly < len(x) < 1.2 * 4 * Lx * Ly
    u = fo.libc_rand(2)
    assert x[0] == u[0] / fo.RAND_MAX * Lx and y[0] == u[1] / fo.RAND_MAX * Ly


# --- product vs oracle -----------------------------------------------------------------


@pytest.mark.parametrize("dx,Lx,Ly", [(1.0, 12.0, 9.0), (5.0, 60.0, 35.0), (2.5, 7.5, 40.0)])
def test_poisson_matches_oracle_bitwise(dx, Lx, Ly):
    x, y = frontend.poisson_disk(dx, Lx, Ly, 4)
    ox, oy = fo.poisson_disk(dx, Lx, Ly, 4)
    assert len(x) == len(ox)
    assert np.array_equal(x, ox) and np.array_equal(y, oy)


def test_poisson_does_not_touch_callers_rand():
    fo.libc_rand(5)
    frontend.poisson_disk(1.0, 5.0, 5.0, 4)
    # the product draws from a private state: the global stream continues
    import ctypes
    assert ctypes.CDLL("libc.so.6").rand() == fo.libc_rand(6)[5]


def test_polygon_area_and_pip_match_oracle():
    vx, vy = fo.read_shapefile_ring(os.path.join(FIX, "brienz_release_polygon.shp"))
    assert frontend.polygon_area(vx, vy) == fo.polygon_area(vx, vy)
    g = np.random.default_rng(1)
    for px, py in zip(g.uniform(vx.min(), vx.max(), 300), g.uniform(vy.min(), vy.max(), 300)):
        assert frontend.point_in_polygon(px, py, vx, vy) == fo.point_in_polygon(px, py, vx, vy)


def test_shapefile_reader_matches_oracle():
    x, y = frontend.read_shapefile_polygon(os.path.join(FIX, "brienz_release_polygon.shp"))
    ox, oy = fo.read_shapefile_ring(os.path.join(FIX, "brienz_release_polygon.shp"))
    assert np.array_equal(x, ox) and np.array_equal(y, oy)
    assert x[0] == x[-1] and y[0] == y[-1] and len(x) == 440
    # the .shx index names the same record offset and length as the .shp walk
    import struct
    shx = open(os.path.join(FIX, "brienz_release_polygon.shx"), "rb").read()
    off, length = struct.unpack(">ii", shx[100:108])
    shp = open(os.path.join(FIX, "brienz_release_polygon.shp"), "rb").read()
    assert off * 2 == 100 and struct.unpack(">ii", shp[100:108])[1] == length


def test_load_json_and_initialize(tmp_path):
    cfg, p = _example("brienz_polygon", tmp_path)
    s = _sim(p)
    s.initialize()
    d, nfiles, pad = s.output()
    assert d == os.path.join(cfg["project_dir"], cfg["sim_name"])
    assert nfiles == 6 and pad == 1
    assert os.path.isdir(os.path.join(d, "rasters")) and os.path.isdir(os.path.join(d, "particles"))
    assert json.load(open(os.path.join(d, "brienz_polygon.json"))) == cfg
    cfg2, p2 = _example("brienz_polygon", tmp_path, **{"time_params.output_dt": 0.5, "sim_name": "b2"})
    s2 = _sim(p2)
    s2.initialize()
    assert s2.output()[1:] == (120, 3)


@pytest.mark.parametrize("field,value,msg", [
    ("friction.type", "COULOMBX", "Unknown friction type: COULOMBX"),
    ("terrain.type", "MESH", "Unknown terrain type: MESH"),
    ("release.type", "CIRCLE", "Unknown release shape type: CIRCLE"),
    ("numerical_params.sampling_method", "GRID", "Unknown sampling method: GRID"),
])
def test_load_json_rejects_unknown_names(tmp_path, field, value, msg):
    _, p = _example("brienz_polygon", tmp_path, **{field: value})
    with pytest.raises(abi.SlabError, match=msg):
        frontend.Simulation.load_from_json(p)


def test_load_json_missing_key_and_file(tmp_path):
    cfg, p = _example("brienz_polygon", tmp_path)
    del cfg["physical_params"]["rho"]
    json.dump(cfg, open(p, "w"))
    with pytest.raises(abi.SlabError, match="rho"):
        frontend.Simulation.load_from_json(p)
    with pytest.raises(abi.SlabError, match="JSON file not found"):
        frontend.Simulation.load_from_json(tmp_path / "none.json")


def test_dem_terrain_matches_oracle_reader(tmp_path):
    _, p = _example("brienz_polygon", tmp_path)
    s = _sim(p)
    s.initialize_terrain()
    t = s.terrain()
    r = fo.read_asc(os.path.join(FIX, "brienz_terrain_5m.asc"))
    assert np.array_equal(t["z"], r["z"])
    assert (t["dx"], t["xllcorner"], t["yllcorner"]) == (5.0, 2764175.0, 1170855.0)


def test_analytical_terrain_closed_form(tmp_path):
    _, p = _example("brienz_polygon", tmp_path, **{"terrain.type": "ANALYTICAL"})
    s = _sim(p)
    s.initialize_terrain()
    t = s.terrain()
    # Lx=10, Ly=20, dx=1: nodes 0..10 and 0..20; z = -(x + y) tan 5deg
    assert t["z"].shape == (11, 21)
    X, Y = np.meshgrid(np.arange(11.0), np.arange(21.0), indexing="ij")
    tn = math.tan(math.radians(5.0))
    assert np.allclose(t["z"], -(X + Y) * tn, rtol=0, atol=1e-12)
    assert (t["xllcorner"], t["yllcorner"]) == (0.0, 0.0)


def test_c0_from_json_equals_the_c0_workload(tmp_path):
    """The 4,800-particle C0 case (35-degree plane, arc, flat runout; 60x80 m
    rectangle) written as a JSON file gives the same terrain and particles as
    the workload generator used by the tests and the benchmark."""
    from slabpro import configs
    edits = {"terrain.type": "ANALYTICAL", "terrain.analytical": {
                 "Lx": 1200.0, "Ly": 300.0, "dx": 2.0, "profile": "PLANE_ARC_FLAT",
                 "theta_deg": 35.0, "x_t": 600.0, "R": 300.0, "z_f": 0.0},
             "release.type": "RECTANGLE", "release.h0": 2.0,
             "release.rectangle": {"center": [100.0, 150.0], "Lx": 60.0, "Ly": 80.0, "rotation_deg": 0.0},
             "numerical_params.sampling_method": "UNIFORM", "physical_params.rho": 300.0}
    _, p = _example("brienz_polygon", tmp_path, **edits)
    s = _sim(p)
    s.initialize_terrain()
    s.initialize_particles()
    c0 = configs.c0()
    assert np.array_equal(s.terrain()["z"], c0["z"])
    q = s.particles()
    assert len(q["x"]) == 4800
    for k in ("x", "y", "h", "A", "m"):
        assert np.array_equal(q[k], c0[k]), k
    _, p2 = _example("brienz_polygon", tmp_path, **{"terrain.type": "ANALYTICAL",
                                                      "terrain.analytical.profile": "CONE"})
    with pytest.raises(abi.SlabError, match="Unknown analytical profile: CONE"):
        frontend.Simulation.load_from_json(p2)


def _compare_init(s, ref):
    p = s.particles()
    assert len(p["x"]) == len(ref["x"]) > 0
    for k in ("x", "y", "h", "A", "m"):
        assert np.array_equal(p[k], ref[k]), k
    area, vol = s.release()
    assert area == ref["area"] and vol == ref["volume"]


@pytest.mark.parametrize("release,method", [
    ("RECTANGLE", "UNIFORM"), ("ELLIPSE", "UNIFORM"), ("RECTANGLE", "POISSON"), ("ELLIPSE", "POISSON"),
])
def test_analytical_shapes_match_oracle(tmp_path, release, method):
    edits = {"terrain.type": "ANALYTICAL", "terrain.analytical.Lx": 80.0, "terrain.analytical.Ly": 60.0,
             "release.type": release, "numerical_params.sampling_method": method,
             "release.rectangle.center": [30.0, 25.0], "release.rectangle.Lx": 14.0,
             "release.rectangle.Ly": 6.0, "release.rectangle.rotation_deg": 30.0,
             "release.ellipse.center": [30.0, 25.0], "release.ellipse.rx": 9.0, "release.ellipse.ry": 4.0,
             "release.ellipse.rotation_deg": -20.0}
    cfg, p = _example("brienz_polygon", tmp_path, **edits)
    s = _sim(p)
    s.initialize_terrain()
    s.initialize_particles()
    ref = fo.initialize_particles(cfg, 0.0, 0.0)
    _compare_init(s, ref)
    # closed-form release volume; the scaled areas give it back exactly as sum h A
    area = 14.0 * 6.0 if release == "RECTANGLE" else math.pi * 9.0 * 4.0
    assert ref["area"] == area and abs(np.sum(ref["h"] * ref["A"]) - area * 22.59) < 1e-9 * area * 22.59


def test_brienz_polygon_uniform_matches_oracle(tmp_path):
    cfg, p = _example("brienz_polygon", tmp_path, **{"numerical_params.sampling_method": "UNIFORM"})
    cfg["_dx"] = 5.0
    s = _sim(p)
    s.initialize_terrain()
    s.initialize_particles()
    _compare_init(s, fo.initialize_particles(cfg, 2764175.0, 1170855.0))


def test_brienz_polygon_poisson_matches_oracle(tmp_path):
    """The example as shipped (POISSON).  The sampler over the example's box
    is checked by the invariants; its bitwise parity is checked on the small
    boxes above.  Filtering, area and volume are checked here."""
    cfg, p = _example("brienz_polygon", tmp_path)
    cfg["_dx"] = 5.0
    s = _sim(p)
    s.initialize_terrain()
    s.initialize_particles()
    vx, vy = fo.read_shapefile_ring(cfg["release"]["polygon"]["filepath"])
    vx, vy = vx - 2764175.0, vy - 1170855.0
    x0, x1 = math.floor(vx.min() / 5.0) * 5.0, math.ceil(vx.max() / 5.0) * 5.0
    y0, y1 = math.floor(vy.min() / 5.0) * 5.0, math.ceil(vy.max() / 5.0) * 5.0
    sx, sy = frontend.poisson_disk(5.0, x1 - x0, y1 - y0, 4)
    r = math.sqrt(15.0 / 24.0 * 25.0 / 4)
    from scipy.spatial import cKDTree
    dd, _ = cKDTree(np.c_[sx, sy]).query(np.c_[sx, sy], k=2)
    assert dd[:, 1].min() >= r
    ref = fo.initialize_particles(cfg, 2764175.0, 1170855.0, samples=(sx, sy))
    _compare_init(s, ref)
    assert len(ref["x"]) == 7700


@pytest.mark.parametrize("method", ["UNIFORM", "POISSON"])
def test_brienz_raster_matches_oracle(tmp_path, method):
    cfg, p = _example("brienz_raster", tmp_path, **{"numerical_params.sampling_method": method})
    cfg["_dx"] = 5.0
    s = _sim(p)
    s.initialize_terrain()
    s.initialize_particles()
    samples = None
    if method == "POISSON":
        r = fo.read_asc(cfg["release"]["raster"]["filepath"])
        ox, oy = r["xllcorner"] - 2764175.0, r["yllcorner"] - 1170855.0
        x0, y0 = math.floor(ox / 5.0) * 5.0, math.floor(oy / 5.0) * 5.0
        x1 = math.ceil((ox + 5.0 * (r["ncols"] - 1)) / 5.0) * 5.0
        y1 = math.ceil((oy + 5.0 * (r["nrows"] - 1)) / 5.0) * 5.0
        samples = frontend.poisson_disk(5.0, x1 - x0, y1 - y0, 4)
    ref = fo.initialize_particles(cfg, 2764175.0, 1170855.0, samples=samples)
    _compare_init(s, ref)
    # raster volume = sum of the cells at or above h_min, and sum h A returns it
    assert abs(np.sum(ref["h"] * ref["A"]) - ref["volume"]) < 1e-9 * ref["volume"]


def test_initialize_particles_needs_terrain(tmp_path):
    _, p = _example("brienz_polygon", tmp_path)
    s = _sim(p)
    with pytest.raises(abi.SlabError, match="initialize_terrain"):
        s.initialize_particles()


def test_cli_usage_and_bad_config(tmp_path):
    """The command fails before touching a device when the input is wrong."""
    import subprocess
    cli = os.path.join(os.path.dirname(abi.PKG), "slab-pro_amd", "bin", "slab-pro")
    r = subprocess.run([cli], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 1 and "Usage:" in r.stdout
    _, p = _example("brienz_polygon", tmp_path, **{"friction.type": "DRY"})
    r = subprocess.run([cli, str(p)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 1 and "Unknown friction type: DRY" in r.stdout


# --- writers ----------------------------------------------------------------------------


def test_ply_writer_round_trip(tmp_path):
    g = np.random.default_rng(3)
    n = 1001
    cols = [g.normal(size=n) * 10 ** k for k in range(8)]
    path = tmp_path / "p.ply"
    frontend.write_ply(path, *cols)
    data, header = frontend.read_ply(path)
    assert header == ["ply", "format binary_little_endian 1.0", f"element vertex {n}", "property float x",
                      "property float y", "property float vx", "property float vy", "property float vz",
                      "property float z", "property float h", "property float A", "end_header"]
    for name, c in zip(("x", "y", "vx", "vy", "vz", "z", "h", "A"), cols):
        assert np.array_equal(data[name], c.astype(np.float32))
    assert os.path.getsize(path) == len("\n".join(header)) + 1 + 32 * n


def test_geotiff_writer_round_trip(tmp_path):
    g = np.random.default_rng(4)
    nx, ny = 37, 23
    bands = [g.normal(size=(nx, ny)).astype(np.float32) for _ in range(6)]
    bands[0][3, 4] = -9999.0
    names = ["h", "v_norm", "m", "v_x", "v_y", "v_z"]
    path = tmp_path / "r.tif"
    frontend.write_geotiff(path, bands, names, 2764175.0 + 35.0, 1170855.0 + 5.0 * 40, 5.0)
    img, tags = frontend.read_geotiff(path)
    assert img.shape == (6, ny, nx)
    for b in range(6):
        # stored north-up: file row r is grid row j = ny-1-r
        assert np.array_equal(img[b], bands[b].T[::-1, :])
    assert tags[33550] == [5.0, 5.0, 0.0]
    assert tags[33922] == [0.0, 0.0, 0.0, 2764175.0 + 35.0, 1170855.0 + 200.0, 0.0]
    assert tags[42113] == "-9999"
    for b, nm in enumerate(names):
        assert f'sample="{b}" role="description">{nm}<' in tags[42112]


def test_geotiff_many_strips(tmp_path):
    nx, ny = 600, 700   # > 1 MB of pixels: several strips
    a = np.arange(nx * ny, dtype=np.float32).reshape(nx, ny)
    path = tmp_path / "big.tif"
    frontend.write_geotiff(path, [a, -a], ["h_max", "v_norm_max"], 0.0, 10.0, 1.0)
    img, tags = frontend.read_geotiff(path)
    assert len(tags[273]) > 1
    assert np.array_equal(img[0], a.T[::-1, :]) and np.array_equal(img[1], -a.T[::-1, :])
