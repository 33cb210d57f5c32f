"""Workload generators: the numpy restatement of the UNIFORM sampler and the
shape filters agrees bit for bit with the oracle's C restatement of
initializeParticles (src/simulation.cpp:259-466), and closed-form counts hold."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from slabpro import configs


def test_c0_counts_and_init():
    c = configs.c0()
    assert c["z"].shape == (601, 151)
    assert len(c["x"]) == 4800
    assert c["x"].min() == 70.5 and c["x"].max() == 129.5
    assert np.all(c["A"] == 1.0) and np.all(c["m"] == 600.0)


@pytest.mark.parametrize("kind,a,b,rot", [(0, 60.0, 80.0, 0.0), (0, 33.0, 17.0, 0.7), (1, 25.0, 12.0, 1.1),
                                          (1, 9.0, 9.0, 0.0)])
def test_uniform_shape_matches_oracle(kind, a, b, rot):
    dx = 2.0
    if kind == 0:
        px, py, _ = configs.rectangle_release(dx, 101.3, 77.7, a, b, math.degrees(rot))
        rot_rad = math.degrees(rot) * math.pi / 180.0
    else:
        px, py, _ = configs.ellipse_release(dx, 101.3, 77.7, a, b, math.degrees(rot))
        rot_rad = math.degrees(rot) * math.pi / 180.0
    ox, oy = orc.uniform_shape(dx, kind, 101.3, 77.7, a, b, rot_rad)
    assert np.array_equal(px, ox) and np.array_equal(py, oy)


def test_ellipse_count_close_to_area():
    px, _, area = configs.ellipse_release(1.0, 500.0, 500.0, 120.0, 80.0, 30.0)
    assert abs(len(px) - 4 * area) / (4 * area) < 0.01


def test_polygon_matches_oracle_and_area():
    g = np.random.default_rng(3)
    vx, vy = configs.star_polygon(300.0, 250.0, 100.0, 32, 0.2, g)
    px, py, area = configs.polygon_release(1.0, vx, vy)
    ox, oy = orc.uniform_polygon(1.0, vx, vy)
    assert np.array_equal(px, ox) and np.array_equal(py, oy)
    assert abs(len(px) - 4 * area) / (4 * area) < 0.01
    # shoelace on a unit square
    assert configs.polygon_area(np.array([0, 1, 1, 0.0]), np.array([0, 0, 1, 1.0])) == 1.0


def test_raster_height_bilinear_exact_and_guards():
    ncols, nrows, cs = 30, 20, 5.0
    x = np.arange(ncols) * cs
    y = np.arange(nrows) * cs
    z = 1.0 + 0.01 * x[:, None] + 0.02 * y[None, :]
    g = np.random.default_rng(0)
    px = g.uniform(0, x[-1] - 1e-6, 200)
    py = g.uniform(0, y[-1] - 1e-6, 200)
    h = configs.raster_height(z, 0.0, 0.0, cs, -9999.0, px, py)
    assert np.allclose(h, 1.0 + 0.01 * px + 0.02 * py, rtol=1e-13)
    for k in range(20):
        assert h[k] == orc.raster_height(z, 0.0, 0.0, cs, -9999.0, px[k], py[k])
    z2 = z.copy()
    z2[3, 4] = -9999.0
    hh = configs.raster_height(z2, 0.0, 0.0, cs, -9999.0, np.array([3.5 * cs, -1.0]), np.array([3.5 * cs, 5.0]))
    assert hh[0] == 0.0 and hh[1] == 0.0


def test_seeds_are_stable():
    assert configs.subseed(1, 0) == configs.splitmix64(configs.MASTER_SEED ^ (1 << 32))
    a = configs.rng(3, 7).uniform(size=4)
    b = configs.rng(3, 7).uniform(size=4)
    assert np.array_equal(a, b)


def test_synthetic_dem_statistics():
    z = configs.synthetic_dem(256, 5.0, valley_y=640.0, seed=configs.subseed(1, 0))
    x = np.arange(256) * 5.0
    y = np.arange(256) * 5.0
    trend = (2500.0 - 0.3 * x)[:, None] + (8e-5 * (y - 640.0) ** 2)[None, :]
    noise = z - trend
    assert abs(noise.std() - 15.0) < 1e-9
    assert abs(noise.mean()) < 3.0
