"""slab_batch_download_particles_noalloc against slab_batch_download_particles
(needs an MI355X).  Tolerance: exact (both read the same FP32 state and form
the FP64 outputs with the same operations).  Two calls in a row, the second
on a smaller slot, then a slot above SLAB_NOALLOC_MAX_PARTICLES is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "vx", "vy", "vz", "z", "h", "A", "m")


def _terrain():
    nx, ny, dx = 96, 80, 2.0
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    z = 500.0 - 0.45 * x[:, None] + 0.002 * (y[None, :] - 80.0) ** 2
    return np.ascontiguousarray(z), dx


def _release(cx, cy, rx, ry, seed):
    from slabpro import configs
    g = np.random.default_rng(seed)
    px, py, area = configs.ellipse_release(2.0, cx, cy, rx, ry, g.uniform(0, 180))
    h = 1.0 + g.uniform(0.0, 1.0, len(px))
    return configs.finish_particles(px, py, h, area * 1.5, 300.0)


def test_noalloc_download_equals_download():
    from slabpro import abi
    z, dx = _terrain()
    big = _release(70.0, 80.0, 30.0, 20.0, 1)
    small = _release(60.0, 70.0, 12.0, 9.0, 2)
    assert len(big["x"]) > len(small["x"])
    b = abi.Batch(z, dx, 2, len(big["x"]) + len(small["x"]))
    prm = abi.params(rho=300.0, total_time=1.0, output_dt=0.5, precision=abi.FP32)
    fr = abi.friction(type=abi.VOELLMY, mu=0.155, xi=2000.0)
    sb = b.admit(prm, fr, big["x"], big["y"], big["h"], big["A"], big["m"])
    ss = b.admit(prm, fr, small["x"], small["y"], small["h"], small["A"], small["m"])
    b.step(30)
    for slot in (sb, ss, sb):  # larger, then smaller, then larger again
        ref = b.particles(slot)
        got = b.particles_noalloc(slot)
        for k in FIELDS:
            assert np.array_equal(got[k], ref[k]), (slot, k)
    while b.step(64) > 0:
        pass
    got = b.particles_noalloc(ss)
    ref = b.particles(ss)
    for k in FIELDS:
        assert np.array_equal(got[k], ref[k]), ("finished", k)
    b.close()


def test_noalloc_download_refuses_a_slot_above_the_maximum():
    from slabpro import abi, configs
    nx = ny = 600
    z = np.ascontiguousarray(np.zeros((nx, ny)) + np.arange(nx)[:, None] * -0.2)
    px, py, area = configs.rectangle_release(1.0, 300.0, 300.0, 270.0, 270.0, 0.0)
    assert len(px) > 262144
    part = configs.finish_particles(px, py, np.ones(len(px)), area, 300.0)
    b = abi.Batch(z, 1.0, 1, len(px))
    prm = abi.params(rho=300.0, total_time=1.0, output_dt=0.5, precision=abi.FP32)
    s = b.admit(prm, abi.friction(type=abi.COULOMB, mu=0.3), part["x"], part["y"], part["h"], part["A"], part["m"])
    with pytest.raises(abi.SlabError) as ei:
        b.particles_noalloc(s)
    assert ei.value.status == abi.EDOMAIN
    b.close()
