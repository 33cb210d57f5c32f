"""Pins the CPU FP64 oracle against closed forms and invariants (no GPU).

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it:
  - B-spline values, partition of unity, zero first moment
  - exact plane gradients; last row/column zero (complex_terrain.cpp:84-111)
  - remesh window and the t=0 mass identity of C0
  - P2G against a dense all-pairs evaluation (pins stencil choice/offsets/cutoff)
  - friction laws against their physical steady states: sliding-block
    acceleration g(sin - mu cos), static Coulomb, Voellmy terminal speed
    s^2 = xi h cos (sin - mu cos), mu(I) steady state mu(I*) = tan(theta)
  - G2P: rigid translation of a uniform block interior, mass invariance,
    sum h_p A_p = sum m_p / rho; the height update under a linear velocity
    field (h / (1 + (a+b) dt)); the FLIP/PIC weights (U + r delta)
  - the depth-averaged pressure: momentum balance of a frictionless dam break
    (flat ground and a 45-degree cross-slope)
  - updateDt / frame snapping; sort = permutation with invariant grid sums
"""
import math

import numpy as np
import pytest

from conftest import oracle_sim
from oracle import oracle as orc


# ---------------------------------------------------------------- shape functions

def test_N_known_values():
    assert orc.N(0.0) == 0.75
    assert orc.N(0.5) == 0.5 and orc.N(-0.5) == 0.5
    assert orc.N(1.0) == 0.125 and orc.N(-1.0) == 0.125
    assert orc.N(1.5) == 0.0 and orc.N(-1.5) == 0.0 and orc.N(3.0) == 0.0
    assert orc.dNdu(0.0) == 0.0
    assert orc.dNdu(0.25) == -0.5 and orc.dNdu(-0.25) == 0.5
    assert orc.dNdu(1.0) == -0.5 and orc.dNdu(-1.0) == 0.5
    assert orc.dNdu(1.5) == 0.0 and orc.dNdu(2.0) == 0.0


@pytest.mark.parametrize("seed", range(5))
def test_N_partition_of_unity_and_moments(seed):
    g = np.random.default_rng(seed)
    for _ in range(200):
        x = g.uniform(10.0, 20.0)  # in cell units
        base = int(x - 0.5)
        us = [x - (base + d) for d in range(3)]
        w = [orc.N(u) for u in us]
        dw = [orc.dNdu(u) for u in us]
        assert abs(sum(w) - 1.0) < 1e-14
        assert abs(sum(dw)) < 1e-14
        # first moment: sum N(u_d) * (X_d - x) = 0
        assert abs(sum(wi * (-u) for wi, u in zip(w, us))) < 1e-14
        # derivative is the derivative (central difference)
        for u in us:
            if abs(abs(u) - 0.5) > 1e-3 and abs(abs(u) - 1.5) > 1e-3:
                fd = (orc.N(u + 1e-6) - orc.N(u - 1e-6)) / 2e-6
                assert abs(fd - orc.dNdu(u)) < 1e-6


# ---------------------------------------------------------------- terrain

def _plane(nx, ny, dx, a, b, c=10.0):
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    return a * x[:, None] + b * y[None, :] + c


def test_plane_gradient_exact_and_last_row_zero():
    z = _plane(12, 9, 2.0, 0.25, -0.5)
    s = oracle_sim({"z": z, "dx": 2.0, "params": dict(rho=300.0), "friction": {"type": 0, "mu": 0.3},
                    "x": [5.0], "y": [5.0], "h": [1.0], "A": [1.0]})
    gx, gy, th = s.terrain()
    assert np.all(gx[:-1, :-1] == 0.25) and np.all(gy[:-1, :-1] == -0.5)
    assert np.all(gx[-1, :] == 0) and np.all(gx[:, -1] == 0)
    assert np.all(gy[-1, :] == 0) and np.all(gy[:, -1] == 0)
    assert np.allclose(th[:-1, :-1], math.atan(math.hypot(0.25, 0.5)), rtol=0, atol=1e-15)
    assert np.all(th[-1, :] == 0)


def test_reference_analytical_plane():
    # the reference's hard-coded alt_z: z = -x tan5 - y tan5 (simulation.cpp:182-185)
    t = math.tan(math.radians(5.0))
    xs = np.arange(20) * 1.0
    z = -xs[:, None] * t - xs[None, :] * t
    s = oracle_sim({"z": z, "dx": 1.0, "params": dict(rho=300.0), "friction": {"type": 0, "mu": 0.3},
                    "x": [5.0], "y": [5.0], "h": [1.0], "A": [1.0]})
    gx, gy, th = s.terrain()
    assert np.allclose(gx[:-1, :-1], -t, rtol=1e-13) and np.allclose(gy[:-1, :-1], -t, rtol=1e-13)
    assert np.allclose(th[:-1, :-1], math.atan(math.sqrt(2.0) * t), rtol=1e-13)


# ---------------------------------------------------------------- C0 closed forms

def test_c0_window_and_t0_mass_identity():
    from slabpro import configs
    cfg = configs.c0()
    assert len(cfg["x"]) == 4800
    assert np.all(cfg["A"] == 1.0) and np.all(cfg["m"] == 600.0)
    s = oracle_sim(cfg)
    s.prime()
    assert s.window() == (34, 54, 33, 43)
    g = s.grid()
    assert g["mass"].sum() == (1 - 2.0 ** -10) * 2.88e6  # exactly 2,877,187.5 kg
    assert abs((g["mass"] * g["vx"]).sum()) == 0.0
    hm, _ = s.max_maps()
    assert hm.max() == pytest.approx(2.0, rel=1e-12)  # interior height = h0


def test_c0_profile_continuity():
    from slabpro import configs
    th = math.radians(35.0)
    xc = 600.0 + 300.0 * math.sin(th)
    for x0 in (600.0, xc):
        zl = configs.c0_profile(np.array([x0 - 1e-6]))[0]
        zr = configs.c0_profile(np.array([x0 + 1e-6]))[0]
        assert abs(zl - zr) < 1e-5
        sl = (configs.c0_profile(np.array([x0 - 1e-6]))[0] - configs.c0_profile(np.array([x0 - 2e-6]))[0]) / 1e-6
        sr = (configs.c0_profile(np.array([x0 + 2e-6]))[0] - configs.c0_profile(np.array([x0 + 1e-6]))[0]) / 1e-6
        assert abs(sl - sr) < 1e-3
    assert configs.c0_profile(np.array([0.0]))[0] == pytest.approx(300 * (1 - math.cos(th)) + 600 * math.tan(th))


# ---------------------------------------------------------------- P2G dense check

def _dense_p2g(z, dx, rho, g, x, y, vx, vy, h, A, m, window):
    """All-pairs evaluation of the scatter over every window node, with the
    w > 3e-3 cutoff; stencil membership is implied by N(u) > 0."""
    iL, jB, gnx, gny = window
    inv = 1.0 / dx
    nx, ny = z.shape
    # terrain gradient / theta at the floor cell, from the definition
    mass = np.zeros((gnx, gny))
    px = np.zeros((gnx, gny))
    py = np.zeros((gnx, gny))
    fx = np.zeros((gnx, gny))
    fy = np.zeros((gnx, gny))
    for p in range(len(x)):
        ci, cj = int(math.floor(x[p] * inv)), int(math.floor(y[p] * inv))
        Gx = (z[ci + 1, cj] + z[ci + 1, cj + 1] - z[ci, cj] - z[ci, cj + 1]) / 2 * inv
        Gy = (z[ci, cj + 1] + z[ci + 1, cj + 1] - z[ci, cj] - z[ci + 1, cj]) / 2 * inv
        c = math.cos(math.atan(math.hypot(Gx, Gy)))
        for a in range(gnx):
            for b in range(gny):
                ux = (x[p] - (iL + a) * dx) * inv
                uy = (y[p] - (jB + b) * dx) * inv
                w = orc.N(ux) * orc.N(uy)
                if w <= 3e-3:
                    continue
                gwx = orc.dNdu(ux) * inv * orc.N(uy)
                gwy = orc.dNdu(uy) * inv * orc.N(ux)
                sig = -0.5 * rho * g * c * c * h[p]
                mass[a, b] += m[p] * w
                px[a, b] += m[p] * vx[p] * w
                py[a, b] += m[p] * vy[p] * w
                fx[a, b] += -m[p] * g * c * c * Gx * w - h[p] * A[p] * sig * gwx
                fy[a, b] += -m[p] * g * c * c * Gy * w - h[p] * A[p] * sig * gwy
    return mass, px, py, fx, fy


@pytest.mark.parametrize("seed", range(4))
def test_p2g_matches_dense_all_pairs(seed):
    gen = np.random.default_rng(100 + seed)
    nx, ny, dx = 24, 20, 1.5
    z = gen.normal(0, 1.0, (nx, ny)) + _plane(nx, ny, dx, -0.4, 0.1)
    n = int(gen.integers(1, 17))
    x = gen.uniform(6 * dx, 16 * dx, n)
    y = gen.uniform(6 * dx, 13 * dx, n)
    h = gen.uniform(0.5, 2.0, n)
    A = gen.uniform(0.3, 1.0, n)
    m = 300.0 * h * A
    vx = gen.normal(0, 3, n)
    vy = gen.normal(0, 3, n)
    cfg = {"z": z, "dx": dx, "params": dict(rho=300.0, gravity=9.81), "friction": {"type": 0, "mu": 0.3},
           "x": x, "y": y, "h": h, "A": A, "m": m}
    s = oracle_sim(cfg)
    s.set_velocity(vx, vy)
    s.prime()  # dt = 0: grid v = P / m exactly as scattered
    grid = s.grid()
    mass, px, py, fx, fy = _dense_p2g(z, dx, 300.0, 9.81, x, y, vx, vy, h, A, m, grid["window"])
    scale = np.abs(mass).max()
    assert np.max(np.abs(grid["mass"] - mass)) <= 1e-13 * scale
    act = mass > 1e-6
    pscale = max(np.abs(px).max(), np.abs(py).max())
    assert np.max(np.abs(grid["vx"][act] * grid["mass"][act] - px[act])) <= 1e-12 * pscale
    assert np.max(np.abs(grid["vy"][act] * grid["mass"][act] - py[act])) <= 1e-12 * pscale
    fscale = max(np.abs(fx).max(), np.abs(fy).max())
    assert np.max(np.abs(grid["Fx"] - fx)) <= 1e-12 * fscale
    assert np.max(np.abs(grid["Fy"] - fy)) <= 1e-12 * fscale


# ---------------------------------------------------------------- friction laws

def _block_on_plane(theta_deg, fric, v_along=0.0, dt_max=1e-3, n=80, dx=1.0, h0=1.0, rho=300.0):
    from slabpro import configs
    t = math.tan(math.radians(theta_deg))
    xs = np.arange(n) * dx
    z = np.repeat((-t * xs)[:, None], n, axis=1) + 1000.0
    px, py, area = configs.rectangle_release(dx, n * dx / 2, n * dx / 2, 0.6 * n * dx, 0.6 * n * dx, 0.0)
    part = configs.finish_particles(px, py, np.full(len(px), h0), area * h0, rho)
    cfg = {"z": z, "dx": dx, "params": dict(rho=rho, gravity=9.81, dt_max=dt_max, output_dt=1e9, total_time=1e9,
                                            C_cfl=0.5, flip_ratio=0.95),
           "friction": fric, **part}
    s = oracle_sim(cfg)
    if v_along:
        c = math.cos(math.radians(theta_deg))
        vx = np.full(len(px), v_along * c)  # horizontal projection of the along-slope speed
        s.set_velocity(vx, np.zeros(len(px)))
    return s, cfg


def _interior(cfg, p, margin):
    cx = cy = cfg["z"].shape[0] * cfg["dx"] / 2
    half = 0.3 * cfg["z"].shape[0] * cfg["dx"] - margin
    return (np.abs(np.asarray(cfg["x"]) - cx) < half) & (np.abs(np.asarray(cfg["y"]) - cy) < half)


def test_sliding_block_coulomb_acceleration():
    theta, mu = 35.0, 0.3
    s, cfg = _block_on_plane(theta, {"type": 0, "mu": mu}, dt_max=1e-3)
    s.prime()
    nsteps = 3  # edge pressure reaches ~3 cells further per step
    for _ in range(nsteps):
        s.step()
    p = s.particles()
    inner = _interior(cfg, p, 12.0)
    c = math.cos(math.radians(theta))
    speed = np.hypot(p["vx"][inner], p["vy"][inner]) / c  # along-slope speed
    expect = 9.81 * (math.sin(math.radians(theta)) - mu * c) * nsteps * 1e-3
    assert np.allclose(speed, expect, rtol=1e-10)
    assert np.all(p["vx"][inner] > 0)  # downhill is +x


def test_static_coulomb_stays_at_rest():
    theta, mu = 20.0, 0.6  # tan 20 = 0.364 < 0.6
    s, cfg = _block_on_plane(theta, {"type": 0, "mu": mu}, dt_max=1e-3, h0=0.05)
    s.prime()
    for _ in range(10):
        s.step()
    p = s.particles()
    inner = _interior(cfg, p, 4.0)
    assert np.all(p["vx"][inner] == 0.0) and np.all(p["vy"][inner] == 0.0)
    assert np.array_equal(p["x"][inner], np.asarray(cfg["x"])[inner])


def _along_slope_accel(theta, fric, v_along, h0=1.0, dt=1e-6):
    s, cfg = _block_on_plane(theta, fric, v_along=v_along, dt_max=dt, h0=h0)
    s.prime()
    s.step()
    p = s.particles()
    inner = _interior(cfg, p, 10.0)
    c = math.cos(math.radians(theta))
    speed = np.hypot(p["vx"][inner], p["vy"][inner]) / c
    return (speed - v_along) / dt


def test_voellmy_terminal_speed():
    theta, mu, xi, h0 = 30.0, 0.155, 2000.0, 1.0
    th = math.radians(theta)
    s_inf = math.sqrt(xi * h0 * math.cos(th) * (math.sin(th) - mu * math.cos(th)))
    g_par = 9.81 * math.sin(th)
    acc = _along_slope_accel(theta, {"type": 1, "mu": mu, "xi": xi}, s_inf, h0)
    assert np.all(np.abs(acc) < 1e-4 * g_par)
    # and 10% faster decelerates, 10% slower accelerates
    assert np.all(_along_slope_accel(theta, {"type": 1, "mu": mu, "xi": xi}, 1.1 * s_inf, h0) < 0)
    assert np.all(_along_slope_accel(theta, {"type": 1, "mu": mu, "xi": xi}, 0.9 * s_inf, h0) > 0)


def test_mu_I_steady_state():
    theta, h0 = 25.0, 1.0
    mu1, mu2, I0, d, phi = 0.38, 0.64, 0.279, 0.5, 0.6
    th = math.radians(theta)
    t = math.tan(th)
    I_star = I0 * (t - mu1) / (mu2 - t)
    s_star = I_star * 2 * h0 * math.cos(th) ** 2 * math.sqrt(9.81 * h0 * phi) / (5 * d)
    fr = {"type": 2, "mu1": mu1, "mu2": mu2, "I0": I0, "d": d, "phi": phi}
    g_par = 9.81 * math.sin(th)
    acc = _along_slope_accel(theta, fr, s_star, h0)
    assert np.all(np.abs(acc) < 1e-4 * g_par)
    assert np.all(_along_slope_accel(theta, fr, 1.2 * s_star, h0) < 0)
    assert np.all(_along_slope_accel(theta, fr, 0.8 * s_star, h0) > 0)


# ---------------------------------------------------------------- G2P invariants

def test_mass_and_volume_invariants():
    from slabpro import configs
    cfg = configs.tiny(fric={"type": 1, "mu": 0.2, "xi": 1000.0})
    s = oracle_sim(cfg)
    s.prime()
    for _ in range(40):
        s.step()
    p = s.particles()
    assert np.array_equal(p["m"], cfg["m"])
    rho = cfg["params"]["rho"]
    assert np.allclose(p["h"] * p["A"], p["m"] / rho, rtol=1e-15, atol=0)


def test_uniform_field_translates_rigidly():
    # interior of a uniformly moving block on flat ground with no friction:
    # partition of unity keeps v, and x advances by dt * v
    s, cfg = _block_on_plane(0.0, {"type": 0, "mu": 0.0}, v_along=3.0, dt_max=1e-3, h0=1e-3)
    s.prime()
    s.step()
    p = s.particles()
    inner = _interior(cfg, p, 6.0)
    assert np.allclose(p["vx"][inner], 3.0, rtol=1e-12)
    assert np.allclose(p["x"][inner] - np.asarray(cfg["x"])[inner], 3.0e-3, rtol=1e-9)


# ---------------------------------------------------------------- driver

def test_update_dt_and_frame_snapping():
    from slabpro import configs
    cfg = configs.tiny()
    cfg["params"].update(dt_max=0.03, output_dt=0.1, C_cfl=0.5)
    s = oracle_sim(cfg)
    s.prime()
    assert s.step() == 0
    assert s.scalars()["dt"] == 0.03  # at rest: dt = dt_max
    s.step()
    s.step()
    r = s.step()  # 0.09 + 0.03 > 0.1 - 1e-12 -> snapped to 0.01
    sc = s.scalars()
    assert r == 1 and sc["frame"] == 1
    assert abs(sc["time"] - 0.1) < 1e-15
    # CFL: dt = C_cfl * dx / vmax when smaller than dt_max
    cfg2 = configs.tiny()
    cfg2["params"].update(dt_max=1.0, output_dt=100.0)
    s2 = oracle_sim(cfg2)
    vx = np.zeros(len(cfg2["x"]))
    vx[3] = 40.0
    s2.set_velocity(vx, np.zeros_like(vx))
    s2.prime()
    s2.step()
    assert s2.scalars()["dt"] == 0.5 * cfg2["dx"] / 40.0


def test_sort_is_permutation_with_invariant_sums():
    from slabpro import configs
    cfg = configs.tiny(fric={"type": 0, "mu": 0.2})
    a = oracle_sim(cfg, sort=True)
    b = oracle_sim(cfg, sort=False)
    a.prime()
    b.prime()
    for _ in range(60):  # crosses the sort at step 50
        a.step()
        b.step()
    pa, pb = a.particles(), b.particles()
    for k in ("x", "y", "vx", "vy", "h"):
        assert np.allclose(pa[k], pb[k], rtol=1e-12, atol=1e-12 * np.abs(pb[k]).max())
    ga, gb = a.grid(), b.grid()
    assert abs(ga["mass"].sum() - gb["mass"].sum()) <= 1e-14 * gb["mass"].sum()


# ---------------------------------------------------------------- G2P height update, FLIP/PIC blend, pressure

def _flat_block(vel, dx=1.0, n=60, half=20.0, h0=1.0, rho=300.0, dt_max=1e-2, flip_ratio=0.95):
    """UNIFORM 2x2 block of side 2*half on flat ground with no friction;
    vel(px, py) -> (vx, vy) sets the initial particle velocities."""
    from slabpro import configs
    z = np.full((n, n), 50.0)
    c = n * dx / 2
    px, py, area = configs.rectangle_release(dx, c, c, 2 * half, 2 * half, 0.0)
    part = configs.finish_particles(px, py, np.full(len(px), h0), area * h0, rho)
    cfg = {"z": z, "dx": dx, "params": dict(rho=rho, gravity=9.81, dt_max=dt_max, output_dt=1e9, total_time=1e9,
                                            C_cfl=0.5, flip_ratio=flip_ratio),
           "friction": {"type": 0, "mu": 0.0}, **part}
    s = oracle_sim(cfg)
    vx, vy = vel(px, py)
    s.set_velocity(vx, vy)
    inner = (np.abs(px - c) < half - 4 * dx) & (np.abs(py - c) < half - 4 * dx)
    return s, cfg, inner, vx, vy


def test_height_update_follows_the_strain_trace():
    # v = (a (x - xc), b (y - yc)) on flat frictionless ground.  The quadratic
    # B-spline reproduces linear fields, so interior grid velocities are the
    # same linear field, sum_i grad w_i . v_i = a + b, and G2P's height update
    # (simulation.cpp:924-937) must give h / (1 + (a + b) dt) exactly.
    a, b = 2.0, -0.5
    s, cfg, inner, vx, vy = _flat_block(lambda px, py: (a * (px - 30.0), b * (py - 30.0)))
    s.prime()
    s.step()
    dt = s.scalars()["dt"]
    assert dt == 0.01  # CFL 0.5 dx / 41 m/s > dt_max
    p = s.particles()
    expect = 1.0 / (1.0 + (a + b) * dt)
    assert np.max(np.abs(p["h"][inner] / expect - 1.0)) < 1e-12
    assert np.allclose(p["A"][inner], p["m"][inner] / (p["h"][inner] * 300.0), rtol=1e-15)
    # velocities are reproduced and positions advect with them
    assert np.allclose(p["vx"][inner], vx[inner], rtol=0, atol=1e-12)
    assert np.allclose(p["x"][inner], np.asarray(cfg["x"])[inner] + dt * vx[inner], rtol=0, atol=1e-12)


def test_flip_pic_blend_weights():
    # Velocities alternate by +-delta around a mean U with the particle's
    # in-cell offset (0.25 vs 0.75 of a cell).  The B-spline weights of the
    # two offsets balance at every node (also after the w > 3e-3 cutoff), so
    # the grid velocity is exactly U, nothing accelerates (flip = 0) and
    # simulation.cpp:932 gives v_p = r (v_p + 0) + (1 - r) U = U + r delta_p.
    U = (1.5, -0.7)
    dlt = (0.4, 0.3)
    r = 0.8

    def vel(px, py):
        sx = np.where(np.isclose(np.mod(px, 1.0), 0.25), 1.0, -1.0)
        sy = np.where(np.isclose(np.mod(py, 1.0), 0.25), 1.0, -1.0)
        return U[0] + dlt[0] * sx, U[1] + dlt[1] * sy

    s, cfg, inner, vx, vy = _flat_block(vel, flip_ratio=r, dt_max=1e-3)
    s.prime()
    g = s.grid()
    act = g["mass"] > 1e-6
    iL, jB = g["window"][:2]
    gi = np.arange(g["mass"].shape[0])[:, None] + iL
    gj = np.arange(g["mass"].shape[1])[None, :] + jB
    core = act & (np.abs(gi - 30) < 16) & (np.abs(gj - 30) < 16)
    assert np.allclose(g["vx"][core], U[0], rtol=0, atol=1e-13)
    assert np.allclose(g["vy"][core], U[1], rtol=0, atol=1e-13)
    s.step()
    p = s.particles()
    assert np.allclose(p["vx"][inner] - U[0], r * (vx[inner] - U[0]), rtol=0, atol=1e-12)
    assert np.allclose(p["vy"][inner] - U[1], r * (vy[inner] - U[1]), rtol=0, atol=1e-12)
    dt = s.scalars()["dt"]
    assert np.allclose(p["x"][inner], np.asarray(cfg["x"])[inner] + dt * U[0], rtol=0, atol=1e-12)
    assert np.allclose(p["h"][inner], 1.0, rtol=0, atol=1e-12)


def _dam_break(theta_deg, h0, T=1.0, dt_max=2e-3, rho=300.0):
    """Frictionless dam break across y: reservoir y in [10, 40] (dam at y = 40),
    x in [5, 45], depth h0, on a plane tilted by theta along x.  Returns the
    y-momentum per unit width and per rho of the particles that start beyond
    the section y = 25 in the columns x in [12, 38], at times 0.5 and T."""
    from slabpro import configs
    dx = 1.0
    t = math.tan(math.radians(theta_deg))
    nx, ny = 56, 60
    z = np.repeat((-t * np.arange(nx) * dx)[:, None], ny, axis=1) + 100.0
    px, py = configs.uniform_samples(dx, 5.0, 45.0 - 1e-9, 10.0, 40.0 - 1e-9)
    part = configs.finish_particles(px, py, np.full(len(px), h0), 40.0 * 30.0 * h0, rho)
    cfg = {"z": z, "dx": dx, "params": dict(rho=rho, gravity=9.81, dt_max=dt_max, output_dt=1e9, total_time=1e9,
                                            C_cfl=0.5, flip_ratio=0.95),
           "friction": {"type": 0, "mu": 0.0}, **part}
    s = oracle_sim(cfg)
    s.prime()
    sel = (px >= 12.0) & (px <= 38.0) & (py > 25.0)
    width = 26.0
    out = []
    for t_end in (0.5, T):
        while s.scalars()["time"] < t_end - 1e-9:
            s.step()
        p = s.particles()
        out.append((s.scalars()["time"], (p["m"][sel] * p["vy"][sel]).sum() / rho / width))
    return out


@pytest.mark.parametrize("theta_deg,tol", [(0.0, 0.01), (45.0, 0.03)])
def test_pressure_term_momentum_balance_of_a_dam_break(theta_deg, tol):
    # Shallow-water momentum balance behind a dam break (Ritter's problem):
    # until the rarefaction from the dam reaches a section inside the resting
    # reservoir, the only force on the fluid beyond that section is the
    # hydrostatic thrust there, 1/2 rho g cos^2(theta) h0^2 per unit width
    # (the depth-averaged pressure sigma = -1/2 rho g cos^2 h of
    # simulation.cpp:753 and its divergence at :792).  So the fluid beyond it
    # gains y-momentum 1/2 g cos^2 h0^2 t per unit width and per rho.  On the
    # tilted plane the fluid also slides along x, which does not enter y.
    h0 = 2.0
    c2 = math.cos(math.radians(theta_deg)) ** 2
    for t, P in _dam_break(theta_deg, h0):
        expect = 0.5 * 9.81 * c2 * h0 * h0 * t
        assert abs(P / expect - 1.0) < tol, (t, P, expect)
