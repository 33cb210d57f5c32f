"""Seeded synthetic inputs for the benchmark workloads and the parity tests.

Workloads (all UNIFORM 2x2 particles per cell):
  c0  analytical 35 deg plane -> arc R=300 m -> flat runout, 1200x300 m, dx=2 m,
      rectangle 60x80 m at x=100 m, h0=2 m (4,800 particles), Coulomb tan 25 deg
  c1  2000x2000-node DEM at 5 m with k^-3 noise, raster release of 40 ellipses
      (~800k particles), Voellmy mu=0.155, xi=2000
  c2  8000x8000-node DEM at 1 m, 32-vertex star polygon release, r=800 m,
      h0=4 m (~8M particles), mu(I)
The particle initialisation restates initializeParticles
(src/simulation.cpp:259-466): dx-snapped bounding box, samples at
((i + d) * dx + x_min) with d in {.25,.75}, i-outer / j / offset order, shape
filters, A_init = V / sum(h) with a sequential sum, m = rho * h * A.
"""
import math

import numpy as np

MASTER_SEED = 20261203
MASK64 = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def subseed(config_index, scenario_index=0):
    return splitmix64(MASTER_SEED ^ ((config_index << 32) | scenario_index))


def rng(config_index, scenario_index=0):
    return np.random.default_rng(subseed(config_index, scenario_index))


def generate_grid(length, dx):
    """Node coordinates covering [0, length]: src/complex_terrain.cpp:67-82."""
    xs = [0.0]
    idx = 0
    x = 0.0
    while x < length:
        idx += 1
        x = float(idx) * dx
        xs.append(x)
    return np.array(xs)


# ---------------------------------------------------------------- terrains

def c0_profile(x, theta_deg=35.0, x_t=600.0, R=300.0, z_f=0.0):
    """35 deg plane for x <= 600 m, circular arc of radius 300 m, flat runout.
    z and slope are continuous at x_t and at x_c = x_t + R sin(theta)."""
    th = theta_deg * math.pi / 180.0
    x_c = x_t + R * math.sin(th)
    x = np.asarray(x, dtype=np.float64)
    z = np.empty_like(x)
    a = x <= x_t
    b = (x > x_t) & (x < x_c)
    c = x >= x_c
    z[a] = z_f + R * (1.0 - math.cos(th)) + (x_t - x[a]) * math.tan(th)
    z[b] = z_f + R - np.sqrt(R * R - (x[b] - x_c) ** 2)
    z[c] = z_f
    return z


def synthetic_dem(n, dx, valley_y, seed, rms=15.0, lambda_min=20.0, z0=2500.0, slope=0.3, curv=8e-5):
    """z = z0 - slope*x + curv*(y - valley_y)^2 + noise; noise is a Gaussian
    random field with isotropic power ~ k^-3 for wavelengths >= lambda_min,
    FFT-synthesised and scaled to the given RMS.  Returns z[i, j] (i along x)."""
    g = np.random.default_rng(seed)
    white = g.standard_normal((n, n))
    F = np.fft.rfft2(white)
    kx = np.fft.fftfreq(n, d=dx) * 2 * np.pi
    ky = np.fft.rfftfreq(n, d=dx) * 2 * np.pi
    k = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    amp = np.zeros_like(k)
    keep = (k > 0) & (k <= 2 * np.pi / lambda_min)
    amp[keep] = k[keep] ** -1.5
    F *= amp
    noise = np.fft.irfft2(F, s=(n, n))
    del F, white
    noise *= rms / max(noise.std(), 1e-30)
    x = np.arange(n, dtype=np.float64) * dx
    y = np.arange(n, dtype=np.float64) * dx
    z = (z0 - slope * x)[:, None] + (curv * (y - valley_y) ** 2)[None, :] + noise
    return np.ascontiguousarray(z)


# ---------------------------------------------------------------- sampler

def _snap_box(dx, xmin, xmax, ymin, ymax):
    inv = 1.0 / dx
    return (math.floor(xmin * inv) * dx, math.ceil(xmax * inv) * dx,
            math.floor(ymin * inv) * dx, math.ceil(ymax * inv) * dx)


def uniform_samples(dx, xmin, xmax, ymin, ymax):
    """UNIFORM sampler of src/simulation.cpp:355-385 (bbox snapped at 355-358)."""
    xmin, xmax, ymin, ymax = _snap_box(dx, xmin, xmax, ymin, ymax)
    inv = 1.0 / dx
    ncx = int(round((xmax - xmin) * inv))
    ncy = int(round((ymax - ymin) * inv))
    ddx = np.array([0.25, 0.75, 0.25, 0.75])
    ddy = np.array([0.25, 0.25, 0.75, 0.75])
    i = np.arange(ncx, dtype=np.float64)[:, None, None]
    j = np.arange(ncy, dtype=np.float64)[None, :, None]
    px = np.broadcast_to((i + ddx[None, None, :]) * dx + xmin, (ncx, ncy, 4)).ravel()
    py = np.broadcast_to((j + ddy[None, None, :]) * dx + ymin, (ncx, ncy, 4)).ravel()
    return px.copy(), py.copy()


def _rotate(px, py, ang):
    c, s = math.cos(ang), math.sin(ang)
    return c * px - s * py, s * px + c * py


def rectangle_release(dx, cx, cy, Lx, Ly, rot_deg=0.0):
    """Samples kept by the rotated-rectangle test, simulation.cpp:268-275, 400-404."""
    rot = rot_deg * math.pi / 180.0
    diag = math.sqrt(Lx * Lx + Ly * Ly) / 2.0
    px, py = uniform_samples(dx, cx - diag, cx + diag, cy - diag, cy + diag)
    lx, ly = _rotate(px - cx, py - cy, -rot)
    keep = (np.abs(lx) <= Lx / 2.0) & (np.abs(ly) <= Ly / 2.0)
    return px[keep], py[keep], Lx * Ly


def ellipse_keep(px, py, cx, cy, rx, ry, rot):
    lx, ly = _rotate(px - cx, py - cy, -rot)
    return (lx / rx) ** 2 + (ly / ry) ** 2 <= 1.0


def ellipse_release(dx, cx, cy, rx, ry, rot_deg=0.0):
    """simulation.cpp:278-285, 405-409."""
    rot = rot_deg * math.pi / 180.0
    big = max(rx, ry)
    px, py = uniform_samples(dx, cx - big, cx + big, cy - big, cy + big)
    keep = ellipse_keep(px, py, cx, cy, rx, ry, rot)
    return px[keep], py[keep], math.pi * rx * ry


def point_in_polygon(x, y, vx, vy):
    """Even-odd test of src/polygon_tools.cpp:87-101, vectorised over points."""
    inside = np.zeros(x.shape, dtype=bool)
    n = len(vx)
    j = n - 1
    for i in range(n):
        xi, yi, xj, yj = vx[i], vy[i], vx[j], vy[j]
        cond = ((yi > y) != (yj > y)) & (x < (xj - xi) * (y - yi) / ((yj - yi) + 1e-12) + xi)
        inside ^= cond
        j = i
    return inside


def polygon_area(vx, vy):
    """Shoelace area of the closed ring (GDAL OGRPolygon::get_Area for a simple ring)."""
    x = np.append(vx, vx[0])
    y = np.append(vy, vy[0])
    return abs(0.5 * float(np.sum(x[:-1] * y[1:] - x[1:] * y[:-1])))


def polygon_release(dx, vx, vy, chunk=1 << 22):
    vx = np.asarray(vx, dtype=np.float64)
    vy = np.asarray(vy, dtype=np.float64)
    px, py = uniform_samples(dx, vx.min(), vx.max(), vy.min(), vy.max())
    keep = np.zeros(px.shape, dtype=bool)
    for a in range(0, len(px), chunk):
        keep[a:a + chunk] = point_in_polygon(px[a:a + chunk], py[a:a + chunk], vx, vy)
    return px[keep], py[keep], polygon_area(vx, vy)


def raster_height(z, ox, oy, cs, nodata, px, py):
    """Bilinear release height of src/complex_terrain.cpp:132-183 (vectorised)."""
    ncols, nrows = z.shape
    xr = px - ox
    yr = py - oy
    xback = ox + (ncols - 1) * cs
    yback = oy + (nrows - 1) * cs
    dxl = (ox + cs) - ox
    dyl = (oy + cs) - oy
    out = np.zeros(px.shape)
    inb = (xr >= 0) & (yr >= 0) & (xr < xback) & (yr < yback)
    i = np.floor(np.where(inb, xr, 0.0) / dxl).astype(np.int64)
    j = np.floor(np.where(inb, yr, 0.0) / dyl).astype(np.int64)
    inb &= (i >= 0) & (j >= 0) & (i < ncols - 1) & (j < nrows - 1)
    ii, jj = i[inb], j[inb]
    fx = (xr[inb] - ii * dxl) / dxl
    fy = (yr[inb] - jj * dyl) / dyl
    z00, z10 = z[ii, jj], z[ii + 1, jj]
    z01, z11 = z[ii, jj + 1], z[ii + 1, jj + 1]
    bad = ((np.abs(z00 - nodata) < 1e-6) | (np.abs(z10 - nodata) < 1e-6) |
           (np.abs(z01 - nodata) < 1e-6) | (np.abs(z11 - nodata) < 1e-6))
    z0 = z00 * (1.0 - fx) + z10 * fx
    z1 = z01 * (1.0 - fx) + z11 * fx
    hgt = z0 * (1.0 - fy) + z1 * fy
    hgt = np.where(bad, 0.0, np.where(hgt > 0.0, hgt, 0.0))
    out[inb] = hgt
    return out


def raster_release(dx, hr, cs, h_min, ox=0.0, oy=0.0, nodata=-9999.0):
    """RASTER release, simulation.cpp:310-353 and 411-425: bbox = raster nodes,
    volume = sum over nodes with h >= h_min and h > 0 of h*cs^2, particles kept
    where the bilinear height is >= h_min and > 0."""
    ncols, nrows = hr.shape
    keep_nodes = (hr >= h_min) & (hr > 0.0)
    cell_area = cs * cs
    # sequential sum in the reference's j-outer, i-inner order
    vals = (hr.T[keep_nodes.T] * cell_area)
    volume = float(np.cumsum(vals)[-1]) if vals.size else 0.0
    x_min, y_min = ox, oy
    x_max, y_max = x_min + cs * (ncols - 1), y_min + cs * (nrows - 1)
    px, py = uniform_samples(dx, x_min, x_max, y_min, y_max)
    hp = raster_height(hr, ox, oy, cs, nodata, px, py)
    keep = (hp >= h_min) & (hp > 0.0)
    return px[keep], py[keep], hp[keep], volume


def finish_particles(px, py, h, volume, rho):
    """A_init = V / sum(h) (sequential sum, simulation.cpp:430-441); m = rho*h*A."""
    h = np.ascontiguousarray(h, dtype=np.float64)
    sum_h = float(np.cumsum(h)[-1])
    A_init = volume / sum_h
    A = np.full(len(px), A_init)
    m = rho * h * A
    return {"x": np.ascontiguousarray(px), "y": np.ascontiguousarray(py), "h": h, "A": A, "m": m}


# ---------------------------------------------------------------- workloads

def _params(gravity=9.81, rho=300.0, C_cfl=0.5, dt_max=0.01, flip_ratio=0.95, output_dt=1.0, total_time=60.0):
    return dict(gravity=gravity, rho=rho, C_cfl=C_cfl, dt_max=dt_max, flip_ratio=flip_ratio, output_dt=output_dt,
                total_time=total_time)


def c0():
    dx = 2.0
    xs = generate_grid(1200.0, dx)
    ys = generate_grid(300.0, dx)
    z = np.ascontiguousarray(np.repeat(c0_profile(xs)[:, None], len(ys), axis=1))
    px, py, area = rectangle_release(dx, 100.0, 150.0, 60.0, 80.0, 0.0)
    h0, rho = 2.0, 300.0
    part = finish_particles(px, py, np.full(len(px), h0), area * h0, rho)
    return {"name": "c0", "z": z, "dx": dx, "params": _params(rho=rho, total_time=60.0, output_dt=1.0),
            "friction": {"type": 0, "mu": math.tan(math.radians(25.0))}, **part}


def c1(release_ellipses=40):
    n, dx = 2000, 5.0
    z = synthetic_dem(n, dx, valley_y=5000.0, seed=subseed(1, 0))
    g = rng(1, 1)
    hr = np.zeros((n, n))
    xs = np.arange(n) * dx
    for _ in range(release_ellipses):
        cx, cy = g.uniform(500.0, 3000.0), g.uniform(1000.0, 9000.0)
        rx, ry = g.uniform(50.0, 300.0), g.uniform(50.0, 300.0)
        rot = g.uniform(0.0, math.pi)
        hgt = g.uniform(0.5, 3.0)
        big = max(rx, ry)
        i0, i1 = max(0, int((cx - big) / dx) - 1), min(n, int((cx + big) / dx) + 2)
        j0, j1 = max(0, int((cy - big) / dx) - 1), min(n, int((cy + big) / dx) + 2)
        X, Y = np.meshgrid(xs[i0:i1], xs[j0:j1], indexing="ij")
        inside = ellipse_keep(X, Y, cx, cy, rx, ry, rot)
        sub = hr[i0:i1, j0:j1]
        np.maximum(sub, np.where(inside, hgt, 0.0), out=sub)
    rho = 300.0
    px, py, hp, vol = raster_release(dx, hr, dx, 0.1)
    part = finish_particles(px, py, hp, vol, rho)
    return {"name": "c1", "z": z, "dx": dx, "params": _params(rho=rho, total_time=180.0),
            "friction": {"type": 1, "mu": 0.155, "xi": 2000.0}, **part}


def star_polygon(cx, cy, r_mean, nv, jitter, g):
    ang = np.arange(nv) * (2 * np.pi / nv)
    r = r_mean * g.uniform(1.0 - jitter, 1.0 + jitter, nv)
    return cx + r * np.cos(ang), cy + r * np.sin(ang)


def c2(n=8000, dx=1.0, r_mean=800.0, center=(1500.0, 4000.0)):
    z = synthetic_dem(n, dx, valley_y=n * dx / 2.0, seed=subseed(2, 0))
    vx, vy = star_polygon(center[0], center[1], r_mean, 32, 0.2, rng(2, 1))
    px, py, area = polygon_release(dx, vx, vy)
    h0, rho = 4.0, 2000.0
    part = finish_particles(px, py, np.full(len(px), h0), area * h0, rho)
    return {"name": "c2", "z": z, "dx": dx, "params": _params(rho=rho, total_time=120.0),
            "friction": {"type": 2, "mu1": 0.38, "mu2": 0.64, "I0": 0.279, "d": 0.5, "phi": 0.6},
            "polygon": (vx, vy), **part}


def tiny(nx=48, ny=40, dx=2.0, theta_deg=30.0, fric=None, seed=7, jitter=True, rho=300.0):
    """A small tilted rough terrain with a jittered rectangular release, for
    fast parity checks of every code path."""
    g = np.random.default_rng(seed)
    xs = np.arange(nx) * dx
    ys = np.arange(ny) * dx
    z = (-(xs * math.tan(math.radians(theta_deg)))[:, None] + 0.05 * ys[None, :] ** 1.2 +
         g.normal(0.0, 0.3, (nx, ny)))
    px, py, area = rectangle_release(dx, 0.3 * nx * dx, 0.5 * ny * dx, 0.25 * nx * dx, 0.35 * ny * dx, 10.0)
    if jitter:
        px = px + g.uniform(-0.2, 0.2, px.shape) * dx
        py = py + g.uniform(-0.2, 0.2, py.shape) * dx
    h = 1.0 + g.uniform(0.0, 1.0, px.shape)
    part = finish_particles(px, py, h, area * 1.5, rho)
    fr = fric or {"type": 0, "mu": 0.3}
    return {"name": "tiny", "z": np.ascontiguousarray(z), "dx": dx,
            "params": _params(rho=rho, total_time=5.0, output_dt=0.5), "friction": fr, **part}


# ---------------------------------------------------------------- ensemble (c3) / requests (c4)

def _scenario_release(g, dx, n_target, x_band=(600.0, 3000.0), y_band=(1500.0, 8500.0), extent=10000.0,
                      margin=250.0):
    """1-10 ellipses clustered around a random centre, semi-axes scaled so the
    UNIFORM sample count is within 5% of n_target.  Overlaps take the larger
    height (as a raster release would); each particle covers dx^2/4.  The
    shapes are then translated, if needed, so every particle lies at least
    `margin` inside the [0, extent]^2 DEM."""
    k = int(g.integers(1, 11))
    cx, cy = g.uniform(*x_band), g.uniform(*y_band)
    shapes = []
    for _ in range(k):
        ex, ey = cx + g.uniform(-300, 300), cy + g.uniform(-300, 300)
        rx, ry = g.uniform(0.5, 1.5), g.uniform(0.5, 1.5)
        shapes.append((ex, ey, rx, ry, g.uniform(0.0, math.pi), g.uniform(0.5, 3.0)))
    unit_area = sum(math.pi * s[2] * s[3] for s in shapes)
    scale = math.sqrt(n_target * dx * dx / 4.0 / unit_area)
    for _ in range(8):
        xs, ys, hs = _ellipse_union(dx, shapes, scale)
        if len(xs) == 0:
            scale *= 1.5
            continue
        ratio = len(xs) / n_target
        if abs(ratio - 1.0) <= 0.05:
            break
        scale /= math.sqrt(ratio)
    sx = sy = 0.0
    if len(xs):
        sx = max(0.0, margin - xs.min()) - max(0.0, xs.max() - (extent - margin))
        sy = max(0.0, margin - ys.min()) - max(0.0, ys.max() - (extent - margin))
    if sx or sy:
        # whole cells, so the UNIFORM lattice and the sample count are kept
        sx, sy = math.ceil(abs(sx) / dx) * dx * math.copysign(1, sx), math.ceil(abs(sy) / dx) * dx * math.copysign(1, sy)
        shapes = [(ex + sx, ey + sy, rx, ry, rot, hgt) for (ex, ey, rx, ry, rot, hgt) in shapes]
        xs, ys, hs = _ellipse_union(dx, shapes, scale)
    return xs, ys, hs


def _ellipse_union(dx, shapes, scale):
    xmin = min(s[0] - max(s[2], s[3]) * scale for s in shapes)
    xmax = max(s[0] + max(s[2], s[3]) * scale for s in shapes)
    ymin = min(s[1] - max(s[2], s[3]) * scale for s in shapes)
    ymax = max(s[1] + max(s[2], s[3]) * scale for s in shapes)
    px, py = uniform_samples(dx, xmin, xmax, ymin, ymax)
    h = np.zeros(len(px))
    for (ex, ey, rx, ry, rot, hgt) in shapes:
        inside = ellipse_keep(px, py, ex, ey, rx * scale, ry * scale, rot)
        np.maximum(h, np.where(inside, hgt, 0.0), out=h)
    keep = h > 0
    return px[keep], py[keep], h[keep]


def c3_scenario(i, dem_n=2000, dx=5.0):
    """Scenario i of the 256-scenario ensemble on the c1 DEM: log-uniform particle
    count in [2k, 1M], friction mix 50% Voellmy / 30% Coulomb / 20% mu(I)."""
    g = rng(3, i)
    n_target = int(round(math.exp(g.uniform(math.log(2e3), math.log(1e6)))))
    u = g.uniform()
    if u < 0.5:
        fr = {"type": 1, "mu": float(g.uniform(0.1, 0.3)), "xi": float(g.uniform(500.0, 4000.0))}
        rho = 300.0
    elif u < 0.8:
        fr = {"type": 0, "mu": float(g.uniform(0.3, 0.6))}
        rho = 300.0
    else:
        fr = {"type": 2, "mu1": 0.38, "mu2": 0.64, "I0": 0.279, "d": 0.5, "phi": 0.6}
        rho = 2000.0
    px, py, h = _scenario_release(g, dx, n_target)
    A = np.full(len(px), dx * dx / 4.0)
    part = {"x": px, "y": py, "h": h, "A": A, "m": rho * h * A}
    return {"name": f"c3_{i}", "dx": dx, "params": _params(rho=rho, total_time=120.0), "friction": fr,
            "n_target": n_target, **part}


def c1_dem():
    return synthetic_dem(2000, 5.0, valley_y=5000.0, seed=subseed(1, 0))


def c4_requests(n=2000, lam=5.0, seed_index=0):
    """Open-loop Poisson arrival times (s) of n requests at rate lam, and the
    scenario index each draws from the c3 distribution (request k uses c3
    scenario generator stream 1000 + k)."""
    g = rng(4, seed_index)
    gaps = g.exponential(1.0 / lam, n)
    return np.cumsum(gaps), [1000 + k for k in range(n)]
