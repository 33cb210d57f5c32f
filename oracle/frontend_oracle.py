"""CPU restatement of the reference's front end: file readers, Poisson-disk
sampling and particle initialisation.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the product's
front end (slab-pro_amd/host/*.cpp through include/slab_frontend.h).  The
product never imports it.

What follows the reference, and what pins each part:
  read_asc            src/complex_terrain.cpp:31-65; pinned by the fixture
                      files' own header values and row order (tests).
  read_shapefile_ring src/polygon_tools.cpp:22-65 via OGR; ESRI Shapefile
                      Technical Description (1998), record layout restated;
                      pinned by the ring closing on itself and the .shx
                      index agreeing with the .shp record offsets.
  polygon_area        OGRPolygon::get_Area (GDAL, version not pinned by the
                      reference's CMakeLists.txt: find_package(GDAL)) as the
                      absolute shoelace sum; pinned by closed forms
                      (rectangles, triangles) and an independent numpy
                      shoelace to 1e-12.
  libc_rand           glibc rand() of a fresh process (srand(1)); pinned by
                      the documented first outputs 1804289383, 846930886,
                      1681692777.
  poisson_disk        src/poisson_disk_sampling.hpp (Bridson 2007); this
                      restatement finds neighbours in a hash of cells of
                      side r (3x3 block), not the reference's r/sqrt(2)
                      background grid, so the two neighbour searches check
                      each other; pinned by the minimum-distance and in-box
                      invariants.
  initialize_particles src/simulation.cpp:259-476 (UNIFORM through the C
                      oracle's orc_uniform_*; POISSON samples filtered here);
                      pinned by closed-form areas / volumes and
                      sum h A = V0 exactly as the reference scales it.
Pure Python loops: small cases only.
"""
import ctypes
import math
import os
import struct

import numpy as np

from . import oracle

# --- files ---------------------------------------------------------------------


def read_asc(path):
    """-> dict(ncols, nrows, xllcorner, yllcorner, cellsize, nodata, z) with
    z[i, j] (i along x, j along y, row order reversed so y points up)."""
    with open(path) as f:
        toks = f.read().split()
    hdr = {}
    for k in range(6):
        hdr[toks[2 * k].lower()] = float(toks[2 * k + 1])
    ncols, nrows = int(hdr["ncols"]), int(hdr["nrows"])
    vals = np.array(toks[12:12 + ncols * nrows], dtype=np.float64).reshape(nrows, ncols)
    z = np.ascontiguousarray(vals[::-1, :].T)
    return {"ncols": ncols, "nrows": nrows, "xllcorner": hdr["xllcorner"], "yllcorner": hdr["yllcorner"],
            "cellsize": hdr["cellsize"], "nodata": hdr["nodata_value"], "z": z}


def read_shapefile_ring(path):
    """Outer ring (first part) of the first non-null polygon record."""
    with open(path, "rb") as f:
        b = f.read()
    code, = struct.unpack(">i", b[0:4])
    assert code == 9994, "not a shapefile"
    shape_type, = struct.unpack("<i", b[32:36])
    assert shape_type in (5, 15, 25), "not a polygon shapefile"
    off = 100
    while off + 8 <= len(b):
        _, words = struct.unpack(">ii", b[off:off + 8])
        body = b[off + 8:off + 8 + 2 * words]
        off += 8 + 2 * words
        st, = struct.unpack("<i", body[:4])
        if st == 0:
            continue
        nparts, npts = struct.unpack("<ii", body[36:44])
        parts = struct.unpack("<%di" % nparts, body[44:44 + 4 * nparts])
        pts = np.frombuffer(body, dtype="<f8", count=2 * npts, offset=44 + 4 * nparts).reshape(npts, 2)
        end = parts[1] if nparts > 1 else npts
        ring = pts[parts[0]:end]
        return ring[:, 0].copy(), ring[:, 1].copy()
    raise ValueError("no polygon record")


# --- geometry ----------------------------------------------------------------------


def polygon_area(vx, vy):
    """|shoelace| of the ring, closed if it is not."""
    x, y = list(map(float, vx)), list(map(float, vy))
    if len(x) < 3:
        return 0.0
    if x[0] != x[-1] or y[0] != y[-1]:
        x.append(x[0])
        y.append(y[0])
    n = len(x)
    s = x[0] * (y[1] - y[n - 1])
    for i in range(1, n - 1):
        s += x[i] * (y[i + 1] - y[i - 1])
    s += x[n - 1] * (y[0] - y[n - 2])
    return 0.5 * abs(s)


def point_in_polygon(px, py, vx, vy):
    """Even-odd crossing test with the 1e-12 guard of src/polygon_tools.cpp:87-100."""
    n = len(vx)
    inside = False
    j = n - 1
    for i in range(n):
        xi, yi, xj, yj = vx[i], vy[i], vx[j], vy[j]
        if (yi > py) != (yj > py) and px < (xj - xi) * (py - yi) / ((yj - yi) + 1e-12) + xi:
            inside = not inside
        j = i
    return inside


# --- random stream and Poisson-disk sampling ----------------------------------------

_libc = None
_libm = None
RAND_MAX = 2147483647


def sincos(a):
    """(sin a, cos a) from glibc's sincos(): a GCC build of the reference
    computes std::cos/std::sin of one angle with one sincos call."""
    global _libm
    if _libm is None:
        _libm = ctypes.CDLL("libm.so.6")
        _libm.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        _libm.sincos.restype = None
    s, c = ctypes.c_double(), ctypes.c_double()
    _libm.sincos(float(a), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def libc_rand(n):
    """The first n values of rand() after srand(1) (a fresh process's stream)."""
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL("libc.so.6")
    _libc.srand(1)
    r = _libc.rand
    return [r() for _ in range(n)]


class _Stream:
    def __init__(self, chunk=1 << 16):
        self.buf, self.pos, self.chunk = [], 0, chunk
        self.total = 0

    def next(self):
        if self.pos == len(self.buf):
            # the stream is regenerated from the start, longer, when it runs out
            self.total += self.chunk
            self.buf = libc_rand(self.total)
            self.chunk *= 2
        v = self.buf[self.pos]
        self.pos += 1
        return v


def poisson_disk(dx, Lx, Ly, ppc=4):
    r = math.sqrt(15.0 / 24.0 * dx * dx / ppc)
    rr = r * r
    rs = _Stream()
    unit = lambda: float(rs.next()) / RAND_MAX  # noqa: E731
    grid = {}

    def key(x, y):
        return (int(math.floor(x / r)), int(math.floor(y / r)))

    def ok(x, y):
        gx, gy = key(x, y)
        for a in (gx - 1, gx, gx + 1):
            for b in (gy - 1, gy, gy + 1):
                for (qx, qy) in grid.get((a, b), ()):
                    if (qx - x) * (qx - x) + (qy - y) * (qy - y) < rr:
                        return False
        return True

    x0 = unit() * Lx
    y0 = unit() * Ly
    pts = [(x0, y0)]
    grid.setdefault(key(x0, y0), []).append((x0, y0))
    active = [0]
    while active:
        k = rs.next() % len(active)
        bx, by = pts[active[k]]
        hit = None
        for _ in range(30):
            rho = unit() * r + r
            th = unit() * 2 * math.pi
            sn, cs = sincos(th)
            cx, cy = bx + rho * cs, by + rho * sn
            if 0 <= cx <= Lx and 0 <= cy <= Ly and ok(cx, cy):
                hit = (cx, cy)
                break
        if hit is None:
            del active[k]
        else:
            pts.append(hit)
            active.append(len(pts) - 1)
            grid.setdefault(key(*hit), []).append(hit)
    a = np.array(pts)
    return a[:, 0].copy(), a[:, 1].copy()


# --- initializeParticles ---------------------------------------------------------------


def release_height(rast, ox, oy, px, py):
    return oracle.raster_height(rast["z"], ox, oy, rast["cellsize"], rast["nodata"], px, py)


def initialize_particles(cfg, xll, yll, samples=None):
    """cfg: the JSON dict (paths already resolved); xll/yll: terrain corner.
    samples: Poisson samples (x, y) of the snapped box to use instead of
    drawing them here (for boxes too large for the Python sampler).
    -> dict(x, y, h, A, m, area, volume)."""
    dx = float(cfg["terrain"]["analytical"]["dx"]) if cfg["terrain"]["type"] == "ANALYTICAL" else cfg["_dx"]
    inv = 1.0 / dx
    rel = cfg["release"]
    kind = rel["type"]
    h0 = float(rel["h0"])
    rho = float(cfg["physical_params"]["rho"])
    rast = None
    if kind == "RECTANGLE":
        c = rel["rectangle"]
        cx, cy, a, b = float(c["center"][0]), float(c["center"][1]), float(c["Lx"]), float(c["Ly"])
        rot = float(c["rotation_deg"]) * math.pi / 180.0
        diag = math.sqrt(a * a + b * b) / 2.0
        box = (cx - diag, cx + diag, cy - diag, cy + diag)
        area = a * b
    elif kind == "ELLIPSE":
        c = rel["ellipse"]
        cx, cy, a, b = float(c["center"][0]), float(c["center"][1]), float(c["rx"]), float(c["ry"])
        rot = float(c["rotation_deg"]) * math.pi / 180.0
        big = max(a, b)
        box = (cx - big, cx + big, cy - big, cy + big)
        area = math.pi * a * b
    elif kind == "POLYGON":
        vx, vy = read_shapefile_ring(rel["polygon"]["filepath"])
        vx, vy = vx - xll, vy - yll
        box = (vx.min(), vx.max(), vy.min(), vy.max())
        area = polygon_area(vx, vy)
    else:
        rast = read_asc(rel["raster"]["filepath"])
        ox, oy = rast["xllcorner"] - xll, rast["yllcorner"] - yll
        cs = rast["cellsize"]
        box = (ox, ox + cs * (rast["ncols"] - 1), oy, oy + cs * (rast["nrows"] - 1))
        hmin = float(rel["raster"]["h_min"])
        zz = rast["z"]
        sel = (zz >= hmin) & (zz > 0.0)
        vol = 0.0
        for j in range(rast["nrows"]):          # same accumulation order: j outer, i inner
            for i in range(rast["ncols"]):
                if sel[i, j]:
                    vol += zz[i, j] * (cs * cs)
        area = int(sel.sum()) * (cs * cs)
    volume = vol if kind == "RASTER" else area * h0
    x_min = math.floor(box[0] * inv) * dx
    x_max = math.ceil(box[1] * inv) * dx
    y_min = math.floor(box[2] * inv) * dx
    y_max = math.ceil(box[3] * inv) * dx
    method = cfg["numerical_params"]["sampling_method"]
    if method == "UNIFORM" and kind in ("RECTANGLE", "ELLIPSE"):
        px, py = oracle.uniform_shape(dx, 0 if kind == "RECTANGLE" else 1, cx, cy, a, b, rot)
        h = np.full(len(px), h0)
    elif method == "UNIFORM" and kind == "POLYGON":
        px, py = oracle.uniform_polygon(dx, vx, vy)
        h = np.full(len(px), h0)
    else:
        if method == "UNIFORM":
            ncx, ncy = int(round((x_max - x_min) * inv)), int(round((y_max - y_min) * inv))
            ox4, oy4 = (0.25, 0.75, 0.25, 0.75), (0.25, 0.25, 0.75, 0.75)
            sx = [((i + ox4[d]) * dx) for i in range(ncx) for j in range(ncy) for d in range(4)]
            sy = [((j + oy4[d]) * dx) for i in range(ncx) for j in range(ncy) for d in range(4)]
        elif samples is not None:
            sx, sy = samples
        else:
            sx, sy = poisson_disk(dx, x_max - x_min, y_max - y_min, 4)
        xs, ys, hs = [], [], []
        for qx, qy in zip(sx, sy):
            px_, py_ = qx + x_min, qy + y_min
            if kind == "RASTER":
                hv = release_height(rast, ox, oy, px_, py_)
                keep = hv >= hmin and hv > 0.0
            elif kind == "POLYGON":
                hv, keep = h0, point_in_polygon(px_, py_, vx, vy)
            else:
                s_, c_ = sincos(-rot)
                lx0, ly0 = px_ - cx, py_ - cy
                lx, ly = c_ * lx0 - s_ * ly0, s_ * lx0 + c_ * ly0
                hv = h0
                if kind == "RECTANGLE":
                    keep = abs(lx) <= a / 2.0 and abs(ly) <= b / 2.0
                else:
                    keep = (lx / a) ** 2 + (ly / b) ** 2 <= 1.0
            if keep:
                xs.append(px_)
                ys.append(py_)
                hs.append(hv)
        px, py, h = np.array(xs), np.array(ys), np.array(hs)
    sum_h = 0.0
    for v in h:
        sum_h += float(v)
    A0 = volume / sum_h
    A = np.full(len(px), A0)
    m = rho * h * A
    return {"x": np.asarray(px), "y": np.asarray(py), "h": h, "A": A, "m": m, "area": area, "volume": volume}


def fixture_dir():
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden", "brienz")
