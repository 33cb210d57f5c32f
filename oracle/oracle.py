"""ctypes binding of the CPU FP64 oracle (oracle/slab_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker / the timed CPU baseline.
The product library never loads it.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

COULOMB, VOELLMY, MU_I = 0, 1, 2


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("gravity", ctypes.c_double), ("rho", ctypes.c_double), ("C_cfl", ctypes.c_double),
        ("dt_max", ctypes.c_double), ("flip_ratio", ctypes.c_double), ("output_dt", ctypes.c_double),
        ("total_time", ctypes.c_double), ("fric_type", ctypes.c_int32),
        ("mu", ctypes.c_double), ("xi", ctypes.c_double), ("mu1", ctypes.c_double),
        ("mu2", ctypes.c_double), ("I0", ctypes.c_double), ("d", ctypes.c_double),
        ("phi", ctypes.c_double),
    ]


def build(force=False):
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return LIB_PATH


def build_mutant(n):
    """Compile the oracle with planted mistake number n (ORC_MUTANT in
    slab_oracle.c); returns the library path.  For the mutation tests only."""
    path = os.path.join(_HERE, "_build", "liboracle_mut%d.so" % int(n))
    subprocess.check_call(["make", "-s", "-C", _HERE, os.path.join("_build", os.path.basename(path))])
    return path


_lib = None
_lib_path = None


@contextlib.contextmanager
def using_library(path):
    """Routes every oracle call inside the block to the library at `path`
    (a mutant build); the ordinary library is restored afterwards."""
    global _lib, _lib_path
    saved = (_lib, _lib_path)
    _lib, _lib_path = None, path
    try:
        yield
    finally:
        _lib, _lib_path = saved


def lib():
    global _lib
    if _lib is None:
        path = _lib_path or LIB_PATH
        if path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_N.restype = ctypes.c_double
        L.orc_N.argtypes = [ctypes.c_double]
        L.orc_dNdu.restype = ctypes.c_double
        L.orc_dNdu.argtypes = [ctypes.c_double]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, dp,
                                 ctypes.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_particles.argtypes = [ctypes.c_void_p, ctypes.c_int32, dp, dp, dp, dp, dp]
        L.orc_prime.argtypes = [ctypes.c_void_p]
        L.orc_step.argtypes = [ctypes.c_void_p]
        L.orc_step.restype = ctypes.c_int
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.orc_run.restype = ctypes.c_int64
        L.orc_scalars.argtypes = [ctypes.c_void_p, dp]
        L.orc_window.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
        L.orc_get_grid.argtypes = [ctypes.c_void_p] + [dp] * 9
        L.orc_get_particles.argtypes = [ctypes.c_void_p] + [dp] * 9
        L.orc_set_velocity.argtypes = [ctypes.c_void_p, dp, dp]
        L.orc_get_max.argtypes = [ctypes.c_void_p, dp, dp]
        L.orc_get_terrain.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.orc_set_sort.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]
        L.orc_uniform_shape.restype = ctypes.c_int64
        L.orc_uniform_shape.argtypes = [ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp,
                                        ctypes.c_int64]
        L.orc_uniform_polygon.restype = ctypes.c_int64
        L.orc_uniform_polygon.argtypes = [ctypes.c_double, dp, dp, ctypes.c_int32, dp, dp, ctypes.c_int64]
        L.orc_raster_height.restype = ctypes.c_double
        L.orc_raster_height.argtypes = [dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def params(gravity=9.81, rho=300.0, C_cfl=0.5, dt_max=0.01, flip_ratio=0.95, output_dt=1.0,
           total_time=60.0, fric_type=COULOMB, mu=0.0, xi=0.0, mu1=0.0, mu2=0.0, I0=0.0, d=0.0, phi=0.0):
    return OrcParams(gravity, rho, C_cfl, dt_max, flip_ratio, output_dt, total_time, fric_type,
                     mu, xi, mu1, mu2, I0, d, phi)


class OracleSim:
    """One simulation on the CPU FP64 oracle (terrain + particles + driver state)."""

    def __init__(self, z, dx, prm, x, y, h, A, m=None):
        z = _f64(z)
        assert z.ndim == 2
        self.nx, self.ny = z.shape
        self.dx = float(dx)
        self._z = z
        L = self._L = lib()  # the library this simulation was created on
        self.h = L.orc_create(self.nx, self.ny, self.dx, _p(z), ctypes.byref(prm))
        if not self.h:
            raise ValueError("orc_create rejected the terrain")
        self.np = len(x)
        xs, ys, hs, As = _f64(x), _f64(y), _f64(h), _f64(A)
        ms = None if m is None else _f64(m)
        L.orc_set_particles(self.h, self.np, _p(xs), _p(ys), _p(hs), _p(As), _p(ms))

    def close(self):
        if self.h:
            self._L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prime(self):
        self._L.orc_prime(self.h)

    def step(self):
        return self._L.orc_step(self.h)

    def run(self, max_steps=1 << 62):
        return self._L.orc_run(self.h, max_steps)

    def set_sort(self, on):
        self._L.orc_set_sort(self.h, 1 if on else 0)

    def set_time(self, t, frame):
        self._L.orc_set_time(self.h, t, frame)

    def scalars(self):
        out = np.zeros(10)
        self._L.orc_scalars(self.h, _p(out))
        return {"time": out[0], "dt": out[1], "frame": int(out[2]), "step": int(out[3]),
                "phases": dict(zip(["updateDt", "remesh", "sort", "P2G", "grid", "G2P"], out[4:10]))}

    def window(self):
        w = (ctypes.c_int32 * 4)()
        self._L.orc_window(self.h, w)
        return tuple(w)

    def grid(self):
        iL, jB, gnx, gny = self.window()
        names = ["mass", "vx", "vy", "vz", "h", "flipx", "flipy", "Fx", "Fy"]
        arrs = {k: np.zeros((gnx, gny)) for k in names}
        self._L.orc_get_grid(self.h, *[_p(arrs[k]) for k in names])
        arrs["window"] = (iL, jB, gnx, gny)
        return arrs

    def particles(self):
        names = ["x", "y", "vx", "vy", "vz", "z", "h", "A", "m"]
        arrs = {k: np.zeros(self.np) for k in names}
        self._L.orc_get_particles(self.h, *[_p(arrs[k]) for k in names])
        return arrs

    def set_velocity(self, vx, vy):
        self._L.orc_set_velocity(self.h, _p(_f64(vx)), _p(_f64(vy)))

    def max_maps(self):
        hm = np.zeros((self.nx, self.ny))
        vm = np.zeros((self.nx, self.ny))
        self._L.orc_get_max(self.h, _p(hm), _p(vm))
        return hm, vm

    def terrain(self):
        gx = np.zeros((self.nx, self.ny))
        gy = np.zeros((self.nx, self.ny))
        th = np.zeros((self.nx, self.ny))
        self._L.orc_get_terrain(self.h, _p(gx), _p(gy), _p(th))
        return gx, gy, th


def N(u):
    return lib().orc_N(float(u))


def dNdu(u):
    return lib().orc_dNdu(float(u))


def uniform_shape(dx, kind, cx, cy, a, b, rot_rad):
    L = lib()
    n = L.orc_uniform_shape(dx, kind, cx, cy, a, b, rot_rad, None, None, 0)
    x = np.zeros(max(n, 1))
    y = np.zeros(max(n, 1))
    L.orc_uniform_shape(dx, kind, cx, cy, a, b, rot_rad, _p(x), _p(y), n)
    return x[:n], y[:n]


def uniform_polygon(dx, vx, vy):
    L = lib()
    vx, vy = _f64(vx), _f64(vy)
    n = L.orc_uniform_polygon(dx, _p(vx), _p(vy), len(vx), None, None, 0)
    x = np.zeros(max(n, 1))
    y = np.zeros(max(n, 1))
    L.orc_uniform_polygon(dx, _p(vx), _p(vy), len(vx), _p(x), _p(y), n)
    return x[:n], y[:n]


def raster_height(z, ox, oy, cs, nodata, px, py):
    z = _f64(z)
    ncols, nrows = z.shape
    return lib().orc_raster_height(_p(z), ncols, nrows, ox, oy, cs, nodata, px, py)
