"""Python side of the configuration front end (include/slab_frontend.h).

Simulation mirrors the reference's driver sequence (src/main.cpp:29-58):

    sim = Simulation.load_from_json("case.json")
    sim.initialize(); sim.initialize_terrain(); sim.initialize_particles()
    sim.simulate()          # GPU: frames + max_values.tif

Readers for the files the solver writes (binary PLY, the uncompressed
GeoTIFFs) are included so outputs can be inspected without GDAL.
"""
import ctypes
import struct

import numpy as np

from . import abi
from .abi import _check, _DP, _IP, _f64, _p, lib


class Simulation:
    def __init__(self, handle):
        self.h = handle

    @classmethod
    def load_from_json(cls, path):
        h = ctypes.c_void_p()
        _check(lib().slab_sim_load_json(str(path).encode(), ctypes.byref(h)))
        return cls(h)

    def close(self):
        if getattr(self, "h", None):
            lib().slab_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_verbose(self, verbose=True, timing=False):
        _check(lib().slab_sim_set_verbose(self.h, int(verbose), int(timing)))

    def set_precision(self, precision):
        _check(lib().slab_sim_set_precision(self.h, int(precision)))

    def initialize(self):
        _check(lib().slab_sim_initialize(self.h))

    def initialize_terrain(self):
        _check(lib().slab_sim_initialize_terrain(self.h))

    def initialize_particles(self):
        _check(lib().slab_sim_initialize_particles(self.h))

    def simulate(self):
        n = ctypes.c_int64()
        _check(lib().slab_sim_simulate(self.h, ctypes.byref(n)))
        return n.value

    def terrain(self):
        nx, ny = ctypes.c_int32(), ctypes.c_int32()
        dx, xll, yll = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().slab_sim_terrain(self.h, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(dx),
                                      ctypes.byref(xll), ctypes.byref(yll), None))
        z = np.zeros((nx.value, ny.value))
        _check(lib().slab_sim_terrain(self.h, None, None, None, None, None, _p(z)))
        return {"z": z, "dx": dx.value, "xllcorner": xll.value, "yllcorner": yll.value}

    def particles(self):
        n = lib().slab_sim_num_particles(self.h)
        if n < 0:
            raise abi.SlabError(abi.ESTATE, "particles not initialized")
        out = {k: np.zeros(n) for k in ("x", "y", "h", "A", "m")}
        _check(lib().slab_sim_particles(self.h, *(_p(out[k]) for k in ("x", "y", "h", "A", "m"))))
        return out

    def release(self):
        a, v = ctypes.c_double(), ctypes.c_double()
        _check(lib().slab_sim_release(self.h, ctypes.byref(a), ctypes.byref(v)))
        return a.value, v.value

    def output(self):
        buf = ctypes.create_string_buffer(4096)
        nf, pad = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().slab_sim_output(self.h, buf, 4096, ctypes.byref(nf), ctypes.byref(pad)))
        return buf.value.decode(), nf.value, pad.value


def poisson_disk(dx, Lx, Ly, ppc=4):
    n = lib().slab_poisson_disk(dx, Lx, Ly, ppc, None, None, 0)
    if n < 0:
        raise abi.SlabError(abi.EINVAL, lib().slab_last_error().decode())
    x, y = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    lib().slab_poisson_disk(dx, Lx, Ly, ppc, _p(x), _p(y), n)
    return x[:n], y[:n]


def polygon_area(vx, vy):
    vx, vy = _f64(vx), _f64(vy)
    return lib().slab_polygon_area(_p(vx), _p(vy), len(vx))


def point_in_polygon(px, py, vx, vy):
    vx, vy = _f64(vx), _f64(vy)
    return bool(lib().slab_point_in_polygon(float(px), float(py), _p(vx), _p(vy), len(vx)))


def read_shapefile_polygon(path):
    n = ctypes.c_int32()
    _check(lib().slab_read_shapefile_polygon(str(path).encode(), None, None, 0, ctypes.byref(n)))
    x, y = np.zeros(n.value), np.zeros(n.value)
    _check(lib().slab_read_shapefile_polygon(str(path).encode(), _p(x), _p(y), n.value, ctypes.byref(n)))
    return x, y


def write_ply(path, x, y, vx, vy, vz, z, h, A):
    arrs = [_f64(a) for a in (x, y, vx, vy, vz, z, h, A)]
    _check(lib().slab_write_ply(str(path).encode(), len(arrs[0]), *(_p(a) for a in arrs)))


def write_geotiff(path, bands, names, x0, ytop, dx, nodata=-9999.0):
    """bands: list of (nx, ny) arrays in grid node order (i along x, j along y)."""
    bs = [np.ascontiguousarray(b, dtype=np.float32) for b in bands]
    nx, ny = bs[0].shape
    fp = ctypes.POINTER(ctypes.c_float)
    arr = (fp * len(bs))(*[b.ctypes.data_as(fp) for b in bs])
    nm = (ctypes.c_char_p * len(bs))(*[n.encode() for n in names])
    _check(lib().slab_write_geotiff(str(path).encode(), nx, ny, x0, ytop, dx, len(bs), nm, arr, nodata))


# --- readers ------------------------------------------------------------------

def read_ply(path):
    """Binary little-endian PLY with float/double vertex properties -> dict of arrays."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    header = data[:end].decode().splitlines()
    if header[0] != "ply" or header[1] != "format binary_little_endian 1.0":
        raise ValueError("not a binary little-endian PLY")
    n, props = 0, []
    for ln in header[2:]:
        t = ln.split()
        if t[0] == "element":
            n = int(t[2])
        elif t[0] == "property":
            props.append((t[2], {"float": "<f4", "double": "<f8"}[t[1]]))
    rec = np.frombuffer(data, dtype=np.dtype(props), count=n, offset=end)
    return {k: np.array(rec[k]) for k, _ in props}, header


_TIFF_TYPES = {1: ("B", 1), 2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 12: ("d", 8)}


def read_geotiff(path):
    """Uncompressed little-endian TIFF (chunky Float32 bands) -> (bands, tags).
    bands has shape (nbands, rows, cols), rows north-up as stored."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != b"II*\x00":
        raise ValueError("not a little-endian TIFF")
    (ifd,) = struct.unpack_from("<I", data, 4)
    (ntag,) = struct.unpack_from("<H", data, ifd)
    tags = {}
    for k in range(ntag):
        tid, typ, cnt, val = struct.unpack_from("<HHII", data, ifd + 2 + 12 * k)
        code, size = _TIFF_TYPES[typ]
        off = ifd + 2 + 12 * k + 8 if size * cnt <= 4 else val
        if typ == 2:
            tags[tid] = data[off:off + cnt].rstrip(b"\x00").decode()
        else:
            tags[tid] = list(struct.unpack_from("<" + code * cnt, data, off))
    cols, rows, nb = tags[256][0], tags[257][0], tags[277][0]
    if tags[259][0] != 1 or tags[284][0] != 1 or set(tags[258]) != {32} or set(tags[339]) != {3}:
        raise ValueError("only uncompressed chunky Float32 is read here")
    buf = b"".join(data[o:o + c] for o, c in zip(tags[273], tags[279]))
    img = np.frombuffer(buf, dtype="<f4").reshape(rows, cols, nb)
    return np.transpose(img, (2, 0, 1)).copy(), tags
