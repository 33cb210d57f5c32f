"""ctypes binding of libslabpro.so (the C ABI of include/slab_abi.h).

Host-side mirror of the reference's Simulation step interface for Python
drivers, tests and the benchmark.  Loads only the in-tree library and raises
if it is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG, "lib", "libslabpro.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "slab_abi.h")
FRONTEND_HEADER = os.path.join(os.path.dirname(PKG), "include", "slab_frontend.h")

OK, EINVAL, ENOMEM, EDEVICE, EDOMAIN, ENAN, ESTATE = range(7)
STATUS_NAMES = ["SLAB_OK", "SLAB_EINVAL", "SLAB_ENOMEM", "SLAB_EDEVICE", "SLAB_EDOMAIN", "SLAB_ENAN",
                "SLAB_ESTATE"]
FP64, FP32 = 0, 1
COULOMB, VOELLMY, MU_I = 0, 1, 2


class SlabError(RuntimeError):
    def __init__(self, status, msg):
        self.status = status
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"{name}: {msg}")


class Friction(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("mu", ctypes.c_double), ("xi", ctypes.c_double),
                ("mu1", ctypes.c_double), ("mu2", ctypes.c_double), ("I0", ctypes.c_double),
                ("d", ctypes.c_double), ("phi", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_double), ("rho", ctypes.c_double), ("C_cfl", ctypes.c_double),
                ("dt_max", ctypes.c_double), ("flip_ratio", ctypes.c_double), ("output_dt", ctypes.c_double),
                ("total_time", ctypes.c_double), ("precision", ctypes.c_int32)]


class Terrain(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("dx", ctypes.c_double),
                ("z", ctypes.POINTER(ctypes.c_double))]


class SlotInfo(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("dt", ctypes.c_double), ("status", ctypes.c_int32),
                ("frame", ctypes.c_int32), ("step", ctypes.c_int32), ("np", ctypes.c_int32),
                ("frame_hit", ctypes.c_int32), ("err", ctypes.c_int32)]


SLOT_FREE, SLOT_PRIMING, SLOT_RUNNING, SLOT_DONE, SLOT_FAILED = range(5)


class StepInfo(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("dt", ctypes.c_double), ("frame", ctypes.c_int32),
                ("step", ctypes.c_int32), ("frame_hit", ctypes.c_int32), ("finished", ctypes.c_int32),
                ("steps_done", ctypes.c_int32), ("reserved", ctypes.c_int32)]


# Every function declared in include/slab_abi.h, with its ctypes signature.
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p
SIGNATURES = {
    "slab_run_create": (ctypes.c_int, [ctypes.POINTER(Terrain), ctypes.POINTER(Params), ctypes.POINTER(Friction),
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "slab_run_upload_particles": (ctypes.c_int, [_VP, ctypes.c_int64] + [_DP] * 8),
    "slab_run_prime": (ctypes.c_int, [_VP]),
    "slab_run_step": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.POINTER(StepInfo)]),
    "slab_run_download_particles": (ctypes.c_int, [_VP] + [_DP] * 9),
    "slab_run_download_grid": (ctypes.c_int, [_VP, _IP, _IP, _IP, _IP] + [_DP] * 5),
    "slab_run_keep_grid": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "slab_run_download_max": (ctypes.c_int, [_VP, _DP, _DP]),
    "slab_run_diag": (ctypes.c_int, [_VP, _DP]),
    "slab_run_timing": (ctypes.c_int, [_VP, ctypes.c_int32, _DP, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]),
    "slab_kernel_name": (ctypes.c_char_p, [ctypes.c_int32]),
    "slab_kernel_count": (ctypes.c_int32, []),
    "slab_run_set_velocity": (ctypes.c_int, [_VP, _DP, _DP]),
    "slab_run_set_time": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int32]),
    "slab_last_error": (ctypes.c_char_p, []),
    "slab_run_destroy": (None, [_VP]),
    "slab_device_info": (ctypes.c_int, [_IP]),
    "slab_stream_probe": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _DP]),
    "slab_batch_create": (ctypes.c_int, [ctypes.POINTER(Terrain), ctypes.c_int32, ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "slab_batch_admit": (ctypes.c_int, [_VP, ctypes.POINTER(Params), ctypes.POINTER(Friction), ctypes.c_int64]
                         + [_DP] * 7 + [_IP]),
    "slab_batch_step": (ctypes.c_int, [_VP, ctypes.c_int32, _IP]),
    "slab_batch_slot_info": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.POINTER(SlotInfo)]),
    "slab_batch_statuses": (ctypes.c_int, [_VP, _IP, _DP]),
    "slab_batch_download_max": (ctypes.c_int, [_VP, ctypes.c_int32, _DP, _DP]),
    "slab_batch_download_particles": (ctypes.c_int, [_VP, ctypes.c_int32] + [_DP] * 9),
    "slab_batch_download_particles_noalloc": (ctypes.c_int, [_VP, ctypes.c_int32] + [_DP] * 9),
    "slab_batch_release": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "slab_batch_footprint": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int64)]),
    "slab_batch_use_graph": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "slab_batch_tiles": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int64)]),
    "slab_batch_set_tile_growth": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "slab_batch_timing": (ctypes.c_int, [_VP, ctypes.c_int32, _DP, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]),
    "slab_batch_destroy": (None, [_VP]),
    "slab_batch_set_resort": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "slab_run_set_resort": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "slab_batch_autotune": (ctypes.c_int, [_VP, ctypes.c_int32, _IP]),
    "slab_run_autotune": (ctypes.c_int, [_VP, ctypes.c_int32, _IP]),
    "slab_batch_set_sort_capacity": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "slab_run_set_sort_capacity": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "slab_batch_launch_shape": (ctypes.c_int, [_VP, _IP]),
    "slab_run_launch_shape": (ctypes.c_int, [_VP, _IP]),
    # include/slab_frontend.h
    "slab_sim_load_json": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "slab_sim_initialize": (ctypes.c_int, [_VP]),
    "slab_sim_initialize_terrain": (ctypes.c_int, [_VP]),
    "slab_sim_initialize_particles": (ctypes.c_int, [_VP]),
    "slab_sim_simulate": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int64)]),
    "slab_sim_set_verbose": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32]),
    "slab_sim_set_precision": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "slab_sim_terrain": (ctypes.c_int, [_VP, _IP, _IP, _DP, _DP, _DP, _DP]),
    "slab_sim_num_particles": (ctypes.c_int64, [_VP]),
    "slab_sim_particles": (ctypes.c_int, [_VP] + [_DP] * 5),
    "slab_sim_release": (ctypes.c_int, [_VP, _DP, _DP]),
    "slab_sim_output": (ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_int32, _IP, _IP]),
    "slab_sim_destroy": (None, [_VP]),
    "slab_poisson_disk": (ctypes.c_int64, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                           _DP, _DP, ctypes.c_int64]),
    "slab_polygon_area": (ctypes.c_double, [_DP, _DP, ctypes.c_int32]),
    "slab_point_in_polygon": (ctypes.c_int32, [ctypes.c_double, ctypes.c_double, _DP, _DP, ctypes.c_int32]),
    "slab_read_shapefile_polygon": (ctypes.c_int, [ctypes.c_char_p, _DP, _DP, ctypes.c_int32, _IP]),
    "slab_write_ply": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64] + [_DP] * 8),
    "slab_write_geotiff": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_char_p),
                                          ctypes.POINTER(ctypes.POINTER(ctypes.c_float)), ctypes.c_float]),
}

_lib = None


def lib():
    """Load the in-tree libslabpro.so; raise when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(st):
    if st != OK:
        raise SlabError(st, lib().slab_last_error().decode(errors="replace"))


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def device_info():
    out = (ctypes.c_int32 * 3)()
    _check(lib().slab_device_info(out))
    return {"device": out[0], "cus": out[1], "gfx950": bool(out[2])}


PROBE_COPY, PROBE_READ, PROBE_WRITE, PROBE_PARTICLES, PROBE_COPY_NT, PROBE_WRITE_NT = range(6)
# eight dependent copy phases: one launch each / one cooperative launch with
# grid barriers; four workgroups per CU, then one
PROBE_PHASES_LAUNCHES, PROBE_PHASES_BARRIERS, PROBE_PHASES_LAUNCHES_1, PROBE_PHASES_BARRIERS_1 = 6, 7, 8, 9


def stream_probe(kind, n, iters=20):
    """Mean kernel time (ms) of a streaming kernel of known size: PROBE_COPY /
    _READ / _WRITE / _COPY_NT / _WRITE_NT over n bytes, PROBE_PARTICLES over n
    particles (28 B read, 32 B written in place)."""
    ms = ctypes.c_double(0.0)
    _check(lib().slab_stream_probe(int(kind), int(n), int(iters), ctypes.byref(ms)))
    return ms.value


def kernel_names():
    L = lib()
    return [L.slab_kernel_name(k).decode() for k in range(L.slab_kernel_count())]


class Run:
    """One simulation on the device: terrain + particles + driver state.

    Mirrors the step part of the reference's Simulation class
    (src/simulation.hpp:79-286): prime() = simulate() lines 493-495,
    step() = the while loop at src/simulation.cpp:500-510.
    """

    def __init__(self, z, dx, params, friction):
        z = _f64(z)
        assert z.ndim == 2
        self.nx, self.ny = z.shape
        self.dx = float(dx)
        t = Terrain(self.nx, self.ny, self.dx, _p(z))
        h = ctypes.c_void_p()
        _check(lib().slab_run_create(ctypes.byref(t), ctypes.byref(params), ctypes.byref(friction),
                                     ctypes.byref(h)))
        self.h = h
        self.np = 0

    def close(self):
        if getattr(self, "h", None):
            lib().slab_run_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, x, y, h, A, m=None, vx=None, vy=None, vz=None):
        x, y, h, A, m, vx, vy, vz = map(_f64, (x, y, h, A, m, vx, vy, vz))
        self.np = len(x)
        _check(lib().slab_run_upload_particles(self.h, self.np, _p(x), _p(y), _p(vx), _p(vy), _p(vz),
                                               _p(h), _p(A), _p(m)))

    def prime(self):
        _check(lib().slab_run_prime(self.h))

    def step(self, max_steps=1):
        info = StepInfo()
        _check(lib().slab_run_step(self.h, int(max_steps), ctypes.byref(info)))
        return {"time": info.time, "dt": info.dt, "frame": info.frame, "step": info.step,
                "frame_hit": bool(info.frame_hit), "finished": bool(info.finished),
                "steps_done": info.steps_done}

    def run(self, max_steps=1 << 30, on_frame=None):
        """Steps until total_time (or max_steps); calls on_frame(info) at output frames."""
        done = 0
        info = None
        while done < max_steps:
            info = self.step(min(max_steps - done, 1 << 20))
            done += info["steps_done"]
            if info["frame_hit"] and on_frame is not None:
                on_frame(info)
            if info["finished"] or info["steps_done"] == 0:
                break
        return done, info

    def keep_grid(self, on=True):
        _check(lib().slab_run_keep_grid(self.h, 1 if on else 0))

    def particles(self):
        names = ["x", "y", "vx", "vy", "vz", "z", "h", "A", "m"]
        arrs = {k: np.zeros(self.np) for k in names}
        _check(lib().slab_run_download_particles(self.h, *[_p(arrs[k]) for k in names]))
        return arrs

    def grid(self):
        i0, j0, nx, ny = (ctypes.c_int32() for _ in range(4))
        L = lib()
        _check(L.slab_run_download_grid(self.h, ctypes.byref(i0), ctypes.byref(j0), ctypes.byref(nx),
                                        ctypes.byref(ny), None, None, None, None, None))
        shape = (nx.value, ny.value)
        names = ["h", "vx", "vy", "vz", "mass"]
        arrs = {k: np.zeros(shape) for k in names}
        _check(L.slab_run_download_grid(self.h, None, None, None, None, *[_p(arrs[k]) for k in names]))
        arrs["window"] = (i0.value, j0.value, nx.value, ny.value)
        return arrs

    def max_maps(self):
        hm = np.zeros((self.nx, self.ny))
        vm = np.zeros((self.nx, self.ny))
        _check(lib().slab_run_download_max(self.h, _p(hm), _p(vm)))
        return hm, vm

    def diag(self):
        out = np.zeros(8)
        _check(lib().slab_run_diag(self.h, _p(out)))
        keys = ["grid_mass", "grid_px", "grid_py", "particle_mass", "particle_volume", "dt", "vmax", "n_active"]
        return dict(zip(keys, out.tolist()))

    def timing(self, enable):
        n = lib().slab_kernel_count()
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        _check(lib().slab_run_timing(self.h, 1 if enable else 0, _p(ms),
                                     cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), n))
        return {name: (ms[k], int(cnt[k])) for k, name in enumerate(kernel_names())}

    def set_velocity(self, vx, vy):
        _check(lib().slab_run_set_velocity(self.h, _p(_f64(vx)), _p(_f64(vy))))

    def set_resort(self, steps):
        _check(lib().slab_run_set_resort(self.h, int(steps)))

    def autotune(self, trial_steps=8):
        out = (ctypes.c_int32 * 2)()
        _check(lib().slab_run_autotune(self.h, int(trial_steps), out))
        return self.launch_shape()

    def set_sort_capacity(self, particles):
        _check(lib().slab_run_set_sort_capacity(self.h, int(particles)))

    def launch_shape(self):
        out = (ctypes.c_int32 * 3)()
        _check(lib().slab_run_launch_shape(self.h, out))
        return {"resort_period": out[0], "workgroups": out[1], "sort_capacity": out[2]}

    def set_time(self, t, frame):
        _check(lib().slab_run_set_time(self.h, float(t), int(frame)))


class Batch:
    """Many scenarios on one shared DEM (FP32 form), advanced together."""

    def __init__(self, z, dx, max_slots, max_particles):
        z = _f64(z)
        self.nx, self.ny = z.shape
        self.dx = float(dx)
        self.max_slots = int(max_slots)
        t = Terrain(self.nx, self.ny, self.dx, _p(z))
        h = ctypes.c_void_p()
        _check(lib().slab_batch_create(ctypes.byref(t), self.max_slots, int(max_particles), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().slab_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, params, friction, x, y, h, A, m=None, vx=None, vy=None):
        """Returns the slot id, or None when no slot / particle capacity is free."""
        x, y, h, A, m, vx, vy = map(_f64, (x, y, h, A, m, vx, vy))
        slot = ctypes.c_int32(-1)
        st = lib().slab_batch_admit(self.h, ctypes.byref(params), ctypes.byref(friction), len(x), _p(x), _p(y),
                                    _p(vx), _p(vy), _p(h), _p(A), _p(m), ctypes.byref(slot))
        if st == ENOMEM:
            return None
        _check(st)
        return slot.value

    def step(self, max_steps=1):
        n = ctypes.c_int32(0)
        _check(lib().slab_batch_step(self.h, int(max_steps), ctypes.byref(n)))
        return n.value

    def statuses(self):
        st = np.zeros(self.max_slots, dtype=np.int32)
        tm = np.zeros(self.max_slots)
        _check(lib().slab_batch_statuses(self.h, st.ctypes.data_as(_IP), _p(tm)))
        return st, tm

    def slot_info(self, slot):
        i = SlotInfo()
        _check(lib().slab_batch_slot_info(self.h, int(slot), ctypes.byref(i)))
        return {"time": i.time, "dt": i.dt, "status": i.status, "frame": i.frame, "step": i.step, "np": i.np,
                "frame_hit": bool(i.frame_hit), "err": i.err}

    def max_maps(self, slot):
        hm = np.zeros((self.nx, self.ny))
        vm = np.zeros((self.nx, self.ny))
        _check(lib().slab_batch_download_max(self.h, int(slot), _p(hm), _p(vm)))
        return hm, vm

    def particles(self, slot):
        n = self.slot_info(slot)["np"]
        names = ["x", "y", "vx", "vy", "vz", "z", "h", "A", "m"]
        arrs = {k: np.zeros(n) for k in names}
        _check(lib().slab_batch_download_particles(self.h, int(slot), *[_p(arrs[k]) for k in names]))
        return arrs

    def particles_noalloc(self, slot):
        """As particles(), through the library's fixed scratch (no device allocation)."""
        n = self.slot_info(slot)["np"]
        names = ["x", "y", "vx", "vy", "vz", "z", "h", "A", "m"]
        arrs = {k: np.zeros(n) for k in names}
        _check(lib().slab_batch_download_particles_noalloc(self.h, int(slot), *[_p(arrs[k]) for k in names]))
        return arrs

    def release(self, slot):
        _check(lib().slab_batch_release(self.h, int(slot)))

    def footprint(self):
        out = (ctypes.c_int64 * 4)()
        _check(lib().slab_batch_footprint(self.h, out))
        return {"bytes_per_particle": out[0], "bytes_per_slot": out[1], "free_particles": out[2],
                "free_slots": out[3]}

    def tiles(self):
        out = (ctypes.c_int64 * 9)()
        _check(lib().slab_batch_tiles(self.h, out))
        return {"grid_tile_capacity": out[0], "grid_tiles": out[1], "particle_tile_capacity": out[2],
                "particle_tiles": out[3], "bytes_per_grid_tile": out[4], "bytes_per_particle_tile": out[5],
                "max_map_tile_capacity": out[6], "max_map_tiles": out[7], "bytes_per_max_map_tile": out[8]}

    def set_tile_growth(self, factor):
        _check(lib().slab_batch_set_tile_growth(self.h, int(factor)))

    def use_graph(self, on=True):
        _check(lib().slab_batch_use_graph(self.h, 1 if on else 0))

    def set_resort(self, steps):
        _check(lib().slab_batch_set_resort(self.h, int(steps)))

    def autotune(self, trial_steps=8):
        out = (ctypes.c_int32 * 2)()
        _check(lib().slab_batch_autotune(self.h, int(trial_steps), out))
        return self.launch_shape()

    def set_sort_capacity(self, particles):
        _check(lib().slab_batch_set_sort_capacity(self.h, int(particles)))

    def launch_shape(self):
        out = (ctypes.c_int32 * 3)()
        _check(lib().slab_batch_launch_shape(self.h, out))
        return {"resort_period": out[0], "workgroups": out[1], "sort_capacity": out[2]}

    def timing(self, enable):
        n = lib().slab_kernel_count()
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        _check(lib().slab_batch_timing(self.h, 1 if enable else 0, _p(ms),
                                       cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), n))
        return {name: (ms[k], int(cnt[k])) for k, name in enumerate(kernel_names())}


def params(gravity=9.81, rho=300.0, C_cfl=0.5, dt_max=0.01, flip_ratio=0.95, output_dt=1.0, total_time=60.0,
           precision=FP64):
    return Params(gravity, rho, C_cfl, dt_max, flip_ratio, output_dt, total_time, precision)


def friction(type=COULOMB, mu=0.0, xi=0.0, mu1=0.0, mu2=0.0, I0=0.0, d=0.0, phi=0.0):
    return Friction(type, mu, xi, mu1, mu2, I0, d, phi)
