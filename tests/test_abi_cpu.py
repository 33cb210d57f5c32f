"""C-ABI library checks that need no GPU: it loads, exports every function
include/slab_abi.h and include/slab_frontend.h declare, and fails cleanly (no
CPU fallback) without a device."""
import os
import re
import subprocess

import pytest

from slabpro import abi


def _declared_functions():
    names = set()
    for h in (abi.HEADER, abi.FRONTEND_HEADER):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(slab_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exists_and_loads():
    assert os.path.exists(abi.LIB_PATH), "run __graft_entry__.build() first"
    abi.lib()


def test_exports_every_declared_symbol():
    declared = _declared_functions()
    assert len(declared) >= 15
    out = subprocess.check_output(["nm", "-D", "--defined-only", abi.LIB_PATH], text=True)
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared if n not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(_declared_functions()) == set(abi.SIGNATURES)


def test_kernel_names():
    names = abi.kernel_names()
    assert {"p2g", "grid", "g2p", "tile_sort", "ctl_begin"} <= set(names)


def test_null_and_bad_arguments_are_rejected():
    import ctypes
    L = abi.lib()
    assert L.slab_run_create(None, None, None, None) == abi.EINVAL
    assert L.slab_run_prime(None) == abi.EINVAL
    assert L.slab_run_step(None, 1, None) == abi.EINVAL
    t = abi.Terrain(2, 2, 1.0, None)
    p = abi.params()
    f = abi.friction()
    h = ctypes.c_void_p()
    assert L.slab_run_create(ctypes.byref(t), ctypes.byref(p), ctypes.byref(f), ctypes.byref(h)) == abi.EINVAL
    assert b"4x4" in L.slab_last_error()


def test_no_device_fails_loudly():
    """Without a GPU the library reports SLAB_EDEVICE instead of computing on the CPU."""
    try:
        info = abi.device_info()
    except abi.SlabError as e:
        assert e.status == abi.EDEVICE
        return
    if info["gfx950"]:
        pytest.skip("a gfx950 device is present")


def test_batch_and_probe_entry_points_reject_bad_arguments():
    import ctypes
    L = abi.lib()
    d = (ctypes.c_double * 4)()
    out = (ctypes.c_int64 * 9)()
    ms = ctypes.c_double(0.0)
    assert L.slab_batch_tiles(None, out) == abi.EINVAL
    assert L.slab_batch_set_tile_growth(None, 4) == abi.EINVAL
    assert L.slab_batch_download_particles_noalloc(None, 0, d, d, d, d, d, d, d, d, d) == abi.EINVAL
    assert L.slab_stream_probe(10, 1 << 20, 1, ctypes.byref(ms)) == abi.EINVAL  # unknown kind
    assert L.slab_stream_probe(0, 1000, 1, ctypes.byref(ms)) == abi.EINVAL      # not a multiple of 16 bytes
    assert L.slab_stream_probe(0, 1 << 20, 0, ctypes.byref(ms)) == abi.EINVAL   # no iterations
    assert L.slab_stream_probe(0, 1 << 20, 1, None) == abi.EINVAL


def test_launch_shape_entry_points_reject_bad_arguments():
    import ctypes
    L = abi.lib()
    shape = (ctypes.c_int32 * 3)()
    assert L.slab_batch_set_sort_capacity(None, 2048) == abi.EINVAL
    assert L.slab_run_set_sort_capacity(None, 2048) == abi.EINVAL
    assert L.slab_batch_launch_shape(None, shape) == abi.EINVAL
    assert L.slab_run_launch_shape(None, shape) == abi.EINVAL
