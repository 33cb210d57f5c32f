"""Python host side of the MI355X depth-averaged MPM step (slab-pro_amd).

abi      ctypes binding of libslabpro.so (include/slab_abi.h)
configs  seeded synthetic workloads c0, c1, c2 and small parity inputs
"""
from . import abi, configs  # noqa: F401


def make_run(cfg, precision=None, keep_grid=False):
    """Create a device run from a config dict of slabpro.configs and upload its particles."""
    p = dict(cfg["params"])
    if precision is not None:
        p["precision"] = precision
    fr = dict(cfg["friction"])
    run = abi.Run(cfg["z"], cfg["dx"], abi.params(**p), abi.friction(**fr))
    run.upload(cfg["x"], cfg["y"], cfg["h"], cfg["A"], cfg.get("m"))
    if keep_grid:
        run.keep_grid(True)
    return run
