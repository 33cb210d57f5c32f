"""Error of the FP32 form against the FP64 form of the same run.

Both forms run on the GPU through the C ABI; the FP64 form is the one pinned
to the CPU oracle.  Reported: max-height map L1 (relative to the FP64 map's
L1 norm) and L-inf (relative to its maximum), max-velocity map L1, runout
(largest particle x, and its difference in cells), deposited mass.
"""
import time

import numpy as np

from . import abi


def error_report(h32, v32, p32, h64, v64, p64, dx, x_foot=None, rest_speed=0.1):
    """Compare two finished runs: max maps (h_max, v_max^2) and final particles.
    Deposited mass: the mass of particles slower than rest_speed at the end, or,
    with x_foot, the mass beyond x_foot."""
    w = np.abs(h64).sum()
    rep = {
        "hmax_L1": float(np.abs(h32 - h64).sum() / w),
        "hmax_Linf": float(np.abs(h32 - h64).max() / np.abs(h64).max()),
        "vmax_L1": float(np.abs(np.sqrt(v32) - np.sqrt(v64)).sum() / np.sqrt(v64).sum()),
        "runout_x_fp32": float(p32["x"].max()),
        "runout_x_fp64": float(p64["x"].max()),
        "mass_fp32": float(np.sum(p32["m"])),
        "mass_fp64": float(np.sum(p64["m"])),
    }
    rep["runout_diff_cells"] = abs(rep["runout_x_fp32"] - rep["runout_x_fp64"]) / dx
    for tag, p in (("fp32", p32), ("fp64", p64)):
        if x_foot is not None:
            dep = p["x"] > x_foot
        else:
            dep = np.hypot(p["vx"], p["vy"]) < rest_speed
        rep[f"deposit_frac_{tag}"] = float(np.sum(p["m"][dep]) / np.sum(p["m"]))
    return rep


def run_to_end(cfg, precision):
    """One run of a workload dict (configs.*) to its total_time on the GPU.
    -> (h_max, v_max2, particles, last step info, wall seconds)."""
    import slabpro
    t0 = time.perf_counter()
    r = slabpro.make_run(cfg, precision=precision)
    r.prime()
    _, info = r.run()
    hm, vm = r.max_maps()
    p = r.particles()
    r.close()
    return hm, vm, p, info, time.perf_counter() - t0


def compare_forms(cfg, x_foot=None):
    """Runs cfg in both forms and returns (report, timings)."""
    h64, v64, p64, i64, t64 = run_to_end(cfg, abi.FP64)
    h32, v32, p32, i32, t32 = run_to_end(cfg, abi.FP32)
    rep = error_report(h32, v32, p32, h64, v64, p64, cfg["dx"], x_foot)
    rep["steps_fp32"], rep["steps_fp64"] = i32["step"], i64["step"]
    return rep, {"fp64_s": t64, "fp32_s": t32}
