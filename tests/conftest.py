import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "slab-pro_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def oracle_sim(cfg, sort=True):
    """Build an oracle simulation from a slabpro.configs dict (test helper)."""
    from oracle import oracle as orc
    p = dict(cfg["params"])
    fr = dict(cfg["friction"])
    typ = fr.pop("type")
    prm = orc.params(fric_type=typ, **p, **fr)
    s = orc.OracleSim(cfg["z"], cfg["dx"], prm, cfg["x"], cfg["y"], cfg["h"], cfg["A"], cfg.get("m"))
    s.set_sort(sort)
    return s
