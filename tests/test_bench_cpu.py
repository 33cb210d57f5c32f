"""bench.py contract pieces that need no GPU: the one-GPU guard, the roofline
traffic lookup from the committed rocprof summaries, and the algorithmic-byte
tables matching the per-step totals."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "single MI355X" in r.stderr and r.stdout == ""


def test_traffic_lookup_reads_the_committed_summary():
    t, src = bench.measured_traffic("fp32", "c2", "g2p")
    assert src is not None and src.startswith("profiles/")
    assert 1e8 < t < 5e9          # bytes per launch for 8M particles
    assert bench.measured_traffic("fp32", "c2", "no_such_kernel") == (None, None)
    assert bench.measured_traffic("fp32", "c9", "g2p") == (None, None)


def test_step_bytes_are_the_sum_of_the_persistent_fields():
    # FP32: read cell ox oy vx vy h m (28 B) + write cell ox oy vx vy h (24 B) per particle
    # (vz and z only around output steps); per node: terrain 12 B, max maps 16 B read+write,
    # vf 16 B written, vf re-read by G2P 16 B
    assert bench.STEP_BYTES_FP32 == {"particle": 52, "node": 60}
    assert bench.ALG_BYTES_FP32["g2p"]["particle"] + bench.ALG_BYTES_FP32["p2g"]["particle"] <= 104
    assert bench.STEP_BYTES_FP64 == {"particle": 104, "node": 120}
