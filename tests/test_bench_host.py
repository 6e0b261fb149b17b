"""Host-side pieces of bench.py: the algorithmic byte model, the traffic read
back from the committed PMC summary, and the one-GPU contract."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_byte_model_film():
    # 256x256x32 FP64 with Px, Py, Pz = 512, 512, 64: S1 = 257*256*32, S2 = 257*512*32
    nx, ny, nz, w = 256, 256, 32, 8
    b, total = bench.byte_model(nx, ny, nz, 512, 512, 64, w, 2, chained=True)
    N = nx * ny * nz
    S1, S2 = 257 * ny * nz, 257 * 512 * nz
    assert b["P2"] == 3 * 2 * w * S1 + 3 * 2 * w * S2
    assert b["P3"] == 2 * 3 * 2 * w * S2 + 6 * w * 257 * 257 * 33
    assert b["P4"] == b["P2"]
    assert b["P1"] == 4 * w * N + 3 * 2 * w * S1
    # chained: the last pass also writes the next x-spectrum; P1 is not in the total
    assert b["P5"] == 2 * 3 * 2 * w * S1 + 7 * w * N + N
    assert total == b["P2"] + b["P3"] + b["P4"] + b["P5"]
    # 1.437 GB per chained evaluation at this size
    assert abs(total - 1.4368e9) < 1e6
    # RK4 stage vectors: y0, k1..k3 read, stage vector and k written, 30 w N per step
    assert bench.rk4_stage_extra(N, w) == 30 * w * N


@pytest.mark.parametrize("p", ["P1", "P2", "P3", "P4", "P5"])
def test_traffic_read_from_profile(p):
    t = bench.read_traffic(p)
    assert t is not None and 1e8 < t < 1e9


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
