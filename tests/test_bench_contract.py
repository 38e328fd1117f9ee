"""bench.py's contract on the CPU: more than one GPU is refused with a
message before anything touches a GPU, and the algorithmic bytes behind
roofline.achieved are the per-node and per-grid-point figures
M (D s_k + s_c) + Q s_c (201.3 MB for the 3-D 128^3, M = 2^21 FP32 workload)."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_of_the_3d_workload():
    import bench
    M, Q, P = 2 ** 21, 256 ** 3, 128 ** 3
    sb = bench.stage_bytes(3, P, Q, M, 8)
    assert sb["spreading"] == M * (3 * 8 + 8) + Q * 8 == 201326592
    assert sb["interpolation"] == sb["spreading"]
    assert sb["deconvolve"] == P * 8 + Q * 8
