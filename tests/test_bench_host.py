"""Host logic of bench.py (no GPU): the per-launch byte model behind the roofline,
the chunk accounting for passes that run chunk by chunk over z, and the one-GPU
contract of --gpus."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


class _Geo:
    def __init__(self, n):
        self.nx = self.ny = self.nz = n
        self.nki, self.nkj, self.nkk = n // 3 + 1, 2 * (n // 3) + 1, 2 * (n // 3) + 1
        self.mw = self.nki * self.nkj * self.nkk
        self.m1 = self.nki * self.nkj * n


def test_xpass_bytes_at_512_fp32_with_tracers():
    # 9 x-pencil fields of nz ny nki complex FP32 values, plus the tracer field
    # (3 FP32 per grid point) written for the tracer advance
    g = _Geo(512)
    m2 = 512 * 512 * 171
    assert bench.kernel_bytes(g, "xpass", 8, 4, 1 << 20) == 9 * m2 * 8 + 3 * 512**3 * 4 == 4838129664
    assert bench.kernel_bytes(g, "xpass", 8, 4, 0) == 9 * m2 * 8
    assert bench.kernel_bytes(g, "unknown", 8, 4, 0) is None


@pytest.mark.parametrize("name,launches,steps,expect", [
    ("xpass", 60, 10, 2.0),   # two chunks per stage
    ("xpass", 30, 10, 1.0),   # whole box
    ("yinv", 120, 10, 4.0),
    ("yfwd", 9, 3, 1.0),
    ("zinv", 60, 10, 1.0),    # never chunked: launches per step only
    ("zfwdformf", 60, 10, 1.0),
])
def test_chunks_per_stage(name, launches, steps, expect):
    assert bench.chunks_per_stage(name, launches, steps) == expect


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
