"""The bench's command line on the CPU: a request for more than one GPU exits
with status 2 and says that only one GPU is measured (nothing is launched)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_several_gpus(n):
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n), "--steps", "1"],
                         capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert out.returncode == 2, out.stdout + out.stderr
    assert "one GPU only" in out.stderr
    assert out.stdout == ""
