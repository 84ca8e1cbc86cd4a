"""The differentiable SSIM family on the GPU: runs tests/grad_checks.py in a child process (it
imports torch, whose HIP runtime then serves the engine in that process; the main GPU test process
keeps the engine on its own runtime and never loads torch)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_gradient_checks_in_child_process():
    r = subprocess.run([sys.executable, "-u", "-m", "tests.grad_checks"], cwd=ROOT, capture_output=True, text=True,
                       timeout=600)
    print(r.stdout[-6000:])
    print(r.stderr[-3000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
