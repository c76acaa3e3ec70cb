"""Config 5 at full size (256x256x1024 cells, 134,217,728 particles, three
RK2 steps with face deposition) against the oracle on a random sample of
131,072 particles, and through face-count conservation over the whole
swarm, in both precision modes: tests/config5_full_size.py holds the checks
and the sample's parity bar.

It runs as a child process.  The workload is generated on the device with
torch (the 64-mode Fourier field over 67M cells takes minutes in numpy), and
torch, first imported in a process where the library has already used the
GPU (as the earlier GPU tests leave this one), reports that no HIP GPU is
available; the child imports torch first, as bench.py does.  The child's
output is printed with the test's."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

SCRIPT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config5_full_size.py")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config5_full_size_against_oracle_sample(precision):
    print("[gpu-case] config5 full size (child process)", precision, flush=True)
    r = subprocess.run([sys.executable, "-u", SCRIPT, precision], capture_output=True, text=True, timeout=180)
    print(r.stdout[-6000:])
    print(r.stderr[-6000:], file=sys.stderr)
    assert r.returncode == 0, f"child exited with {r.returncode}"
    assert r.stdout.rstrip().endswith("ok")
