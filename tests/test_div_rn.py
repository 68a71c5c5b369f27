"""The kernels compute the voxel coordinates (x' - origin) / spacing with one multiply and
two fmas from RN(1/spacing) (div_rn in the HIP source) and fall back to a true division for
offsets below 2^-99.  This checks, exhaustively over every float a in [2^-99, 4096), that the
three-operation form equals IEEE single division bit for bit for the spacings of the
workloads and a few others (including divisors next to powers of two), so voxel indices and
trilinear weights are those of test-grid-energrad.py:216-218's division."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path_factory.mktemp("divrn") / "div_rn_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", os.path.join(HERE, "csrc", "div_rn_check.c"),
                    "-o", exe, "-lm"], check=True)
    return exe


def test_div_rn_equals_ieee_division(checker):
    divisors = ["0.9", "0.375", "1.2", "0.99999994", "1.99999988", "0.45"]
    out = subprocess.run([checker, "0x1p-99", "4096", *divisors], capture_output=True, text=True, timeout=600)
    rows = [line.split() for line in out.stdout.strip().splitlines()]
    assert len(rows) == len(divisors)
    for b, bad, first in rows:
        assert int(bad) == 0, f"divisor {b}: {bad} mismatches, first at bit pattern {first}"
    assert out.returncode == 0


def test_checker_detects_underflow_region(checker):
    """Below 2^-99 the residual underflows and the short form differs from a true division
    for some a: the reason the kernels take the true division there."""
    out = subprocess.run([checker, "1.1754944e-38", "1.2e-38", "0.9"], capture_output=True, text=True, timeout=60)
    b, bad, first = out.stdout.split()
    assert int(bad) > 0
