"""The fluid sweep divides by the pencil-constant speed with a reciprocal and
one FMA residual step instead of an IEEE division; this checks on the CPU that
the two agree bit for bit over every divisor significand (tests/c/markstein_div.c).
The GPU parity tests then confirm it on real data."""
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc missing")
def test_reciprocal_fma_division_is_ieee(tmp_path):
    exe = tmp_path / "markstein_div"
    # -mfma: the host's hardware FMA, so fmaf is a single fused rounding
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", str(ROOT / "tests" / "c" / "markstein_div.c"),
                    "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe), "2"], capture_output=True, text=True, timeout=600)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
