"""The host code that runs without a GPU, under AddressSanitizer and
UndefinedBehaviorSanitizer: the library's text-format reader and writers
(mhd_io.hip) and seeded host IC generator (mhd_ic.hip), and the oracle
(mhd_ref.c, FP32 and FP64), driven by tests/c/sanitize_harness.cpp over small
cubic and non-cubic boxes and good and malformed files.  Built by
tests/c/sanitize.mk (hipcc with the sanitizers on the host pass only, ROCm's
clang++ to link); make rebuilds only what changed."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CLANG = Path("/opt/rocm/llvm/bin/clang++")


@pytest.mark.skipif(shutil.which("hipcc") is None and not Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc missing")
@pytest.mark.skipif(not CLANG.exists(), reason="ROCm clang++ missing")
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_host_code_and_oracle_under_asan_ubsan(tmp_path, prec):
    subprocess.run(["make", "-s", "-f", str(ROOT / "tests" / "c" / "sanitize.mk")], check=True,
                   capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(ROOT / "tests" / "c" / "_san" / f"sanitize_{prec}"), str(tmp_path)], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitize harness ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
