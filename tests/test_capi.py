"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/mhd_hip.h declares, and the header compiles as C for a reference-style
driver.  No GPU compute is called here."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "mhd_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?\**\s*\**\b([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for ref in ("fluidx", "copy_matrix", "advectbyzx", "transpose12", "transpose13", "calcfl"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    from heterogeneous_mhd_amd._lib import LIB_PATH, SIGNATURES
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    names = declared_functions()
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert sorted(SIGNATURES) == names


def test_library_loads_without_gpu_and_reports_version():
    from heterogeneous_mhd_amd import lib
    L = lib()
    assert b"gfx950" in L.mhd_version()
    assert L.mhd_max_pencil() >= 1024
    assert L.mhd_ic_mode_count(4) == 128  # |m| <= 4: 256 lattice points, one of each +-m


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc missing")
def test_reference_style_c_driver_compiles_and_links(tmp_path):
    """A C file written against the reference prototypes links against
    libmhd_hip.so unchanged (tests/c/ref_driver.c, run on the GPU by
    tests/test_gpu_c_driver.py)."""
    src = ROOT / "tests" / "c" / "ref_driver.c"
    lib = ROOT / "heterogeneous_mhd_amd"
    exe = tmp_path / "ref_driver"
    subprocess.run(["gcc", "-O2", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src), "-o",
                    str(exe), "-L", str(lib), "-lmhd_hip", f"-Wl,-rpath,{lib}", "-lm"], check=True)
    assert exe.exists()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; N > 1 exits with a message, before any GPU work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
