"""The reference's step loop as a plain C program on host arrays, linked
against libmhd_hip.so (tests/c/ref_driver.c): every step is bitwise the
compiled reference's (golden trajectory digests)."""
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_traj, sha

pytestmark = pytest.mark.gpu


def test_reference_style_c_driver_matches_golden(gpu, tmp_path):
    from heterogeneous_mhd_amd import solver
    lib = ROOT / "heterogeneous_mhd_amd"
    exe = tmp_path / "ref_driver"
    subprocess.run(["gcc", "-O2", "-I", str(ROOT / "include"), str(ROOT / "tests" / "c" / "ref_driver.c"),
                    "-o", str(exe), "-L", str(lib), "-lmhd_hip", f"-Wl,-rpath,{lib}", "-lm"], check=True)
    tr = load_traj("24x16x16")
    nx, ny, nz = tr["dims"]
    u, b = solver.ic_host(nx, ny, nz, solver.ICSpec(**tr["spec"]))
    (tmp_path / "in.bin").write_bytes(u.tobytes() + b.tobytes())
    out = subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), str(nx), str(ny),
                          str(nz), str(tr["nsteps"])], check=True, capture_output=True, text=True,
                         timeout=120).stdout.split("\n")
    for s, ref in enumerate(tr["f32"]["steps"]):
        dt = float.fromhex(out[s].split()[3])
        assert np.float32(dt) == np.float32(ref["dt"]), s
    raw = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.float32)
    n = nx * ny * nz
    assert sha(raw[:5 * n]) == tr["f32"]["steps"][-1]["sha_u"]
    assert sha(raw[5 * n:]) == tr["f32"]["steps"][-1]["sha_b"]
