"""The reference's text formats on the GPU path: an IC file in the format
C/init.c:34-52 reads, stepped by the device plan, dumped in the format of
C/check_value.c:53-69, byte for byte equal to the dumps of the golden final
state (the compiled reference's own 10-step trajectory at 16^3)."""
import hashlib

import numpy as np
import pytest

from conftest import GOLDEN, bits_equal, load_traj

pytestmark = pytest.mark.gpu

DUMPS = [f"C_version_check_u{c}.dat" for c in range(1, 6)] + [f"C_version_check_b{c}.dat" for c in range(1, 4)]


def test_source_init_file_stepped_on_gpu_dumps_match_golden(gpu, tmp_path):
    import torch
    from heterogeneous_mhd_amd import solver, textio
    tr = load_traj("16")
    nx, ny, nz = tr["dims"]
    u0, b0 = solver.ic_host(nx, ny, nz, solver.ICSpec(**tr["spec"]))
    src = tmp_path / "source_init_alfvenlinear.dat"
    textio.write_source_init(src, u0, b0, nx, ny, nz)
    print("read the IC file, 10 steps on the GPU, dumps", flush=True)
    u, b = textio.read_source_init(src, nx, ny, nz)
    assert bits_equal(u, u0) and bits_equal(b, b0)
    box = solver.Box(nx, ny, nz, cfl=tr["cfl"])
    box.set_state(torch.from_numpy(u), torch.from_numpy(b))
    box.run(tr["nsteps"], graph=True)
    ug, bg = box.host_state()
    box.close()
    last = tr["f32"]["steps"][-1]
    assert hashlib.sha256(ug.tobytes()).hexdigest() == last["sha_u"]
    assert hashlib.sha256(bg.tobytes()).hexdigest() == last["sha_b"]
    (tmp_path / "gpu").mkdir()
    (tmp_path / "golden").mkdir()
    textio.write_check_dumps(tmp_path / "gpu", ug, bg, nx, ny, nz)
    z = np.load(GOLDEN / "traj_16_final.npz", allow_pickle=False)
    textio.write_check_dumps(tmp_path / "golden", z["u"], z["b"], nx, ny, nz)
    for name in DUMPS:
        got = (tmp_path / "gpu" / name).read_bytes()
        want = (tmp_path / "golden" / name).read_bytes()
        assert len(got) > 0 and got == want, name
    # the dump holds every value: one %8.6e token per cell
    assert len((tmp_path / "gpu" / DUMPS[0]).read_text().split()) == nx * ny * nz
