"""The full step on the GPU (native plan, eager and HIP-graph replay) against
the reference's golden trajectories, plus the IC and FP64 diagnostics."""
import numpy as np
import pytest

from conftest import TRAJ_CASES, bits_equal, load_traj, sha

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", TRAJ_CASES)
@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_trajectory_matches_reference_digests(gpu, case, graph):
    """Every step's u, b and dt are bitwise those of the compiled reference."""
    import torch
    from heterogeneous_mhd_amd import solver
    tr = load_traj(case)
    nx, ny, nz = tr["dims"]
    print(f"trajectory {case} graph={graph}", flush=True)
    spec = solver.ICSpec(**tr["spec"])
    u0, b0 = solver.ic_host(nx, ny, nz, spec)
    assert sha(u0) == tr["ic_sha_u"] and sha(b0) == tr["ic_sha_b"]
    box = solver.Box(nx, ny, nz, cfl=tr["cfl"])
    box.set_state(torch.from_numpy(u0), torch.from_numpy(b0))
    for s, ref in enumerate(tr["f32"]["steps"]):
        box.run(1, graph=graph)
        t, dt = box.time()
        assert np.float32(dt) == np.float32(ref["dt"]), f"dt at step {s}"
        uh, bh = box.host_state()
        assert sha(uh) == ref["sha_u"], f"u at step {s}"
        assert sha(bh) == ref["sha_b"], f"b at step {s}"
    assert np.float32(t) == np.float32(tr["f32"]["t"])
    box.close()


def test_graph_replays_match_eager(gpu):
    """Several graph replays with eager work and syncs in between match the
    eager path bit for bit."""
    import torch
    from heterogeneous_mhd_amd import solver
    nx = ny = nz = 48
    spec = solver.ICSpec(seed=9, amp_v=0.3, amp_a=0.1)
    a = solver.Box(nx, ny, nz, cfl=0.7)
    g = solver.Box(nx, ny, nz, cfl=0.7)
    a.init(spec)
    g.init(spec)
    for _ in range(4):
        a.run(1, graph=False)
        g.run(1, graph=True)
        torch.cuda.synchronize()
        _ = (g.u * 2).sum().item()  # unrelated eager work between replays
        assert bits_equal(a.u.cpu().numpy(), g.u.cpu().numpy())
        assert bits_equal(a.b.cpu().numpy(), g.b.cpu().numpy())
        assert a.time() == g.time()


def test_ic_device_equals_host(gpu):
    from heterogeneous_mhd_amd import solver
    for dims, spec in [((32, 24, 24), solver.ICSpec(seed=3)),
                       ((64, 64, 64), solver.ICSpec(seed=1, amp_v=0.3, amp_a=0.05))]:
        uh, bh = solver.ic_host(*dims, spec)
        ud, bd = solver.ic_device(*dims, spec)
        assert bits_equal(ud.cpu().numpy(), uh) and bits_equal(bd.cpu().numpy(), bh)


def test_diagnostics_match_oracle(gpu):
    from heterogeneous_mhd_amd import solver
    from oracle import Oracle
    nx = ny = nz = 40
    u, b = solver.ic_host(nx, ny, nz, solver.ICSpec(seed=4, amp_v=0.2, amp_a=0.1))
    ud, bd = solver.ic_device(nx, ny, nz, solver.ICSpec(seed=4, amp_v=0.2, amp_a=0.1))
    g = solver.diagnostics(ud, bd, nx, ny, nz)
    o = Oracle("f32").diagnostics(u, b, nx, ny, nz)
    assert g["mass"] == pytest.approx(o["mass"], rel=1e-13)
    assert g["energy"] == pytest.approx(o["energy"], rel=1e-13)
    for a, c in zip(g["mom"], o["mom"]):
        assert a == pytest.approx(c, rel=1e-9, abs=1e-9)
    assert g["divb_max"] == pytest.approx(o["divb_max"], rel=1e-9, abs=1e-15)
    assert g["divb_max"] < 1e-6


def test_conservation_and_divb_256(gpu):
    """256^3, 3 steps: mass, momentum and energy sums move only by FP32
    round-off (the scheme is in flux-difference form) and div b stays at
    round-off (constrained transport); two runs are bitwise identical."""
    from heterogeneous_mhd_amd import solver
    n = 256
    spec = solver.ICSpec(seed=1, amp_v=0.1, amp_a=0.05)
    box = solver.Box(n, n, n, cfl=0.7)
    box.init(spec)
    d0 = box.diagnostics()
    box.run(3, graph=True)
    d1 = box.diagnostics()
    ncell = n ** 3
    assert abs(d1["mass"] - d0["mass"]) / ncell < 1e-6
    assert abs(d1["energy"] - d0["energy"]) / d0["energy"] < 1e-6
    for a, c in zip(d0["mom"], d1["mom"]):
        assert abs(a - c) / ncell < 1e-6
    assert d1["divb_max"] < 1e-5
    u1, b1 = box.host_state()
    box2 = solver.Box(n, n, n, cfl=0.7)
    box2.init(spec)
    box2.run(3, graph=False)
    u2, b2 = box2.host_state()
    assert bits_equal(u1, u2) and bits_equal(b1, b2)
