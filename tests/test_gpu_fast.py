"""The FP32 tolerance mode (MHD_PREC_FAST) against the reference.

The reference's arithmetic mixes float storage with double-promoted
subexpressions; the fast mode computes in FP32 only (hardware reciprocal and
square root, FMA).  Stated tolerances, relative to the field group's largest
magnitude (groups: rho, momentum, E, B):

  * one step (64^3, the CPU-runnable workload: v amp 0.1, A amp 0.05,
    CFL 0.7): L-inf <= 2e-6 against the reference FP32 path (measured
    <= 5.4e-7 on MI355X);
  * 100 steps of the same box: L-inf <= 1e-2 and L2 <= 1e-4 against the
    reference FP32 path, and the fast mode's L2 error against the FP64 build
    of the reference is at most 3x the reference FP32 path's own L2 error
    against FP64 (measured about 2x).

The oracle FP32 path is bitwise the compiled reference (see test_oracle.py)
and the FP64 oracle is the reference's arithmetic with float read as double.
"""
import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu

N = 64
SPEC = dict(seed=1, amp_v=0.1, amp_a=0.05)


def _groups(u, b):
    u = np.asarray(u).reshape(-1, 5).astype(np.float64)
    b = np.asarray(b).reshape(-1, 3).astype(np.float64)
    return {"rho": u[:, 0], "mom": u[:, 1:4], "E": u[:, 4], "B": b}


def rel_errors(x, ref):
    """{group: (L-inf, L2)} of x against ref, relative to ref's magnitude."""
    gx, gr = _groups(*x), _groups(*ref)
    return {k: (float(np.abs(gx[k] - gr[k]).max() / np.abs(gr[k]).max()),
                float(np.linalg.norm(gx[k] - gr[k]) / np.linalg.norm(gr[k]))) for k in gx}


def _gpu_run(u0, b0, steps, precision, graph=True):
    import torch
    from heterogeneous_mhd_amd import solver
    box = solver.Box(N, N, N, cfl=0.7, precision=precision)
    box.set_state(torch.from_numpy(u0), torch.from_numpy(b0))
    box.run(steps, graph=graph)
    out = box.host_state()
    t = box.time()[0]
    box.close()
    return out, t


@pytest.fixture(scope="module")
def ic():
    from heterogeneous_mhd_amd import solver
    return solver.ic_host(N, N, N, solver.ICSpec(**SPEC))


def test_fast_one_step_within_fp32_noise_of_reference(gpu, ic):
    from oracle import Oracle
    u0, b0 = ic
    print("fast mode, 64^3, 1 step", flush=True)
    ref = Oracle("f32", threads=16).run(u0, b0, N, N, N, 1)
    for graph in (False, True):
        (u, b), t = _gpu_run(u0, b0, 1, "fast", graph)
        assert np.float32(t) == np.float32(ref[2])  # dt comes from the exact calcfl of the IC
        err = rel_errors((u, b), ref[:2])
        print(graph, err, flush=True)
        for k, (linf, _) in err.items():
            assert linf <= 2e-6, (k, linf)


def test_fast_100_steps_tolerance_and_fp64_error(gpu, ic):
    from oracle import Oracle
    u0, b0 = ic
    print("fast mode, 64^3, 100 steps, against FP32 and FP64 oracles", flush=True)
    r32 = Oracle("f32", threads=16).run(u0, b0, N, N, N, 100)
    r64 = Oracle("f64", threads=16).run(u0.astype(np.float64), b0.astype(np.float64), N, N, N, 100)
    (u, b), _ = _gpu_run(u0, b0, 100, "fast")
    e_ref = rel_errors((u, b), r32[:2])
    e_fast64 = rel_errors((u, b), r64[:2])
    e_ref64 = rel_errors(r32[:2], r64[:2])
    print("fast vs f32", e_ref, "\nfast vs f64", e_fast64, "\nf32 vs f64", e_ref64, flush=True)
    for k in e_ref:
        assert e_ref[k][0] <= 1e-2, (k, e_ref[k])
        assert e_ref[k][1] <= 1e-4, (k, e_ref[k])
        assert e_fast64[k][1] <= 3.0 * e_ref64[k][1] + 1e-7, (k, e_fast64[k], e_ref64[k])


def test_exact_mode_after_fast_is_still_bitwise(gpu, ic):
    """Switching a plan back to exact re-captures its graph: the steps are
    bitwise the reference again."""
    import torch
    from heterogeneous_mhd_amd import solver
    from oracle import Oracle
    u0, b0 = ic
    ref = Oracle("f32", threads=16).run(u0, b0, N, N, N, 2)
    box = solver.Box(N, N, N, cfl=0.7, precision="fast")
    box.set_state(torch.from_numpy(u0), torch.from_numpy(b0))
    box.run(1, graph=True)
    box.set_precision("exact")
    box.set_state(torch.from_numpy(u0), torch.from_numpy(b0))
    box.run(2, graph=True)
    u, b = box.host_state()
    box.close()
    assert bits_equal(u, ref[0]) and bits_equal(b, ref[1])


def test_fast_conserves_mass_and_keeps_divb_at_roundoff(gpu):
    from heterogeneous_mhd_amd import solver
    box = solver.Box(N, N, N, cfl=0.7, precision="fast")
    box.init(solver.ICSpec(seed=4, amp_v=0.3, amp_a=0.1))
    d0 = box.diagnostics()
    box.run(20, graph=True)
    d1 = box.diagnostics()
    box.close()
    print(d0, d1, flush=True)
    assert abs(d1["mass"] - d0["mass"]) <= 1e-6 * abs(d0["mass"])
    assert abs(d1["energy"] - d0["energy"]) <= 1e-5 * abs(d0["energy"])
    assert d1["divb_max"] <= 1e-5


def test_fast_ensemble_boxes_match_single_box_fast_runs(gpu):
    import torch
    from heterogeneous_mhd_amd import solver
    sides, specs = solver.ensemble_specs(6, seed=3, sides=(16, 32, 48))
    ens = solver.Ensemble(sides, cfl=0.7, precision="fast")
    ens.init(specs)
    ens.run(3, graph=True)
    for i, (n, sp) in enumerate(zip(sides, specs)):
        box = solver.Box(n, n, n, cfl=0.7, precision="fast")
        box.init(sp)
        box.run(3)
        bu, bb = box.host_state()
        box.close()
        eu, eb = ens.box_state(i)
        assert bits_equal(bu, eu) and bits_equal(bb, eb), f"box {i} side {n}"
    ens.close()
    torch.cuda.synchronize()


def _dt_of(u, b, n, cfl=0.7):
    """dt the reference's driver takes from this state (C/mhd.c:62-63)."""
    from oracle import Oracle
    return np.float32(cfl * float(Oracle("f32", threads=16).calcfl(u, b, n, n, n)))


def _check_fused_cfl(u0, b0, n, steps):
    """Step a fast-mode plan one step at a time: every step's dt, which from
    the second step on comes from the cfl the previous step's last sweep
    fused, equals the reference calcfl of the state it starts from."""
    import torch
    from heterogeneous_mhd_amd import solver
    box = solver.Box(n, n, n, cfl=0.7, precision="fast")
    box.set_state(torch.from_numpy(u0), torch.from_numpy(b0))
    u, b = u0, b0
    for s in range(steps):
        want = _dt_of(u, b, n)
        box.run(1, graph=s > 0)
        got = np.float32(box.time()[1])
        assert got.tobytes() == want.tobytes(), f"step {s}: dt {got!r} vs calcfl {want!r}"
        u, b = box.host_state()
    box.close()


def test_fast_fused_cfl_is_the_exact_calcfl(gpu, ic):
    u0, b0 = ic
    print("fast mode, 64^3: fused cfl against the reference calcfl, 6 steps", flush=True)
    _check_fused_cfl(u0, b0, N, 6)


def test_fast_fused_cfl_exact_on_unscreened_and_near_uniform_states(gpu):
    """Two states that take the other branches of the screened maximum: a
    cell outside the screened domain (huge energy) makes its workgroup
    evaluate every cell exactly; a nearly uniform state puts every cell
    within the screen's keep band."""
    from heterogeneous_mhd_amd import solver
    n = 32
    u, b = solver.ic_host(n, n, n, solver.ICSpec(seed=3, amp_v=0.1, amp_a=0.05))
    u = u.copy()
    u[(((5 * n) + 7) * n + 9) * 5 + 4] = np.float32(3.0e9)
    print("fast mode, 32^3 with a huge cell: fused cfl", flush=True)
    _check_fused_cfl(u, b, n, 3)
    uu, bu = solver.ic_host(n, n, n, solver.ICSpec(seed=5, amp_v=1e-6, amp_a=1e-6))
    print("fast mode, 32^3 nearly uniform: fused cfl", flush=True)
    _check_fused_cfl(uu, bu, n, 3)


def test_fast_to_exact_switch_recomputes_calcfl(gpu, ic):
    """A plan switched from fast to exact without a new state takes its next
    dt from calcfl of the state it holds, and its step is bitwise the
    reference's step from that state."""
    import torch
    from heterogeneous_mhd_amd import solver
    from oracle import Oracle
    u0, b0 = ic
    box = solver.Box(N, N, N, cfl=0.7, precision="fast")
    box.set_state(torch.from_numpy(u0), torch.from_numpy(b0))
    box.run(2, graph=True)
    uf, bf = box.host_state()
    t_fast = np.float32(box.time()[0])
    box.set_precision("exact")
    box.run(1, graph=True)
    u, b = box.host_state()
    t, dt = box.time()
    box.close()
    ref = Oracle("f32", threads=16).run(uf, bf, N, N, N, 1, t=float(t_fast))
    assert np.float32(dt).tobytes() == ref[3][0].tobytes()
    assert np.float32(t).tobytes() == np.float32(ref[2]).tobytes()
    assert bits_equal(u, ref[0]) and bits_equal(b, ref[1])
