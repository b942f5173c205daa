"""Boxes with ny != nz (the next-row extension "non-cubic boxes").

The reference's transpose13 refuses nx != nz (C/transpose13.c:16), so its
driver (C/mhd.c:88, :102) steps only boxes with ny == nz.  The general x<->z
swap keeps the reference's component permutation and maps dims (nx, ny, nz) to
(nz, ny, nx); the step is otherwise unchanged.  There is no reference result
for such boxes, so the oracle's general transpose is pinned by its index
formula and by being an involution, and its general step by agreeing with the
reference step where both apply (ny == nz) and by the scheme's invariants
(conserved mass, div b at round-off).  The GPU is then checked against it
bit for bit.
"""
import numpy as np
import pytest

from conftest import bits_equal

PU13, PB13 = (0, 3, 2, 1, 4), (2, 1, 0)
SPEC = dict(seed=5, amp_v=0.2, amp_a=0.05)


def _state(nx, ny, nz, seed=3):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    return rng.standard_normal(5 * n).astype(np.float32), rng.standard_normal(3 * n).astype(np.float32)


def _closed_form_t13(u, b, nx, ny, nz):
    """out cell (k, j, i) of dims (nz, ny, nx) = in cell (i, j, k), permuted."""
    U = u.reshape(nz, ny, nx, 5)
    B = b.reshape(nz, ny, nx, 3)
    uo = np.transpose(U, (2, 1, 0, 3))[..., list(PU13)]  # indices (i, j, k) -> memory (x'=k fastest)
    bo = np.transpose(B, (2, 1, 0, 3))[..., list(PB13)]
    return np.ascontiguousarray(uo).reshape(-1), np.ascontiguousarray(bo).reshape(-1)


@pytest.mark.parametrize("dims", [(6, 4, 9), (9, 4, 6), (5, 3, 5), (32, 2, 17)])
def test_oracle_transpose13_any_closed_form_and_involution(dims):
    from oracle import Oracle
    nx, ny, nz = dims
    u, b = _state(nx, ny, nz)
    o = Oracle("f32")
    tu, tb = o.transpose13_any(u, b, nx, ny, nz)
    cu, cb = _closed_form_t13(u, b, nx, ny, nz)
    assert bits_equal(tu, cu) and bits_equal(tb, cb)
    bu, bb = o.transpose13_any(tu, tb, nz, ny, nx)  # dims (nz, ny, nx) back to (nx, ny, nz)
    assert bits_equal(bu, u) and bits_equal(bb, b)
    if nx == nz:
        ru, rb = o.transpose13(u, b, nx, ny, nz)
        assert bits_equal(ru, tu) and bits_equal(rb, tb)


def test_oracle_reference_transpose13_still_refuses_nx_ne_nz():
    from oracle import Oracle
    u, b = _state(6, 4, 9)
    with pytest.raises(ValueError):
        Oracle("f32").transpose13(u, b, 6, 4, 9)


def test_oracle_general_step_equals_reference_step_when_ny_eq_nz():
    from heterogeneous_mhd_amd import solver
    from oracle import Oracle
    nx, ny, nz = 24, 16, 16
    u0, b0 = solver.ic_host(nx, ny, nz, solver.ICSpec(**SPEC))
    o = Oracle("f32")
    a = o.run(u0, b0, nx, ny, nz, 3)
    g = o.run(u0, b0, nx, ny, nz, 3, any_shape=True)
    assert bits_equal(a[0], g[0]) and bits_equal(a[1], g[1]) and a[2] == g[2]


def _mass_divb(u, b, nx, ny, nz):
    U = u.reshape(nz, ny, nx, 5).astype(np.float64)
    B = b.reshape(nz, ny, nx, 3).astype(np.float64)
    div = (np.roll(B[..., 0], -1, 2) - B[..., 0]) + (np.roll(B[..., 1], -1, 1) - B[..., 1]) + \
          (np.roll(B[..., 2], -1, 0) - B[..., 2])
    return U[..., 0].sum(), U[..., 4].sum(), np.abs(div).max()


@pytest.mark.parametrize("dims", [(24, 16, 32), (32, 24, 16)])
def test_oracle_general_step_invariants(dims):
    """Mass and energy are conserved to round-off and div b stays at
    round-off on boxes with ny != nz (pins the general step's sweep order and
    transposes: a wrong permutation or dimension breaks these)."""
    from heterogeneous_mhd_amd import solver
    from oracle import Oracle
    nx, ny, nz = dims
    u0, b0 = solver.ic_host(nx, ny, nz, solver.ICSpec(**SPEC))
    m0, e0, d0 = _mass_divb(u0, b0, nx, ny, nz)
    u, b, t, dts = Oracle("f32").run(u0, b0, nx, ny, nz, 4, any_shape=True)
    m1, e1, d1 = _mass_divb(u, b, nx, ny, nz)
    assert d0 < 1e-5 and d1 < 1e-5
    assert abs(m1 - m0) <= 1e-5 * abs(m0)
    assert abs(e1 - e0) <= 1e-5 * abs(e0)
    assert all(np.isfinite(dts)) and t > 0


# ---------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(40, 24, 56), (56, 24, 40), (33, 5, 70), (64, 3, 64)])
def test_gpu_transpose13_any_matches_oracle(gpu, dims):
    import torch
    from heterogeneous_mhd_amd import ops
    from oracle import Oracle
    nx, ny, nz = dims
    print(f"transpose13_any {dims}", flush=True)
    u, b = _state(nx, ny, nz, seed=nx + nz)
    ud, bd = torch.from_numpy(u).to(gpu), torch.from_numpy(b).to(gpu)
    tu, tb = torch.empty_like(ud), torch.empty_like(bd)
    ops.transpose13_any(tu, tb, ud, bd, nx, ny, nz)
    torch.cuda.synchronize()
    ou, ob = Oracle("f32").transpose13_any(u, b, nx, ny, nz)
    assert bits_equal(tu.cpu().numpy(), ou) and bits_equal(tb.cpu().numpy(), ob)


@pytest.mark.gpu
def test_gpu_reference_transpose13_refuses_nx_ne_nz(gpu):
    import torch
    from heterogeneous_mhd_amd import ops
    from heterogeneous_mhd_amd._lib import MHDError
    u = torch.zeros(5 * 6 * 4 * 9, device=gpu)
    b = torch.zeros(3 * 6 * 4 * 9, device=gpu)
    with pytest.raises(MHDError):
        ops.transpose13(u.clone(), b.clone(), u, b, 6, 4, 9)


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(48, 32, 40), (40, 24, 64)])
def test_gpu_plan_steps_noncubic_box_bitwise(gpu, dims):
    """The device plan on a box with ny != nz equals the oracle's general
    step bit for bit (exact mode, graph and eager), and the fast mode stays
    within the stated one-step tolerance of it."""
    import torch
    from heterogeneous_mhd_amd import solver
    from oracle import Oracle
    nx, ny, nz = dims
    u0, b0 = solver.ic_host(nx, ny, nz, solver.ICSpec(**SPEC))
    ref = Oracle("f32", threads=16).run(u0, b0, nx, ny, nz, 5, any_shape=True)
    for graph in (True, False):
        print(f"noncubic plan {dims} graph={graph}", flush=True)
        box = solver.Box(nx, ny, nz, cfl=0.7)
        box.set_state(torch.from_numpy(u0), torch.from_numpy(b0))
        box.run(5, graph=graph)
        u, b = box.host_state()
        t = box.time()[0]
        box.close()
        assert bits_equal(u, ref[0]) and bits_equal(b, ref[1])
        assert np.float32(t) == np.float32(ref[2])
    ref1 = Oracle("f32", threads=16).run(u0, b0, nx, ny, nz, 1, any_shape=True)
    box = solver.Box(nx, ny, nz, cfl=0.7, precision="fast")
    box.set_state(torch.from_numpy(u0), torch.from_numpy(b0))
    box.run(1)
    u, b = box.host_state()
    box.close()
    for got, want in ((u, ref1[0]), (b, ref1[1])):
        assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()
