"""CPU checks of the seeded initial condition (library host form).  Its
device form is bitwise equal (tests/test_gpu_step.py)."""
import numpy as np
import pytest

from heterogeneous_mhd_amd import solver


def div_b(b, nx, ny, nz):
    B = b.astype(np.float64).reshape(nz, ny, nx, 3)
    return ((np.roll(B[..., 0], -1, 2) - B[..., 0]) + (np.roll(B[..., 1], -1, 1) - B[..., 1])
            + (np.roll(B[..., 2], -1, 0) - B[..., 2]))


@pytest.mark.parametrize("dims", [(32, 32, 32), (40, 24, 24)])
def test_ic_closed_forms(dims):
    nx, ny, nz = dims
    spec = solver.ICSpec(seed=3, amp_v=0.1, amp_a=0.05)
    u, b = solver.ic_host(nx, ny, nz, spec)
    U = u.reshape(-1, 5).astype(np.float64)
    B = b.reshape(-1, 3).astype(np.float64)
    assert np.all(U[:, 0] == 1.0)
    # per-component RMS of v is amp_v (discrete orthogonality of distinct modes)
    rms = np.sqrt((U[:, 1:4] ** 2).mean(0))
    assert np.allclose(rms, spec.amp_v, rtol=1e-5)
    # mean of every Fourier field is zero
    assert np.abs(U[:, 1:4].mean(0)).max() < 1e-7
    # b = discrete curl of A: div b vanishes to FP32 round-off of b
    assert np.abs(div_b(b, nx, ny, nz)).max() < 4 * np.finfo(np.float32).eps * np.abs(B).max()
    assert np.abs(B).max() > 0.05
    # E makes the fluid sweep's pressure equal p0
    bc = b.reshape(nz, ny, nx, 3)
    h = np.float32(0.5)
    c1 = h * bc[..., 0] + h * np.roll(bc[..., 0], -1, 2)
    c2 = h * bc[..., 1] + np.roll(bc[..., 1], -1, 1) / np.float32(2)
    c3 = h * bc[..., 2] + np.roll(bc[..., 2], -1, 0) / np.float32(2)
    bsq = (c1.astype(np.float64) ** 2 + c2.astype(np.float64) ** 2 + c3.astype(np.float64) ** 2).reshape(-1)
    p = (U[:, 4] - 0.5 * (U[:, 1:4] ** 2).sum(1) / U[:, 0] - 0.5 * bsq) * (spec.gamma - 1)
    assert np.allclose(p, spec.p0, rtol=1e-6)


def test_ic_b_amplitude_independent_of_resolution():
    spec = solver.ICSpec(seed=5, amp_v=0.1, amp_a=0.05)
    r = []
    for n in (16, 32, 64):
        _, b = solver.ic_host(n, n, n, spec)
        r.append(np.sqrt((b.astype(np.float64) ** 2).mean()))
    assert abs(r[1] / r[2] - 1) < 0.02 and abs(r[0] / r[2] - 1) < 0.1


def test_ic_is_deterministic_and_seeded():
    a = solver.ic_host(16, 12, 12, solver.ICSpec(seed=1))
    b = solver.ic_host(16, 12, 12, solver.ICSpec(seed=1))
    c = solver.ic_host(16, 12, 12, solver.ICSpec(seed=2))
    assert a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()
    assert a[0].tobytes() != c[0].tobytes()


def test_ic_same_physical_field_across_resolutions():
    """Modes are in box units: the 64^3 field block-averaged 2x matches the
    32^3 field sampled at the matching positions up to discretisation error."""
    spec = solver.ICSpec(seed=9, amp_v=0.1, amp_a=0.05)
    u32, _ = solver.ic_host(32, 32, 32, spec)
    u64, _ = solver.ic_host(64, 64, 64, spec)
    v32 = u32.reshape(32, 32, 32, 5)[..., 1]
    v64 = u64.reshape(64, 64, 64, 5)[..., 1][::2, ::2, ::2]   # same cell centres (i*2 -> i)
    assert np.abs(v64 - v32).max() < 1e-6
