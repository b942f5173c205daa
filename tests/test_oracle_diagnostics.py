"""CPU pins of the two checker functions that have no counterpart in the
reference: the oracle's FP64 diagnostics (conserved sums and div b,
oracle/mhd_ref.c mhdref_diagnostics) and the corrector oracle's features
(cell-centred B, oracle/corrector_ref.py).

Each is checked against what the mathematics fixes, computed independently
of the function under test:
* sums against math.fsum (exactly rounded) of the float64 values;
* div b against a numpy forward-difference divergence, and against b = curl A
  built with forward differences, whose forward divergence vanishes because
  forward differences commute.  The backward-difference divergence of the same
  field does not vanish, so an oracle that took the wrong neighbour fails;
* a planted perturbation of one face or one cell moves the diagnostics by the
  closed-form amount, at the right place;
* the cell-centred B of a field that varies only along its own axis as
  cos(2 pi i / n) is cos(pi / n) cos(2 pi (i + 1/2) / n) (a product-to-sum
  identity); a shift the wrong way or along the wrong axis gives a different
  field.
Dimensions are non-cubic throughout, so a swapped axis fails too.
"""
import math

import numpy as np
import pytest

from oracle import Oracle
from oracle import corrector_ref as R

O32 = Oracle("f32")
O64 = Oracle("f64")
DIMS = [(7, 5, 3), (12, 10, 6), (4, 9, 11)]


def _grid(a, nx, ny, nz, c):
    return np.asarray(a, np.float64).reshape(nz, ny, nx, c)


def _fwd_div(b):
    """[nz, ny, nx, 3] face field -> forward divergence per cell (float64)."""
    return ((np.roll(b[..., 0], -1, axis=2) - b[..., 0]) + (np.roll(b[..., 1], -1, axis=1) - b[..., 1])
            + (np.roll(b[..., 2], -1, axis=0) - b[..., 2]))


def _bwd_div(b):
    return ((b[..., 0] - np.roll(b[..., 0], 1, axis=2)) + (b[..., 1] - np.roll(b[..., 1], 1, axis=1))
            + (b[..., 2] - np.roll(b[..., 2], 1, axis=0)))


def _curl_fwd(A):
    """b = curl A with forward differences, A on cell edges ([nz, ny, nx, 3])."""
    def D(f, ax):
        return np.roll(f, -1, axis=ax) - f
    ax_, ay_, az_ = A[..., 0], A[..., 1], A[..., 2]
    bx = D(az_, 1) - D(ay_, 0)   # d/dy Az - d/dz Ay
    by = D(ax_, 0) - D(az_, 2)   # d/dz Ax - d/dx Az
    bz = D(ay_, 2) - D(ax_, 1)   # d/dx Ay - d/dy Ax
    return np.stack([bx, by, bz], axis=-1)


def _state(nx, ny, nz, seed, dtype):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-2.0, 3.0, size=(nz, ny, nx, 5)).astype(dtype)
    b = rng.normal(size=(nz, ny, nx, 3)).astype(dtype)
    return u.reshape(-1), b.reshape(-1)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("O", [O32, O64], ids=["f32", "f64"])
def test_sums_match_exactly_rounded_fsum(O, dims):
    nx, ny, nz = dims
    u, b = _state(nx, ny, nz, 3 + nx, O.dtype)
    d = O.diagnostics(u, b, nx, ny, nz)
    ug = _grid(u, nx, ny, nz, 5)
    want = [math.fsum(ug[..., c].ravel()) for c in range(5)]
    got = [d["mass"], *d["mom"], d["energy"]]
    scale = [math.fsum(np.abs(ug[..., c]).ravel()) for c in range(5)]
    for c in range(5):
        # long double accumulation of N values: |error| <= N * 2^-64 * sum|x|
        assert abs(got[c] - want[c]) <= 1e-15 * scale[c] + 1e-300, (c, got[c], want[c])
    # planted: +1 on one cell's energy moves only the energy sum, by 1
    u2 = u.copy()
    u2[(((nz - 1) * ny + 1) * nx + 2) * 5 + 4] += O.dtype(1.0)
    d2 = O.diagnostics(u2, b, nx, ny, nz)
    assert d2["mass"] == d["mass"] and d2["mom"] == d["mom"]
    e_before = float(u[(((nz - 1) * ny + 1) * nx + 2) * 5 + 4])
    e_after = float(u2[(((nz - 1) * ny + 1) * nx + 2) * 5 + 4])
    assert abs((d2["energy"] - d["energy"]) - (e_after - e_before)) <= 1e-15 * scale[4]


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("O", [O32, O64], ids=["f32", "f64"])
def test_divb_matches_forward_difference_divergence(O, dims):
    nx, ny, nz = dims
    u, b = _state(nx, ny, nz, 11 + ny, O.dtype)
    d = O.diagnostics(u, b, nx, ny, nz)
    div = _fwd_div(_grid(b, nx, ny, nz, 3))
    assert d["divb_max"] == pytest.approx(np.abs(div).max(), rel=1e-14)
    assert d["divb_rms"] == pytest.approx(np.sqrt(np.mean(div * div)), rel=1e-14)
    # the backward divergence of a random field has another maximum: the
    # comparison above distinguishes the two neighbour conventions
    assert abs(np.abs(_bwd_div(_grid(b, nx, ny, nz, 3))).max() - np.abs(div).max()) > 1e-6


@pytest.mark.parametrize("dims", DIMS)
def test_divb_of_forward_curl_vanishes_and_planted_face_shows(dims):
    nx, ny, nz = dims
    rng = np.random.default_rng(5)
    A = rng.normal(size=(nz, ny, nx, 3))
    bg = _curl_fwd(A)
    b = bg.reshape(-1).copy()
    u = np.ones(5 * nx * ny * nz)
    d = O64.diagnostics(u, b, nx, ny, nz)
    assert d["divb_max"] < 1e-13                      # commuting forward differences
    assert np.abs(_bwd_div(bg)).max() > 1e-2          # the wrong neighbour would not vanish
    # planted: delta on one x face moves div at the two cells that share it
    i0, j0, k0, delta = 2, 1, nz - 1, 0.25
    b2 = b.copy()
    b2[((k0 * ny + j0) * nx + i0) * 3 + 0] += delta
    d2 = O64.diagnostics(u, b2, nx, ny, nz)
    n = nx * ny * nz
    assert d2["divb_max"] == pytest.approx(delta, abs=1e-13)
    assert d2["divb_rms"] == pytest.approx(delta * math.sqrt(2.0 / n), rel=1e-10)
    # and on a z face (the third axis), same closed form
    b3 = b.copy()
    b3[((0 * ny + j0) * nx + i0) * 3 + 2] -= delta
    d3 = O64.diagnostics(u, b3, nx, ny, nz)
    assert d3["divb_max"] == pytest.approx(delta, abs=1e-13)
    assert d3["divb_rms"] == pytest.approx(delta * math.sqrt(2.0 / n), rel=1e-10)


def test_divb_in_f32_is_roundoff_of_the_f64_curl():
    nx, ny, nz = 12, 10, 6
    A = np.random.default_rng(9).normal(size=(nz, ny, nx, 3))
    b = _curl_fwd(A).astype(np.float32).reshape(-1)
    u = np.ones(5 * nx * ny * nz, np.float32)
    d = O32.diagnostics(u, b, nx, ny, nz)
    div = _fwd_div(_grid(b, nx, ny, nz, 3))      # exact in float64 from float32 faces
    assert d["divb_max"] == pytest.approx(np.abs(div).max(), rel=1e-14)
    assert 0 < d["divb_max"] < 1e-5


@pytest.mark.parametrize("dims", DIMS)
def test_features_centre_average_closed_form(dims):
    nx, ny, nz = dims
    rng = np.random.default_rng(1)
    u = rng.normal(size=(nz, ny, nx, 5))
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    b = np.stack([np.cos(2 * np.pi * ii / nx), np.cos(2 * np.pi * jj / ny),
                  np.cos(2 * np.pi * kk / nz)], axis=-1)
    f = R.features(u, b)
    assert f.shape == (nz, ny, nx, 8)
    assert np.array_equal(f[..., :5], u)
    want = np.stack([np.cos(np.pi / nx) * np.cos(2 * np.pi * (ii + 0.5) / nx),
                     np.cos(np.pi / ny) * np.cos(2 * np.pi * (jj + 0.5) / ny),
                     np.cos(np.pi / nz) * np.cos(2 * np.pi * (kk + 0.5) / nz)], axis=-1)
    assert np.abs(f[..., 5:] - want).max() < 1e-14
    # the low-face average, (i - 1/2), and the average along another axis both
    # differ from it by O(1 / n)
    wrong_dir = np.cos(np.pi / nx) * np.cos(2 * np.pi * (ii - 0.5) / nx)
    assert np.abs(wrong_dir - want[..., 0]).max() > 0.1
    assert np.abs(b[..., 0] - want[..., 0]).max() > 0.05   # averaging along y or z leaves b_x as is


def test_features_uniform_field_and_pack_layout():
    nx, ny, nz = 5, 4, 3
    u = np.random.default_rng(2).normal(size=(nz, ny, nx, 5))
    b = np.broadcast_to(np.array([0.3, -0.7, 1.1]), (nz, ny, nx, 3)).copy()
    f = R.features(u, b)
    assert np.abs(f[..., 5:] - np.array([0.3, -0.7, 1.1])).max() < 1e-15
    mean, inv_std = np.zeros(8), np.ones(8)
    x = R.pack(u, b, mean, inv_std, bf=False)
    assert x.shape == (nz, ny, nx, R.C)
    assert np.array_equal(x[..., :8], f) and not x[..., 8:].any()
