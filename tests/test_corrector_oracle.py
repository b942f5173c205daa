"""Pins the CPU restatement of the learned corrector (oracle/corrector_ref.py)
against a library routine, adjoint identities, finite differences and closed
forms, so that the GPU kernels can be checked against it."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import corrector_ref as R

RNG = np.random.default_rng(1234)


def rand(*shape):
    return RNG.standard_normal(shape)


def test_conv_matches_scipy_correlate_wrap():
    x, w, bias = rand(5, 6, 7, 3), rand(27, 4, 3), rand(4)
    out = R.conv(x, w, bias)
    for o in range(4):
        ref = np.full(x.shape[:3], bias[o])
        for c in range(3):
            k = w[:, o, c].reshape(3, 3, 3)  # [dz+1][dy+1][dx+1]
            ref += ndimage.correlate(x[..., c], k, mode="wrap")
        np.testing.assert_allclose(out[..., o], ref, rtol=1e-12, atol=1e-12)


def test_conv_gradients_are_adjoints():
    """<conv(x,w), dy> = <x, conv_input_grad(dy, w)> = <w, conv_weight_grad(dy, x)>."""
    x, w, dy = rand(4, 5, 6, 3), rand(27, 2, 3), rand(4, 5, 6, 2)
    lhs = np.sum(R.conv(x, w) * dy)
    assert np.isclose(lhs, np.sum(x * R.conv_input_grad(dy, w)), rtol=1e-12)
    dw, db = R.conv_weight_grad(dy, x)
    assert np.isclose(lhs, np.sum(w * dw), rtol=1e-12)
    np.testing.assert_allclose(db, dy.reshape(-1, 2).sum(axis=0))


def test_a_planted_tap_mistake_fails_the_adjoint_check():
    """A transposed offset in the input gradient (a plausible slip) breaks it."""
    x, w, dy = rand(4, 5, 6, 3), rand(27, 2, 3), rand(4, 5, 6, 2)
    wrong = np.zeros(x.shape)
    for t, (dz, dyy, dx) in enumerate(R.TAPS):
        wrong += R._shift(dy, dz, dyy, dx) @ w[t]  # sign of the offset not flipped
    assert not np.isclose(np.sum(R.conv(x, w) * dy), np.sum(x * wrong), rtol=1e-6)


def test_gelu_grad_is_the_derivative():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (R.gelu(z + h) - R.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(R.gelu_grad(z), fd, rtol=1e-7, atol=1e-9)
    assert R.gelu(0.0) == 0.0 and np.isclose(R.gelu(10.0), 10.0)
    from scipy.special import erf
    np.testing.assert_allclose(R.gelu(z), 0.5 * z * (1 + erf(z / np.sqrt(2))), rtol=1e-12, atol=1e-15)


def test_curl_is_div_free_and_adjoint_is_its_transpose():
    a, r = rand(6, 5, 4, 3), rand(6, 5, 4, 3)
    assert np.abs(R.div_faces(R.curl(a))).max() < 1e-12
    assert np.isclose(np.sum(R.curl(a) * r), np.sum(a * R.curl_adjoint(r)), rtol=1e-12)


def test_downsample_block_means_and_div_free():
    nz, ny, nx, r = 8, 12, 16, 4
    u = rand(nz, ny, nx, 5)
    b = R.curl(rand(nz, ny, nx, 3))  # div-free fine field
    uc, bc = R.downsample(u, b, r)
    assert uc.shape == (2, 3, 4, 5) and bc.shape == (2, 3, 4, 3)
    np.testing.assert_allclose(uc[1, 2, 3], u[4:8, 8:12, 12:16].mean(axis=(0, 1, 2)))
    np.testing.assert_allclose(bc[1, 2, 3, 0], b[4:8, 8:12, 12, 0].mean())
    np.testing.assert_allclose(bc[1, 2, 3, 1], b[4:8, 8, 12:16, 1].mean())
    np.testing.assert_allclose(bc[1, 2, 3, 2], b[4, 8:12, 12:16, 2].mean())
    np.testing.assert_allclose(uc.sum(axis=(0, 1, 2)) * r ** 3, u.sum(axis=(0, 1, 2)))
    assert np.abs(R.div_faces(bc)).max() < 1e-12


def test_network_gradient_matches_finite_differences():
    """Whole loss (4 conv layers, GELU, curl) in FP64, central differences."""
    shape = (4, 4, 6)
    params = [(0.2 * rand(27, 32, 32), 0.1 * rand(32)) for _ in range(4)]
    net = R.Net(params)
    x0 = np.zeros(shape + (32,))
    x0[..., :8] = rand(*shape, 8)
    tgt = rand(*shape, 8)

    def loss_of(n):
        y, _ = n.forward(x0, bf=False)
        return R.loss_and_grad(y, tgt)[0]

    y, cache = net.forward(x0, bf=False)
    _, g = R.loss_and_grad(y, tgt)
    grads = net.backward(g, cache, bf=False)
    h = 1e-6
    for li, idx in [(0, (13, 2, 5)), (1, (0, 31, 7)), (2, (26, 4, 4)), (3, (13, 6, 9)), (3, (5, 1, 0))]:
        w = net.params[li][0]
        old = w[idx]
        w[idx] = old + h
        lp = loss_of(net)
        w[idx] = old - h
        lm = loss_of(net)
        w[idx] = old
        fd = (lp - lm) / (2 * h)
        assert np.isclose(grads[li][0][idx], fd, rtol=1e-5, atol=1e-9), (li, idx, grads[li][0][idx], fd)
    for li in range(4):
        bb = net.params[li][1]
        old = bb[3]
        bb[3] = old + h
        lp = loss_of(net)
        bb[3] = old - h
        lm = loss_of(net)
        bb[3] = old
        assert np.isclose(grads[li][1][3], (lp - lm) / (2 * h), rtol=1e-5, atol=1e-9)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0], np.float32)
    np.testing.assert_array_equal(R.bf16(x), [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 0.0])


def test_apply_correction_checks():
    shape = (4, 4, 4)
    u = np.zeros(shape + (5,))
    u[..., 0] = 1.0
    u[..., 4] = 1.0
    b = np.zeros(shape + (3,))
    y = np.zeros(shape + (32,))
    scale = np.ones(6)
    y[..., 0] = 0.01  # a uniform correction is removed by the zero-total projection
    uu, _, ok = R.apply_correction(u, b, y, scale, etol=1e-3)
    assert ok and np.array_equal(uu, u)
    y[..., 0] = 0.01 * RNG.standard_normal(shape)
    uu, _, ok = R.apply_correction(u, b, y, scale, etol=1e-3)
    assert ok and abs(uu[..., 0].sum() - u[..., 0].sum()) < 1e-12
    y[0, 0, 0, 0] = -2.0  # rho < 0 in one cell
    uu, _, ok = R.apply_correction(u, b, y, scale, etol=1e-3)
    assert not ok and uu is u
    y[..., 0] = 0.0
    y[1, 2, 3, 4] = -5.0  # p < 0 in one cell
    assert not R.apply_correction(u, b, y, scale, etol=1e-3)[2]
    # a momentum correction that turns 0.5% of the energy into kinetic energy:
    # positivity holds, the total of E does not move, the energy check refuses it
    y[...] = 0.0
    y[..., 1] = 0.1 * np.where(np.arange(4) % 2 == 0, 1.0, -1.0)[None, None, :]
    uu, _, ok = R.apply_correction(u, b, y, scale, etol=1e-3)
    kin = 0.5 * 0.1 ** 2  # per cell, rho = 1, E = 1
    assert not ok and uu is u
    assert R.apply_correction(u, b, y, scale, etol=1.1 * kin)[2]
    assert not R.apply_correction(u, b, y, scale, etol=0.9 * kin)[2]
    # a field correction converts magnetic energy likewise
    y[...] = 0.0
    y[..., 7] = 0.2 * np.sin(2 * np.pi * np.arange(4) / 4)[None, :, None]
    _, bb, ok = R.apply_correction(u, b, y, scale, etol=1e-4)
    assert not ok and not bb.any()


def _random_state(nz=6, ny=5, nx=8):
    u = RNG.standard_normal((nz, ny, nx, 5)).astype(np.float32)
    u[..., 0] += 4.0
    b = RNG.standard_normal((nz, ny, nx, 3)).astype(np.float32)
    mean = RNG.standard_normal(8).astype(np.float32) * 0.1
    inv = (1.0 + RNG.random(8)).astype(np.float32)
    return u, b, mean, inv


def _normalised(u, b, mean, inv):
    # an independent FP64 statement of the normalised inputs
    bc = 0.5 * (b + np.stack([np.roll(b[..., 0], -1, 2), np.roll(b[..., 1], -1, 1),
                              np.roll(b[..., 2], -1, 0)], axis=-1))
    return (np.concatenate([u, bc], axis=-1).astype(np.float64) - mean) * inv


def test_pack_features_derived_channels_match_scipy():
    """Laplacian and per-axis second differences against scipy.ndimage
    (periodic), the low parts against the closed form hi + lo/256 = f, to the
    bf16 rounding of the stored values."""
    u, b, mean, inv = _random_state()
    f = _normalised(u, b, mean, inv)
    dn = (0.5 + RNG.random(24)).astype(np.float32)
    x0 = R.pack_features(u, b, mean, inv, dn, 0)
    np.testing.assert_allclose(x0[..., :8], f, rtol=2 ** -8, atol=1e-6)
    assert not x0[..., 8:].any()
    x1 = R.pack_features(u, b, mean, inv, dn, 1)
    np.testing.assert_array_equal(x1[..., :8], x0[..., :8])
    for c in range(8):
        lap = ndimage.laplace(f[..., c], mode="wrap") * dn[c]
        np.testing.assert_allclose(x1[..., 8 + c], lap, rtol=2 ** -7, atol=1e-5)
    x2 = R.pack_features(u, b, mean, inv, dn, 2)
    for c in range(8):
        for d, axis in enumerate((2, 1, 0)):  # x, y, z
            sd = ndimage.correlate1d(f[..., c], [1.0, -2.0, 1.0], axis=axis, mode="wrap") * dn[3 * c + d]
            np.testing.assert_allclose(x2[..., 8 + 3 * c + d], sd, rtol=2 ** -7, atol=1e-5)
    x3 = R.pack_features(u, b, mean, inv, dn, 3)
    rec = x3[..., :8] + x3[..., 8:16] / 256.0
    np.testing.assert_allclose(rec, f, rtol=2 ** -15, atol=1e-9)
    # |f - bf16(f)| <= half a bf16 ulp <= 2^-8 |f|, so the scaled low part stays within |f|
    assert np.all(np.abs(x3[..., 8:16]) <= np.abs(f) * (1 + 1e-6))


def test_pack_features_axis_order():
    """A field that varies along x only has zero y and z second differences:
    a swapped axis (a plausible slip) fails here."""
    nz, ny, nx = 4, 4, 8
    u = np.zeros((nz, ny, nx, 5), np.float32)
    u[..., 0] = 1.0 + 0.5 * np.sin(2 * np.pi * np.arange(nx) / nx)[None, None, :]
    b = np.zeros((nz, ny, nx, 3), np.float32)
    x2 = R.pack_features(u, b, np.zeros(8, np.float32), np.ones(8, np.float32),
                         np.ones(24, np.float32), 2)
    assert np.abs(x2[..., 8 + 0]).max() > 0.1        # rho along x
    assert not x2[..., 8 + 1].any() and not x2[..., 8 + 2].any()
