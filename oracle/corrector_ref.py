"""CPU restatement (numpy, FP64) of the learned coarse corrector -- test
infrastructure only: the product path never imports it.

It mirrors heterogeneous_mhd_amd/csrc/mhd_corrector.hip: periodic 3x3x3
convolutions over [nz][ny][nx][C] arrays (tap t = (dz+1)*9 + (dy+1)*3 +
(dx+1), weights [27][out][in]), GELU (erf form), the vector-potential curl
(b = curl A with A on cell edges, forward differences, the convention of
the seeded initial condition) and its adjoint, the loss and its gradient,
the block-average downsampling and the fallback checks.  With bf16=True the
functions round to bf16 exactly where the GPU kernels store bf16.

The corrector has no counterpart in the reference, so there is no reference
code to follow; each function is pinned by tests/test_corrector_oracle.py
against a library routine (scipy.ndimage.correlate for the convolution),
adjoint (dot-product) identities, finite differences and closed forms
(div curl = 0, block means).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erfc

C = 32
TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]  # (dz, dy, dx)


def bf16(x):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(a)
    out[nan] = np.nan
    return out


def _shift(x, dz, dy, dx):
    """y[z, y, x] = x[z+dz, y+dy, x+dx] (periodic) over the first three axes."""
    return np.roll(x, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def conv(x, w, bias=None):
    """out[p, o] = sum_t sum_c x[p + off_t, c] * w[t, o, c] (+ bias[o])."""
    out = np.zeros(x.shape[:3] + (w.shape[1],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(x, dz, dy, dx) @ w[t].T
    if bias is not None:
        out += bias
    return out


def conv_input_grad(dy, w):
    """dx[p, c] = sum_t sum_o dy[p - off_t, o] * w[t, o, c]."""
    out = np.zeros(dy.shape[:3] + (w.shape[2],))
    for t, (dz, dyy, dx) in enumerate(TAPS):
        out += _shift(dy, -dz, -dyy, -dx) @ w[t]
    return out


def conv_weight_grad(dy, x):
    """(dW[t, o, c] = sum_p dy[p, o] * x[p + off_t, c], db[o] = sum_p dy[p, o])."""
    d2 = dy.reshape(-1, dy.shape[-1])
    dw = np.stack([d2.T @ _shift(x, dz, dyy, dx).reshape(-1, x.shape[-1]) for dz, dyy, dx in TAPS])
    return dw, d2.sum(axis=0)


def gelu(z):
    """0.5 z (1 + erf(z/sqrt2)), written with erfc (no cancellation for z < 0)."""
    return 0.5 * z * erfc(-z / np.sqrt(2.0))


def gelu_grad(z):
    return 0.5 * erfc(-z / np.sqrt(2.0)) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def curl(a):
    """b = curl A, A [nz,ny,nx,3] on edges, b [nz,ny,nx,3] on low faces."""
    ax, ay, az = a[..., 0], a[..., 1], a[..., 2]
    sh = lambda f, dz, dy, dx: _shift(f, dz, dy, dx)  # noqa: E731
    bx = (sh(az, 0, 1, 0) - az) - (sh(ay, 1, 0, 0) - ay)
    by = (sh(ax, 1, 0, 0) - ax) - (sh(az, 0, 0, 1) - az)
    bz = (sh(ay, 0, 0, 1) - ay) - (sh(ax, 0, 1, 0) - ax)
    return np.stack([bx, by, bz], axis=-1)


def curl_adjoint(r):
    """The transpose of curl: <curl a, r> = <a, curl_adjoint(r)>."""
    rx, ry, rz = r[..., 0], r[..., 1], r[..., 2]
    sh = lambda f, dz, dy, dx: _shift(f, dz, dy, dx)  # noqa: E731
    gx = (sh(ry, -1, 0, 0) - ry) - (sh(rz, 0, -1, 0) - rz)
    gy = (rx - sh(rx, -1, 0, 0)) + (sh(rz, 0, 0, -1) - rz)
    gz = (sh(rx, 0, -1, 0) - rx) - (sh(ry, 0, 0, -1) - ry)
    return np.stack([gx, gy, gz], axis=-1)


def div_faces(b):
    """Discrete div of a low-face field [nz,ny,nx,3] at cell centres."""
    return ((_shift(b[..., 0], 0, 0, 1) - b[..., 0]) + (_shift(b[..., 1], 0, 1, 0) - b[..., 1])
            + (_shift(b[..., 2], 1, 0, 0) - b[..., 2]))


def grid(u, b, nx, ny, nz):
    """Flat AoS state -> ([nz,ny,nx,5], [nz,ny,nx,3]) float64."""
    return (np.asarray(u, np.float64).reshape(nz, ny, nx, 5),
            np.asarray(b, np.float64).reshape(nz, ny, nx, 3))


def features(u, b):
    """rho, m1..m3, E and the cell-centred B of the fluid sweep."""
    bc = 0.5 * (b + np.stack([_shift(b[..., 0], 0, 0, 1), _shift(b[..., 1], 0, 1, 0),
                              _shift(b[..., 2], 1, 0, 0)], axis=-1))
    return np.concatenate([u, bc], axis=-1)


def pack(u, b, mean, inv_std, bf=True):
    f = (features(u, b) - mean) * inv_std
    x = np.zeros(f.shape[:3] + (C,))
    x[..., :8] = bf16(f) if bf else f
    return x


def pack_features(u, b, mean, inv_std, dnorm, kind):
    """The network input of mhd_corr_pack_features, in FP32 in the kernel's
    order: channels 0..7 bf16 of the normalised fields f; kind 1 adds the
    Laplacian of f (channel 8+c), kind 2 its second differences along x, y
    and z (channel 8+3c+d), each times dnorm; kind 3 adds (f - bf16(f)) * 256.
    Returns float64 values of the bf16 channels [nz,ny,nx,32]."""
    f32 = np.float32
    u32, b32 = np.asarray(u, f32), np.asarray(b, f32)
    bc = f32(0.5) * (b32 + np.stack([_shift(b32[..., 0], 0, 0, 1), _shift(b32[..., 1], 0, 1, 0),
                                     _shift(b32[..., 2], 1, 0, 0)], axis=-1))
    f = (np.concatenate([u32, bc], axis=-1) - np.asarray(mean, f32)) * np.asarray(inv_std, f32)
    dn = np.asarray(dnorm, f32)
    x = np.zeros(f.shape[:3] + (C,))
    x[..., :8] = bf16(f)
    if kind == 3:
        x[..., 8:16] = bf16((f - bf16(f).astype(f32)) * f32(256.0))
    elif kind in (1, 2):
        steps = ((0, 0, 1), (0, 1, 0), (1, 0, 0))  # +e_x, +e_y, +e_z as (dz, dy, dx)
        sd = [(_shift(f, -a, -bq, -c) + _shift(f, a, bq, c)) - f32(2.0) * f for a, bq, c in steps]
        if kind == 1:
            x[..., 8:16] = bf16((((f32(0.0) + sd[0]) + sd[1]) + sd[2]) * dn[:8])  # the kernel's sum
        else:
            for c in range(8):
                for d in range(3):
                    x[..., 8 + 3 * c + d] = bf16(sd[d][..., c] * dn[3 * c + d])
    return x


def downsample(u, b, r):
    """Block means by r: u over r^3 cells; b_d over the r^2 fine faces on the
    coarse cell's low d-face."""
    nz, ny, nx = u.shape[:3]
    cz, cy, cx = nz // r, ny // r, nx // r
    uc = u.reshape(cz, r, cy, r, cx, r, 5).mean(axis=(1, 3, 5))
    bx = b[:, :, ::r, 0].reshape(cz, r, cy, r, cx).mean(axis=(1, 3))
    by = b[:, ::r, :, 1].reshape(cz, r, cy, cx, r).mean(axis=(1, 4))
    bz = b[::r, :, :, 2].reshape(cz, cy, r, cx, r).mean(axis=(2, 4))
    return uc, np.stack([bx, by, bz], axis=-1)


class Net:
    """Parameters: list of 4 (w [27,32,32], bias [32]) in FP64."""

    def __init__(self, params):
        self.params = [(np.asarray(w, np.float64), np.asarray(bb, np.float64)) for w, bb in params]

    def forward(self, x0, bf=True):
        """Returns y (output layer, FP64 [.., 32]) and the cache for backward."""
        rw = (lambda a: bf16(a)) if bf else (lambda a: a)
        a, cache = x0, []
        for li, (w, bb) in enumerate(self.params):
            z = conv(a, rw(w), bb)
            cache.append((a, z))
            if li < len(self.params) - 1:
                a = rw(gelu(z))
        return z, cache

    def backward(self, dy, cache, bf=True):
        """Gradients [(dW, db)] for dy = dLoss/dy."""
        rw = (lambda a: bf16(a)) if bf else (lambda a: a)
        grads = [None] * len(self.params)
        d = rw(dy)
        for li in range(len(self.params) - 1, -1, -1):
            a_in, _ = cache[li]
            grads[li] = conv_weight_grad(d, a_in)
            if li > 0:
                w = rw(self.params[li][0])
                d = rw(conv_input_grad(d, w) * gelu_grad(cache[li - 1][1]))
        return grads


def loss_and_grad(y, target):
    """Loss = mean over voxels of |y_u - t_u|^2 + |curl(y_A) - t_B|^2 and
    dLoss/dy (zero on the padding channels)."""
    n = y.shape[0] * y.shape[1] * y.shape[2]
    ru = y[..., :5] - target[..., :5]
    rb = curl(y[..., 5:8]) - target[..., 5:8]
    loss = (np.sum(ru * ru) + np.sum(rb * rb)) / n
    g = np.zeros_like(y)
    g[..., :5] = 2.0 * ru / n
    g[..., 5:8] = curl_adjoint(2.0 * rb / n)
    return loss, g


def kinetic_magnetic(u, b):
    """Per cell: m^2 / (2 rho) + |B_c|^2 / 2 with the cell-centred B."""
    f = features(u, b)
    return 0.5 * np.sum(f[..., 1:4] ** 2, axis=-1) / f[..., 0] + 0.5 * np.sum(f[..., 5:8] ** 2, axis=-1)


def apply_correction(u, b, y, scale, etol, limit=0.0, gamma=5.0 / 3.0):
    """(u', b', accepted): the corrected state if it passes the positivity
    and energy checks, else the input state.  Each output value is limited to
    [-limit, limit] (limit > 0); the u part of the correction is projected to
    zero box total (the scheme conserves the totals of u), so the total of E
    does not move and the energy check bounds the kinetic plus magnetic
    energy the correction converts to or from heat: |delta sum(K + M)| <=
    etol |sum E|."""
    if limit > 0:
        y = np.clip(y, -limit, limit)
    yu = y[..., :5] - y[..., :5].mean(axis=(0, 1, 2))
    uc = u + scale[:5] * yu
    bcand = b + scale[5] * curl(y[..., 5:8])
    km1 = kinetic_magnetic(uc, bcand)
    rho = uc[..., 0]
    p = (gamma - 1.0) * (uc[..., 4] - km1)
    e0, e1 = kinetic_magnetic(u, b).sum(), km1.sum()
    ok = bool(np.all(rho > 0) and np.all(p > 0) and abs(e1 - e0) <= etol * abs(u[..., 4].sum()))
    return (uc, bcand, True) if ok else (u, b, False)
