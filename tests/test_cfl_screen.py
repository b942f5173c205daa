"""CPU check of the error bound behind the tolerance mode's exact fused CFL
(cfl_screen in heterogeneous_mhd_amd/csrc/mhd_cfl_common.h).

The kernel screens each cell with an FP32 estimate f of the reference's
per-cell bound e (C/calcfl.c:45-54) and evaluates e exactly only where f is
within 2^-8 of the workgroup's largest f.  That gives the exact maximum if,
for every cell in the screened domain, |f - e| <= 2^-12 e.  This test
emulates f in numpy float32 with the hardware reciprocal and square root
taken 1 ulp off in either direction (their documented accuracy), takes e
from the oracle (bitwise the reference, pinned by the golden vectors), and
checks the bound over physical states, states at the edge of the domain's
no-cancellation test and magnitudes from 2^-20 to 2^20.  It also checks
that NaN, inf, non-positive or out-of-range density and huge values leave
the domain (the kernel then evaluates every cell of the workgroup exactly).
"""
import numpy as np
import pytest

from oracle import Oracle

O32 = Oracle("f32")
F = np.float32
G = F(5.0 / 3.0)


def _fma(a, b, c):
    return (np.float64(a) * np.float64(b) + np.float64(c)).astype(F)


def _ulp_shift(x, k):
    """x moved by k ulps (float32)."""
    return np.nextafter(x, np.where(k > 0, F(np.inf), F(-np.inf))).astype(F) if k else x


def screen(q, bb, rcp_off=0, sqrt_off=0):
    """Emulation of cfl_screen: (f, in_domain)."""
    q = q.astype(F)
    bb = bb.astype(F)
    rho = q[:, 0]
    with np.errstate(all="ignore"):
        rr = _ulp_shift((F(1) / rho).astype(F), rcp_off)
        a1, a2, a3, a4 = (np.abs(q[:, c]) for c in range(1, 5))
        v = (np.maximum(np.maximum(a1, a2), a3) * rr).astype(F)
        msq = _fma(q[:, 1], q[:, 1], _fma(q[:, 2], q[:, 2], (q[:, 3] * q[:, 3]).astype(F)))
        bsq = _fma(bb[:, 0], bb[:, 0], _fma(bb[:, 1], bb[:, 1], (bb[:, 2] * bb[:, 2]).astype(F)))
        ps = _fma(_fma((F(-0.5) * msq).astype(F), rr, q[:, 4]), G - F(1), ((F(1) - F(0.5) * G) * bsq).astype(F))
        p = _fma(F(-0.5), bsq, ps)
        x = np.abs((_fma(F(2), bsq, (G * p).astype(F)) * rr).astype(F))
        w = (F(2.0 / 3.0) * (_fma((F(0.5) * msq).astype(F), rr, a4) + bsq).astype(F)).astype(F)
        S = (_fma(G, w, (F(2) * bsq).astype(F)) * rr).astype(F)
        mag = (((a1 + a2).astype(F) + (a3 + a4).astype(F)).astype(F)
               + ((np.abs(bb[:, 0]) + np.abs(bb[:, 1])).astype(F) + np.abs(bb[:, 2])).astype(F)).astype(F)
        inside = ((rho >= F(2.0 ** -30)) & (rho <= F(2.0 ** 30)) & (mag <= F(2.0 ** 30))
                  & (S >= F(2.0 ** -90)) & (x >= (F(2.0 ** -6) * S).astype(F)))
        s = _ulp_shift(np.sqrt(x).astype(F), sqrt_off)
        f = (v + s).astype(F)
    return f, inside


def exact(q, bb):
    return np.array([O32.cfl_cell(q[i], bb[i]) for i in range(len(q))], dtype=F)


def _physical(rng, n, scale=1.0):
    rho = rng.uniform(0.05, 20.0, n)
    vel = rng.normal(0.0, 1.0, (n, 3)) * rng.uniform(0.0, 3.0, (n, 1))
    B = rng.normal(0.0, 1.0, (n, 3)) * rng.uniform(0.0, 3.0, (n, 1))
    p = rng.uniform(1e-3, 10.0, n)
    m = rho[:, None] * vel
    E = p / (5.0 / 3.0 - 1.0) + 0.5 * np.sum(m * m, axis=1) / rho + 0.5 * np.sum(B * B, axis=1)
    q = np.column_stack([rho, m, E]) * scale
    return q.astype(F), (B * np.sqrt(scale)).astype(F)


def _edge(rng, n):
    """Internal energy small against the kinetic one: |X| near S/8."""
    rho = rng.uniform(0.1, 5.0, n)
    vel = rng.normal(0.0, 1.0, (n, 3))
    m = rho[:, None] * vel
    K = 0.5 * np.sum(m * m, axis=1) / rho
    B = rng.normal(0.0, 0.05, (n, 3)) * rng.uniform(0, 1, (n, 1))
    frac = 10.0 ** rng.uniform(-3, 0, n)       # thermal / kinetic energy
    E = K * (1 + frac) + 0.5 * np.sum(B * B, axis=1)
    return np.column_stack([rho, m, E]).astype(F), B.astype(F)


@pytest.mark.parametrize("kind", ["physical", "edge", "tiny", "huge"])
def test_screen_error_bound(kind):
    rng = np.random.default_rng({"physical": 1, "edge": 2, "tiny": 3, "huge": 4}[kind])
    n = 4000
    if kind == "physical":
        q, bb = _physical(rng, n)
    elif kind == "edge":
        q, bb = _edge(rng, n)
    elif kind == "tiny":
        q, bb = _physical(rng, n, scale=2.0 ** -20)
    else:
        q, bb = _physical(rng, n, scale=2.0 ** 20)
    e = exact(q, bb).astype(np.float64)
    covered = 0
    for ro in (-1, 0, 1):
        for so in (-1, 0, 1):
            f, inside = screen(q, bb, ro, so)
            rel = np.abs(f.astype(np.float64) - e) / e
            assert np.all(rel[inside] <= 2.0 ** -12), (kind, ro, so, rel[inside].max())
            covered = max(covered, int(inside.sum()))
    # the domain is not vacuous: most cells are screened, not sent to the exact path
    assert covered >= (0.5 if kind == "edge" else 0.9) * n


def test_screen_domain_excludes_bad_cells():
    q, bb = _physical(np.random.default_rng(7), 8)
    q[0, 0] = np.nan
    q[1, 2] = np.inf
    q[2, 0] = 0.0
    q[3, 0] = -1.0
    q[4, 4] = F(2.0 ** 31)
    bb[5, 1] = np.nan
    q[6, 0] = F(2.0 ** -31)
    _, inside = screen(q, bb)
    assert not inside[:7].any() and inside[7]


def test_threshold_keeps_every_maximiser():
    """T = F (1 - 2^-8) against the bound: a cell with e = max e has
    f >= E (1 - 2^-12) > F (1 - 2^-8) whenever every f is within 2^-12."""
    E = 1.0
    Fmax = E * (1 + 2.0 ** -12)
    assert E * (1 - 2.0 ** -12) > float(F(Fmax) * F(1.0 - 2.0 ** -8))
