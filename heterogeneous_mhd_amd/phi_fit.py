"""Fit of the standard normal CDF used by the corrector's GELU epilogue.

The kernels evaluate erfc(x), x >= 0, in the form of Numerical Recipes'
erfcc: erfc(x) = t exp(-x^2 + P(t)) with t = 1/(1 + x/2), t in (0, 1].  P is
our own degree-9 polynomial fit of log(exp(x^2) erfc(x) / t), made here by
least squares on Chebyshev-spaced t with Lawson reweighting toward the minimax
error.  ``python -m heterogeneous_mhd_amd.phi_fit`` prints the coefficients
(kPhiP in csrc/mhd_corrector.hip, lowest degree first) and their error.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erfcx


def target(t):
    """P(t) = log(erfcx(x) / t), x = 2 (1/t - 1)."""
    t = np.asarray(t, np.float64)
    return np.log(erfcx(2.0 * (1.0 / t - 1.0)) / t)


def fit(deg: int = 9, iters: int = 30):
    """Coefficients (lowest degree first) and max |error| in P on [0, 1]."""
    t = 0.5 - 0.5 * np.cos(np.linspace(0.0, np.pi, 20001))
    t = t[t > 1e-6]
    y = target(t)
    w = np.ones_like(t)
    for _ in range(iters):
        c = np.polynomial.polynomial.polyfit(t, y, deg, w=np.sqrt(w))
        e = np.abs(np.polynomial.polynomial.polyval(t, c) - y)
        w = w * (e / e.max() + 1e-3)
        w /= w.mean()
    c = np.polynomial.polynomial.polyfit(t, y, deg, w=np.sqrt(w))
    return c, float(np.abs(np.polynomial.polynomial.polyval(t, c) - y).max())


def erfc_approx(x, c):
    """erfc(x) for x >= 0 from coefficients c (float64 evaluation)."""
    x = np.asarray(x, np.float64)
    t = 1.0 / (1.0 + 0.5 * x)
    return t * np.exp(np.polynomial.polynomial.polyval(t, c) - x * x)


if __name__ == "__main__":
    coef, err = fit()
    print("kPhiP =", ", ".join(f"{v:.9e}f" for v in coef))
    print(f"max |error in P| = {err:.3e}")
