"""The fitted normal CDF of the corrector's GELU epilogue (CPU only).

The coefficients are read from the kernel source itself, so a mistyped or
reordered constant fails here: erfc to FP32 accuracy over the whole range
the activations use, and GELU, GELU' against scipy.
"""
import re
from pathlib import Path

import numpy as np
from scipy.special import erfc

from heterogeneous_mhd_amd.phi_fit import erfc_approx, fit
from oracle import corrector_ref as R

SRC = Path(__file__).resolve().parents[1] / "heterogeneous_mhd_amd" / "csrc" / "mhd_corrector.hip"


def kernel_coefficients():
    m = re.search(r"constexpr float kPhiP\[10\] = \{([^}]*)\}", SRC.read_text())
    assert m, "kPhiP not found in the kernel source"
    return np.array([float(v.strip().rstrip("f")) for v in m.group(1).split(",")])


def phi(z, c):
    z = np.asarray(z, np.float64)
    half = 0.5 * erfc_approx(np.abs(z) / np.sqrt(2.0), c)
    return np.where(z < 0, half, 1.0 - half)


def test_erfc_to_fp32_accuracy():
    c = kernel_coefficients()
    x = np.linspace(0.0, 9.0, 200001)  # erfc(9) ~ 4e-37, the FP32 range
    rel = np.abs(erfc_approx(x, c) - erfc(x)) / erfc(x)
    assert rel.max() < 2.5e-7, rel.max()


def test_gelu_and_its_derivative():
    c = kernel_coefficients()
    z = np.linspace(-12.0, 12.0, 240001)
    g = z * phi(z, c)
    ref = R.gelu(z)
    nz = np.abs(ref) > 1e-300
    assert np.max(np.abs(g - ref)[nz] / np.abs(ref)[nz]) < 2.5e-7
    gg = phi(z, c) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)
    assert np.max(np.abs(gg - R.gelu_grad(z))) < 2.5e-7


def test_a_wrong_coefficient_fails():
    c = kernel_coefficients()
    for q in (0, 4, 9):
        bad = c.copy()
        bad[q] *= 1.001
        x = np.linspace(0.0, 9.0, 20001)
        assert np.max(np.abs(erfc_approx(x, bad) - erfc(x)) / erfc(x)) > 1e-5


def test_the_fit_reproduces_the_kernel_coefficients():
    c, err = fit()
    assert err < 1.1e-7
    assert np.allclose(c, kernel_coefficients(), rtol=1e-6, atol=1e-8)
