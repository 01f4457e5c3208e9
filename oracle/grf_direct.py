"""Direct-sum restatement of the reference's GRF generator -- TEST INFRASTRUCTURE ONLY.

generate_GRF (src/mhd_init.f90:617-710) builds spectral coefficients from Box-Muller
noise and evaluates the inverse DFT as an explicit O(n^4) sum of cosines and sines.
``grf_field_direct`` repeats that sum, mode by mode in the reference's loop order
(kx inner, ky outer) for every output point, so the FFT-based generator of the
package can be checked against it on small grids.  Only tests/ use this module.

Arrays are numpy float64 shaped (n, n) as [iy, ix] (== Fortran (ix, iy)); the
uniforms are [ky, kx, (u1, u2)].
"""
from __future__ import annotations

import math

import numpy as np

TWO_PI = 6.2831853071795865   # src/mhd_config.f90:55


def grf_coefficients(u: np.ndarray, ell: float) -> tuple[np.ndarray, np.ndarray]:
    """Steps 1-2 of generate_GRF (src/mhd_init.f90:640-679): fhat_re, fhat_im [ky, kx]."""
    n = u.shape[0]
    re = np.empty((n, n))
    im = np.empty((n, n))
    for ky_idx in range(n):
        for kx_idx in range(n):
            kxs = kx_idx if kx_idx <= n // 2 else kx_idx - n
            kys = ky_idx if ky_idx <= n // 2 else ky_idx - n
            kx, ky = TWO_PI * kxs, TWO_PI * kys
            s_k = math.exp(-0.5 * ell * ell * (kx * kx + ky * ky))
            u1 = max(u[ky_idx, kx_idx, 0], 1.0e-15)
            u2 = u[ky_idx, kx_idx, 1]
            amp = math.sqrt(-2.0 * math.log(u1))
            re[ky_idx, kx_idx] = s_k * amp * math.cos(TWO_PI * u2)
            im[ky_idx, kx_idx] = s_k * amp * math.sin(TWO_PI * u2)
    h = n // 2
    for ky, kx in ((0, 0), (0, h), (h, 0), (h, h)):
        im[ky, kx] = 0.0
    return re, im


def grf_field_direct(u: np.ndarray, ell: float) -> np.ndarray:
    """generate_GRF with the reference's direct inverse DFT (src/mhd_init.f90:694-706):
    f(ix, iy) = sum_{kx,ky} re cos(phi) - im sin(phi), phi = 2 pi ((kx)(ix) + (ky)(iy)) / n,
    divided by n^2, then the mean removed (:708)."""
    n = u.shape[0]
    re, im = grf_coefficients(u, ell)
    f = np.empty((n, n))
    for iy in range(n):
        for ix in range(n):
            acc = 0.0
            for ky_idx in range(n):
                for kx_idx in range(n):
                    phi = TWO_PI * (float(kx_idx * ix) + float(ky_idx * iy)) / float(n)
                    acc = acc + re[ky_idx, kx_idx] * math.cos(phi) - im[ky_idx, kx_idx] * math.sin(phi)
            f[iy, ix] = acc / float(n * n)
    return f - np.sum(f) / float(n * n)
