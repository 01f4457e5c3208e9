"""Seeded test inputs shared by the CPU and GPU parity tests."""
from __future__ import annotations

import numpy as np

BC_SETS = {
    "periodic": (1, 1, 1, 1),
    "outflow": (3, 3, 3, 3),
    "fixed_inflow": (2, 2, 4, 4),        # both act as zero-gradient without an ambient
    "x_periodic": (1, 1, 3, 3),
    "y_periodic": (3, 3, 1, 1),
}
VARIANTS = [(s, l) for s in (1, 2, 3) for l in (1, 2)]


def random_prims(ny: int, nx: int, seed: int, rough: float = 1.0) -> np.ndarray:
    """(6, ny, nx) = rho, vx, vy, P(total), Bx, By with smooth structure plus noise and a
    few hostile cells (near-vacuum, low thermal pressure, strong field)."""
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.linspace(0, 1, ny, endpoint=False), np.linspace(0, 1, nx, endpoint=False),
                       indexing="ij")
    rho = 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + rough * 0.2 * rng.random((ny, nx))
    vx = 0.5 * np.sin(2 * np.pi * y) + rough * 0.3 * rng.standard_normal((ny, nx))
    vy = 0.5 * np.cos(2 * np.pi * x) + rough * 0.3 * rng.standard_normal((ny, nx))
    Bx = 0.3 + 0.4 * np.cos(2 * np.pi * y) + rough * 0.2 * rng.standard_normal((ny, nx))
    By = 0.2 * np.sin(2 * np.pi * x) + rough * 0.2 * rng.standard_normal((ny, nx))
    pth = 0.6 + 0.3 * rng.random((ny, nx))
    k = rng.integers(0, ny * nx, size=max(3, ny * nx // 40))
    rho.flat[k[: len(k) // 3]] = 1e-6                       # near vacuum
    pth.flat[k[len(k) // 3: 2 * len(k) // 3]] = 1e-9         # thermal pressure at the floor
    Bx.flat[k[2 * len(k) // 3:]] *= 8.0                      # strong field
    P = pth + 0.5 * (Bx * Bx + By * By)
    return np.ascontiguousarray(np.stack([rho, vx, vy, P, Bx, By]))


def random_face_states(n: int, seed: int) -> np.ndarray:
    """(12, n) Riemann inputs (rho_L, rho_R, vx_L, vx_R, vy_L, vy_R, P_L, P_R, Bx_L, Bx_R,
    By_L, By_R) covering subsonic, supersonic, degenerate-|Bn| and capped-c_f faces."""
    rng = np.random.default_rng(seed)
    rL, rR = rng.uniform(0.05, 3, n), rng.uniform(0.05, 3, n)
    uL, uR = rng.normal(0, 1.5, n), rng.normal(0, 1.5, n)
    wL, wR = rng.normal(0, 1, n), rng.normal(0, 1, n)
    BL, BR = rng.normal(0, 1, n), rng.normal(0, 1, n)
    CL, CR = rng.normal(0, 1, n), rng.normal(0, 1, n)
    pL, pR = rng.uniform(0.01, 2, n), rng.uniform(0.01, 2, n)
    m = n // 8
    uL[:m] += 20.0; uR[:m] += 20.0                       # supersonic to the right
    uL[m:2 * m] -= 20.0; uR[m:2 * m] -= 20.0             # supersonic to the left
    BL[2 * m:3 * m] = 0.0; BR[2 * m:3 * m] = 0.0         # |Bn| = 0
    rL[3 * m:4 * m] = 1e-7                               # c_f beyond the caps
    pL[4 * m:5 * m] = -0.5                               # negative thermal pressure -> floor
    BR[5 * m:6 * m] = BL[5 * m:6 * m]                    # continuous normal field
    PL = pL + 0.5 * (BL * BL + CL * CL)
    PR = pR + 0.5 * (BR * BR + CR * CR)
    return np.ascontiguousarray(np.stack([rL, rR, uL, uR, wL, wR, PL, PR, BL, BR, CL, CR]))
