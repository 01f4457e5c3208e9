"""Test states on which the discrete step is differentiable, so that central
differences are a valid check of its gradient:

* ``smooth_loop``: a Gaussian magnetic loop plus a smooth periodic Az on a smooth,
  nowhere-uniform background (no limiter tie between equal neighbours, no kink at a
  loop edge); used with the Van Leer limiter, whose faces stay strictly inside the MOOD
  bounds, so MOOD never fires;
* ``monotone_outflow``: fields increasing along x and y with gentle curvature on an
  outflow grid, where MC takes its central branch in every interior cell (MC's
  2*min branches put a face value exactly on the MOOD bound, a discontinuity).

The seeds are chosen so that ``min_neighbour_gap`` stays above 1e-6.
"""
from __future__ import annotations

import numpy as np
import torch

from chimera_amd import problems
from tests.adjoint_reference import Geometry, get_primitive

# (Az seed, background seed) per grid with a neighbour gap above 1e-6
LOOP_SEEDS = {(24, 20): (1, 16), (40, 40): (1, 20)}


def smooth_loop_az(nx: int, ny: int, seed: int = 1):
    """Gaussian Az (amplitude 3e-4, width 0.2) + 1e-4 x a smooth periodic field on the
    field-loop grid (dx = 2 / nx, centred at the origin)."""
    dx = 2.0 / nx
    x = -1.0 + (np.arange(nx) + 0.5) * dx
    y = -0.5 + (np.arange(ny) + 0.5) * dx
    X, Y = np.meshgrid(x, y, indexing="xy")
    Az = 3.0e-4 * np.exp(-(X * X + Y * Y) / 0.04) + 1.0e-4 * problems.smooth_periodic(nx, ny, seed=seed)
    return Az, dx


def smooth_loop_parts(nx: int, ny: int):
    az_seed, bg_seed = LOOP_SEEDS[(nx, ny)]
    Az, dx = smooth_loop_az(nx, ny, az_seed)
    return Az, dx, problems.field_loop_background(nx, ny, seed=bg_seed)


def smooth_loop(nx: int, ny: int, scheme):
    Az, dx, hyd = smooth_loop_parts(nx, ny)
    return problems.field_loop_from_az(Az, dx, scheme=scheme, hydro=hyd)


def monotone_outflow(n: int, scheme, seed: int = 0):
    rng = np.random.default_rng(seed)
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(x, x, indexing="xy")

    def ramp(base, slope):
        a, b = rng.uniform(0.5, 1.0, 2) * slope
        return base + a * X + b * Y + 0.05 * slope * (X * X - 0.3 * Y * Y)

    rho = ramp(1.0, 0.5)
    vx, vy = ramp(0.2, 0.5), ramp(-0.3, 0.5)
    pth = ramp(1.0, 0.5)
    Bx, By = ramp(0.3, 0.3), ramp(0.2, 0.3)
    gamma = 5.0 / 3.0
    st = problems._pack(*problems.prim_to_cons(rho, vx, vy, pth + 0.5 * (Bx * Bx + By * By),
                                               Bx, By, gamma, dx * dx), Bx, By)
    return problems.Problem("monotone", st, dx, gamma, problems.OUTFLOW, tuple(scheme))


def min_neighbour_gap(p) -> float:
    """Smallest |f(i) - f(i-1)| / max|f| over the cell primitives, both directions:
    a limiter switches (dfL * dfR = 0) where this vanishes."""
    U = torch.from_numpy(p.state)
    G = Geometry.of(p)
    imx = torch.arange(p.nx) - 1
    imy = torch.arange(p.ny) - 1
    imx = imx % p.nx if G.px else imx.clamp_min(0)
    imy = imy % p.ny if G.py else imy.clamp_min(0)
    Bx = 0.5 * (U[4] + U[4][:, imx])
    By = 0.5 * (U[5] + U[5][imy, :])
    prims = get_primitive(U[0], U[1], U[2], U[3], Bx, By, p.gamma, p.dx * p.dx) + (Bx, By)
    gap = 1.0
    for f in prims:
        f = f.numpy()
        s = np.abs(f).max()
        for ax in (0, 1):
            d = np.abs(np.diff(f, axis=ax)) / s
            gap = min(gap, float(d.min()))
    return gap
