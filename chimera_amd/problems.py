"""Initial conditions (host side, numpy) for the benchmark and test problems.

These restate the reference's problem setups (src/mhd_init.f90) so that the same
FP64 input arrays can be fed to the GPU path and to the CPU oracle; they are set
up once per run and are not part of the per-step hot path.  States are returned
as a C-contiguous float64 array of shape (6, ny, nx) holding
(Mass, Momx, Momy, Energy, b_x, b_y): conserved cell integrals and face B.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PI = 3.1415926535897932       # src/mhd_config.f90:54-56
TWO_PI = 6.2831853071795865
FOUR_PI = 12.566370614359173
P_FLOOR = 1.0e-12
RHO_FLOOR = 1.0e-12

PERIODIC = (1, 1, 1, 1)
OUTFLOW = (3, 3, 3, 3)


@dataclass
class Problem:
    """A ready-to-run grid: state (6, ny, nx) plus the parameters the stage needs."""
    name: str
    state: np.ndarray
    dx: float
    gamma: float
    bc: tuple
    scheme: tuple          # (riemann_solver, slope_limiter, use_slope_limiting, mood)
    meta: dict = field(default_factory=dict)

    @property
    def nx(self) -> int:
        return self.state.shape[2]

    @property
    def ny(self) -> int:
        return self.state.shape[1]


# ---- field operations used by the setups (src/mhd_field_ops.f90) --------------

def curl_2d(Az: np.ndarray, dx: float, bc) -> tuple[np.ndarray, np.ndarray]:
    """Backward-difference curl of a corner potential (field_ops:15-57)."""
    inv_dx = 1.0 / dx
    bx = np.empty_like(Az)
    by = np.empty_like(Az)
    bx[1:, :] = (Az[1:, :] - Az[:-1, :]) * inv_dx
    bx[0, :] = (Az[0, :] - Az[-1, :]) * inv_dx if bc[2] == 1 else bx[1, :]
    by[:, 1:] = -((Az[:, 1:] - Az[:, :-1]) * inv_dx)
    by[:, 0] = -((Az[:, 0] - Az[:, -1]) * inv_dx) if bc[0] == 1 else by[:, 1]
    return bx, by


def face_to_cell_B(bx: np.ndarray, by: np.ndarray, bc) -> tuple[np.ndarray, np.ndarray]:
    """Cell-centred B from face B (field_ops:120-156)."""
    Bx = np.empty_like(bx)
    By = np.empty_like(by)
    Bx[:, 1:] = 0.5 * (bx[:, 1:] + bx[:, :-1])
    Bx[:, 0] = 0.5 * (bx[:, 0] + bx[:, -1]) if bc[0] == 1 else bx[:, 0]
    By[1:, :] = 0.5 * (by[1:, :] + by[:-1, :])
    By[0, :] = 0.5 * (by[0, :] + by[-1, :]) if bc[2] == 1 else by[0, :]
    return Bx, By


def prim_to_cons(rho, vx, vy, P, Bx, By, gamma, vol):
    """get_conserved (src/mhd_change_states.f90:16-58)."""
    r = np.maximum(rho, RHO_FLOOR)
    p = np.maximum(P, P_FLOOR)
    M = r * vol
    half_b2 = 0.5 * (Bx * Bx + By * By)
    E = ((p - half_b2) / (gamma - 1.0) + 0.5 * r * (vx * vx + vy * vy) + half_b2) * vol
    return M, M * vx, M * vy, E


def _pack(M, Mx, My, E, bx, by) -> np.ndarray:
    return np.ascontiguousarray(np.stack([M, Mx, My, E, bx, by]).astype(np.float64))


def _libm(name: str, a) -> np.ndarray:
    """Elementwise exp/sin/cos through the C library's libm (Python's math module), as
    the reference's compiled Fortran evaluates them.  numpy's vectorised versions can
    differ in the last bit, and one bit is enough to flip a MOOD or limiter tie: the
    Kelvin-Helmholtz set-up with numpy's exp drifts 2e-4 from the reference by t = 0.01."""
    f = np.frompyfunc(getattr(math, name), 1, 1)
    return np.asarray(f(np.asarray(a, dtype=np.float64)), dtype=np.float64)


def _centres(n: int, dx: float, lo: float = 0.0) -> np.ndarray:
    # xlin = (i - 0.5) * dx, src/main.f90:36
    return lo + (np.arange(n) + 0.5) * dx


# ---- problem setups ------------------------------------------------------------

def orszag_tang(n: int, scheme=(1, 2, 1, 1)) -> Problem:
    """Orszag-Tang vortex, src/mhd_init.f90:296-311 (Az at cell centres)."""
    dx = 1.0 / n
    x = _centres(n, dx)
    X, Y = np.meshgrid(x, x, indexing="xy")   # X[j, i] = x_i, Y[j, i] = y_j
    rho = np.full((n, n), 25.0 / (36.0 * PI))
    P = np.full((n, n), 5.0 / (12.0 * PI))
    # each transcendental term depends on x or y alone: evaluate it once per column/row
    sx, sy = _libm("sin", TWO_PI * x), _libm("sin", TWO_PI * x)
    c4x, c2y = _libm("cos", FOUR_PI * x), _libm("cos", TWO_PI * x)
    vx = -np.broadcast_to(sy[:, None], (n, n))
    vy = np.broadcast_to(sx[None, :], (n, n)).copy()
    Az = c4x[None, :] / (FOUR_PI * np.sqrt(FOUR_PI)) + c2y[:, None] / (TWO_PI * np.sqrt(FOUR_PI))
    bx, by = curl_2d(Az, dx, PERIODIC)
    Bx, By = face_to_cell_B(bx, by, PERIODIC)
    P = P + 0.5 * (Bx * Bx + By * By)
    gamma = 5.0 / 3.0
    st = _pack(*prim_to_cons(rho, vx, vy, P, Bx, By, gamma, dx * dx), bx, by)
    return Problem("orszag_tang", st, dx, gamma, PERIODIC, tuple(scheme))


def rotor(n: int, scheme=(2, 1, 1, 1)) -> Problem:
    """MHD rotor, src/mhd_init.f90:399-450 (outflow on all sides, gamma = 1.4)."""
    dx = 1.0 / n
    x = _centres(n, dx)
    X, Y = np.meshgrid(x, x, indexing="xy")
    rho = np.ones((n, n))
    vx = np.zeros((n, n))
    vy = np.zeros((n, n))
    P = np.ones((n, n))
    Bx = np.full((n, n), 5.0 / np.sqrt(4.0 * PI))
    By = np.zeros((n, n))
    r0, r1, u0 = 0.1, 0.115, 2.0
    rx, ry = X - 0.5, Y - 0.5
    r = np.sqrt(rx * rx + ry * ry)
    core = r <= r0
    rho[core] = 10.0
    vx[core] = -u0 * (ry[core] / r0)
    vy[core] = u0 * (rx[core] / r0)
    tap = (~core) & (r < r1)
    f = (r1 - r[tap]) / (r1 - r0)
    rho[tap] = 1.0 + 9.0 * f
    vx[tap] = -f * u0 * (ry[tap] / r[tap])
    vy[tap] = f * u0 * (rx[tap] / r[tap])
    P = P + 0.5 * (Bx * Bx + By * By)
    gamma = 1.4
    st = _pack(*prim_to_cons(rho, vx, vy, P, Bx, By, gamma, dx * dx), Bx.copy(), np.zeros((n, n)))
    return Problem("rotor", st, dx, gamma, OUTFLOW, tuple(scheme))


def kelvin_helmholtz(n: int, scheme=(1, 2, 1, 1)) -> Problem:
    """Kelvin-Helmholtz shear layer, src/mhd_init.f90:319-352 (B = 0, gamma = 1.4)."""
    dx = 1.0 / n
    x = _centres(n, dx)
    X, Y = np.meshgrid(x, x, indexing="xy")
    inner = np.abs(Y - 0.5) <= 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    sigma, x0, y0, amp = 0.05, 0.5, 0.25, 0.5
    vy = amp * _libm("exp", -((X - x0) ** 2 + (Y - y0) ** 2) / (2.0 * sigma ** 2))
    P = np.full((n, n), 2.5)
    z = np.zeros((n, n))
    gamma = 1.4
    st = _pack(*prim_to_cons(rho, vx, vy, P, z, z, gamma, dx * dx), z.copy(), z.copy())
    return Problem("kelvin_helmholtz", st, dx, gamma, PERIODIC, tuple(scheme))


def field_loop(nx: int = 512, ny: int = 256, Az_scale=None, scheme=(2, 2, 1, 1),
               return_az: bool = False):
    """Field-loop advection on [-1,1] x [-0.5,0.5] (dx = 2/nx), v = (2, 1), rho = p = 1,
    Az = 1e-3 (0.3 - r) inside r < 0.3 sampled at cell centres (as
    src/mhd_init.f90:372-386 samples it), gamma = 5/3.

    ``Az_scale`` (optional array (ny, nx)) multiplies Az, for perturbed targets."""
    dx = 2.0 / nx
    x = _centres(nx, dx, -1.0)
    y = _centres(ny, dx, -0.5)
    X, Y = np.meshgrid(x, y, indexing="xy")
    r = np.sqrt(X * X + Y * Y)
    Az = np.where(r < 0.3, 1.0e-3 * (0.3 - r), 0.0)
    if Az_scale is not None:
        Az = Az * Az_scale
    prob = field_loop_from_az(Az, dx, scheme)
    return (prob, Az) if return_az else prob


def field_loop_unit(n: int, scheme=(1, 2, 1, 1)) -> Problem:
    """The reference's field-loop problem (problem type 3), src/mhd_init.f90:360-391:
    unit box, loop of radius 0.3 and amplitude 1e-3 centred at (0.5, 0.5), Az at cell
    centres, rho = p = 1, v = (-2, -1), gamma = 5/3, periodic."""
    dx = 1.0 / n
    x = _centres(n, dx)
    X, Y = np.meshgrid(x, x, indexing="xy")
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
    Az = np.where(r < 0.3, 1.0e-3 * (0.3 - r), 0.0)
    bx, by = curl_2d(Az, dx, PERIODIC)
    Bx, By = face_to_cell_B(bx, by, PERIODIC)
    rho = np.ones((n, n))
    vx = np.full((n, n), -2.0)
    vy = np.full((n, n), -1.0)
    P = 1.0 + 0.5 * (Bx * Bx + By * By)
    gamma = 5.0 / 3.0
    st = _pack(*prim_to_cons(rho, vx, vy, P, Bx, By, gamma, dx * dx), bx, by)
    return Problem("field_loop_unit", st, dx, gamma, PERIODIC, tuple(scheme))


# the reference's problem types 1-4 (src/mhd_init.f90:259-288) and their host set-ups
SETUP_NAMES = {1: "orszag_tang", 2: "kelvin_helmholtz", 3: "field_loop_unit", 4: "rotor"}
HOST_SETUPS = {1: orszag_tang, 2: kelvin_helmholtz, 3: field_loop_unit, 4: rotor}


def setup_device(problem_type: int, n: int, scheme=(1, 2, 1, 1)) -> Problem:
    """Problem type 1-4 on N x N cells with the state generated on the current GPU by
    the library (chimera_setup_problem, on torch's current stream): the same bits as
    the host set-up of the same problem (``HOST_SETUPS``).  The returned Problem's
    ``state`` is a float64 tensor (6, n, n) on that GPU."""
    import torch
    from . import _lib
    from .solver import problem_info
    dx, gamma, bc = problem_info(problem_type, n)
    st = torch.empty((6, n, n), dtype=torch.float64, device="cuda")
    ptrs = _lib.state_ptrs(st, (6, n, n), "setup_device")
    _lib.check(_lib.lib().chimera_setup_problem(int(problem_type), int(n), ptrs,
                                                torch.cuda.current_stream().cuda_stream),
               "chimera_setup_problem")
    return Problem(SETUP_NAMES[problem_type], st, dx, gamma, tuple(bc), tuple(scheme))


def realized_parameters(prob: Problem) -> tuple[float, float]:
    """(M_s, beta) of an initial state as the reference reports them for its random
    problem (src/mhd_init.f90:833-841): rms |v| over the sound speed of the mean state,
    and 2 <p> / B_rms^2 (p thermal, cell B)."""
    st = np.asarray(prob.state)
    vol = prob.dx * prob.dx
    bx, by = st[4], st[5]
    Bx, By = face_to_cell_B(bx, by, prob.bc)
    rho = st[0] / vol
    vx, vy = st[1] / st[0], st[2] / st[0]
    half_b2 = 0.5 * (Bx * Bx + By * By)
    p = (st[3] / vol - 0.5 * rho * (vx * vx + vy * vy) - half_b2) * (prob.gamma - 1.0)
    n = rho.size
    v_rms = np.sqrt(np.sum(vx * vx + vy * vy) / n)
    b_rms = np.sqrt(np.sum(Bx * Bx + By * By) / n)
    p_avg = np.sum(p) / n
    c_s = np.sqrt(prob.gamma * p_avg / (np.sum(rho) / n))
    return float(v_rms / (c_s + 1.0e-16)), float(2.0 * p_avg / (b_rms * b_rms + 1.0e-16))


def field_loop_from_az(Az: np.ndarray, dx: float, scheme=(2, 2, 1, 1), hydro=None) -> Problem:
    """Field-loop state for a given corner potential Az (periodic).  ``hydro``
    (optional (rho, vx, vy, p_thermal) arrays) replaces the uniform rho = p = 1,
    v = (2, 1) background."""
    ny, nx = Az.shape
    bx, by = curl_2d(Az, dx, PERIODIC)
    Bx, By = face_to_cell_B(bx, by, PERIODIC)
    if hydro is None:
        rho = np.ones((ny, nx))
        vx = np.full((ny, nx), 2.0)
        vy = np.full((ny, nx), 1.0)
        pth = np.ones((ny, nx))
    else:
        rho, vx, vy, pth = hydro
    P = pth + 0.5 * (Bx * Bx + By * By)
    gamma = 5.0 / 3.0
    st = _pack(*prim_to_cons(rho, vx, vy, P, Bx, By, gamma, dx * dx), bx, by)
    return Problem("field_loop", st, dx, gamma, PERIODIC, tuple(scheme))


def smooth_periodic(nx: int, ny: int, seed: int, modes: int = 3) -> np.ndarray:
    """A smooth periodic field of unit-order amplitude: a few Fourier modes with random
    amplitudes and phases, so that no two neighbouring cells are equal by symmetry."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="xy")
    f = np.zeros((ny, nx))
    for kx in range(0, modes + 1):
        for ky in range(-modes, modes + 1):
            if kx == 0 and ky <= 0:
                continue
            a = rng.standard_normal() / (kx * kx + ky * ky)
            f += a * np.cos(TWO_PI * (kx * X + ky * Y) + rng.uniform(0.0, TWO_PI))
    return f / np.abs(f).max()


def field_loop_background(nx: int, ny: int, seed: int = 11, amp: float = 0.1):
    """(rho, vx, vy, p_thermal) = (1, 2, 1, 1) each times (1 + amp * smooth field):
    a background with no uniform field, so that the discrete map is differentiable
    everywhere (no limiter ties between equal neighbours)."""
    f = [smooth_periodic(nx, ny, seed + q) for q in range(4)]
    return (1.0 + amp * f[0], 2.0 * (1.0 + amp * f[1]), 1.0 * (1.0 + amp * f[2]),
            1.0 + amp * f[3])


# ---- Gaussian-random-field ensemble -------------------------------------------

def rbf_field(n: int, ell: float, rng: np.random.Generator) -> np.ndarray:
    """White noise convolved (periodically, by FFT) with exp(-r^2 / (2 ell^2)) on the
    unit torus, normalised to unit rms."""
    w = rng.standard_normal((n, n))
    k = np.fft.fftfreq(n, d=1.0 / n) * TWO_PI
    KX, KY = np.meshgrid(k, k, indexing="xy")
    # Fourier transform of the periodic Gaussian kernel (up to a constant)
    S = np.exp(-0.5 * ell * ell * (KX * KX + KY * KY))
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * S))
    f -= f.mean()
    rms = np.sqrt(np.mean(f * f))
    return f / rms if rms > 0 else f


def grf_member(n: int, seed: int, scheme=(1, 2, 1, 1), gamma_range=(1.4, 5.0 / 3.0),
               v0_range=(0.1, 1.0), beta_range=(0.1, 10.0), ell_range=(0.05, 0.15),
               cf_cap: float = 3.0) -> Problem:
    """One GRF ensemble member on the periodic unit square.

    psi and Az are RBF-smoothed white noise (length ell ~ U[ell_range]); v is the
    centred curl of psi (as src/mhd_init.f90:764-773), face B the backward curl of Az
    (compute_curl_2d), rho = p = 1, gamma ~ U[gamma_range], rms |v| ~ U[v0_range],
    beta ~ U[beta_range] fixes B_rms = sqrt(2 p / beta), and v alone is scaled down so
    max(c_f + |v|) <= cf_cap (src/mhd_init.f90:815-831)."""
    rng = np.random.default_rng(seed)
    gamma = rng.uniform(*gamma_range)
    v0 = rng.uniform(*v0_range)
    beta = rng.uniform(*beta_range)
    ell = rng.uniform(*ell_range)
    dx = 1.0 / n
    psi = rbf_field(n, ell, rng)
    Az = rbf_field(n, ell, rng)
    vx = (np.roll(psi, -1, axis=0) - np.roll(psi, 1, axis=0)) / (2.0 * dx)
    vy = -(np.roll(psi, -1, axis=1) - np.roll(psi, 1, axis=1)) / (2.0 * dx)
    v_rms = np.sqrt(np.mean(vx * vx + vy * vy))
    vx *= v0 / v_rms
    vy *= v0 / v_rms
    bx, by = curl_2d(Az, dx, PERIODIC)
    Bx, By = face_to_cell_B(bx, by, PERIODIC)
    b_rms = np.sqrt(np.mean(Bx * Bx + By * By))
    b0 = np.sqrt(2.0 * 1.0 / beta)
    s = b0 / b_rms
    bx, by, Bx, By = bx * s, by * s, Bx * s, By * s
    rho = np.ones((n, n))
    p = np.ones((n, n))
    cf = np.sqrt(gamma * p / rho + (Bx * Bx + By * By) / rho)
    vmag = np.sqrt(vx * vx + vy * vy)
    tmax = float(np.max(cf + vmag))
    if tmax > cf_cap:
        sc = cf_cap / tmax
        vx *= sc
        vy *= sc
    P = p + 0.5 * (Bx * Bx + By * By)
    st = _pack(*prim_to_cons(rho, vx, vy, P, Bx, By, gamma, dx * dx), bx, by)
    return Problem(f"grf_{n}_{seed}", st, dx, gamma, PERIODIC, tuple(scheme),
                   {"seed": seed, "ell": ell, "v0": v0, "beta": beta})


def grf_sizes(members: int, choices=(128, 256, 384, 512), seed: int = 20261203) -> list[int]:
    """Grid size of each ensemble member, drawn uniformly from ``choices``."""
    rng = np.random.default_rng(seed)
    return [int(c) for c in rng.choice(choices, size=members)]


def grf_params(seed: int, gamma_range=(1.4, 5.0 / 3.0), v0_range=(0.1, 1.0),
               beta_range=(0.1, 10.0), ell_range=(0.05, 0.15), cf_cap: float = 3.0) -> dict:
    """The physical parameters of GRF member ``seed`` (the numpy draw grf_member makes
    first): gamma, rms |v|, plasma beta, RBF length and the signal-speed cap."""
    rng = np.random.default_rng(seed)
    return {"gamma": rng.uniform(*gamma_range), "v0": rng.uniform(*v0_range),
            "beta": rng.uniform(*beta_range), "ell": rng.uniform(*ell_range), "cf_cap": cf_cap}


def grf_state_from_noise(w, n: int, params: dict):
    """GRF member state (6, n, n) from two white-noise fields w = (w_psi, w_Az) in torch,
    differentiable in w: RBF smoothing by FFT, v = centred curl of psi scaled to rms v0,
    face B = backward curl of Az scaled to B_rms = sqrt(2 p / beta), v scaled so that
    max(c_f + |v|) <= cf_cap, rho = p = 1 (as grf_member)."""
    import torch
    dev = w.device
    gamma, v0, beta, ell, cf_cap = (params[k] for k in ("gamma", "v0", "beta", "ell", "cf_cap"))
    dx = 1.0 / n
    k = torch.fft.fftfreq(n, d=1.0 / n, device=dev, dtype=torch.float64) * TWO_PI
    S = torch.exp(-0.5 * ell * ell * (k[None, :] ** 2 + k[:, None] ** 2))

    def smooth(x):
        f = torch.fft.ifft2(torch.fft.fft2(x) * S).real
        f = f - f.mean()
        return f / torch.sqrt(torch.mean(f * f))

    psi, Az = smooth(w[0]), smooth(w[1])
    vx = (torch.roll(psi, -1, 0) - torch.roll(psi, 1, 0)) / (2.0 * dx)
    vy = -(torch.roll(psi, -1, 1) - torch.roll(psi, 1, 1)) / (2.0 * dx)
    v_rms = torch.sqrt(torch.mean(vx * vx + vy * vy))
    vx, vy = vx * (v0 / v_rms), vy * (v0 / v_rms)
    inv_dx = 1.0 / dx
    bx = (Az - torch.roll(Az, 1, 0)) * inv_dx
    by = -((Az - torch.roll(Az, 1, 1)) * inv_dx)
    Bx = 0.5 * (bx + torch.roll(bx, 1, 1))
    By = 0.5 * (by + torch.roll(by, 1, 0))
    s = np.sqrt(2.0 / beta) / torch.sqrt(torch.mean(Bx * Bx + By * By))
    bx, by, Bx, By = bx * s, by * s, Bx * s, By * s
    cf = torch.sqrt(gamma + Bx * Bx + By * By)
    tmax = torch.max(cf + torch.sqrt(vx * vx + vy * vy))
    sc = torch.clamp(cf_cap / tmax, max=1.0)
    vx, vy = vx * sc, vy * sc
    vol = dx * dx
    half_b2 = 0.5 * (Bx * Bx + By * By)
    M = torch.full_like(vx, vol)
    E = (1.0 / (gamma - 1.0) + 0.5 * (vx * vx + vy * vy) + half_b2) * vol
    return torch.stack([M, M * vx, M * vy, E, bx, by]).contiguous()


# ---- the reference's random problem (problem type 5) ----------------------------

def grf_problem_uniforms(n: int, seed: int) -> dict:
    """The uniform draws of the reference's random problem, in the order it consumes
    them (src/mhd_init.f90:504-710): three for gamma, M_s and beta, one for the RBF
    length, then for each of the four fields two per Fourier mode, ky outer and kx
    inner ((n, n, 2) as [ky, kx, (u1, u2)]).  numpy's PCG64 seeded with ``seed``
    stands in for the Fortran compiler's random_number, which no other build
    reproduces."""
    rng = np.random.default_rng(seed)
    return {"params": rng.random(3), "ell": rng.random(),
            "fields": [rng.random((n, n, 2)) for _ in range(4)]}


def grf_problem_parameters(u3, target_M_s: float = -1.0, target_beta: float = -1.0) -> dict:
    """sample_physics_parameters (src/mhd_init.f90:504-586): gamma ~ U(1.2, 1.7),
    log10 M_s ~ U(log10 0.2, log10 2), log10 beta ~ U(log10 0.5, log10 5), overridden
    by positive targets; rho = p = 1 give c_s = sqrt(gamma), V_0 = M_s c_s and
    B_0 = sqrt(2 / beta)."""
    gamma = 1.2 + 0.5 * u3[0]
    if target_M_s > 0.0:
        m_s = target_M_s
    else:
        m_s = 10.0 ** (np.log10(0.2) + (np.log10(2.0) - np.log10(0.2)) * u3[1])
        m_s = min(max(m_s, 0.2), 2.0)
    if target_beta > 0.0:
        beta = target_beta
    else:
        beta = 10.0 ** (np.log10(0.5) + (np.log10(5.0) - np.log10(0.5)) * u3[2])
        beta = min(max(beta, 0.5), 5.0)
    c_s = np.sqrt(gamma)
    return {"gamma": float(gamma), "M_s": float(m_s), "beta": float(beta),
            "V_0": float(m_s * c_s), "B_0": float(np.sqrt(2.0 / beta))}


def grf_spectral_field(u, ell: float, device=None):
    """generate_GRF (src/mhd_init.f90:617-710) with an inverse FFT in place of the
    direct O(n^4) sum: complex Gaussian noise from Box-Muller on the uniforms
    u = [ky, kx, (u1, u2)] (u1 floored at 1e-15), amplitude exp(-ell^2 |k|^2 / 2) with
    k = 2 pi (signed index), imaginary parts of the (0,0), (n/2,0), (0,n/2) and
    (n/2,n/2) modes cleared, real part of the inverse DFT (exponent +i, 1/n^2), mean
    removed.  Returns an (n, n) float64 tensor [iy, ix] on ``device``."""
    import torch
    u = torch.as_tensor(u, dtype=torch.float64, device=device)
    n = u.shape[0]
    dev = u.device
    u1 = torch.clamp(u[..., 0], min=1.0e-15)
    r = torch.sqrt(-2.0 * torch.log(u1))
    ph = TWO_PI * u[..., 1]
    kk = torch.arange(n, device=dev, dtype=torch.float64)
    kk = torch.where(kk > n // 2, kk - n, kk) * TWO_PI
    S = torch.exp(-0.5 * ell * ell * (kk[None, :] ** 2 + kk[:, None] ** 2))
    re, im = S * r * torch.cos(ph), S * r * torch.sin(ph)
    h = n // 2
    for (ky, kx) in ((0, 0), (0, h), (h, 0), (h, h)):
        im[ky, kx] = 0.0
    f = torch.fft.ifft2(torch.complex(re, im)).real
    return f - f.sum() / (n * n)


def grf_problem(n: int, seed: int, target_M_s: float = -1.0, target_beta: float = -1.0,
                scheme=(1, 2, 1, 1), device=None) -> tuple[Problem, dict]:
    """The reference's random problem, setup_Monte_Carlo and setup_GRF_fields
    (src/mhd_init.f90:461-486, 731-860), on ``device`` (torch; the FFTs run there):

      psi, A, g_rho, g_p = four GRFs with one RBF length ell = 0.15 + 0.1 U;
      v = 0.05 (d psi/dy, -d psi/dx) (centred) scaled to rms V_0;
      Az = 0.001 A, face B its backward curl, cell B the face average, both scaled
      to rms |B| = B_0;
      rho = max(1 + 0.1 g_rho / max|g_rho|, 0.1), p = max(1 + 0.05 g_p / max|g_p|, 0.1);
      v scaled down (B is not) so that max(c_f + |v|) <= 3;
      realized M_s = v_rms / c_s(mean p, mean rho), beta = 2 <p> / B_rms^2.

    Returns (Problem with a float64 tensor state (6, n, n), info) where info holds the
    sampled and the realized parameters.  The reference's fallback for a vanishing
    potential (rms B <= 1e-10) is not representable with face B and is rejected."""
    import torch
    u = grf_problem_uniforms(n, seed)
    par = grf_problem_parameters(u["params"], target_M_s, target_beta)
    gamma, V_0, B_0 = par["gamma"], par["V_0"], par["B_0"]
    ell = 0.15 + 0.1 * u["ell"]
    psi, A, g_rho, g_p = (grf_spectral_field(w, ell, device) for w in u["fields"])
    dx = 1.0 / n
    nn = float(n * n)
    vx = 0.05 * (torch.roll(psi, -1, 0) - torch.roll(psi, 1, 0)) / (2.0 * dx)
    vy = -0.05 * (torch.roll(psi, -1, 1) - torch.roll(psi, 1, 1)) / (2.0 * dx)
    v_rms = torch.sqrt(torch.sum(vx * vx + vy * vy) / nn)
    if float(v_rms) > 1.0e-10:
        vx, vy = vx * (V_0 / v_rms), vy * (V_0 / v_rms)
    Az = 0.001 * A
    inv_dx = 1.0 / dx
    bx = (Az - torch.roll(Az, 1, 0)) * inv_dx
    by = -((Az - torch.roll(Az, 1, 1)) * inv_dx)
    Bx = 0.5 * (bx + torch.roll(bx, 1, 1))
    By = 0.5 * (by + torch.roll(by, 1, 0))
    b_rms = torch.sqrt(torch.sum(Bx * Bx + By * By) / nn)
    if not float(b_rms) > 1.0e-10:
        raise ValueError("GRF potential vanishes (rms B <= 1e-10)")
    sc = B_0 / b_rms
    bx, by, Bx, By = bx * sc, by * sc, Bx * sc, By * sc
    for g in (g_rho, g_p):
        m = torch.max(torch.abs(g))
        if float(m) > 1.0e-10:
            g /= m
    rho = torch.clamp(1.0 + 0.1 * g_rho, min=0.1)
    p = torch.clamp(1.0 + 0.05 * g_p, min=0.1)
    b2 = Bx * Bx + By * By
    c_s = torch.sqrt(gamma * p / rho)
    c_a = torch.sqrt(b2 / rho)
    total_max = float(torch.max(torch.sqrt(c_s * c_s + c_a * c_a) + torch.sqrt(vx * vx + vy * vy)))
    if total_max > 3.0:
        vx, vy = vx * (3.0 / total_max), vy * (3.0 / total_max)
    v_rms = float(torch.sqrt(torch.sum(vx * vx + vy * vy) / nn))
    b_rms = float(torch.sqrt(torch.sum(b2) / nn))
    p_avg = float(torch.sum(p) / nn)
    c_s_avg = float(np.sqrt(gamma * p_avg / (float(torch.sum(rho)) / nn)))
    info = dict(par, ell=ell, seed=seed, M_s_actual=v_rms / (c_s_avg + 1.0e-16),
                beta_actual=2.0 * p_avg / (b_rms * b_rms + 1.0e-16), cf_max=total_max)
    # get_conserved with total pressure (src/mhd_change_states.f90:16-58)
    vol = dx * dx
    r = torch.clamp(rho, min=RHO_FLOOR)
    ptot = torch.clamp(p + 0.5 * b2, min=P_FLOOR)
    half_b2 = 0.5 * b2
    M = r * vol
    E = ((ptot - half_b2) / (gamma - 1.0) + 0.5 * r * (vx * vx + vy * vy) + half_b2) * vol
    st = torch.stack([M, M * vx, M * vy, E, bx, by]).contiguous()
    return Problem(f"grf_problem_{n}_{seed}", st, dx, gamma, PERIODIC, tuple(scheme), info), info


def grf_member_device(n: int, seed: int, scheme=(1, 2, 1, 1), device="cuda", **kw):
    """The ensemble member of ``grf_member`` generated on the GPU (torch FFTs): the
    physical parameters come from the same numpy draw, the two white-noise fields from
    a torch generator seeded with ``seed``.  The returned Problem's ``state`` is a
    float64 tensor (6, n, n) on ``device``."""
    import torch
    params = grf_params(seed, **kw)
    gen = torch.Generator(device=device).manual_seed(int(seed))
    w = torch.stack([torch.randn((n, n), generator=gen, device=device, dtype=torch.float64)
                     for _ in range(2)])
    st = grf_state_from_noise(w, n, params)
    return Problem(f"grf_{n}_{seed}", st, 1.0 / n, params["gamma"], PERIODIC, tuple(scheme),
                   {"seed": seed, "ell": params["ell"], "v0": params["v0"], "beta": params["beta"]})
