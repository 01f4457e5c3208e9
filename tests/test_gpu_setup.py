"""Problem set-ups on the device (chimera_setup_problem / chimera_ctx_setup_problem)
against the host set-ups, which tests/test_problem_setups.py pins to the reference
program's own initial state (src/mhd_init.f90:296-450 + get_conserved), and against
the closed forms of the initial conditions."""
from __future__ import annotations

import math

import numpy as np
import pytest

from chimera_amd import problems
from chimera_amd.solver import Grid, problem_info

pytestmark = pytest.mark.gpu

SCHEME = (1, 2, 1, 1)     # the reference's mhd_config scheme: Rusanov, MC, limiting, MOOD


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def device_state(problem_type: int, n: int) -> np.ndarray:
    return problems.setup_device(problem_type, n, scheme=SCHEME).state.cpu().numpy()


@pytest.mark.parametrize("n", [4, 37, 64, 100, 256, 1024])
@pytest.mark.parametrize("problem_type", [1, 2, 3, 4])
def test_device_setup_bits_equal_host_setup(problem_type, n, gpu_lib):
    print(f"set-up {problem_type} N={n}", flush=True)
    host = problems.HOST_SETUPS[problem_type](n, scheme=SCHEME)
    dev = problems.setup_device(problem_type, n, scheme=SCHEME)
    assert (dev.dx, dev.gamma, tuple(dev.bc)) == (host.dx, host.gamma, tuple(host.bc))
    d = dev.state.cpu().numpy()
    for q in range(6):
        diff = int(np.count_nonzero(bits(d[q]) != bits(host.state[q])))
        assert diff == 0, f"field {q}: {diff} of {n * n} cells differ"


@pytest.mark.parametrize("problem_type", [1, 4])
def test_device_setup_bits_at_2048(problem_type, gpu_lib):
    """The rotor at its benchmark size (2048^2) and the Orszag-Tang vortex (its
    sine/cosine tables) at the same size."""
    print(f"set-up {problem_type} N=2048", flush=True)
    host = problems.HOST_SETUPS[problem_type](2048, scheme=SCHEME).state
    d = device_state(problem_type, 2048)
    assert np.array_equal(bits(d), bits(host))


@pytest.mark.parametrize("problem_type,n", [(1, 64), (2, 64), (3, 48), (4, 64)])
def test_context_setup_steps_like_upload(problem_type, n, gpu_lib):
    """Grid.from_setup writes the context's state in place of an upload: the state and
    20 reference-policy steps from it equal those of the uploaded host set-up."""
    print(f"ctx set-up {problem_type} N={n}", flush=True)
    host = problems.HOST_SETUPS[problem_type](n, scheme=SCHEME)
    with Grid.from_setup(problem_type, n, SCHEME) as a, Grid.from_problem(host) as b:
        assert np.array_equal(bits(a.download()), bits(host.state))
        a.steps(20)
        b.steps(20)
        assert a.time() == b.time()
        assert np.array_equal(bits(a.download()), bits(b.download()))


def test_context_setup_rejects_another_grid(gpu_lib):
    from chimera_amd._lib import ChimeraError, check
    dx, gamma, _ = problem_info(4, 64)
    with Grid(64, 64, dx, gamma, (1, 1, 1, 1), SCHEME) as g:      # the rotor needs outflow
        with pytest.raises(ChimeraError, match="outflow"):
            check(g._L.chimera_ctx_setup_problem(g._h, 4), "chimera_ctx_setup_problem")
    with Grid(64, 32, 1.0 / 64, 5.0 / 3.0, (1, 1, 1, 1), SCHEME) as g:   # not N x N
        with pytest.raises(ChimeraError, match="N x N"):
            check(g._L.chimera_ctx_setup_problem(g._h, 1), "chimera_ctx_setup_problem")


# ---- closed forms of the initial conditions (src/mhd_init.f90:296-450) -------------------

def primitives(st: np.ndarray, dx: float):
    vol = dx * dx
    rho = st[0] / vol
    return rho, st[1] / st[0], st[2] / st[0]


def centres(n: int):
    x = (np.arange(n) + 0.5) / n
    return np.meshgrid(x, x, indexing="xy")


def test_orszag_tang_closed_form(gpu_lib):
    n = 256
    st = device_state(1, n)
    X, Y = centres(n)
    rho, vx, vy = primitives(st, 1.0 / n)
    assert np.allclose(rho, 25.0 / (36.0 * math.pi), rtol=1e-14, atol=0)
    assert np.allclose(vx, -np.sin(2 * np.pi * Y), rtol=0, atol=1e-14)
    assert np.allclose(vy, np.sin(2 * np.pi * X), rtol=0, atol=1e-14)
    # face B is a discrete curl: its periodic divergence vanishes to round-off, and it
    # samples B = (-sin(2 pi y), sin(4 pi x)) / sqrt(4 pi) at the faces to O(dx^2)
    bx, by = st[4], st[5]
    div = (bx - np.roll(bx, 1, 1) + by - np.roll(by, 1, 0)) * n     # compute_divB's stencil
    assert np.max(np.abs(div)) < 1e-11
    xf, yf = X - 0.5 / n, Y - 0.5 / n      # the potential sits at cell centres
    assert np.max(np.abs(bx - (-np.sin(2 * np.pi * yf) / math.sqrt(4 * math.pi)))) < 1e-3
    assert np.max(np.abs(by - np.sin(4 * np.pi * xf) / math.sqrt(4 * math.pi))) < 2e-3


def test_kelvin_helmholtz_closed_form(gpu_lib):
    n = 128
    st = device_state(2, n)
    X, Y = centres(n)
    rho, vx, vy = primitives(st, 1.0 / n)
    inner = np.abs(Y - 0.5) <= 0.25
    assert np.array_equal(rho, np.where(inner, 2.0, 1.0))
    assert np.array_equal(vx, np.where(inner, 0.5, -0.5))
    g = 0.5 * np.exp(-((X - 0.5) ** 2 + (Y - 0.25) ** 2) / (2 * 0.05 ** 2))
    assert np.allclose(vy, g, rtol=1e-15, atol=0)     # numpy exp and M vy / M: ulps
    assert np.array_equal(st[4], np.zeros_like(st[4])) and np.array_equal(st[5], st[4])
    # E = P / (gamma - 1) + kinetic energy, P = 2.5
    e_th = st[3] - 0.5 * st[0] * (vx * vx + vy * vy)
    assert np.allclose(e_th * n * n, 2.5 / 0.4, rtol=1e-14)


def test_field_loop_closed_form(gpu_lib):
    n = 256
    st = device_state(3, n)
    X, Y = centres(n)
    rho, vx, vy = primitives(st, 1.0 / n)
    assert np.array_equal(rho, np.ones_like(rho))
    assert np.array_equal(vx, np.full_like(vx, -2.0)) and np.array_equal(vy, np.full_like(vy, -1.0))
    bx, by = st[4], st[5]
    div = (bx - np.roll(bx, 1, 1) + by - np.roll(by, 1, 0)) * n
    assert np.max(np.abs(div)) < 1e-14
    # |B| = A0 = 1e-3 inside the loop (Az = A0 (R - r)), 0 well outside
    Bx = 0.5 * (bx + np.roll(bx, 1, 1))
    By = 0.5 * (by + np.roll(by, 1, 0))
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
    ring = (r > 0.05) & (r < 0.25)
    assert np.allclose(np.hypot(Bx, By)[ring], 1e-3, rtol=2e-2)
    assert np.max(np.abs(bx[r > 0.32])) == 0.0 and np.max(np.abs(by[r > 0.32])) == 0.0


def test_rotor_closed_form(gpu_lib):
    n = 512
    st = device_state(4, n)
    X, Y = centres(n)
    rho, vx, vy = primitives(st, 1.0 / n)
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2)
    assert np.all(rho[r <= 0.1] == 10.0) and np.all(rho[r >= 0.115] == 1.0)
    taper = (r > 0.1) & (r < 0.115)
    assert np.allclose(rho[taper], 1.0 + 9.0 * (0.115 - r[taper]) / 0.015, rtol=1e-13)
    # solid-body rotation at angular speed u0 / r0 = 20 in the core
    core = r <= 0.1
    assert np.allclose(vx[core], -20.0 * (Y[core] - 0.5), atol=1e-13)
    assert np.allclose(vy[core], 20.0 * (X[core] - 0.5), atol=1e-13)
    assert np.all(vx[r >= 0.115] == 0.0) and np.all(vy[r >= 0.115] == 0.0)
    b0 = 5.0 / math.sqrt(4.0 * math.pi)
    assert np.all(st[4] == b0) and np.all(st[5] == 0.0)
    # zero net momentum: the rotation is odd about the centre
    assert abs(st[1].sum()) < 1e-12 * np.abs(st[1]).sum()
    assert abs(st[2].sum()) < 1e-12 * np.abs(st[2]).sum()
