"""The reference's random problem (problem type 5) as generated by problems.grf_problem.

The compiler's random_number makes the reference's own initial conditions
unreproducible, so the generator is pinned in three ways:
  * its FFT field equals the reference's direct O(n^4) inverse-DFT sum
    (oracle/grf_direct.py, src/mhd_init.f90:617-710) on the same uniforms, and that
    restatement is itself pinned to numpy's inverse FFT and to a closed-form delta;
  * the parameter sampler (src/mhd_init.f90:504-586) against its ranges, its
    log-uniform medians and the stress-test overrides;
  * the assembled state against invariants the set-up fixes (src/mhd_init.f90:731-860):
    discrete div v = 0 and div B = 0, rms v = V_0 and rms B = B_0, the perturbation
    bounds, the fast-speed cap, and the realized M_s and beta.
Runs on the CPU (torch's CPU FFT); the GPU copy is compared in test_gpu_cli.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from chimera_amd import problems
from oracle import grf_direct


def _field(u, ell):
    return problems.grf_spectral_field(u, ell, "cpu").numpy()


# ---- the direct-sum restatement, pinned ------------------------------------------

@pytest.mark.parametrize("n", [4, 5, 8])
def test_direct_sum_is_the_inverse_dft(n):
    u = np.random.default_rng(100 + n).random((n, n, 2))
    re, im = grf_direct.grf_coefficients(u, 0.2)
    want = np.real(np.fft.ifft2(re + 1j * im))
    want -= want.mean()
    assert np.max(np.abs(grf_direct.grf_field_direct(u, 0.2) - want)) < 1e-15


def test_direct_sum_delta():
    """r = 1, phase 0 and ell = 0 give fhat = 1 on every mode: the inverse DFT is a
    unit spike at the origin, minus its mean 1/n^2."""
    n = 6
    u = np.zeros((n, n, 2))
    u[..., 0] = np.exp(-0.5)
    f = grf_direct.grf_field_direct(u, 0.0)
    want = np.full((n, n), -1.0 / n**2)
    want[0, 0] += 1.0
    assert np.max(np.abs(f - want)) < 1e-14


def test_coefficients_clear_the_real_modes():
    n = 8
    u = np.random.default_rng(3).random((n, n, 2))
    _, im = grf_direct.grf_coefficients(u, 0.2)
    for ky, kx in ((0, 0), (0, 4), (4, 0), (4, 4)):
        assert im[ky, kx] == 0.0
    assert np.count_nonzero(im) == n * n - 4


# ---- the FFT generator against the direct sum -------------------------------------

@pytest.mark.parametrize("n,ell,seed", [(4, 0.15, 0), (5, 0.2, 1), (8, 0.25, 2), (12, 0.18, 3)])
def test_fft_field_matches_direct_sum(n, ell, seed):
    u = np.random.default_rng(seed).random((n, n, 2))
    d = grf_direct.grf_field_direct(u, ell)
    f = _field(u, ell)
    assert np.max(np.abs(f - d)) <= 1e-14 * max(1.0, np.max(np.abs(d)))


def test_spectrum_matches_rbf_length():
    """E|F(f)_k|^2 is proportional to exp(-ell^2 |k|^2): a least-squares fit of the log
    power against |k|^2 over the resolved modes recovers ell."""
    n, ell = 64, 0.2
    k = np.fft.fftfreq(n, d=1.0 / n) * problems.TWO_PI
    k2 = (k[None, :] ** 2 + k[:, None] ** 2).ravel()
    power = np.zeros(n * n)
    for s in range(16):
        f = _field(np.random.default_rng(s).random((n, n, 2)), ell)
        power += np.abs(np.fft.fft2(f)).ravel() ** 2
    sel = (k2 > 0) & (k2 * ell * ell < 20.0)
    slope = np.polyfit(k2[sel], np.log(power[sel]), 1)[0]
    assert abs(np.sqrt(-slope) / ell - 1.0) < 0.05


# ---- parameter sampler ------------------------------------------------------------

def test_parameter_ranges_and_log_uniform_medians():
    draws = [problems.grf_problem_parameters(problems.grf_problem_uniforms(4, s)["params"])
             for s in range(400)]
    g = np.array([d["gamma"] for d in draws])
    m = np.array([d["M_s"] for d in draws])
    b = np.array([d["beta"] for d in draws])
    assert g.min() >= 1.2 and g.max() <= 1.7
    assert m.min() >= 0.2 and m.max() <= 2.0
    assert b.min() >= 0.5 and b.max() <= 5.0
    # log-uniform: half the draws lie below the geometric midpoint
    assert abs(np.mean(m < np.sqrt(0.2 * 2.0)) - 0.5) < 0.08
    assert abs(np.mean(b < np.sqrt(0.5 * 5.0)) - 0.5) < 0.08
    for d in draws[:20]:
        assert d["V_0"] == pytest.approx(d["M_s"] * np.sqrt(d["gamma"]), rel=1e-15)
        assert d["B_0"] == pytest.approx(np.sqrt(2.0 / d["beta"]), rel=1e-15)


def test_overrides_keep_the_other_draws():
    u = problems.grf_problem_uniforms(4, 7)["params"]
    base = problems.grf_problem_parameters(u)
    over = problems.grf_problem_parameters(u, 5.0, 0.1)
    assert over["M_s"] == 5.0 and over["beta"] == 0.1
    assert over["gamma"] == base["gamma"]
    only_m = problems.grf_problem_parameters(u, 3.0, -1.0)
    assert only_m["beta"] == base["beta"] and only_m["M_s"] == 3.0


# ---- the assembled initial state ---------------------------------------------------

def _prims(prob):
    st = prob.state.numpy()
    vol = prob.dx * prob.dx
    bx, by = st[4], st[5]
    Bx, By = problems.face_to_cell_B(bx, by, problems.PERIODIC)
    rho = st[0] / vol
    vx, vy = st[1] / st[0], st[2] / st[0]
    b2 = Bx * Bx + By * By
    p = (st[3] / vol - 0.5 * rho * (vx * vx + vy * vy) - 0.5 * b2) * (prob.gamma - 1.0)
    return rho, vx, vy, p, Bx, By, bx, by


@pytest.mark.parametrize("seed", [1, 42, 43])
def test_state_invariants(seed):
    n = 48
    prob, info = problems.grf_problem(n, seed, device="cpu")
    rho, vx, vy, p, Bx, By, bx, by = _prims(prob)
    dx = prob.dx
    # face B from a potential: backward-difference divergence vanishes
    divb = (bx - np.roll(bx, 1, axis=1)) + (by - np.roll(by, 1, axis=0))
    assert np.max(np.abs(divb)) / dx < 1e-13 * max(1.0, np.max(np.abs(bx)) / dx) * n
    # v from a stream function: centred divergence vanishes
    divv = (np.roll(vx, -1, 1) - np.roll(vx, 1, 1)) + (np.roll(vy, -1, 0) - np.roll(vy, 1, 0))
    assert np.max(np.abs(divv)) < 1e-13 * np.max(np.abs(vx))
    # rms |B| = B_0; perturbation bounds with max|g| normalisation
    assert np.sqrt(np.mean(Bx * Bx + By * By)) == pytest.approx(info["B_0"], rel=1e-12)
    assert 0.9 - 1e-12 <= rho.min() and rho.max() <= 1.1 + 1e-12
    assert np.isclose(rho.max(), 1.1) or np.isclose(rho.min(), 0.9)
    assert 0.95 - 1e-12 <= p.min() and p.max() <= 1.05 + 1e-12
    assert np.isclose(p.max(), 1.05) or np.isclose(p.min(), 0.95)
    # fast-speed cap: v (only v) is scaled by 3 / max(c_f + |v|) when that exceeds 3
    cf = np.sqrt(prob.gamma * p / rho + (Bx * Bx + By * By) / rho)
    top = np.max(cf + np.sqrt(vx * vx + vy * vy))
    scale = min(1.0, 3.0 / info["cf_max"])
    v_rms = np.sqrt(np.mean(vx * vx + vy * vy))
    assert v_rms == pytest.approx(info["V_0"] * scale, rel=1e-12)
    if info["cf_max"] > 3.0:
        assert top <= info["cf_max"] and top >= 3.0 - 1e-12
    else:
        assert top == pytest.approx(info["cf_max"], rel=1e-12)
    # realized diagnostics as the reference defines them
    c_s = np.sqrt(prob.gamma * np.mean(p) / np.mean(rho))
    assert info["M_s_actual"] == pytest.approx(v_rms / c_s, rel=1e-9)
    assert info["beta_actual"] == pytest.approx(2.0 * np.mean(p) / info["B_0"] ** 2, rel=1e-9)
    assert info["M_s_actual"] == pytest.approx(info["M_s"] * scale, rel=0.01)
    assert info["beta_actual"] == pytest.approx(info["beta"], rel=0.01)


def test_high_mach_target_hits_the_cap():
    """target M_s = 5 drives max(c_f + |v|) past 3, so v alone is scaled by
    s = 3 / max(c_f + |v|); the realized M_s falls below the target
    (tests/stress/test_stress.py:126-177).  Since c_f is not scaled, the new maximum
    is only bounded by 3 + (1 - s) max c_f."""
    prob, info = problems.grf_problem(64, 42, 5.0, -1.0, device="cpu")
    rho, vx, vy, p, Bx, By, _, _ = _prims(prob)
    cf = np.sqrt(prob.gamma * p / rho + (Bx * Bx + By * By) / rho)
    s = 3.0 / info["cf_max"]
    assert s < 1.0
    assert np.sqrt(np.mean(vx * vx + vy * vy)) == pytest.approx(info["V_0"] * s, rel=1e-12)
    assert np.max(cf + np.sqrt(vx * vx + vy * vy)) <= 3.0 + (1.0 - s) * cf.max() + 1e-12
    assert 0.0 < info["M_s_actual"] < 5.0


def test_low_beta_target_is_realized():
    """target beta = 0.1: B is never rescaled by the cap, so beta_actual ~ 0.1
    (tests/stress/test_stress.py:179-236)."""
    _, info = problems.grf_problem(64, 42, -1.0, 0.1, device="cpu")
    assert info["beta_actual"] == pytest.approx(0.1, rel=0.01)


def test_same_seed_same_state_other_seed_differs():
    a, _ = problems.grf_problem(16, 5, device="cpu")
    b, _ = problems.grf_problem(16, 5, device="cpu")
    c, _ = problems.grf_problem(16, 6, device="cpu")
    assert np.array_equal(a.state.numpy(), b.state.numpy())
    assert not np.array_equal(a.state.numpy(), c.state.numpy())
