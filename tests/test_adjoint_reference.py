"""CPU checks of the torch restatement used to verify the adjoint (adjoint_reference):

* its forward step equals the oracle's bit for bit (problems x solver x limiter x BCs),
  so every branch it takes is the oracle's;
* its autograd derivative equals central differences of the oracle wherever the
  discrete map is differentiable: smooth states without ties (Van Leer, where MOOD
  never fires, periodic) and monotone states (MC on its central branch, outflow)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from chimera_amd import problems
from tests.adjoint_cases import min_neighbour_gap, monotone_outflow, smooth_loop
from tests.adjoint_reference import Geometry, step
from tests.inputs import VARIANTS


def _problem(name, s, l):
    sch = (s, l, 1, 1)
    if name == "ot":
        return problems.orszag_tang(16, scheme=sch)
    if name == "rotor":
        return problems.rotor(20, scheme=sch)
    if name == "kh":
        return problems.kelvin_helmholtz(16, scheme=sch)
    return problems.field_loop(24, 16, scheme=sch)


@pytest.mark.parametrize("s,l", VARIANTS)
@pytest.mark.parametrize("name", ["ot", "rotor", "kh", "fl"])
def test_restatement_bitexact_vs_oracle(orc, name, s, l):
    p = _problem(name, s, l)
    dts = np.array([0.3 * p.dx / 8, 0.2 * p.dx / 8])
    st = p.state.copy()
    orc.run(st, p.dx, p.gamma, p.bc, p.scheme, 2, dt_mode=2, dt_in=dts)
    G = Geometry.of(p)
    U = torch.from_numpy(p.state.copy())
    with torch.no_grad():
        for dt in dts:
            U = step(U, float(dt), G)
    assert np.array_equal(U.numpy(), st)


def _jvp_vs_oracle_fd(orc, p, nsteps, seed):
    G = Geometry.of(p)
    dt = 0.25 * p.dx / 4.0
    U = torch.from_numpy(p.state.copy())
    scale = np.abs(p.state).reshape(6, -1).max(1)[:, None, None]
    v = np.stack([problems.smooth_periodic(p.nx, p.ny, seed + q) for q in range(6)]) * scale

    def f(x):
        for _ in range(nsteps):
            x = step(x, dt, G)
        return x

    _, jv = torch.autograd.functional.jvp(f, U, torch.from_numpy(v))
    jv = jv.numpy()
    # a small relative step: with ~1e4 max/min/limiter decisions per step, steps of
    # 1e-6 and above cross a switch point in a few of them; rounding error of the
    # differences at 1e-8 is ~3e-8 of the derivative
    h = 1e-8
    sp, sm = p.state + h * v, p.state - h * v
    sp, sm = np.ascontiguousarray(sp), np.ascontiguousarray(sm)
    dts = np.full(nsteps, dt)
    orc.run(sp, p.dx, p.gamma, p.bc, p.scheme, nsteps, dt_mode=2, dt_in=dts)
    orc.run(sm, p.dx, p.gamma, p.bc, p.scheme, nsteps, dt_mode=2, dt_in=dts)
    fd = (sp - sm) / (2 * h)
    errs = [np.abs(jv[q] - fd[q]).max() / np.abs(fd[q]).max() for q in range(6)]
    return errs


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_restatement_derivative_smooth_periodic(orc, solver):
    p = smooth_loop(24, 20, (solver, 1, 1, 1))
    # the state keeps every neighbour difference far above the step h below
    assert min_neighbour_gap(p) > 1e-6
    errs = _jvp_vs_oracle_fd(orc, p, 2, seed=100 + solver)
    assert max(errs) <= 1e-6, errs


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_restatement_derivative_monotone_mc(orc, solver):
    p = monotone_outflow(16, (solver, 2, 1, 1))
    errs = _jvp_vs_oracle_fd(orc, p, 2, seed=200 + solver)
    assert max(errs) <= 1e-6, errs
