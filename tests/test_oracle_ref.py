"""Pin the C oracle to the reference itself: every routine and whole runs, bit for bit,
against the amdflang build of the reference Fortran modules (oracle/_ref).

Skipped where oracle/_ref has not been built (it needs the reference sources)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle.oracle import RefVariant, ref_available
from chimera_amd import problems
from tests.inputs import VARIANTS, random_face_states, random_prims

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")

N = 24
BCS = [(1, 1, 1, 1), (3, 3, 3, 3), (1, 1, 3, 3), (2, 2, 4, 4)]


def same(a, b):
    return np.array_equal(a, b)


@pytest.mark.parametrize("bc", BCS)
def test_fill_ghost_cells(orc, bc):
    ref = RefVariant(1, 2)
    f = np.random.default_rng(1).random((N, N))
    amb = np.random.default_rng(2).random((N, N))
    assert same(orc.fill_ghost_cells(f, bc), ref.fill_ghost_cells(f, bc))
    assert same(orc.fill_ghost_cells(f, bc, amb), ref.fill_ghost_cells(f, bc, amb))


@pytest.mark.parametrize("limiter", [1, 2])
def test_slopes(orc, limiter):
    ref = RefVariant(1, limiter)
    pad = np.random.default_rng(3).standard_normal((N + 2, N + 2))
    pad[5, :] = 0.25          # flat rows: zero slopes
    for a, b in zip(orc.apply_slope_limiter(pad, 1 / N, limiter), ref.apply_slope_limiter(pad, 1 / N)):
        assert same(a, b)
    for a, b in zip(orc.compute_gradients(pad, 1 / N), ref.compute_gradients(pad, 1 / N)):
        assert same(a, b)


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("rho_or_p", [0, 1])
def test_reconstruction(orc, bc, rho_or_p):
    ref = RefVariant(1, 2)
    rng = np.random.default_rng(4)
    f = rng.random((N, N)) + (0 if rho_or_p else -0.5)
    gx, gy = rng.standard_normal((N, N)) * 10, rng.standard_normal((N, N)) * 10
    a = orc.reconstruction(f, gx, gy, 1 / N, bc, rho_or_p)
    b = ref.reconstruction(f, gx, gy, 1 / N, bc, rho_or_p)
    for x, y in zip(a, b):
        assert same(x, y)


def test_thermal_pressure_check(orc):
    ref = RefVariant(1, 2)
    rng = np.random.default_rng(5)
    faces = rng.random((24, N, N))
    faces[0:4] = rng.random((4, N, N)) * 0.3     # P faces small: many resets
    cells = rng.random((6, N, N))
    fa, fb = faces.copy(), faces.copy()
    orc.thermal_pressure_check(fa, cells)
    ref.thermal_pressure_check(fb, cells)
    assert same(fa, fb) and not same(fa, faces)


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_compute_fluxes(orc, solver):
    ref = RefVariant(solver, 2)
    fin = random_face_states(4096, 6 + solver)
    assert same(orc.compute_fluxes(fin, 5 / 3, solver), ref.compute_fluxes(fin, 5 / 3))


@pytest.mark.parametrize("bc", BCS)
def test_update_ct_curl_divb_avg(orc, bc):
    ref = RefVariant(1, 2)
    rng = np.random.default_rng(7)
    F, fX, fY = rng.random((N, N)), rng.standard_normal((N, N)), rng.standard_normal((N, N))
    dx, dt = 1 / N, 1e-3
    assert same(orc.update_conserved(F, fX, fY, dx, dt, bc), ref.update_conserved(F, fX, fY, dx, dt, bc))
    bx, by = rng.standard_normal((N, N)), rng.standard_normal((N, N))
    for a, b in zip(orc.constrained_transport(bx, by, fX, fY, dx, dt, bc),
                    ref.constrained_transport(bx, by, fX, fY, dx, dt, bc)):
        assert same(a, b)
    Az = rng.standard_normal((N, N))
    for a, b in zip(orc.compute_curl_2d(Az, dx, bc), ref.compute_curl_2d(Az, dx, bc)):
        assert same(a, b)
    assert same(orc.compute_divB(bx, by, dx, bc), ref.compute_divB(bx, by, dx, bc))
    for a, b in zip(orc.average_face_to_cell_B(bx, by, bc), ref.average_face_to_cell_B(bx, by, bc)):
        assert same(a, b)


def test_change_states(orc):
    ref = RefVariant(1, 2)
    prim = random_prims(N, N, 8)
    prim[0, 0, :4] = [1e-20, -1.0, 0.0, 1e-13]      # density floors
    prim[3, 1, :3] = [-1.0, 0.0, 1e-14]              # pressure floors
    g, vol = 1.4, (1 / N) ** 2
    cons = orc.get_conserved(prim, g, vol)
    assert same(cons, ref.get_conserved(prim, g, vol))
    assert same(orc.get_primitive(cons, prim[4], prim[5], g, vol),
                ref.get_primitive(cons, prim[4], prim[5], g, vol))
    cons[0, 2, :3] = [-1e-9, 1e-30, 0.0]             # mass floor branch
    cons[1, 3, :3] = [1e3, -1e3, 5e2]                # velocity cap branch
    assert same(orc.apply_conserved_floors(cons, prim[4], prim[5], vol, g),
                ref.apply_conserved_floors(cons, prim[4], prim[5], vol, g))


@pytest.mark.parametrize("solver,limiter", VARIANTS)
@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (3, 3, 3, 3)])
def test_stage_fluxes(orc, solver, limiter, bc):
    ref = RefVariant(solver, limiter)
    prim = random_prims(N, N, 9 + solver)
    a = orc.stage_fluxes(prim, 1 / N, 5 / 3, bc, [solver, limiter, 1, 1])
    b, _ = ref.stage_fluxes(prim, 1 / N, 5 / 3, bc)
    assert same(a, b)


@pytest.mark.parametrize("solver,limiter", VARIANTS)
def test_runs_orszag_tang(orc, solver, limiter):
    ref = RefVariant(solver, limiter)
    p = problems.orszag_tang(32)
    a, b = p.state.copy(), p.state.copy()
    ra = orc.run(a, p.dx, p.gamma, p.bc, [solver, limiter, 1, 1], 25)
    rb = ref.run(b, p.dx, p.gamma, p.bc, 25)
    assert same(a, b)
    assert same(ra["dt"], rb["dt"]) and same(ra["divb"], rb["divb"])
    assert ra["t"] == rb["t"] and ra["out_count"] == rb["out_count"]


@pytest.mark.parametrize("solver,limiter", VARIANTS)
def test_runs_rotor_outflow(orc, solver, limiter):
    ref = RefVariant(solver, limiter)
    p = problems.rotor(40)
    a, b = p.state.copy(), p.state.copy()
    orc.run(a, p.dx, p.gamma, p.bc, [solver, limiter, 1, 1], 20, dt_mode=1)
    ref.run(b, p.dx, p.gamma, p.bc, 20, dt_mode=1)
    assert same(a, b)


def test_run_prescribed_dt_and_time_end(orc):
    ref = RefVariant(2, 2)
    p = problems.kelvin_helmholtz(24)
    dts = np.full(8, 2.5e-3)
    a, b = p.state.copy(), p.state.copy()
    ra = orc.run(a, p.dx, p.gamma, p.bc, [2, 2, 1, 1], 8, dt_mode=2, dt_in=dts)
    rb = ref.run(b, p.dx, p.gamma, p.bc, 8, dt_mode=2, dt_in=dts)
    assert same(a, b) and ra["steps"] == rb["steps"] == 8
    a, b = p.state.copy(), p.state.copy()
    ra = orc.run(a, p.dx, p.gamma, p.bc, [2, 2, 1, 1], 1000, t_end=0.02)
    rb = ref.run(b, p.dx, p.gamma, p.bc, 1000, t_end=0.02)
    assert same(a, b) and ra["steps"] == rb["steps"] and ra["t"] == rb["t"]
