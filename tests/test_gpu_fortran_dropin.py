"""The Fortran drop-in boundary: the reference's main-loop structure (oracle's
ref_harness.f90) linked with the product's Fortran modules chimera_amd/fortran/mhd_*.f90
instead of the reference's -- every routine through the C ABI on the GPU -- must
reproduce the pure-reference build bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from chimera_amd import problems
from oracle.oracle import RefVariant, dropin_available, ref_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (ref_available() and dropin_available()),
                                 reason="oracle/_ref reference / drop-in builds not present")]


@pytest.mark.parametrize("solver,limiter", [(1, 2), (2, 1), (3, 2)])
@pytest.mark.parametrize("which", ["ot", "rotor"])
def test_fortran_dropin_matches_reference(gpu_lib, solver, limiter, which):
    prob = problems.orszag_tang(24) if which == "ot" else problems.rotor(24)
    mode = 0 if which == "ot" else 1
    print(f"fortran drop-in {which} s{solver} l{limiter}", flush=True)
    a, b = prob.state.copy(), prob.state.copy()
    ra = RefVariant(solver, limiter, "ref").run(a, prob.dx, prob.gamma, prob.bc, 3, dt_mode=mode)
    rb = RefVariant(solver, limiter, "dropin").run(b, prob.dx, prob.gamma, prob.bc, 3, dt_mode=mode)
    assert np.array_equal(a, b)
    assert np.array_equal(ra["dt"], rb["dt"]) and np.array_equal(ra["divb"], rb["divb"])


def test_fortran_dropin_routines(gpu_lib):
    """Single routines through the Fortran drop-in modules vs the reference modules."""
    ref, dro = RefVariant(3, 1, "ref"), RefVariant(3, 1, "dropin")
    rng = np.random.default_rng(5)
    n = 20
    f = rng.random((n, n))
    for bc in [(1, 1, 1, 1), (3, 3, 3, 3)]:
        assert np.array_equal(ref.fill_ghost_cells(f, bc), dro.fill_ghost_cells(f, bc))
        Az = rng.standard_normal((n, n))
        for x, y in zip(ref.compute_curl_2d(Az, 0.05, bc), dro.compute_curl_2d(Az, 0.05, bc)):
            assert np.array_equal(x, y)
    fin = np.ascontiguousarray(rng.uniform(0.2, 2.0, (12, 64)))
    assert np.array_equal(ref.compute_fluxes(fin, 1.4), dro.compute_fluxes(fin, 1.4))
