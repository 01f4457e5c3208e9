"""The problem set-ups against the reference's own set-up code (src/mhd_init.f90:296-450).

The reference's main program (oracle/_ref/chimera_ref_main) writes its first snapshot
at t = 0.01; the reference modules (oracle/_ref, the same Fortran) stepped from this
package's initial state to the same time with the same dt policy must reproduce that
snapshot bit for bit.  A last-bit difference in the initial state does not survive
this: MOOD and limiter ties amplify it (the Kelvin-Helmholtz set-up with numpy's exp
in place of libm's misses by 2e-4).
"""
from __future__ import annotations

import numpy as np
import pytest

from chimera_amd import h5io, problems
from oracle.oracle import RefVariant, ref_available

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref reference build not present")

SETUPS = {1: problems.orszag_tang, 2: problems.kelvin_helmholtz, 3: problems.field_loop_unit,
          4: problems.rotor}


def first_snapshot(prob):
    ref = RefVariant(1, 2, "ref")      # the reference's mhd_config scheme: Rusanov, MC
    st = prob.state.copy()
    r = ref.run(st, prob.dx, prob.gamma, prob.bc, 100000, dt_mode=0, t_end=0.01)
    Bx, By = ref.average_face_to_cell_B(st[4], st[5], prob.bc)
    pr = ref.get_primitive(st[:4], Bx, By, prob.gamma, prob.dx * prob.dx)
    return r["t"], {"rho": pr[0], "Vx": pr[1], "Vy": pr[2], "P": pr[3] - 0.5 * (Bx * Bx + By * By),
                    "Bx": Bx, "By": By}


@pytest.mark.parametrize("problem_type,n", [(1, 32), (1, 40), (2, 32), (2, 48), (3, 32), (3, 40),
                                            (4, 32), (4, 48)])
def test_setup_matches_reference_setup(problem_type, n, ref_main_run):
    d = h5io.read(ref_main_run(problem_type, n))
    prob = SETUPS[problem_type](n, scheme=(1, 2, 1, 1))
    t, snap = first_snapshot(prob)
    assert t == d["time"][0]
    for f in h5io.FIELDS:
        assert np.array_equal(snap[f], d[f][0]), f
    assert d["parameters"]["gamma"] == prob.gamma


def test_setup_boundaries():
    assert problems.orszag_tang(8).bc == problems.PERIODIC
    assert problems.kelvin_helmholtz(8).bc == problems.PERIODIC
    assert problems.field_loop_unit(8).bc == problems.PERIODIC
    assert problems.rotor(8).bc == problems.OUTFLOW
