"""The reference's end-to-end test suites, restated as checks on snapshot files.

Each check takes ``run(problem_type, N, seed=0, extra=None) -> path`` (a program that
writes the reference's HDF5 layout) and asserts what the reference's suite asserts
about its own binary:

  structural   tests/structural/test_structural.py:18-119
  physics      tests/physics/test_physics.py:40-142
  rk3          tests/physics/test_rk3.py:74-242
  stress       tests/stress/test_stress.py:76-236
  validation   tests/validation/test_validation.py:150-318 (N = 128 and 256)

The thresholds are the reference's.  Files are read with chimera_amd.h5io.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from chimera_amd import h5io

ATHENA_NPZ = Path(__file__).resolve().parent / "golden" / "ot_athena500_t05.npz"


def snaps(path, field: str) -> np.ndarray:
    n = h5io.group_count(path, field)
    return np.stack([h5io.read_dataset(path, f"{field}/SNAPSHOT{k + 1}") for k in range(n)])


def params(path) -> dict:
    return {p: float(h5io.read_dataset(path, f"parameters/{p}")[0]) for p in h5io.PARAMS}


def times(path) -> np.ndarray:
    return h5io.read_dataset(path, "time/sim_time")


def snapshot_near(path, field: str, t: float = 0.5, tol: float = 0.006) -> np.ndarray:
    tt = times(path)
    k = int(np.argmin(np.abs(tt - t)))
    assert abs(tt[k] - t) <= tol, f"no snapshot within {tol} of t={t} (closest {tt[k]})"
    return h5io.read_dataset(path, f"{field}/SNAPSHOT{k + 1}")


# ---- structural ---------------------------------------------------------------------

def structural_layout(run):
    path = run(1, 32)
    counts = {f: h5io.group_count(path, f) for f in h5io.FIELDS}
    assert len(set(counts.values())) == 1, counts
    k = counts["rho"]
    assert k >= 5
    for f in h5io.FIELDS:
        assert h5io.dims(path, f"{f}/SNAPSHOT1") == (32, 32)
    t = times(path)
    assert t.shape == (k,) and np.all(np.diff(t) > 0)


def structural_parameters_present(run):
    path = run(5, 32, seed=1)
    assert h5io.group_count(path, "parameters") == 3
    for p in h5io.PARAMS:
        assert h5io.dims(path, f"parameters/{p}") == (1,)


def structural_thermal_pressure(run):
    """OT: the first snapshot's mean P sits near the thermal 5/(12 pi) ~ 0.133, well
    below the total pressure's ~0.173."""
    P = snaps(run(1, 32), "P")
    assert P[0].min() > 0.0 and P[-1].min() > 0.0
    assert P[0].mean() < 0.16


# ---- physics ------------------------------------------------------------------------

def _mass_drift(path) -> float:
    m = snaps(path, "rho").sum(axis=(1, 2))
    return float(np.max(np.abs(m - m[0])) / m[0])


def physics_mass_conserved(run, problem_type):
    assert _mass_drift(run(problem_type, 32)) < 1e-10


def physics_positivity(run):
    ot = run(1, 32)
    assert snaps(ot, "rho").min() > 0.0 and snaps(ot, "P").min() > 0.0
    assert snaps(run(3, 32), "rho").min() > 0.0


def physics_kh_completes(run):
    assert h5io.group_count(run(2, 32), "rho") >= 5


def _loop_error(path) -> float:
    """L1 change of |B| between the first and the last snapshot over the mean |B| of
    the loop: the loop returns to its start at t = 1 (v = (-2, -1) on the unit box)."""
    Bx, By = snaps(path, "Bx"), snaps(path, "By")
    b0 = np.sqrt(Bx[0] ** 2 + By[0] ** 2)
    bf = np.sqrt(Bx[-1] ** 2 + By[-1] ** 2)
    inside = b0 > 1e-8
    scale = b0[inside].mean() if inside.any() else 1.0
    return float(np.mean(np.abs(bf - b0)) / scale)


def physics_field_loop_convergence(run):
    e32, e64 = _loop_error(run(3, 32)), _loop_error(run(3, 64))
    assert e64 < e32
    assert e32 / e64 >= 1.5, (e32, e64)


# ---- rk3 ----------------------------------------------------------------------------

def _energy(path) -> tuple[np.ndarray, np.ndarray]:
    g = params(path)["gamma"]
    assert g > 1.0
    rho, P = snaps(path, "rho"), snaps(path, "P")
    Bx, By, Vx, Vy = (snaps(path, f) for f in ("Bx", "By", "Vx", "Vy"))
    e = P / (g - 1.0) + 0.5 * rho * (Vx**2 + Vy**2) + 0.5 * (Bx**2 + By**2)
    return times(path), e.mean(axis=(1, 2))


def rk3_field_loop_no_density_spike(run):
    rho = snaps(run(3, 64), "rho")
    r0 = rho[0].max()
    assert rho.max() <= r0 * (1.0 + 1e-3)


def rk3_ot_pressure_floor(run):
    P = snaps(run(1, 64), "P")
    frac = np.max(np.mean(P <= 1e-11, axis=(1, 2)))
    assert frac < 1e-3


def rk3_ot_energy(run):
    t, E = _energy(run(1, 64))
    k = int(np.argmin(np.abs(t - 0.5)))
    assert abs(t[k] - 0.5) <= 0.02
    assert abs(E[k] - E[0]) / abs(E[0]) < 0.03
    early = E[t <= 0.2]
    assert len(early) >= 2
    assert max(0.0, float(np.max(np.diff(early)))) < 1e-6 * abs(E[0])


def rk3_mass_conserved(run, problem_type):
    assert _mass_drift(run(problem_type, 64)) < 1e-10


# ---- stress (random problem) ----------------------------------------------------------

def stress_completes(run, seed):
    assert h5io.group_count(run(5, 64, seed=seed), "rho") >= 5


def stress_parameters_plausible(run):
    p = params(run(5, 64, seed=42))
    assert np.isfinite(p["gamma"]) and p["gamma"] > 1.0
    assert np.isfinite(p["M_s"]) and 0.01 < p["M_s"] < 20.0
    assert np.isfinite(p["beta"]) and 0.01 < p["beta"] < 100.0


def stress_positive(run, extra=None):
    path = run(5, 64, seed=42, extra=extra)
    assert h5io.group_count(path, "rho") >= 5
    assert snaps(path, "rho").min() > 0.0 and snaps(path, "P").min() > 0.0
    return path


def stress_high_mach(run):
    path = stress_positive(run, (5.0, -1.0))
    for v in params(path).values():
        assert np.isfinite(v) and v > 0.0


def stress_low_beta(run):
    path = stress_positive(run, (-1.0, 0.1))
    assert params(path)["beta"] < 0.5


# ---- validation (Orszag-Tang at N = 128, 256) ------------------------------------------

OT_BOUNDS = {"rho_max": (0.42, 0.65), "rho_min": (0.03, 0.12),
             "P_max": (0.38, 0.72), "P_min": (0.008, 0.07)}


def validation_ot_bounds(run):
    path = run(1, 128)
    rho, P = snapshot_near(path, "rho"), snapshot_near(path, "P")
    got = {"rho_max": rho.max(), "rho_min": rho.min(), "P_max": P.max(), "P_min": P.min()}
    for k, (lo, hi) in OT_BOUNDS.items():
        assert lo < got[k] < hi, (k, got[k])


def _coarsen(a, f):
    m = a.shape[0] // f
    return a.reshape(m, f, m, f).mean(axis=(1, 3))


def _bilinear(a, n):
    """Cell-centred bilinear interpolation of a square field onto an n x n grid of the
    unit box (numpy.interp along x, then along y)."""
    src = (np.arange(a.shape[0]) + 0.5) / a.shape[0]
    dst = (np.arange(n) + 0.5) / n
    rows = np.stack([np.interp(dst, src, r) for r in a])
    return np.stack([np.interp(dst, src, rows[:, j]) for j in range(n)], axis=1)


def _l2(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def validation_convergence(run):
    ns = (64, 128, 256)
    fields = {n: {f: snapshot_near(run(1, n), f) for f in ("rho", "P")} for n in ns}
    for f in ("rho", "P"):
        fine = fields[256][f]
        e64 = _l2(fields[64][f], _coarsen(fine, 4))
        e128 = _l2(fields[128][f], _coarsen(fine, 2))
        assert e128 < e64
        assert np.log2(e64 / e128) > 1.0, (f, e64, e128)
    ath = np.load(ATHENA_NPZ)
    for f in ("rho", "P"):
        errs = [_l2(fields[n][f], _bilinear(ath[f], n)) for n in ns]
        assert errs[0] > errs[1] > errs[2], (f, errs)


CHECKS = [
    ("structural_layout", structural_layout, ()),
    ("structural_parameters_present", structural_parameters_present, ()),
    ("structural_thermal_pressure", structural_thermal_pressure, ()),
    ("physics_mass_ot", physics_mass_conserved, (1,)),
    ("physics_mass_field_loop", physics_mass_conserved, (3,)),
    ("physics_positivity", physics_positivity, ()),
    ("physics_kh_completes", physics_kh_completes, ()),
    ("physics_field_loop_convergence", physics_field_loop_convergence, ()),
    ("rk3_field_loop_no_density_spike", rk3_field_loop_no_density_spike, ()),
    ("rk3_ot_pressure_floor", rk3_ot_pressure_floor, ()),
    ("rk3_ot_energy", rk3_ot_energy, ()),
    ("rk3_mass_ot", rk3_mass_conserved, (1,)),
    ("rk3_mass_field_loop", rk3_mass_conserved, (3,)),
    ("rk3_mass_kh", rk3_mass_conserved, (2,)),
    ("stress_completes_42", stress_completes, (42,)),
    ("stress_completes_43", stress_completes, (43,)),
    ("stress_parameters_plausible", stress_parameters_plausible, ()),
    ("stress_positive", stress_positive, ()),
    ("stress_high_mach", stress_high_mach, ()),
    ("stress_low_beta", stress_low_beta, ()),
]
# N = 128 and 256 runs: minutes of CPU for the reference's main, seconds on the GPU
SLOW_CHECKS = [
    ("validation_ot_bounds", validation_ot_bounds, ()),
    ("validation_convergence", validation_convergence, ()),
]


def ids(checks):
    return [c[0] for c in checks]


def apply(check, run):
    _, fn, args = check
    return fn(run, *args)

