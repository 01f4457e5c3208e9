"""Snapshot files in the reference's HDF5 layout (src/mhd_write_h5.f90:12-177) and the
command line (src/mhd_init.f90:167-256), on the CPU.

The layout checks restate the reference's structural suite
(tests/structural/test_structural.py:18-119) and run on files written by
libchimera_h5.so directly and by the reference's own main program linked against it
(oracle/_ref/chimera_ref_main, built by oracle/build_ref.sh).
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from chimera_amd import cli, h5io

from tests.conftest import ref_env

ROOT = Path(__file__).resolve().parent.parent
REF_MAIN = ROOT / "oracle" / "_ref" / "chimera_ref_main"
ATHENA_H5 = Path("/root/reference/tests/validation/ot_reference.h5")
ATHENA_NPZ = ROOT / "tests" / "golden" / "ot_athena500_t05.npz"


def _fields(nsnap, ny, nx, seed=0):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal((nsnap, ny, nx)) for k in h5io.FIELDS}


def check_layout(path, n, nsnap=None, *, min_snaps=5):
    """The reference's structural checks on one file: every field group holds the same
    number of (n, n) snapshots, time/sim_time is as long and strictly increasing,
    the three parameters are present as {1} datasets."""
    counts = {f: h5io.group_count(path, f) for f in h5io.FIELDS}
    assert len(set(counts.values())) == 1, counts
    k = counts["rho"]
    if nsnap is not None:
        assert k == nsnap
    assert k >= min_snaps
    for f in h5io.FIELDS:
        assert h5io.dims(path, f"{f}/SNAPSHOT1") == (n, n)
        assert h5io.dims(path, f"{f}/SNAPSHOT{k}") == (n, n)
    assert h5io.dims(path, "time/sim_time") == (k,)
    t = h5io.read_dataset(path, "time/sim_time")
    assert np.all(np.diff(t) > 0)
    assert h5io.group_count(path, "parameters") == 3
    for p in h5io.PARAMS:
        assert h5io.dims(path, f"parameters/{p}") == (1,)
    return k


# ---- writer / reader ---------------------------------------------------------------

def test_round_trip_non_square(tmp_path):
    path = tmp_path / "snaps.h5"
    f = _fields(3, 5, 7)
    times = np.array([0.01, 0.02, 0.025])
    h5io.write_snapshots(path, f, times, 5.0 / 3.0, 1.25, 0.5)
    d = h5io.read(path)
    for k in h5io.FIELDS:
        assert len(d[k]) == 3
        for s in range(3):
            assert d[k][s].shape == (5, 7)     # C dims {ny, nx}: Fortran (nx, ny)
            assert np.array_equal(d[k][s], f[k][s])
    assert np.array_equal(d["time"], times)
    assert d["parameters"] == {"gamma": 5.0 / 3.0, "M_s": 1.25, "beta": 0.5}
    assert h5io.group_count(path, "time") == 1


def test_rewrite_truncates(tmp_path):
    path = tmp_path / "snaps.h5"
    h5io.write_snapshots(path, _fields(4, 4, 4), np.arange(1, 5) * 0.01, 1.4, 0.0, 0.0)
    h5io.write_snapshots(path, _fields(2, 4, 4, 1), np.array([0.5, 1.0]), 1.4, 0.0, 0.0)
    assert h5io.group_count(path, "rho") == 2


def test_errors_are_reported(tmp_path):
    with pytest.raises(h5io.H5Error, match="Error opening file"):
        h5io.write_snapshots(tmp_path / "no_such_dir" / "x.h5", _fields(1, 4, 4), [0.1], 1.4, 0, 0)
    path = tmp_path / "snaps.h5"
    h5io.write_snapshots(path, _fields(1, 4, 4), [0.1], 1.4, 0, 0)
    with pytest.raises(h5io.H5Error, match="no dataset"):
        h5io.dims(path, "rho/SNAPSHOT2")
    with pytest.raises(h5io.H5Error, match="no group"):
        h5io.group_count(path, "Bz")
    with pytest.raises(h5io.H5Error, match="cannot open"):
        h5io.dims(tmp_path / "missing.h5", "rho/SNAPSHOT1")


@pytest.mark.skipif(not ATHENA_H5.exists(), reason="the reference's validation file is not here")
def test_reads_a_file_from_another_writer():
    """The reference's Athena++ fixture (written by h5py) reads back as the committed
    extraction of it."""
    d = h5io.read(ATHENA_H5)
    g = np.load(ATHENA_NPZ)
    assert np.array_equal(d["rho"][0], g["rho"]) and np.array_equal(d["P"][0], g["P"])
    assert d["time"][0] == 0.5 and d["parameters"]["gamma"] == pytest.approx(5.0 / 3.0)


def test_athena_fixture_shape():
    g = np.load(ATHENA_NPZ)
    assert g["rho"].shape == (500, 500) and g["P"].shape == (500, 500)
    assert float(g["time"][0]) == 0.5
    assert 0.0 < g["rho"].min() < g["rho"].max() < 1.0 and g["P"].min() > 0.0


@pytest.mark.skipif(not Path("/opt/conda/bin/h5dump").exists(), reason="h5dump not available")
def test_writer_layout_equals_the_reference_fixture(tmp_path):
    """A one-snapshot 500 x 500 file from this writer has the structure of the reference's
    own snapshot-format fixture (tests/validation/ot_reference.h5, which its tests read
    like a CHIMERA output): the same groups and datasets, FP64 little-endian, the same
    dataspaces, contiguous storage without filters, the same fill and allocation
    properties (`h5dump -H -p`, recorded by tests/golden/make_ot_athena.py)."""
    sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
    from make_ot_athena import LAYOUT, h5_layout
    path = tmp_path / "one.h5"
    fields = {k: np.full((1, 500, 500), 1.0 + i) for i, k in enumerate(h5io.FIELDS)}
    h5io.write_snapshots(path, fields, np.array([0.5]), 5.0 / 3.0, 0.0, 0.0)
    assert h5_layout(path) == LAYOUT.read_text()


# ---- the reference's main program writing through libchimera_h5 ----------------------

@pytest.fixture(scope="module")
def ref_ot32(tmp_path_factory):
    if not REF_MAIN.exists():
        pytest.skip("oracle/_ref/chimera_ref_main is not built")
    out = tmp_path_factory.mktemp("refmain")
    r = subprocess.run([str(REF_MAIN), "1", "32", "0", f"{out}/", "ot32.h5"], env=ref_env(),
                       capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out / "ot32.h5", r.stdout


def test_reference_main_layout(ref_ot32):
    path, _ = ref_ot32
    assert check_layout(path, 32) == 100
    t = h5io.read_dataset(path, "time/sim_time")
    assert t[0] == pytest.approx(0.01, abs=1e-12) and t[-1] >= 1.0
    p = h5io.read(path)["parameters"]
    assert p == {"gamma": pytest.approx(5.0 / 3.0, rel=1e-15), "M_s": 0.0, "beta": 0.0}


def test_reference_main_stores_thermal_pressure(ref_ot32):
    """test_structural.py:83-119: the mean P of the first OT snapshot is near the
    thermal 5/(12 pi) ~ 0.133, below the total ~0.173."""
    path, _ = ref_ot32
    P0 = h5io.read_dataset(path, "P/SNAPSHOT1")
    assert P0.min() > 0.0 and P0.mean() < 0.16


def test_reference_main_progress_lines(ref_ot32):
    _, out = ref_ot32
    lines = [ln for ln in out.splitlines() if "avg(|divB|)" in ln]
    assert len(lines) == 100
    assert lines[0].startswith("  1%:  t =  0.010")
    assert lines[-1].startswith("100%:  t =  1.000")


# ---- the command line ----------------------------------------------------------------

def test_parse_defaults_and_optionals():
    o = cli.parse_args(["3", "64", "7"])
    assert o == {"problem_type": 3, "N": 64, "seed": 7, "out_path": "./outputs/",
                 "h5_filename": "primitive_snaps.h5", "target_M_s": -1.0, "target_beta": -1.0}
    o = cli.parse_args(["5", "32", "1", "  /tmp/x/ ", "a.h5", "5.0", "bad"])
    assert o["out_path"] == "/tmp/x/" and o["h5_filename"] == "a.h5"
    assert o["target_M_s"] == 5.0 and o["target_beta"] == -1.0


@pytest.mark.parametrize("argv,msg", [
    (["1", "32"], "ERROR: Usage: ./chimera <problem_type> <N> <seed>"),
    (["6", "32", "0"], "ERROR: problem_type must be 1-5, got: '6'"),
    (["x", "32", "0"], "ERROR: problem_type must be 1-5, got: 'x'"),
    (["1", "3", "0"], "ERROR: N must be an integer >= 4, got: '3'"),
    (["1", "32", "s"], "ERROR: Invalid seed: 's'"),
])
def test_parse_errors(argv, msg):
    with pytest.raises(cli.UsageError, match=msg.replace("(", r"\(").replace(")", r"\)")
                       .replace("[", r"\[").replace("]", r"\]").replace("*", r"\*")):
        cli.parse_args(argv)


def test_cli_usage_error_stops_like_the_reference():
    """An invalid command line prints the message and stops with status 0, as the
    reference's `stop` does, before anything touches a GPU."""
    r = subprocess.run([sys.executable, "-m", "chimera_amd.cli", "9", "32", "0"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert "ERROR: problem_type must be 1-5, got: '9'" in r.stdout


def test_executable_wrapper(tmp_path):
    """chimera_amd/bin/chimera, linked where the reference's drivers expect ./chimera."""
    link = tmp_path / "chimera"
    link.symlink_to(ROOT / "chimera_amd" / "bin" / "chimera")
    r = subprocess.run([str(link), "1", "2", "0"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "ERROR: N must be an integer >= 4, got: '2'" in r.stdout


@pytest.mark.skipif(not REF_MAIN.exists(), reason="oracle/_ref/chimera_ref_main is not built")
def test_cli_messages_match_reference_main():
    for argv in (["1", "32"], ["6", "32", "0"], ["1", "2", "0"], ["1", "32", "q"]):
        r = subprocess.run([str(REF_MAIN), *argv], capture_output=True, text=True, timeout=60,
                           env=ref_env())
        assert r.returncode == 0
        with pytest.raises(cli.UsageError) as e:
            cli.parse_args(argv)
        ref_lines = [ln.strip() for ln in r.stdout.splitlines() if ln.strip()]
        ours = [ln.strip() for ln in str(e.value).splitlines() if ln.strip()]
        assert ours == ref_lines
