"""GPU parity of the batched ensemble and of the tile-shape variants: every member of
a batch advanced by the persistent per-stage launches must equal the oracle's run of
that member alone (per-member dt), bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from chimera_amd import problems

pytestmark = pytest.mark.gpu


def _members():
    sch = (1, 2, 1, 1)
    probs = [problems.grf_member(n, 1000 + k, scheme=sch) for k, n in enumerate([32, 64, 48, 96])]
    ot = problems.orszag_tang(40, scheme=sch)
    rot = problems.rotor(36, scheme=sch)
    return probs + [ot, rot]


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("shape,bpc", [(0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (4, 1), (0, 64)])
def test_batch_members_match_oracle(gpu_lib, orc, graph, shape, bpc):
    """bpc: blocks per CU of the stage launch (0: the default, 1: one resident block per
    CU walking many tiles, 64: far more blocks than fit at once)."""
    from chimera_amd.batch import Ensemble, DT_CFL
    probs = _members()
    print(f"batch graph={graph} shape={shape} bpc={bpc}", flush=True)
    with Ensemble(probs) as ens:
        ens.set_tiling(shape, bpc)
        ens.steps(4, dt_mode=DT_CFL, graph=graph)
        outs = [ens.download(m) for m in range(len(probs))]
        times = [ens.member_time(m) for m in range(len(probs))]
    for m, p in enumerate(probs):
        cpu = p.state.copy()
        r = orc.run(cpu, p.dx, p.gamma, p.bc, list(p.scheme), 4, dt_mode=1)
        assert np.array_equal(outs[m], cpu), f"member {m} ({p.name})"
        assert times[m][0] == r["t"] and times[m][2] == r["dt"][-1]


def test_batch_reference_dt_policy_and_autotune(gpu_lib, orc):
    from chimera_amd.batch import Ensemble, DT_REFERENCE
    probs = _members()[:4]
    with Ensemble(probs) as ens:
        best = ens.autotune(reps=2)
        assert 0 <= best[0] <= 4 and best[1] >= 1
        for m, p in enumerate(probs):                 # autotuning leaves the state untouched
            assert np.array_equal(ens.download(m), p.state)
        ens.steps(3, dt_mode=DT_REFERENCE, graph=True)
        outs = [ens.download(m) for m in range(len(probs))]
    for m, p in enumerate(probs):
        cpu = p.state.copy()
        orc.run(cpu, p.dx, p.gamma, p.bc, list(p.scheme), 3, dt_mode=0)
        assert np.array_equal(outs[m], cpu)


@pytest.mark.parametrize("shape,scheme", [(1, (3, 1, 1, 1)), (2, (3, 1, 1, 1)),
                                          (3, (2, 1, 1, 1)), (3, (1, 2, 1, 1)),
                                          (4, (2, 1, 1, 1)), (4, (3, 2, 1, 1)),
                                          (4, (1, 2, 0, 0))])
def test_single_grid_tile_shapes(gpu_lib, shape, scheme):
    """Every tile shape gives the default shape's result bit for bit (72^2: partial
    tiles; the 32 x 32 and 32 x 16 shapes on a grid smaller than three tiles per side)."""
    from chimera_amd.solver import Grid
    prob = problems.rotor(72, scheme=scheme)
    with Grid.from_problem(prob) as g:
        g.steps(3, dt_mode=1)
        ref = g.download()
    with Grid.from_problem(prob) as g:
        g.set_tile_shape(shape)
        assert g.tile_shape()[0] == shape
        g.steps(3, dt_mode=1)
        out = g.download()
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("make,scheme", [("rotor", (2, 1, 1, 1)), ("orszag_tang", (1, 2, 1, 1)),
                                         ("rotor", (3, 2, 1, 1))])
def test_default_tile_on_large_grids_matches_oracle(gpu_lib, orc, make, scheme):
    """A 512^2 grid has a 32 x 16 tile per CU, so a new context runs Rusanov and HLLE
    on 32 x 16 tiles (HLLD on 16 x 16); two CFL steps equal the oracle's bit for bit."""
    from chimera_amd.solver import Grid
    prob = getattr(problems, make)(512, scheme=scheme)
    print(f"default tile {make} scheme={scheme} 512^2", flush=True)
    with Grid.from_problem(prob) as g:
        assert g.tile_shape()[0] == (1 if scheme[0] == 3 else 4)
        g.steps(2, dt_mode=1)
        out = g.download()
    ref = prob.state.copy()
    orc.run(ref, prob.dx, prob.gamma, prob.bc, list(prob.scheme), 2, dt_mode=1)
    assert np.array_equal(out, ref)


def test_batch_large_members_autotuned_match_oracle(gpu_lib, orc):
    """Members of the benchmark's sizes (512^2 and 256^2 GRF, Rusanov / MC) at the tile
    and blocks-per-CU count the autotuner picks: two CFL steps equal the oracle's run
    of each member alone, bit for bit."""
    from chimera_amd.batch import Ensemble, DT_CFL
    probs = [problems.grf_member(n, 3000 + k, scheme=(1, 2, 1, 1)) for k, n in enumerate([512, 256])]
    print("batch large members 512^2 + 256^2, autotuned", flush=True)
    with Ensemble(probs) as ens:
        shape, bpc = ens.autotune(reps=1)
        print(f"autotuned tile {shape} blocks/CU {bpc}", flush=True)
        ens.steps(2, dt_mode=DT_CFL, graph=True)
        outs = [ens.download(m) for m in range(len(probs))]
    for m, p in enumerate(probs):
        ref = p.state.copy()
        orc.run(ref, p.dx, p.gamma, p.bc, list(p.scheme), 2, dt_mode=1)
        assert np.array_equal(outs[m], ref), f"member {m}"


def test_single_grid_autotune_keeps_state(gpu_lib):
    from chimera_amd.solver import Grid
    prob = problems.orszag_tang(128)
    with Grid.from_problem(prob) as g:
        best, ms = g.autotune(reps=2)
        assert 0 <= best <= 4 and len(ms) == 5 and all(t > 0 for t in ms)
        assert np.array_equal(g.download(), prob.state)


def test_batch_hlld_tiles(gpu_lib, orc):
    """HLLD members: the 32 x 16 tile matches the oracle; 32 x 32 is refused."""
    from chimera_amd._lib import ChimeraError
    from chimera_amd.batch import Ensemble, DT_CFL
    probs = [problems.grf_member(n, 2000 + k, scheme=(3, 2, 1, 1)) for k, n in enumerate([40, 72])]
    with Ensemble(probs) as ens:
        with pytest.raises(ChimeraError):
            ens.set_tiling(3, 0)
        ens.set_tiling(4, 0)
        ens.steps(3, dt_mode=DT_CFL)
        outs = [ens.download(m) for m in range(len(probs))]
    for m, p in enumerate(probs):
        cpu = p.state.copy()
        orc.run(cpu, p.dx, p.gamma, p.bc, list(p.scheme), 3, dt_mode=1)
        assert np.array_equal(outs[m], cpu), f"member {m}"


def test_large_tile_unavailable_for_hlld(gpu_lib):
    from chimera_amd._lib import ChimeraError
    from chimera_amd.solver import Grid
    prob = problems.orszag_tang(64, scheme=(3, 2, 1, 1))
    with Grid.from_problem(prob) as g:
        with pytest.raises(ChimeraError):
            g.set_tile_shape(3)
        best, ms = g.autotune(reps=1)
        assert best != 3 and ms[3] == -1.0


@pytest.mark.parametrize("graph", [False, True])
def test_cfl_after_reupload_and_set_time(gpu_lib, orc, graph):
    """CFL steps after a re-upload and a set_time reduce the new state (eager and
    graph replay): the run equals the oracle's on the uploaded state bit for bit."""
    from chimera_amd.solver import Grid
    a = problems.orszag_tang(64, scheme=(1, 2, 1, 1))
    b = problems.rotor(64, scheme=(1, 2, 1, 1))
    b = problems.Problem("rotor_as_ot", b.state, a.dx, a.gamma, a.bc, a.scheme)
    with Grid.from_problem(a) as g:
        g.steps(3, dt_mode=1, graph=graph)
        g.upload(b.state)
        g.steps(3, dt_mode=1, graph=graph)
        g.set_time(0.0, 1)
        g.steps(2, dt_mode=1, graph=graph)
        out = g.download()
    ref = b.state.copy()
    orc.run(ref, b.dx, b.gamma, b.bc, list(b.scheme), 5, dt_mode=1)
    assert np.array_equal(out, ref)
