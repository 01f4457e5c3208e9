"""CPU checks of the C ABI: the library loads, exports every function declared in
include/*.h, and rejects bad arguments before touching a GPU."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


# the library that implements each header's declarations
LIBRARY_OF = {"chimera_hip.h": "libchimera_hip.so", "chimera_h5.h": "libchimera_h5.so"}


def declared_functions(headers=None) -> set[str]:
    names = set()
    for h in headers or HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(chimera_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def L():
    from chimera_amd import _lib
    return _lib.lib()


def test_headers_declare_api():
    names = declared_functions()
    assert {"chimera_compute_fluxes", "chimera_ctx_create", "chimera_ctx_steps"} <= names


def test_every_header_has_a_library():
    assert {h.name for h in HEADERS} == set(LIBRARY_OF)


@pytest.mark.parametrize("header", sorted(LIBRARY_OF))
def test_library_exports_every_declared_symbol(header):
    lib = ROOT / "chimera_amd" / LIBRARY_OF[header]
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = declared_functions([ROOT / "include" / header])
    assert declared
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    handle = C.CDLL(str(lib))
    for name in declared:
        getattr(handle, name)     # resolves through dlsym


def test_headers_compile_as_c_and_link(tmp_path):
    """Both headers are plain C (C99, pedantic, no warnings): a C caller includes them
    and links against the two libraries."""
    src = tmp_path / "caller.c"
    src.write_text("#include \"chimera_hip.h\"\n#include \"chimera_h5.h\"\n"
                   "int main(void) {\n"
                   "    double dx, g; chimera_bc_t bc; chimera_ctx_t *ctx = 0;\n"
                   "    (void)ctx;\n"
                   "    return chimera_problem_info(1, 64, &dx, &g, &bc);\n}\n")
    exe = tmp_path / "caller"
    lib = ROOT / "chimera_amd"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                        f"-I{ROOT / 'include'}", str(src), f"-L{lib}", "-lchimera_hip",
                        "-lchimera_h5", f"-Wl,-rpath,{lib}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # chimera_problem_info is host-only: the C caller runs here without a GPU
    assert subprocess.run([str(exe)], capture_output=True, timeout=60).returncode == 0


def test_invalid_arguments_rejected_on_host(L):
    from chimera_amd._lib import BC, Scheme
    h = C.c_void_p()
    # grid too small
    rc = L.chimera_ctx_create(C.byref(h), 2, 8, 0.1, 1.4, C.byref(BC(1, 1, 1, 1)),
                              C.byref(Scheme(1, 2, 1, 1)), None)
    assert rc == -1 and b"at least 4" in L.chimera_last_error()
    # bad solver / limiter / flags
    assert L.chimera_ctx_create(C.byref(h), 8, 8, 0.1, 1.4, C.byref(BC(1, 1, 1, 1)),
                                C.byref(Scheme(4, 2, 1, 1)), None) == -1
    assert L.chimera_ctx_create(C.byref(h), 8, 8, 0.1, 1.4, C.byref(BC(1, 1, 1, 1)),
                                C.byref(Scheme(1, 3, 1, 1)), None) == -1
    assert L.chimera_ctx_create(C.byref(h), 8, 8, 0.1, 1.4, C.byref(BC(0, 1, 1, 1)),
                                C.byref(Scheme(1, 2, 1, 1)), None) == -1
    # periodic on one side of an axis only
    assert L.chimera_ctx_create(C.byref(h), 8, 8, 0.1, 1.4, C.byref(BC(1, 3, 1, 1)),
                                C.byref(Scheme(1, 2, 1, 1)), None) == -4
    # non-physical gamma / dx
    assert L.chimera_ctx_create(C.byref(h), 8, 8, 0.1, 1.0, C.byref(BC(1, 1, 1, 1)),
                                C.byref(Scheme(1, 2, 1, 1)), None) == -1
    # drop-ins validate before allocating
    z = np.zeros(16)
    p = z.ctypes.data_as(C.POINTER(C.c_double))
    assert L.chimera_compute_fluxes(*([p] * 12), C.c_double(1.4), 4, 4, *([p] * 5), 7) == -1
    assert L.chimera_apply_slope_limiter(p, C.c_double(0.1), 4, 4, p, p, 0) == -1
    assert L.chimera_update_conserved(p, p, p, C.c_double(0.1), C.c_double(0.1), 1, 4,
                                      C.byref(BC(1, 1, 1, 1))) == -1


def test_adjoint_entry_points_reject_invalid_arguments_on_host(L):
    """The sweep, batched-sweep, autotune and tile entry points validate their
    arguments before touching the device."""
    from chimera_amd._lib import BC, Scheme
    ptrs = (C.c_void_p * 6)()
    dts = (C.c_double * 4)(1e-3, 1e-3, 1e-3, 1e-3)
    assert L.chimera_adj_sweep(None, ptrs, dts, 4, 2, ptrs) == -1
    assert b"NULL" in L.chimera_last_error()
    assert L.chimera_badj_sweep(None, ptrs, dts, 4, 2, ptrs) == -1
    assert L.chimera_badj_autotune(None, ptrs, dts, 1, None, None) == -1
    assert L.chimera_badj_set_tiles(None, 0, 0, 0) == -1
    h = C.c_void_p()
    n = (C.c_int * 1)(8)
    d = (C.c_double * 1)(0.1)
    g = (C.c_double * 1)(1.4)
    assert L.chimera_badj_create(C.byref(h), 0, n, n, d, g, C.byref(BC(1, 1, 1, 1)),
                                 C.byref(Scheme(3, 2, 1, 1)), None) == -1
    assert b"members" in L.chimera_last_error()


def test_scheme_and_bc_structs():
    from chimera_amd._lib import BC, Scheme
    assert C.sizeof(BC) == 16 and C.sizeof(Scheme) == 16
    s = Scheme.of((3, 1))
    assert (s.riemann_solver, s.slope_limiter, s.use_slope_limiting, s.fallback_2_mood) == (3, 1, 1, 1)
    with pytest.raises(ValueError):
        Scheme.of((1, 2, 3))


def test_state_pointers_reject_bad_buffers():
    """The Python binding checks shape, dtype and layout of every state it hands to the
    library (which copies nx * ny doubles per field) before any copy."""
    import torch
    from chimera_amd._lib import state_ptrs
    good = np.zeros((6, 8, 12))
    assert len(state_ptrs(good, (6, 8, 12))) == 6
    with pytest.raises(ValueError, match="shape"):
        state_ptrs(np.zeros((6, 8, 11)), (6, 8, 12))
    with pytest.raises(ValueError, match="shape"):
        state_ptrs(np.zeros((5, 8, 12)))
    with pytest.raises(ValueError, match="float64"):
        state_ptrs(np.zeros((6, 8, 12), dtype=np.float32), (6, 8, 12))
    with pytest.raises(ValueError, match="contiguous"):
        state_ptrs(np.zeros((6, 12, 8)).transpose(0, 2, 1), (6, 8, 12))
    with pytest.raises(ValueError, match="CUDA"):
        state_ptrs(torch.zeros((6, 8, 12), dtype=torch.float64), (6, 8, 12))
    with pytest.raises(TypeError):
        state_ptrs([[0.0]], (6, 8, 12))


def test_python_entry_points_reject_bad_arrays(tmp_path):
    """The drop-in, snapshot and adjoint helpers check the arrays the library would read
    or write in full before calling it (and without assert, which -O strips)."""
    from chimera_amd import adjoint, dropin, h5io
    with pytest.raises(ValueError, match="contiguous"):
        dropin._p(np.zeros((8, 12)).T)
    with pytest.raises(ValueError, match="float64"):
        dropin._p(np.zeros((8, 12), dtype=np.float32))
    faces = np.zeros((24, 8, 12))
    with pytest.raises(ValueError, match=r"cells of shape \(6, 7, 12\)"):
        dropin.thermal_pressure_check(faces, np.zeros((6, 7, 12)))
    with pytest.raises(ValueError, match="faces"):
        dropin.thermal_pressure_check(np.zeros((23, 8, 12)), np.zeros((6, 8, 12)))
    f, short = np.zeros((8, 12)), np.zeros((8, 11))
    bc = (1, 1, 1, 1)
    for call, what in ((lambda: dropin.fill_ghost_cells(f, bc, short), "f_ambient"),
                       (lambda: dropin.reconstruction(f, f, short, 0.1, bc, 0), "f_dy"),
                       (lambda: dropin.update_conserved(f, short, f, 0.1, 0.1, bc), "flux_F_X"),
                       (lambda: dropin.constrained_transport(f, short, f, f, 0.1, 0.1, bc), "by"),
                       (lambda: dropin.compute_divB(f, short, 0.1, bc), "by"),
                       (lambda: dropin.average_face_to_cell_B(f, short, bc), "b_y"),
                       (lambda: dropin.compute_gradients(np.zeros(10), 0.1), "f_pad"),
                       (lambda: dropin.get_conserved(np.zeros((5, 8, 12)), 1.4, 0.1), "prim"),
                       (lambda: dropin.get_primitive(np.zeros((4, 8, 12)), f, short, 1.4, 0.1), "By"),
                       (lambda: dropin.apply_conserved_floors(np.zeros((3, 8, 12)), f, f, 0.1, 1.4),
                        "cons"),
                       (lambda: dropin.compute_fluxes(np.zeros((11, 20)), 1.4, 1), "fin"),
                       (lambda: dropin.stage_fluxes(np.zeros((5, 8, 12)), 0.1, 1.4, bc, (1, 2)),
                        "prim")):
        with pytest.raises(ValueError, match=what):
            call()
    fields = {k: np.zeros((2, 8, 12)) for k in h5io.FIELDS}
    fields["Vy"] = np.zeros((2, 8, 11))
    with pytest.raises(ValueError, match="field Vy"):
        h5io.write_snapshots(tmp_path / "x.h5", fields, np.zeros(2), 1.4, 0.0, 0.0)
    fields["Vy"] = np.zeros((2, 8, 12))
    with pytest.raises(ValueError, match="times"):
        h5io.write_snapshots(tmp_path / "x.h5", fields, np.zeros(3), 1.4, 0.0, 0.0)
    assert not (tmp_path / "x.h5").exists()
    badj = adjoint.BatchAdjoint.__new__(adjoint.BatchAdjoint)   # no device context
    badj.members, badj.shapes, badj._h = 3, [(8, 8)] * 3, None
    with pytest.raises(ValueError, match="2 states for 3 members"):
        badj._member_ptrs([None, None], "states")
    assert adjoint._dt_table(np.ones((3, 5)), 3).shape == (3, 5)
    with pytest.raises(ValueError, match="members = 4"):
        adjoint._dt_table(np.ones((3, 5)), 4)
    with pytest.raises(ValueError, match="members = 3"):
        adjoint._dt_table(np.ones(3), 3)


def test_prescribed_steps_need_enough_dts(monkeypatch):
    """Grid.steps rejects a short dt sequence before calling the library."""
    from chimera_amd import solver
    g = solver.Grid.__new__(solver.Grid)   # no device context: the check comes first
    g._L = None
    with pytest.raises(ValueError, match="dts"):
        g.steps(5, dt_mode=solver.DT_PRESCRIBED, dt=np.ones(3))
    with pytest.raises(ValueError):
        g.steps(5, dt_mode=solver.DT_PRESCRIBED, dt=None)
