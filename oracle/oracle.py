"""ctypes loaders for the CPU oracle -- TEST INFRASTRUCTURE ONLY.

``COracle`` wraps ``oracle/liborc.so`` (the C restatement of the reference
hot path, ``oracle/chimera_oracle.c``).  ``RefVariant`` wraps one
``oracle/_ref/libchimera_ref_s<solver>_l<limiter>.so`` -- the reference Fortran
modules compiled by ``oracle/build_ref.sh`` with the bind(C) harness
``oracle/ref_harness.f90``.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module, and only as the checker or the reported CPU baseline.

Arrays are numpy float64, C-contiguous, shaped ``(ny, nx)`` (== Fortran
``(nx, ny)``); stacks of fields are ``(k, ny, nx)``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBORC = HERE / "liborc.so"
REF_DIR = HERE / "_ref"

BC_PERIODIC, BC_FIXED, BC_OUTFLOW, BC_INFLOW = 1, 2, 3, 4
VAN_LEER, MC = 1, 2
RUSANOV, HLLE, HLLD = 1, 2, 3
DT_REFERENCE, DT_CFL, DT_PRESCRIBED = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint)


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _i(seq):
    a = (C.c_int * len(seq))(*seq)
    return a


def _ptrs(arrs):
    return (_dp * len(arrs))(*[_d(a) for a in arrs])


class COracle:
    """The C restatement (oracle/chimera_oracle.c)."""

    def __init__(self, path: Path | str = LIBORC):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = C.CDLL(str(path))
        L = self.lib
        L.orc_cfl_max_speed.restype = C.c_double
        L.orc_run.restype = C.c_int
        L.orc_stage_fluxes.restype = C.c_int
        L.orc_num_threads.restype = C.c_int

    def num_threads(self) -> int:
        return self.lib.orc_num_threads()

    # ---- per-routine restatements -------------------------------------
    def fill_ghost_cells(self, f, bc, f_amb=None):
        ny, nx = f.shape
        pad = np.empty((ny + 2, nx + 2))
        self.lib.orc_fill_ghost_cells(nx, ny, _i(bc), _d(f), _d(pad),
                                      _d(f_amb) if f_amb is not None else None)
        return pad

    def compute_gradients(self, f_pad, dx):
        ny, nx = f_pad.shape[0] - 2, f_pad.shape[1] - 2
        gx, gy = np.empty((ny, nx)), np.empty((ny, nx))
        self.lib.orc_compute_gradients(nx, ny, _d(f_pad), C.c_double(dx), _d(gx), _d(gy))
        return gx, gy

    def apply_slope_limiter(self, f_pad, dx, limiter):
        ny, nx = f_pad.shape[0] - 2, f_pad.shape[1] - 2
        gx, gy = np.empty((ny, nx)), np.empty((ny, nx))
        self.lib.orc_apply_slope_limiter(nx, ny, limiter, _d(f_pad), C.c_double(dx), _d(gx), _d(gy))
        return gx, gy

    def reconstruction(self, f, f_dx, f_dy, dx, bc, rho_or_p, mood=True, field=0, mask=None):
        ny, nx = f.shape
        out = [np.empty((ny, nx)) for _ in range(4)]
        self.lib.orc_reconstruction(nx, ny, _i(bc), int(mood), _d(f), _d(f_dx), _d(f_dy),
                                    C.c_double(dx), *[_d(o) for o in out], int(rho_or_p),
                                    mask.ctypes.data_as(_up) if mask is not None else None, field)
        return out

    def thermal_pressure_check(self, faces, cells, mask=None):
        """faces: (24, ny, nx) in the reference argument order P,Bx,By,rho,vx,vy
        x (XL,XR,YL,YR), modified in place; cells: (6, ny, nx) rho,vx,vy,P,Bx,By."""
        n = faces.shape[1] * faces.shape[2]
        fp = _ptrs([faces[k] for k in range(24)])
        cp = _ptrs([cells[k] for k in range(6)])
        self.lib.orc_thermal_pressure_check(n, fp, cp,
                                            mask.ctypes.data_as(_up) if mask is not None else None)

    def compute_fluxes(self, fin, gamma, solver):
        """fin: (12, ...) -> (5, ...)."""
        shape = fin.shape[1:]
        n = int(np.prod(shape))
        out = np.empty((5,) + shape)
        self.lib.orc_compute_fluxes(n, solver, C.c_double(gamma),
                                    _ptrs([fin[k] for k in range(12)]),
                                    _ptrs([out[k] for k in range(5)]))
        return out

    def update_conserved(self, F, fX, fY, dx, dt, bc):
        ny, nx = F.shape
        F = F.copy()
        self.lib.orc_update_conserved(nx, ny, _i(bc), _d(F), _d(fX), _d(fY), C.c_double(dx),
                                      C.c_double(dt))
        return F

    def constrained_transport(self, bx, by, fByX, fBxY, dx, dt, bc):
        ny, nx = bx.shape
        bx, by = bx.copy(), by.copy()
        self.lib.orc_constrained_transport(nx, ny, _i(bc), _d(bx), _d(by), _d(fByX), _d(fBxY),
                                           C.c_double(dx), C.c_double(dt))
        return bx, by

    def compute_curl_2d(self, Az, dx, bc):
        ny, nx = Az.shape
        bx, by = np.empty((ny, nx)), np.empty((ny, nx))
        self.lib.orc_compute_curl_2d(nx, ny, _i(bc), _d(Az), C.c_double(dx), _d(bx), _d(by))
        return bx, by

    def compute_divB(self, bx, by, dx, bc):
        ny, nx = bx.shape
        d = np.empty((ny, nx))
        self.lib.orc_compute_divB(nx, ny, _i(bc), _d(bx), _d(by), C.c_double(dx), _d(d))
        return d

    def average_face_to_cell_B(self, b_x, b_y, bc):
        ny, nx = b_x.shape
        Bx, By = np.empty((ny, nx)), np.empty((ny, nx))
        self.lib.orc_average_face_to_cell_B(nx, ny, _i(bc), _d(b_x), _d(b_y), _d(Bx), _d(By))
        return Bx, By

    def get_conserved(self, prim, gamma, vol):
        """prim (6, ...) rho,vx,vy,P,Bx,By -> (4, ...) Mass,Momx,Momy,Energy."""
        n = int(np.prod(prim.shape[1:]))
        out = np.empty((4,) + prim.shape[1:])
        self.lib.orc_get_conserved(n, *[_d(prim[k]) for k in range(6)], C.c_double(gamma),
                                   C.c_double(vol), *[_d(out[k]) for k in range(4)])
        return out

    def get_primitive(self, cons, Bx, By, gamma, vol):
        n = int(np.prod(cons.shape[1:]))
        out = np.empty((4,) + cons.shape[1:])
        self.lib.orc_get_primitive(n, *[_d(cons[k]) for k in range(4)], _d(Bx), _d(By),
                                   C.c_double(gamma), C.c_double(vol),
                                   *[_d(out[k]) for k in range(4)])
        return out

    def apply_conserved_floors(self, cons, Bx, By, vol, gamma):
        n = int(np.prod(cons.shape[1:]))
        cons = cons.copy()
        self.lib.orc_apply_conserved_floors(n, *[_d(cons[k]) for k in range(4)], _d(Bx), _d(By),
                                            C.c_double(vol), C.c_double(gamma))
        return cons

    def cfl_max_speed(self, prim, gamma):
        n = int(np.prod(prim.shape[1:]))
        return self.lib.orc_cfl_max_speed(n, C.c_double(gamma), *[_d(prim[k]) for k in range(6)])

    def stage_fluxes(self, prim, dx, gamma, bc, scheme, want_mask=False):
        _, ny, nx = prim.shape
        flux = np.empty((10, ny, nx))
        mask = np.zeros((ny, nx), dtype=np.uint32) if want_mask else None
        rc = self.lib.orc_stage_fluxes(nx, ny, _i(bc), _i(scheme), C.c_double(gamma),
                                       C.c_double(dx), _ptrs([prim[k] for k in range(6)]),
                                       _ptrs([flux[k] for k in range(10)]),
                                       mask.ctypes.data_as(_up) if mask is not None else None)
        assert rc == 0
        return (flux, mask) if want_mask else flux

    def run(self, state, dx, gamma, bc, scheme, nsteps, dt_mode=DT_REFERENCE, dt_in=None,
            t=0.0, out_count=1, t_end=0.0, want_mask=False):
        """Advance state (6, ny, nx) in place; returns a dict of diagnostics."""
        _, ny, nx = state.shape
        assert state.flags.c_contiguous and state.dtype == np.float64
        dt_out = np.zeros(max(nsteps, 1))
        divb = np.zeros((max(nsteps, 1), 2))
        t_c, oc = C.c_double(t), C.c_int(out_count)
        dti = np.ascontiguousarray(dt_in, dtype=np.float64) if dt_in is not None else None
        mask = np.zeros((ny, nx), dtype=np.uint32) if want_mask else None
        taken = self.lib.orc_run(nx, ny, _i(bc), _i(scheme), C.c_double(gamma), C.c_double(dx),
                                 _ptrs([state[k] for k in range(6)]), nsteps, dt_mode,
                                 _d(dti) if dti is not None else None, _d(dt_out), C.byref(t_c),
                                 C.byref(oc), C.c_double(t_end), _d(divb),
                                 mask.ctypes.data_as(_up) if mask is not None else None)
        if taken < 0:
            raise RuntimeError(f"orc_run failed ({taken})")
        return {"steps": taken, "dt": dt_out[:taken], "divb": divb[:taken], "t": t_c.value,
                "out_count": oc.value, "mask": mask}


def on_big_stack(fn, *args, stack_bytes: int = 4 << 30, **kw):
    """Run fn on a thread with a large stack.  The reference keeps its N x N work arrays
    as automatic (stack) arrays, which overflow the default 8 MiB stack beyond ~512^2."""
    out, err = {}, {}

    def body():
        try:
            out["v"] = fn(*args, **kw)
        except BaseException as e:  # re-raised in the caller
            err["e"] = e

    old = threading.stack_size()
    threading.stack_size(stack_bytes)
    try:
        th = threading.Thread(target=body)
        th.start()
        th.join()
    finally:
        threading.stack_size(old)
    if "e" in err:
        raise err["e"]
    return out.get("v")


def ref_available() -> bool:
    return all((REF_DIR / f"libchimera_ref_s{s}_l{l}.so").exists()
               for s in (1, 2, 3) for l in (1, 2))


def dropin_available() -> bool:
    return all((REF_DIR / f"libchimera_dropin_s{s}_l{l}.so").exists()
               for s in (1, 2, 3) for l in (1, 2))


class RefVariant:
    """One amdflang build of the reference modules (fixed solver / limiter).

    kind="ref": the reference modules; kind="dropin": the same harness and the
    reference's mhd_config linked with the product's Fortran drop-in modules
    (chimera_amd/fortran), so every routine runs on the GPU."""

    _cache: dict = {}

    def __new__(cls, solver: int, limiter: int, kind: str = "ref"):
        key = (solver, limiter, kind)
        if key not in cls._cache:
            obj = super().__new__(cls)
            path = REF_DIR / f"libchimera_{kind}_s{solver}_l{limiter}.so"
            if not path.exists():
                raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
            obj.lib = C.CDLL(str(path), mode=os.RTLD_LOCAL)
            obj.lib.ref_run.restype = C.c_int
            obj.solver, obj.limiter = solver, limiter
            cls._cache[key] = obj
        return cls._cache[key]

    def scheme(self):
        out = (C.c_int * 4)()
        self.lib.ref_scheme(out)
        return list(out)

    def set_grid(self, n, bc):
        self.lib.ref_set_grid(n, _i(bc))

    def fill_ghost_cells(self, f, bc, f_amb=None):
        n = f.shape[0]
        self.set_grid(n, bc)
        pad = np.empty((n + 2, n + 2))
        if f_amb is None:
            self.lib.ref_fill_ghost_cells(_d(f), _d(pad))
        else:
            self.lib.ref_fill_ghost_cells_amb(_d(f), _d(pad), _d(f_amb))
        return pad

    def compute_gradients(self, f_pad, dx):
        n = f_pad.shape[0] - 2
        self.set_grid(n, [1, 1, 1, 1])
        gx, gy = np.empty((n, n)), np.empty((n, n))
        self.lib.ref_compute_gradients(_d(f_pad), C.c_double(dx), _d(gx), _d(gy))
        return gx, gy

    def apply_slope_limiter(self, f_pad, dx):
        n = f_pad.shape[0] - 2
        self.set_grid(n, [1, 1, 1, 1])
        gx, gy = np.empty((n, n)), np.empty((n, n))
        self.lib.ref_apply_slope_limiter(_d(f_pad), C.c_double(dx), _d(gx), _d(gy))
        return gx, gy

    def reconstruction(self, f, f_dx, f_dy, dx, bc, rho_or_p):
        n = f.shape[0]
        self.set_grid(n, bc)
        out = [np.empty((n, n)) for _ in range(4)]
        self.lib.ref_reconstruction(_d(f), _d(f_dx), _d(f_dy), C.c_double(dx),
                                    *[_d(o) for o in out], int(rho_or_p))
        return out

    def thermal_pressure_check(self, faces, cells):
        n = faces.shape[1]
        self.set_grid(n, [1, 1, 1, 1])
        g = [np.ascontiguousarray(faces[4 * k:4 * k + 4]) for k in range(6)]
        self.lib.ref_thermal_pressure_check(*[_d(x) for x in g], _d(cells))
        for k in range(6):
            faces[4 * k:4 * k + 4] = g[k]

    def compute_fluxes(self, fin, gamma):
        shape = fin.shape[1:]
        a = fin.reshape(12, -1)
        ny, nx = 1, a.shape[1]
        out = np.empty((5, nx))
        a = np.ascontiguousarray(a)
        self.lib.ref_compute_fluxes(nx, ny, _d(a), C.c_double(gamma), _d(out))
        return out.reshape((5,) + shape)

    def update_conserved(self, F, fX, fY, dx, dt, bc):
        self.set_grid(F.shape[0], bc)
        F = F.copy()
        self.lib.ref_update_conserved(_d(F), _d(fX), _d(fY), C.c_double(dx), C.c_double(dt))
        return F

    def constrained_transport(self, bx, by, fByX, fBxY, dx, dt, bc):
        self.set_grid(bx.shape[0], bc)
        bx, by = bx.copy(), by.copy()
        self.lib.ref_constrained_transport(_d(bx), _d(by), _d(fByX), _d(fBxY), C.c_double(dx),
                                           C.c_double(dt))
        return bx, by

    def compute_curl_2d(self, Az, dx, bc):
        n = Az.shape[0]
        self.set_grid(n, bc)
        bx, by = np.empty((n, n)), np.empty((n, n))
        self.lib.ref_compute_curl_2d(_d(Az), C.c_double(dx), _d(bx), _d(by))
        return bx, by

    def compute_divB(self, bx, by, dx, bc):
        n = bx.shape[0]
        self.set_grid(n, bc)
        d = np.empty((n, n))
        self.lib.ref_compute_divB(_d(bx), _d(by), C.c_double(dx), _d(d))
        return d

    def average_face_to_cell_B(self, b_x, b_y, bc):
        n = b_x.shape[0]
        self.set_grid(n, bc)
        Bx, By = np.empty((n, n)), np.empty((n, n))
        self.lib.ref_average_face_to_cell_B(_d(b_x), _d(b_y), _d(Bx), _d(By))
        return Bx, By

    def get_conserved(self, prim, gamma, vol):
        n = prim.shape[1]
        self.set_grid(n, [1, 1, 1, 1])
        out = np.empty((4, n, n))
        self.lib.ref_get_conserved(_d(np.ascontiguousarray(prim)), C.c_double(gamma),
                                   C.c_double(vol), _d(out))
        return out

    def get_primitive(self, cons, Bx, By, gamma, vol):
        n = cons.shape[1]
        self.set_grid(n, [1, 1, 1, 1])
        out = np.empty((4, n, n))
        self.lib.ref_get_primitive(_d(np.ascontiguousarray(cons)), _d(Bx), _d(By),
                                   C.c_double(gamma), C.c_double(vol), _d(out))
        return out

    def apply_conserved_floors(self, cons, Bx, By, vol, gamma):
        n = cons.shape[1]
        self.set_grid(n, [1, 1, 1, 1])
        cons = np.ascontiguousarray(cons).copy()
        self.lib.ref_apply_conserved_floors(_d(cons), _d(Bx), _d(By), C.c_double(vol),
                                            C.c_double(gamma))
        return cons

    def stage_fluxes(self, prim, dx, gamma, bc):
        n = prim.shape[1]
        self.set_grid(n, bc)
        flux = np.empty((10, n, n))
        faces = np.empty((24, n, n))
        self.lib.ref_stage_fluxes(_d(np.ascontiguousarray(prim)), C.c_double(dx),
                                  C.c_double(gamma), _d(flux), _d(faces))
        return flux, faces

    def run(self, state, dx, gamma, bc, nsteps, dt_mode=DT_REFERENCE, dt_in=None, t=0.0,
            out_count=1, t_end=0.0):
        n = state.shape[1]
        self.set_grid(n, bc)
        assert state.flags.c_contiguous and state.dtype == np.float64
        dt_out = np.zeros(max(nsteps, 1))
        divb = np.zeros((max(nsteps, 1), 2))
        dti = (np.ascontiguousarray(dt_in, dtype=np.float64) if dt_in is not None
               else np.zeros(1))
        t_c, oc = C.c_double(t), C.c_int(out_count)
        taken = on_big_stack(self.lib.ref_run, _d(state), C.c_double(dx), C.c_double(gamma),
                             nsteps, dt_mode, _d(dti), _d(dt_out), C.byref(t_c), C.byref(oc),
                             C.c_double(t_end), _d(divb))
        return {"steps": taken, "dt": dt_out[:taken], "divb": divb[:taken], "t": t_c.value,
                "out_count": oc.value}
