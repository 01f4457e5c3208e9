"""Python forms of the per-routine drop-ins (chimera_* host-pointer entry points).

Each function replaces the reference Fortran routine of the same name and takes
numpy float64 arrays shaped (ny, nx) (== Fortran (nx, ny)).  The BC flags and
scheme choices that the Fortran reads from module mhd_config are explicit
arguments here.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BC, Scheme, check, c_dp


def _p(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError(f"need a C-contiguous float64 array (got {a.dtype}, "
                         f"contiguous={a.flags.c_contiguous})")
    return a.ctypes.data_as(c_dp)


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _shaped(a, shape, what: str) -> np.ndarray:
    """a as a contiguous float64 array, which must have the shape the library reads."""
    a = _f(a)
    if a.shape != tuple(shape):
        raise ValueError(f"{what} of shape {a.shape}: need {tuple(shape)}")
    return a


def _grid(a, what: str) -> tuple[int, int]:
    if a.ndim != 2:
        raise ValueError(f"{what}: need a 2-D (ny, nx) array, got shape {a.shape}")
    return a.shape


def _fields(a, count: int, what: str) -> tuple[int, ...]:
    if a.ndim < 2 or a.shape[0] != count:
        raise ValueError(f"{what}: need {count} fields along the first axis, got shape {a.shape}")
    return a.shape[1:]


def fill_ghost_cells(f, bc, f_ambient=None) -> np.ndarray:
    """src/mhd_bc.f90:36-141 -> (ny+2, nx+2) padded field."""
    f = _f(f)
    ny, nx = _grid(f, "f")
    pad = np.empty((ny + 2, nx + 2))
    amb = _shaped(f_ambient, (ny, nx), "f_ambient") if f_ambient is not None else None
    check(_lib.lib().chimera_fill_ghost_cells(_p(f), _p(pad), _p(amb) if amb is not None else None,
                                             nx, ny, C.byref(BC.of(bc))), "fill_ghost_cells")
    return pad


def compute_gradients(f_pad, dx):
    f_pad = _f(f_pad)
    ny, nx = (m - 2 for m in _grid(f_pad, "f_pad"))
    gx, gy = np.empty((ny, nx)), np.empty((ny, nx))
    check(_lib.lib().chimera_compute_gradients(_p(f_pad), float(dx), nx, ny, _p(gx), _p(gy)),
          "compute_gradients")
    return gx, gy


def apply_slope_limiter(f_pad, dx, limiter):
    f_pad = _f(f_pad)
    ny, nx = (m - 2 for m in _grid(f_pad, "f_pad"))
    gx, gy = np.empty((ny, nx)), np.empty((ny, nx))
    check(_lib.lib().chimera_apply_slope_limiter(_p(f_pad), float(dx), nx, ny, _p(gx), _p(gy),
                                                int(limiter)), "apply_slope_limiter")
    return gx, gy


def reconstruction(f, f_dx, f_dy, dx, bc, rho_or_p, mood=True):
    f = _f(f)
    ny, nx = _grid(f, "f")
    f_dx, f_dy = _shaped(f_dx, (ny, nx), "f_dx"), _shaped(f_dy, (ny, nx), "f_dy")
    out = [np.empty((ny, nx)) for _ in range(4)]
    check(_lib.lib().chimera_reconstruction(_p(f), _p(f_dx), _p(f_dy), float(dx), nx, ny,
                                           *[_p(o) for o in out], int(bool(rho_or_p)),
                                           C.byref(BC.of(bc)), int(bool(mood))), "reconstruction")
    return out


def thermal_pressure_check(faces, cells):
    """faces (24, ny, nx) in the reference order P,Bx,By,rho,vx,vy x (XL,XR,YL,YR),
    updated in place; cells (6, ny, nx) = rho, vx, vy, P, Bx, By."""
    if faces.dtype != np.float64 or not faces.flags.c_contiguous or faces.ndim != 3 \
            or faces.shape[0] != 24:
        raise ValueError("faces: need a C-contiguous float64 array (24, ny, nx)")
    ny, nx = faces.shape[1:]
    cells = _shaped(cells, (6, ny, nx), "cells")
    check(_lib.lib().chimera_thermal_pressure_check(nx, ny, *[_p(faces[k]) for k in range(24)],
                                                   *[_p(cells[k]) for k in range(6)]),
          "thermal_pressure_check")


def compute_fluxes(fin, gamma, solver) -> np.ndarray:
    """fin (12, ...) = rho_L,rho_R,vx_L,vx_R,vy_L,vy_R,P_L,P_R,Bx_L,Bx_R,By_L,By_R ->
    (5, ...) = Mass, Momx, Momy, Energy, By."""
    fin = _f(fin)
    shape = _fields(fin, 12, "fin")
    n = int(np.prod(shape))
    a = fin.reshape(12, n)
    out = np.empty((5, n))
    check(_lib.lib().chimera_compute_fluxes(*[_p(a[k]) for k in range(12)], float(gamma), n, 1,
                                           *[_p(out[k]) for k in range(5)], int(solver)),
          "compute_fluxes")
    return out.reshape((5,) + shape)


def update_conserved(F, fX, fY, dx, dt, bc):
    F = _f(F).copy()
    ny, nx = _grid(F, "F")
    fX, fY = _shaped(fX, (ny, nx), "flux_F_X"), _shaped(fY, (ny, nx), "flux_F_Y")
    check(_lib.lib().chimera_update_conserved(_p(F), _p(fX), _p(fY), float(dx), float(dt), nx, ny,
                                             C.byref(BC.of(bc))), "update_conserved")
    return F


def constrained_transport(bx, by, fByX, fBxY, dx, dt, bc):
    bx = _f(bx).copy()
    ny, nx = _grid(bx, "bx")
    by = _shaped(by, (ny, nx), "by").copy()
    fByX, fBxY = _shaped(fByX, (ny, nx), "flux_By_X"), _shaped(fBxY, (ny, nx), "flux_Bx_Y")
    check(_lib.lib().chimera_constrained_transport(_p(bx), _p(by), _p(fByX), _p(fBxY), float(dx),
                                                  float(dt), nx, ny, C.byref(BC.of(bc))),
          "constrained_transport")
    return bx, by


def compute_curl_2d(Az, dx, bc):
    Az = _f(Az)
    ny, nx = _grid(Az, "Az")
    bx, by = np.empty((ny, nx)), np.empty((ny, nx))
    check(_lib.lib().chimera_compute_curl_2d(_p(Az), float(dx), nx, ny, _p(bx), _p(by),
                                            C.byref(BC.of(bc))), "compute_curl_2d")
    return bx, by


def compute_divB(bx, by, dx, bc):
    bx = _f(bx)
    ny, nx = _grid(bx, "bx")
    by = _shaped(by, (ny, nx), "by")
    d = np.empty((ny, nx))
    check(_lib.lib().chimera_compute_divB(_p(bx), _p(by), float(dx), nx, ny, _p(d),
                                         C.byref(BC.of(bc))), "compute_divB")
    return d


def average_face_to_cell_B(b_x, b_y, bc):
    b_x = _f(b_x)
    ny, nx = _grid(b_x, "b_x")
    b_y = _shaped(b_y, (ny, nx), "b_y")
    Bx, By = np.empty((ny, nx)), np.empty((ny, nx))
    check(_lib.lib().chimera_average_face_to_cell_B(_p(b_x), _p(b_y), nx, ny, _p(Bx), _p(By),
                                                   C.byref(BC.of(bc))), "average_face_to_cell_B")
    return Bx, By


def get_conserved(prim, gamma, vol) -> np.ndarray:
    prim = _f(prim)
    shape = _fields(prim, 6, "prim")
    n = int(np.prod(shape))
    out = np.empty((4,) + shape)
    check(_lib.lib().chimera_get_conserved(*[_p(prim[k]) for k in range(6)], float(gamma),
                                          float(vol), n, 1, *[_p(out[k]) for k in range(4)]),
          "get_conserved")
    return out


def get_primitive(cons, Bx, By, gamma, vol) -> np.ndarray:
    cons = _f(cons)
    shape = _fields(cons, 4, "cons")
    Bx, By = _shaped(Bx, shape, "Bx"), _shaped(By, shape, "By")
    n = int(np.prod(shape))
    out = np.empty((4,) + shape)
    check(_lib.lib().chimera_get_primitive(*[_p(cons[k]) for k in range(4)], _p(Bx), _p(By),
                                          float(gamma), float(vol), n, 1,
                                          *[_p(out[k]) for k in range(4)]), "get_primitive")
    return out


def apply_conserved_floors(cons, Bx, By, vol, gamma) -> np.ndarray:
    cons = _f(cons).copy()
    shape = _fields(cons, 4, "cons")
    Bx, By = _shaped(Bx, shape, "Bx"), _shaped(By, shape, "By")
    n = int(np.prod(shape))
    check(_lib.lib().chimera_apply_conserved_floors(*[_p(cons[k]) for k in range(4)], _p(Bx),
                                                   _p(By), float(vol), float(gamma), n, 1),
          "apply_conserved_floors")
    return cons


def stage_fluxes(prim, dx, gamma, bc, scheme, want_mask=False):
    """compute_stage_fluxes (src/main.f90:346-403) -> (10, ny, nx) fluxes
    (Mass_X, Mass_Y, Momx_X, Momx_Y, Momy_X, Momy_Y, Energy_X, Energy_Y, By_X, Bx_Y)."""
    prim = _f(prim)
    if prim.ndim != 3:
        raise ValueError(f"prim: need (6, ny, nx), got shape {prim.shape}")
    ny, nx = _fields(prim, 6, "prim")
    flux = np.empty((10, ny, nx))
    mask = np.zeros((ny, nx), dtype=np.uint32) if want_mask else None
    check(_lib.lib().chimera_stage_fluxes(_p(prim), _p(flux),
                                         mask.ctypes.data_as(_lib.c_u32p) if want_mask else None,
                                         float(dx), float(gamma), nx, ny, C.byref(BC.of(bc)),
                                         C.byref(Scheme.of(scheme))), "stage_fluxes")
    return (flux, mask) if want_mask else flux
