"""ctypes wrapper of the CPU oracle (liboracle.so, built from gpe_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class orc_params(ct.Structure):
    _fields_ = [
        ("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32),
        ("order", ct.c_int32), ("bcs", ct.c_int32), ("eqn_to_solve", ct.c_int32),
        ("real_time", ct.c_int32), ("pad_", ct.c_int32),
        ("xr", ct.c_double), ("yr", ct.c_double), ("zr", ct.c_double),
        ("Urhs", ct.c_double), ("diss_amp", ct.c_double), ("g", ct.c_double), ("mu", ct.c_double),
        ("nn", ct.c_double), ("omx", ct.c_double), ("omy", ct.c_double), ("omz", ct.c_double),
        ("vscale", ct.c_double),
        ("eps", ct.c_double), ("safety", ct.c_double), ("dt_decrease", ct.c_double), ("dt_increase", ct.c_double),
    ]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ct.CDLL(LIB)
        P = ct.POINTER(orc_params)
        dp = ct.c_void_p
        ip = ct.POINTER(ct.c_int)
        sig = {
            "orc_second_index": (None, [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ip, ip]),
            "orc_fourth_index": (None, [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ip, ip]),
            "orc_grid": (None, [P, dp, dp, dp, dp]),
            "orc_deriv": (ct.c_int, [P, ct.c_int, dp, dp]),
            "orc_rhs": (ct.c_int, [P, dp, dp]),
            "orc_euler": (ct.c_int, [P, ct.c_double, ct.c_double, dp, dp, dp]),
            "orc_rk2": (ct.c_int, [P, ct.c_double, ct.c_double, dp, dp, dp]),
            "orc_rk4": (ct.c_int, [P, ct.c_double, ct.c_double, dp, dp, dp]),
            "orc_rk4_n": (ct.c_int, [P, ct.c_double, ct.c_double, dp, ct.c_int, dp]),
            "orc_rkck": (ct.c_int, [P, ct.c_double, ct.c_double, dp, dp, dp]),
            "orc_rkqs": (ct.c_int, [P, dp, dp, dp, dp, dp]),
            "orc_norm": (ct.c_double, [P, dp]),
            "orc_energy": (ct.c_double, [P, dp]),
            "orc_mass": (ct.c_double, [P, dp]),
            "orc_renormalise": (None, [P, dp, ct.c_double]),
            "orc_linelength": (ct.c_double, [P, dp, dp]),
            "orc_threads": (ct.c_int, []),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def params(p) -> orc_params:
    """orc_params from a gpe_amd.RunParams (or any object with the same fields)."""
    return orc_params(p.nx, p.ny, p.nz, p.order, p.bcs, p.eqn_to_solve, int(bool(p.real_time)), 0,
                      p.xr, p.yr, p.zr, p.Urhs, p.diss_amp, p.g, p.mu, p.nn, p.omx, p.omy, p.omz, p.vscale,
                      p.eps, p.safety, p.dt_decrease, p.dt_increase)


def _c(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def _chk(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed: {rc}")


def deriv(p, op: int, f: np.ndarray) -> np.ndarray:
    f = _c(f)
    out = np.empty_like(f)
    _chk(lib().orc_deriv(ct.byref(params(p)), op, f.ctypes.data, out.ctypes.data), "deriv")
    return out


def rhs(p, f: np.ndarray) -> np.ndarray:
    f = _c(f)
    out = np.empty_like(f)
    _chk(lib().orc_rhs(ct.byref(params(p)), f.ctypes.data, out.ctypes.data), "rhs")
    return out


def step(p, scheme: str, psi: np.ndarray, dt: complex, t=(0.0, 0.0)):
    """One step of euler/rk2/rk4; returns (new, (t, im_t))."""
    psi = _c(psi)
    out = np.empty_like(psi)
    tt = (ct.c_double * 2)(*t)
    fn = {"euler": lib().orc_euler, "rk2": lib().orc_rk2, "rk4": lib().orc_rk4}[scheme]
    _chk(fn(ct.byref(params(p)), dt.real, dt.imag, psi.ctypes.data, out.ctypes.data, tt), scheme)
    return out, (tt[0], tt[1])


def rk4_n(p, psi: np.ndarray, dt: complex, nsteps: int, t=(0.0, 0.0)):
    psi = _c(psi).copy()
    tt = (ct.c_double * 2)(*t)
    _chk(lib().orc_rk4_n(ct.byref(params(p)), dt.real, dt.imag, psi.ctypes.data, nsteps, tt), "rk4_n")
    return psi, (tt[0], tt[1])


def rkck(p, psi: np.ndarray, dt: complex):
    psi = _c(psi)
    out = np.empty_like(psi)
    err = np.empty_like(psi)
    _chk(lib().orc_rkck(ct.byref(params(p)), dt.real, dt.imag, psi.ctypes.data, out.ctypes.data, err.ctypes.data),
         "rkck")
    return out, err


def rkqs(p, psi: np.ndarray, dt: complex, t=(0.0, 0.0)):
    """One accepted rkqs step.  Returns (new, dt_next, (t, im_t), errmax, n_reject, n_rhs)."""
    psi = _c(psi)
    out = np.empty_like(psi)
    d = (ct.c_double * 2)(dt.real, dt.imag)
    tt = (ct.c_double * 2)(*t)
    info = (ct.c_double * 3)()
    _chk(lib().orc_rkqs(ct.byref(params(p)), d, tt, psi.ctypes.data, out.ctypes.data, info), "rkqs")
    return out, complex(d[0], d[1]), (tt[0], tt[1]), info[0], int(info[1]), int(info[2])


def norm(p, psi):
    psi = _c(psi)
    return lib().orc_norm(ct.byref(params(p)), psi.ctypes.data)


def energy(p, psi):
    psi = _c(psi)
    return lib().orc_energy(ct.byref(params(p)), psi.ctypes.data)


def mass(p, psi):
    psi = _c(psi)
    return lib().orc_mass(ct.byref(params(p)), psi.ctypes.data)


def linelength(p, psi):
    """linelength (src/variables.f90:879-1145) times dx, and the counts of
    cells with >= 2 hits, 1 hit, ended by a face, ended by an edge."""
    psi = _c(psi)
    counts = np.zeros(4, dtype=np.int64)
    v = lib().orc_linelength(ct.byref(params(p)), psi.ctypes.data, counts.ctypes.data)
    return v, counts


def renormalise(p, psi, nrm):
    psi = _c(psi).copy()
    lib().orc_renormalise(ct.byref(params(p)), psi.ctypes.data, nrm)
    return psi


def index_map(bcs: int, order: int, indx: int, n: int, x_deriv: bool):
    mi = (ct.c_int * 2)()
    pl = (ct.c_int * 2)()
    fn = lib().orc_second_index if order == 2 else lib().orc_fourth_index
    fn(bcs, indx, n, int(x_deriv), mi, pl)
    k = 1 if order == 2 else 2
    return list(mi)[:k], list(pl)[:k]


def threads() -> int:
    return lib().orc_threads()
