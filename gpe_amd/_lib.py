"""ctypes binding of libgpehip.so (the C ABI declared in include/gpe_hip.h).

The library is the only compute path: if it is missing or cannot be loaded
this module raises, it never falls back to another implementation.
"""
from __future__ import annotations

import ctypes as ct
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpehip.so")

GPE_OK = 0
GPE_C128, GPE_C64, GPE_MIXED = 0, 1, 2
GPE_HOST_C128, GPE_HOST_C64 = 0, 1
GPE_FIELD_STATE, GPE_FIELD_AUX = 0, 1
GPE_ABI_VERSION = 2
GPE_NKINDS = 37  # kernel kinds counted by gpe_prof_read
SCHEMES = {"euler": 0, "rk2": 1, "rk4": 2, "rk45": 3}
DERIV_OPS = {"x": 0, "y": 1, "z": 2, "xx": 3, "yy": 4, "zz": 5, "lap": 6}


class GpeError(RuntimeError):
    """A failed libgpehip call (the reference stops with emergency_stop)."""


class gpe_params(ct.Structure):
    _fields_ = [
        ("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32),
        ("order", ct.c_int32), ("bcs", ct.c_int32), ("eqn_to_solve", ct.c_int32),
        ("real_time", ct.c_int32), ("storage", ct.c_int32),
        ("xr", ct.c_double), ("yr", ct.c_double), ("zr", ct.c_double),
        ("Urhs", ct.c_double), ("diss_amp", ct.c_double), ("g", ct.c_double), ("mu", ct.c_double),
        ("nn", ct.c_double), ("omx", ct.c_double), ("omy", ct.c_double), ("omz", ct.c_double),
        ("vscale", ct.c_double),
        ("eps", ct.c_double), ("safety", ct.c_double), ("dt_decrease", ct.c_double), ("dt_increase", ct.c_double),
    ]


class gpe_line(ct.Structure):
    _fields_ = [(n, ct.c_double) for n in ("x0", "y0", "z0", "amp1", "amp2", "ll", "sgn")] + [
        ("dir", ct.c_int32), ("imprint_phase", ct.c_int32)]


class gpe_step_info(ct.Structure):
    _fields_ = [
        ("t", ct.c_double), ("im_t", ct.c_double), ("dt_re", ct.c_double), ("dt_im", ct.c_double),
        ("errmax", ct.c_double), ("psi2max", ct.c_double),
        ("n_accept", ct.c_int64), ("n_reject", ct.c_int64), ("n_rhs", ct.c_int64),
    ]


# (name, restype, argtypes) of every exported entry point of include/gpe_hip.h
_VP, _I64, _D, _I = ct.c_void_p, ct.c_int64, ct.c_double, ct.c_int
PROTOTYPES = [
    ("gpe_abi_version", _I, []),
    ("gpe_last_error", ct.c_char_p, [_VP]),
    ("gpe_ctx_create", _I, [ct.POINTER(gpe_params), ct.POINTER(_VP)]),
    ("gpe_ctx_destroy", None, [_VP]),
    ("gpe_set_tiles", _I, [_VP, _I, _I, _I]),
    ("gpe_set_fused", _I, [_VP, _I]),
    ("gpe_field_upload", _I, [_VP, _I, _VP, _I, _I64, _I64, _I64]),
    ("gpe_select_field", _I, [_VP, _I]),
    ("gpe_set_real_time", _I, [_VP, _I]),
    ("gpe_energy_noalloc", _I, [ct.POINTER(gpe_params), _VP, _VP, _VP]),
    ("gpe_get_tiles", _I, [_VP, ct.POINTER(_I), ct.POINTER(_I), ct.POINTER(_I)]),
    ("gpe_upload", _I, [_VP, _VP, _I, _I64, _I64, _I64]),
    ("gpe_download", _I, [_VP, _VP, _I]),
    ("gpe_step", _I, [_VP, _I, _I, ct.POINTER(gpe_step_info)]),
    ("gpe_solver", _I, [_VP, _I, _VP, _I64, _I64, _I64, _VP, _I, ct.POINTER(gpe_step_info)]),
    ("gpe_rk45_attempt", _I, [_VP, _D, _D, _D, _D, ct.POINTER(_D)]),
    ("gpe_rk45_accept", _I, [_VP]),
    ("gpe_deriv", _I, [_VP, _I, _VP, _I64, _I64, _I64, _VP, _I]),
    ("gpe_rhs", _I, [_VP, _VP, _I64, _I64, _I64, _VP, _I]),
    ("gpe_norm", _I, [_VP, ct.POINTER(_D)]),
    ("gpe_energy", _I, [_VP, ct.POINTER(_D)]),
    ("gpe_mass", _I, [_VP, ct.POINTER(_D)]),
    ("gpe_momentum", _I, [_VP, ct.POINTER(_D)]),
    ("gpe_imprint_line", _I, [_VP, ct.POINTER(gpe_line), _D]),
    ("gpe_linelength", _I, [_VP, ct.POINTER(_D)]),
    ("gpe_condensed_particles", _I, [_VP, ct.POINTER(_D), ct.POINTER(_D), ct.POINTER(ct.c_int64), ct.POINTER(_D)]),
    ("gpe_peek", _I, [_VP, ct.c_int32, ct.c_int32, ct.c_int32, ct.POINTER(_D)]),
    ("gpe_velocity_pdf", _I, [_VP, _I, ct.POINTER(_I64), ct.POINTER(_D), ct.POINTER(_D), ct.POINTER(_D),
                              ct.POINTER(_D)]),
    ("gpe_vcf", _I, [_VP, ct.POINTER(_D)]),
    ("gpe_hamiltonian", _I, [_VP, ct.POINTER(_D)]),
    ("gpe_renormalise", _I, [_VP, _D]),
    ("gpe_prof_enable", _I, [_VP, _I]),
    ("gpe_prof_read", _I, [_VP, _I, ct.POINTER(_D), ct.POINTER(_I64), ct.POINTER(_D)]),
    ("gpe_kernel_kind_name", ct.c_char_p, [_I]),
    ("gpe_bench_copy", _I, [_VP, _VP, _I64, _VP]),
    ("gpe_bench_stream5", _I, [_VP, _VP, _VP, _VP, _VP, _I64, _VP]),
    ("gpe_ens_create", _I, [ct.POINTER(gpe_params), _I, ct.POINTER(_VP)]),
    ("gpe_ens_destroy", None, [_VP]),
    ("gpe_ens_last_error", ct.c_char_p, [_VP]),
    ("gpe_ens_set_tiles", _I, [_VP, _I, _I, ct.POINTER(_I)]),
    ("gpe_ens_upload", _I, [_VP, _I, _VP, _I]),
    ("gpe_ens_download", _I, [_VP, _I, _VP, _I]),
    ("gpe_ens_step", _I, [_VP, _I, _D, _D, _I]),
    ("gpe_ens_work_items", _I, [_VP]),
    ("gpe_ens_set_grouping", _I, [_VP, _I64, _I]),
    ("gpe_ens_num_groups", _I, [_VP]),
    ("gpe_ens_stream", _VP, [_VP]),
    ("gpe_ens_sync", _I, [_VP]),
    ("gpe_mlp_create", _I, [_I, ct.POINTER(_VP)]),
    ("gpe_mlp_destroy", None, [_VP]),
    ("gpe_mlp_last_error", ct.c_char_p, [_VP]),
    ("gpe_mlp_num_params", _I, [_VP]),
    ("gpe_mlp_set_params", _I, [_VP, _VP, _I64]),
    ("gpe_mlp_get_params", _I, [_VP, _VP, _I64]),
    ("gpe_mlp_forward", _I, [_VP, _VP, _I, _VP]),
    ("gpe_mlp_grad", _I, [_VP, _VP, _VP, _I, _VP, _VP]),
    ("gpe_mlp_train", _I, [_VP, _VP, _VP, _I, _I, _I, ct.c_float, _VP]),
    ("gpe_state_ptr", _VP, [_VP]),
    ("gpe_stream", _VP, [_VP]),
    ("gpe_sync", _I, [_VP]),
]

_lib = None


def load(path: str | None = None) -> ct.CDLL:
    """Load libgpehip.so once and attach prototypes.  Raises GpeError if the
    library has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise GpeError(f"{p} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                       "or `make -C gpe_amd/csrc`")
    lib = ct.CDLL(p)
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc != GPE_OK:
        lib = load()
        msg = lib.gpe_last_error(ctx)
        raise GpeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
