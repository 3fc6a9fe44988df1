"""Device-resident solver context over libgpehip.so.

`Context` owns one whole-domain field on the GPU and advances it with the
HIP kernels; host arrays cross the boundary only in upload/download and the
host-array drop-ins.  Fields are numpy arrays of shape (nz, ny, nx) in C
order (x fastest), complex128 or complex64.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .params import RunParams


def _host_dtype(a: np.ndarray) -> int:
    if a.dtype == np.complex128:
        return _lib.GPE_HOST_C128
    if a.dtype == np.complex64:
        return _lib.GPE_HOST_C64
    raise TypeError(f"field must be complex128 or complex64, got {a.dtype}")


class StepInfo:
    """Python view of gpe_step_info (t, im_t, dt and counters)."""

    def __init__(self, dt: complex = 0j, t: float = 0.0, im_t: float = 0.0):
        self.c = _lib.gpe_step_info(t, im_t, dt.real, dt.imag, 0.0, 0.0, 0, 0, 0)

    @property
    def t(self) -> float:
        return self.c.t

    @property
    def im_t(self) -> float:
        return self.c.im_t

    @property
    def dt(self) -> complex:
        return complex(self.c.dt_re, self.c.dt_im)

    @dt.setter
    def dt(self, v: complex) -> None:
        self.c.dt_re, self.c.dt_im = v.real, v.imag

    def __getattr__(self, k):
        if k in ("errmax", "psi2max", "n_accept", "n_reject", "n_rhs"):
            return getattr(self.c, k)
        raise AttributeError(k)


class Context:
    def __init__(self, params: RunParams):
        self.lib = _lib.load()
        self.params = params
        self._cp = params.to_c()
        h = ct.c_void_p()
        _lib.check(self.lib.gpe_ctx_create(ct.byref(self._cp), ct.byref(h)), None, "gpe_ctx_create")
        self.h = h
        self.storage_dtype = np.complex64 if params.storage == "c64" else np.complex128

    # ---- lifetime
    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gpe_ctx_destroy(self.h)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc: int, what: str) -> None:
        _lib.check(rc, self.h, what)

    # ---- configuration
    def set_tiles(self, tx: int, ty: int, tz: int = 0) -> None:
        self._chk(self.lib.gpe_set_tiles(self.h, tx, ty, tz), "gpe_set_tiles")

    def set_fused(self, on: bool) -> None:
        """Fused stages where they apply (RK4: two kernels per step; RKF45:
        three kernels per Cash-Karp attempt); on by default."""
        self._chk(self.lib.gpe_set_fused(self.h, int(bool(on))), "gpe_set_fused")

    def tiles(self):
        a, b, c = ct.c_int(), ct.c_int(), ct.c_int()
        self._chk(self.lib.gpe_get_tiles(self.h, ct.byref(a), ct.byref(b), ct.byref(c)), "gpe_get_tiles")
        return a.value, b.value, c.value

    # ---- transfers
    def _layout(self, a: np.ndarray):
        p = self.params
        if not a.flags.c_contiguous:
            raise ValueError("field must be C-contiguous (x fastest)")
        if a.shape == p.shape:
            return p.ny, 0, 0
        if a.shape == (p.nz + 4, p.ny + 4, p.nx):
            return p.ny + 4, 2, 2  # the reference's halo array (0:nx1, js-2:je+2, ks-2:ke+2)
        raise ValueError(f"field shape {a.shape} is neither {p.shape} nor the halo shape")

    def upload(self, a: np.ndarray) -> None:
        nya, yo, zo = self._layout(a)
        self._chk(self.lib.gpe_upload(self.h, a.ctypes.data, _host_dtype(a), nya, yo, zo), "gpe_upload")

    def upload_aux(self, a: np.ndarray) -> None:
        """Upload into the context's auxiliary field (never the state)."""
        nya, yo, zo = self._layout(a)
        self._chk(self.lib.gpe_field_upload(self.h, _lib.GPE_FIELD_AUX, a.ctypes.data, _host_dtype(a), nya, yo, zo),
                  "gpe_field_upload")

    def select(self, aux: bool) -> None:
        """Make the reductions, diagnostics, renormalise, imprint_line, peek and
        download act on the auxiliary field (True) or the state (False)."""
        f = _lib.GPE_FIELD_AUX if aux else _lib.GPE_FIELD_STATE
        self._chk(self.lib.gpe_select_field(self.h, f), "gpe_select_field")

    def set_real_time(self, real_time: bool) -> None:
        """The module variable real_time that rkqs reads for the direction
        of a retried step (src/solve.f90:195-201)."""
        self._chk(self.lib.gpe_set_real_time(self.h, int(bool(real_time))), "gpe_set_real_time")

    def download(self, dtype=np.complex128) -> np.ndarray:
        out = np.empty(self.params.shape, dtype=dtype)
        self._chk(self.lib.gpe_download(self.h, out.ctypes.data, _host_dtype(out)), "gpe_download")
        return out

    # ---- stepping
    def step(self, scheme: str, nsteps: int, info: StepInfo) -> StepInfo:
        if scheme not in _lib.SCHEMES:
            raise _lib.GpeError("ERROR: Unrecognised time stepping scheme.")
        self._chk(self.lib.gpe_step(self.h, _lib.SCHEMES[scheme], nsteps, ct.byref(info.c)), "gpe_step")
        return info

    def rk45_attempt(self, dt: complex, dt0: complex):
        out = (ct.c_double * 2)()
        self._chk(self.lib.gpe_rk45_attempt(self.h, dt.real, dt.imag, dt0.real, dt0.imag, out), "gpe_rk45_attempt")
        return out[0], out[1]

    def rk45_accept(self) -> None:
        self._chk(self.lib.gpe_rk45_accept(self.h), "gpe_rk45_accept")

    # ---- operators on host arrays (or on the state when a is None)
    def deriv(self, op: str, a: np.ndarray | None = None) -> np.ndarray:
        dt = a.dtype if a is not None else np.complex128
        out = np.empty(self.params.shape, dtype=dt)
        if a is None:
            rc = self.lib.gpe_deriv(self.h, _lib.DERIV_OPS[op], None, 0, 0, 0, out.ctypes.data, _host_dtype(out))
        else:
            nya, yo, zo = self._layout(a)
            rc = self.lib.gpe_deriv(self.h, _lib.DERIV_OPS[op], a.ctypes.data, nya, yo, zo, out.ctypes.data,
                                    _host_dtype(a))
        self._chk(rc, "gpe_deriv")
        return out

    def rhs(self, a: np.ndarray | None = None) -> np.ndarray:
        dt = a.dtype if a is not None else np.complex128
        out = np.empty(self.params.shape, dtype=dt)
        if a is None:
            rc = self.lib.gpe_rhs(self.h, None, 0, 0, 0, out.ctypes.data, _host_dtype(out))
        else:
            nya, yo, zo = self._layout(a)
            rc = self.lib.gpe_rhs(self.h, a.ctypes.data, nya, yo, zo, out.ctypes.data, _host_dtype(a))
        self._chk(rc, "gpe_rhs")
        return out

    # ---- reductions of the state
    def norm(self) -> float:
        v = ct.c_double()
        self._chk(self.lib.gpe_norm(self.h, ct.byref(v)), "gpe_norm")
        return v.value

    def energy(self) -> float:
        v = ct.c_double()
        self._chk(self.lib.gpe_energy(self.h, ct.byref(v)), "gpe_energy")
        return v.value

    def hamiltonian(self) -> float:
        v = ct.c_double()
        self._chk(self.lib.gpe_hamiltonian(self.h, ct.byref(v)), "gpe_hamiltonian")
        return v.value

    def momentum(self) -> np.ndarray:
        """variables::momentum (src/variables.f90:615-694): the 3 components."""
        m = (ct.c_double * 3)()
        self._chk(self.lib.gpe_momentum(self.h, m), "gpe_momentum")
        return np.array(m[:])

    def velocity_pdf(self, nbins: int = 256) -> dict:
        """get_pdf_velocity + get_pdf (src/variables.f90:377-463): per component
        the valid count, mean, stdev, vmax and the PDF over nbins bins (bin i
        holds (2(i-1) vmax/nbins, 2 i vmax/nbins], i = -nbins/2+1 .. nbins/2).
        The velocity is the reference's, -Im(psi) Re(d psi)/|psi|^2."""
        cnt = (ct.c_int64 * 3)()
        mean, sd, vmax = (ct.c_double * 3)(), (ct.c_double * 3)(), (ct.c_double * 3)()
        pdf = np.zeros((3, nbins), dtype=np.float64)
        self._chk(self.lib.gpe_velocity_pdf(self.h, nbins, cnt, mean, sd, vmax,
                                            pdf.ctypes.data_as(ct.POINTER(ct.c_double))), "gpe_velocity_pdf")
        return dict(counts=np.array(cnt[:]), vmean=np.array(mean[:]), vstdev=np.array(sd[:]),
                    vmax=np.array(vmax[:]), pdf=pdf)

    def vcf(self) -> np.ndarray:
        """get_vcf (src/variables.f90:467-507): f(r), r = 0 .. nx-1."""
        f = np.zeros(self.params.nx, dtype=np.float64)
        self._chk(self.lib.gpe_vcf(self.h, f.ctypes.data_as(ct.POINTER(ct.c_double))), "gpe_vcf")
        return f

    def imprint_line(self, vl, scal: float = 1.0) -> None:
        """psi <- |psi| vortex_line(vl) (imprint_vortex_line,
        src/variables.f90:1148-1162); vl is an ics.LineParam."""
        ln = _lib.gpe_line(vl.x0, vl.y0, vl.z0, vl.amp1, vl.amp2, vl.ll, vl.sgn, "xyz".index(vl.dir),
                           int(bool(vl.imprint_phase)))
        self._chk(self.lib.gpe_imprint_line(self.h, ct.byref(ln), float(scal)), "gpe_imprint_line")

    def linelength(self) -> float:
        """linelength (src/variables.f90:879-1145): total length of the zero
        lines of psi (dx = dy = dz assumed, as the reference does)."""
        v = ct.c_double()
        self._chk(self.lib.gpe_linelength(self.h, ct.byref(v)), "gpe_linelength")
        return v.value

    def condensed_particles(self, spectrum: bool = True) -> dict:
        """condensed_particles (src/io.f90:315-404) and spectrum (632-725) of
        the state: row = M/(8 xr yr zr), n0/V, M, n0, temp, temp2, H/V (the
        mass.dat values after time and im_t); with spectrum, eta and nharm per
        log2 shell (7) and F(kk), kk = 1..nx/2."""
        row = (ct.c_double * 7)()
        if not spectrum:
            self._chk(self.lib.gpe_condensed_particles(self.h, row, None, None, None), "gpe_condensed_particles")
            return dict(row=np.array(row[:]))
        eta = (ct.c_double * 7)()
        nharm = (ct.c_int64 * 7)()
        F = np.zeros(self.params.nx // 2, dtype=np.float64)
        self._chk(self.lib.gpe_condensed_particles(self.h, row, eta, nharm, F.ctypes.data_as(ct.POINTER(ct.c_double))),
                  "gpe_condensed_particles")
        return dict(row=np.array(row[:]), eta=np.array(eta[:]), nharm=np.array(nharm[:], dtype=np.int64), F=F)

    def peek(self, i: int, j: int, k: int) -> complex:
        """The state's value at cell (i, j, k)."""
        v = (ct.c_double * 2)()
        self._chk(self.lib.gpe_peek(self.h, int(i), int(j), int(k), v), "gpe_peek")
        return complex(v[0], v[1])

    def mass(self) -> float:
        v = ct.c_double()
        self._chk(self.lib.gpe_mass(self.h, ct.byref(v)), "gpe_mass")
        return v.value

    def renormalise(self, norm: float) -> None:
        self._chk(self.lib.gpe_renormalise(self.h, norm), "gpe_renormalise")

    # ---- raw handles
    @property
    def stream(self) -> int:
        return self.lib.gpe_stream(self.h)

    @property
    def state_ptr(self) -> int:
        return self.lib.gpe_state_ptr(self.h)

    def sync(self) -> None:
        self._chk(self.lib.gpe_sync(self.h), "gpe_sync")

    def kernel_profile(self, fn) -> dict:
        """Run fn() with the library's per-launch event timing on and return
        {kind: {"launches", "avg_ms", "bytes"}} (bytes = algorithmic field
        bytes of one launch)."""
        self._chk(self.lib.gpe_prof_enable(self.h, 1), "gpe_prof_enable")
        try:
            fn()
            out = {}
            for k in range(_lib.GPE_NKINDS):
                ms, n, b = ct.c_double(), ct.c_int64(), ct.c_double()
                self._chk(self.lib.gpe_prof_read(self.h, k, ct.byref(ms), ct.byref(n), ct.byref(b)), "gpe_prof_read")
                if n.value:
                    out[self.lib.gpe_kernel_kind_name(k).decode()] = {"launches": n.value, "avg_ms": ms.value / n.value,
                                                                      "bytes": b.value}
        finally:
            self.lib.gpe_prof_enable(self.h, 0)
        return out
