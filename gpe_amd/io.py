"""File formats of the reference kept by this package.

* end_state restart file (src/io.f90:840-891, read back by state_restart
  src/ic.f90:117-186): stream-unformatted int32 nx, ny, nz, p; float64 t,
  im_t; complex128 dt; complex128 psi (x fastest).  The Fortran host writes
  its checkpoints to proc00/end_state.dat as the reference does (io.f90:852)
  and restarts from end_state00.dat (the name run.sh -r gives the copy); a
  fresh run (restart = .false.) reads its initial state from ic_state.dat in
  the same layout.
* dens<p>.dat 3D saves of idl_surface (src/io.f90:489-529): float64 t+im_t;
  int32 nx, ny, nz, nyprocs, nzprocs, js, je, ks, ke; complex128 psi of the
  rank's block; float64 x, y, z.
* run.in with the reference's namelists (src/io.f90:78-114).
"""
from __future__ import annotations

import numpy as np


def write_end_state(path: str, psi: np.ndarray, p: int = 0, t: float = 0.0, im_t: float = 0.0,
                    dt: complex = 0j, extra: tuple = ()) -> None:
    nz, ny, nx = psi.shape
    with open(path, "wb") as f:
        np.array([nx, ny, nz, p], dtype="<i4").tofile(f)
        np.array([t, im_t, dt.real, dt.imag], dtype="<f8").tofile(f)
        np.ascontiguousarray(psi, dtype="<c16").tofile(f)
        if extra:
            np.array(extra, dtype="<f8").tofile(f)


def read_end_state(path: str, n_extra: int = 0):
    """Returns (psi, p, t, im_t, dt[, extras])."""
    with open(path, "rb") as f:
        nx, ny, nz, p = (int(v) for v in np.fromfile(f, dtype="<i4", count=4))
        t, im_t, dre, dim = np.fromfile(f, dtype="<f8", count=4)
        psi = np.fromfile(f, dtype="<c16", count=nx * ny * nz).reshape(nz, ny, nx)
        out = (psi, p, float(t), float(im_t), complex(dre, dim))
        if n_extra:
            out = out + (np.fromfile(f, dtype="<f8", count=n_extra),)
    return out


def write_dens(path: str, psi: np.ndarray, p, time: float) -> None:
    """A one-rank dens file as idl_surface writes it; p: RunParams."""
    nz, ny, nx = psi.shape
    x, y, z = p.axes()
    with open(path, "wb") as f:
        np.array([time], dtype="<f8").tofile(f)
        np.array([nx, ny, nz, 1, 1, 0, ny - 1, 0, nz - 1], dtype="<i4").tofile(f)
        np.ascontiguousarray(psi, dtype="<c16").tofile(f)
        for a in (x, y, z):
            np.asarray(a, dtype="<f8").tofile(f)


def read_dens(path: str) -> dict:
    """Reads a dens file of one rank: time, grid, block bounds, psi (z, y, x
    of the block), x, y, z."""
    with open(path, "rb") as f:
        time = float(np.fromfile(f, dtype="<f8", count=1)[0])
        nx, ny, nz, nyp, nzp, js, je, ks, ke = (int(v) for v in np.fromfile(f, dtype="<i4", count=9))
        nyl, nzl = je - js + 1, ke - ks + 1
        psi = np.fromfile(f, dtype="<c16", count=nx * nyl * nzl).reshape(nzl, nyl, nx)
        x = np.fromfile(f, dtype="<f8", count=nx)
        y = np.fromfile(f, dtype="<f8", count=ny)
        z = np.fromfile(f, dtype="<f8", count=nz)
        rest = f.read()
    if rest:
        raise ValueError(f"{path}: {len(rest)} trailing bytes")
    return dict(time=time, nx=nx, ny=ny, nz=nz, nyprocs=nyp, nzprocs=nzp, js=js, je=je, ks=ks, ke=ke, psi=psi,
                x=x, y=y, z=z)


def set_namelist(text: str, **values) -> str:
    """Sets `name = value` entries of a run.in text (the entry must exist);
    values are Python bools, numbers or strings."""
    import re

    for key, v in values.items():
        if isinstance(v, bool):
            lit = ".true." if v else ".false."
        elif isinstance(v, str):
            lit = f"'{v}'"
        else:
            lit = repr(v)
        text, k = re.subn(rf"(?m)^(\s*{re.escape(key)}\s*=\s*).*$", lambda m: m.group(1) + lit, text)
        if k != 1:
            raise KeyError(f"run.in has no single entry {key!r}")
    return text


def run_in_text(p, end_time: float, restart: bool = True, save_rate: int = 10, nbins: int = 256) -> str:
    """run.in with the reference's four namelists (src/io.f90:78-114) plus
    grid_params for the Fortran host of this package."""
    def b(v):
        return ".true." if v else ".false."

    def r(v):
        return repr(float(v))

    return f"""&run_params
  tau = {r(p.tau)}
  end_time = {r(end_time)}
  xr = {r(p.xr)}
  yr = {r(p.yr)}
  zr = {r(p.zr)}
  scheme = '{p.scheme}'
  eqn_to_solve = {p.eqn_to_solve}
  bcs = {p.bcs}
  order = {p.order}
  restart = {b(restart)}
  saved_restart = .false.
  renorm = .false.
  imprint_vl = .false.
  stop_imag = .false.
  real_time = {b(p.real_time)}
/
&eqn_params
  Urhs = {r(p.Urhs)}
  diss_amp = {r(p.diss_amp)}
  scal = 1.0
  nv = 0.75
  enerv = 0.75
  g = {r(p.g)}
  mu = {r(p.mu)}
  nn = {r(p.nn)}
  omx = {r(p.omx)}
  omy = {r(p.omy)}
  omz = {r(p.omz)}
/
&io_params
  save_rate = {save_rate}
  save_rate2 = 1.0e30
  save_rate3 = 1.0e30
  p_save = 1.0e30
  save_contour = .false.
  save_3d = .false.
  save_filter = .false.
  filter_kc = 10.0
  save_average = .false.
  save_spectrum = .false.
  save_pdf = .false.
  save_vcf = .false.
  save_ll = .false.
  save_zeros = .false.
/
&misc_params
  nbins = {nbins}
  diagnostic = .false.
  pp_filtered_surface = .false.
  nlines = 3
  nfilter = 1
  fscale = 1.0
/
&grid_params
  nx = {p.nx}
  ny = {p.ny}
  nz = {p.nz}
  storage = '{p.storage}'
  vscale = {r(p.vscale)}
  host_mode = 'resident'
  save_mom = .false.
/
"""
