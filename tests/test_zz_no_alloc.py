"""gpe_energy_noalloc: gpe_energy's kernels on a caller's device field with
the block partials in the library's static scratch, the result written to a
caller's device double on the caller's stream, no allocation.  Checked
against gpe_energy (bit for bit: same kernels, tiles and summation order)
on calls in a row, the later ones on smaller inputs, and its size limit."""
import ctypes as ct

import numpy as np
import pytest

from gpe_amd import Context, RunParams, _lib
from gpe_amd.hiprt import Hip


def entry():
    lib = _lib.load()
    fn = lib.gpe_energy_noalloc
    fn.restype = ct.c_int
    fn.argtypes = [ct.POINTER(_lib.gpe_params), ct.c_void_p, ct.c_void_p, ct.c_void_p]
    return lib, fn


def test_rejects_grids_beyond_the_scratch():
    """A grid whose tile blocks outnumber the static partials is refused
    before any device call (8192 x 8192 planes: 65536 tile columns)."""
    lib, fn = entry()
    p = RunParams(nx=8192, ny=8200, nz=6, xr=10.0, yr=10.0, zr=1.0).to_c()
    rc = fn(ct.byref(p), ct.c_void_p(16), ct.c_void_p(16), None)
    assert rc == -1  # GPE_EINVAL
    assert b"block partials" in lib.gpe_last_error(None)


@pytest.mark.gpu
def test_energy_noalloc_matches_gpe_energy():
    lib, fn = entry()
    hip = Hip()
    hip.h.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    dout = hip.malloc(8)
    rng = np.random.default_rng(11)
    cases = [RunParams(nx=96, ny=80, nz=72, xr=9.0, yr=8.0, zr=7.0, order=4, bcs=1),
             RunParams(nx=40, ny=36, nz=30, xr=5.0, yr=4.5, zr=4.0, order=2, bcs=2),
             RunParams(nx=33, ny=21, nz=17, xr=4.0, yr=3.0, zr=2.5, order=4, bcs=2, storage="c64"),
             RunParams(nx=12, ny=10, nz=8, xr=2.0, yr=1.5, zr=1.0, order=4, bcs=1)]
    try:
        for p in cases:
            psi = np.ascontiguousarray(1.0 + 0.3 * (rng.standard_normal(p.shape) + 1j * rng.standard_normal(p.shape)))
            with Context(p) as ctx:
                ctx.upload(psi)
                exp = ctx.energy()
                cp = p.to_c()
                rc = fn(ct.byref(cp), ct.c_void_p(ctx.state_ptr), dout, ct.c_void_p(ctx.stream))
                assert rc == 0, lib.gpe_last_error(None)
                ctx.sync()
                got = ct.c_double()
                assert hip.h.hipMemcpy(ct.addressof(got), dout, 8, 2) == 0  # device to host
            print(f"{p.nx}x{p.ny}x{p.nz} order {p.order} {p.storage}: noalloc {got.value!r} gpe_energy {exp!r}")
            assert got.value == exp
    finally:
        hip.free(dout)
