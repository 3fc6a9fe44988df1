"""CPU checks of the C ABI: libgpehip.so loads, exports every entry point
declared in include/gpe_hip.h, and validates parameters before touching a
device.  No compute call is made (there is no GPU here)."""
import ctypes as ct
import os
import re

import pytest

from gpe_amd import _lib
from gpe_amd.params import RunParams, read_namelist, xr_for_spacing

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "gpe_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpe_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert {n for n, _, _ in _lib.PROTOTYPES} == set(names)
    assert lib.gpe_abi_version() == _lib.GPE_ABI_VERSION == 2


def test_kernel_kind_names_match_header_count():
    text = open(os.path.join(ROOT, "include", "gpe_hip.h")).read()
    assert int(re.search(r"#define GPE_NKINDS (\d+)", text).group(1)) == _lib.GPE_NKINDS
    lib = _lib.load()
    names = [lib.gpe_kernel_kind_name(k).decode() for k in range(_lib.GPE_NKINDS)]
    assert all(names) and len(set(names)) == len(names)
    assert {"rk4_stage1", "rk4_stage2", "rk4_stage3", "rk4_stage4"} <= set(names)
    assert lib.gpe_kernel_kind_name(_lib.GPE_NKINDS) == b""


def test_storage_modes_match_header():
    text = open(os.path.join(ROOT, "include", "gpe_hip.h")).read()
    for name in ("GPE_C128", "GPE_C64", "GPE_MIXED"):
        assert int(re.search(rf"#define {name} (\d+)", text).group(1)) == getattr(_lib, name)
    p = RunParams(nx=8, ny=8, nz=8, xr=2.0, yr=2.0, zr=2.0)
    assert [p.replace(storage=s).to_c().storage for s in ("c128", "c64", "mixed")] == [0, 1, 2]
    with pytest.raises(ValueError, match="storage"):
        p.replace(storage="c32").to_c()


def test_struct_layouts_match_header():
    # gpe_params: 8 int32 then 16 doubles; gpe_step_info: 6 doubles + 3 int64
    assert ct.sizeof(_lib.gpe_params) == 8 * 4 + 16 * 8
    assert ct.sizeof(_lib.gpe_step_info) == 6 * 8 + 3 * 8
    # gpe_line: 7 doubles then 2 int32
    assert ct.sizeof(_lib.gpe_line) == 7 * 8 + 2 * 4


@pytest.mark.parametrize("field,value,msg", [
    ("order", 3, "order"), ("bcs", 0, "bcs"), ("eqn_to_solve", 5, "eqn_to_solve"), ("nx", 4, ">= 6"),
    ("xr", -1.0, "xr"),
])
def test_ctx_create_validates_params(field, value, msg):
    lib = _lib.load()
    p = RunParams().replace(**{field: value}).to_c()
    h = ct.c_void_p()
    rc = lib.gpe_ctx_create(ct.byref(p), ct.byref(h))
    assert rc == -1
    assert msg in lib.gpe_last_error(None).decode()
    assert not h.value


def test_null_arguments_are_rejected():
    lib = _lib.load()
    assert lib.gpe_step(None, 2, 1, None) == -1
    assert lib.gpe_norm(None, None) == -1
    assert b"ctx is NULL" in lib.gpe_last_error(None)


def test_namelist_reader_reads_reference_run_in_format():
    text = """
&run_params
  tau           = 0.01
  end_time      = 1.995
  xr            = 15.75
  scheme        = 'rk4'
  eqn_to_solve  = 1
  bcs           = 1
  order         = 2
  real_time     = .true.
/
&eqn_params
  Urhs          = 0.0
  g             = 1.0
  omx           = 0.05
/
&io_params
  save_rate     = 10
/
"""
    p = read_namelist(text)
    assert p.tau == 0.01 and p.end_time == 1.995 and p.xr == 15.75 and p.scheme == "rk4"
    assert p.real_time is True and p.omx == 0.05 and p.order == 2
    assert abs(xr_for_spacing(64, 0.5) - 15.75) < 1e-15
    assert abs(RunParams(nx=64, xr=15.75).dx - 0.5) < 1e-15
