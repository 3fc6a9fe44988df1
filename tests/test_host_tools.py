"""Host-side measurement tools (no GPU): the storage/kernel variant names of
the studies and the labels and algorithmic bytes that the rocprofv3
summaries attach to each kernel."""
import pytest

from gpe_amd.profsum import cell_bytes, kernel_label
from gpe_amd.studies import variant


def test_variant_names():
    assert variant("c128") == ("c128", True)
    assert variant("c128/stages") == ("c128", False)
    assert variant("mixed") == ("mixed", True)
    assert variant("c64/stages") == ("c64", False)
    with pytest.raises(ValueError):
        variant("c128/fast")


# (demangled kernel name as rocprofv3 reports it, label, algorithmic bytes per cell)
CASES = [
    ("void gpe::k_rk4_f12<double, 4, 0>(gpe::Geom, gpe::StW<double>, gpe::RhsC<double>, gpe::cxt<double>::T const*, "
     "gpe::cxt<double>::T*, gpe::cxt<double>::T)", "rk4_fused12 c128 order4 tile68x15", 32),
    ("void gpe::k_rk4_f34<float, 2, 1>(gpe::Geom, gpe::StW<float>, gpe::RhsC<float>, ...)",
     "rk4_fused34 c64 order2 tile66x15 with-trap", 24),
    ("void gpe::k_ck_pair<double, 4, 0, 1>(gpe::Geom, gpe::StW<double>, gpe::RhsC<double>, gpe::CkPairArgs<double>, "
     "double*)", "rkck_fused12 c128 order4 tile68x15", 64),
    ("void gpe::k_ck_pair<float, 4, 0, 3>(gpe::Geom, gpe::StW<float>, gpe::RhsC<float>, gpe::CkPairArgs<float>, "
     "double*)", "rkck_fused56_error c64 order4 tile68x15", 56),
    ("void gpe::k_ckmx_pair<4, 0, 2>(gpe::Geom, gpe::StW<double>, gpe::RhsC<double>, gpe::CkMxArgs, double*)",
     "mx_rkck_fused34 mixed order4 tile68x15", 64),
    ("void gpe::k_ckmx_pair<4, 1, 3>(gpe::Geom, gpe::StW<double>, gpe::RhsC<double>, gpe::CkMxArgs, double*)",
     "mx_rkck_fused56_error mixed order4 tile68x15 with-trap", 72),
]


@pytest.mark.parametrize("name,label,nbytes", CASES, ids=[c[1] for c in CASES])
def test_fused_kernel_labels_and_bytes(name, label, nbytes):
    got = kernel_label(name)
    assert got == label
    kind, prec = got.split()[:2]
    assert cell_bytes(kind, prec) == nbytes


def test_fused_attempt_bytes_add_up():
    """18 field accesses per Cash-Karp attempt (288 B c128, 144 B c64) and
    176 B with mixed storage; 5 accesses per RK4 step."""
    ck = ("rkck_fused12", "rkck_fused34", "rkck_fused56_error")
    assert sum(cell_bytes(k, "c128") for k in ck) == 288
    assert sum(cell_bytes(k, "c64") for k in ck) == 144
    assert sum(cell_bytes("mx_" + k, "mixed") for k in ck) == 176
    assert cell_bytes("rk4_fused12", "c128") + cell_bytes("rk4_fused34", "c128") == 80
