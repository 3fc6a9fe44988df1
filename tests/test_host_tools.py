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


class _St:
    def __init__(self, x):
        self.wall_s = 1.0 / x


@pytest.mark.parametrize("order,x0,slope,slope_max,factor", [(0.8, 1e-6, 0.8, 2.0, 10.0), (4.0, 0.01, 4.0, 6.0, 2.0)])
def test_tune_to_error_lands_under_target(order, x0, slope, slope_max, factor):
    """A method whose error grows as C x^order lands within [target/1.3,
    target]: a tolerance (order ~0.8) and a fixed RK4 step (order 4)."""
    from gpe_amd.studies import tune_to_error

    C = 3.5e-6 / 0.01 ** order if order == 4.0 else 0.2

    def run(x):
        return _St(x), C * x ** order

    x, st, err, tries = tune_to_error(run, 1e-6, eps0=x0, slope=slope, slope_max=slope_max, max_factor=factor)
    assert 1e-6 / 1.3 <= err <= 1e-6
    assert abs(err - C * x ** order) < 1e-18
    assert all(max(a, b) / min(a, b) <= factor * (1 + 1e-12) for (a, _, _), (b, _, _) in zip(tries, tries[1:]))


def test_tune_to_error_unstable_step_and_floor():
    """An unstable step (a non-finite error) counts as 1e3 and the search
    moves back; an error floor above the target returns no landing."""
    from gpe_amd.studies import tune_to_error

    def unstable(dt):
        return _St(dt), float("nan") if dt > 0.03 else 3.5e-6 * (dt / 0.01) ** 4

    dt, st, err, tries = tune_to_error(unstable, 1e-6, eps0=0.05, slope=4.0, slope_max=6.0, max_factor=2.0)
    assert tries[0][1] == 1e3 and 1e-6 / 1.3 <= err <= 1e-6 and dt < 0.01

    def floor(eps):
        return _St(eps), 3.9e-5 + eps

    eps, st, err, tries = tune_to_error(floor, 1e-6)
    assert eps is None and st is None and len(tries) >= 2
