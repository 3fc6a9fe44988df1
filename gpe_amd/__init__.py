"""MI355X (gfx950) implementation of the GPE time-stepping hot path of the
reference Fortran solver liangjj/gpe: FD derivatives, get_rhs and the
euler/rk2/rk4/rkqs(Cash-Karp) steppers with their reductions, as HIP
kernels in libgpehip.so behind the C ABI of include/gpe_hip.h."""
from ._lib import GpeError, LIB_PATH, load  # noqa: F401
from .context import Context, StepInfo  # noqa: F401
from .params import RunParams, read_namelist, xr_for_spacing  # noqa: F401
