// Stencil kernels, complex64 fields, FP32 stencil, order 2: explicit instantiations of the
// launcher for every stage operator.
#include "gpe_stencil_impl.h"

namespace gpe {
#define GPE_INST_HERE(R, OP) GPE_INSTANTIATE(R, OP, 2)
GPE_FOR_EACH_OP(GPE_INST_HERE, float)
}  // namespace gpe
