// Stencil kernels, complex128 fields, FP64 stencil, order 4: explicit instantiations of the
// launcher for every stage operator.
#include "gpe_stencil_impl.h"

namespace gpe {
#define GPE_INST_HERE(R, OP) GPE_INSTANTIATE(R, OP, 4)
GPE_FOR_EACH_OP(GPE_INST_HERE, double)
}  // namespace gpe
