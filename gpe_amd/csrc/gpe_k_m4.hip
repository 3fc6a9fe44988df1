// Stencil kernels of the mixed storage mode (complex128 state, complex64
// stage increments, FP64 stencil), order 4: explicit instantiations.
#include "gpe_stencil_impl.h"

namespace gpe {
#define GPE_INST_PLAIN_HERE(R, OP) GPE_INSTANTIATE(R, OP, 4)
GPE_FOR_EACH_MIXED_PLAIN(GPE_INST_PLAIN_HERE)
#define GPE_INST_SUM_HERE(OP) GPE_INSTANTIATE_SUM(OP, 4)
GPE_FOR_EACH_MIXED_SUM(GPE_INST_SUM_HERE)
}  // namespace gpe
