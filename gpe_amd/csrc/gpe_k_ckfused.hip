// Fused Cash-Karp attempt (three kernels of two stages each, gpe_fused.h,
// gpe_fused_mx.h): launchers and instantiations for complex128, complex64
// and mixed storage.
#include "gpe_fused_mx.h"

namespace gpe {

template <typename R, int ORD, int TERMS>
static hipError_t launch_ck_pair(const Geom& g, const StW<R>& w, const RhsC<R>& k, const CkPairArgs<R>& a, int pair,
                                 double* partial, hipStream_t s) {
    const dim3 grid(g.nblocks), block(1024);
    if (pair == 1) hipLaunchKernelGGL((k_ck_pair<R, ORD, TERMS, 1>), grid, block, 0, s, g, w, k, a, partial);
    else if (pair == 2) hipLaunchKernelGGL((k_ck_pair<R, ORD, TERMS, 2>), grid, block, 0, s, g, w, k, a, partial);
    else hipLaunchKernelGGL((k_ck_pair<R, ORD, TERMS, 3>), grid, block, 0, s, g, w, k, a, partial);
    return hipGetLastError();
}

template <typename R, int ORD>
hipError_t launch_ck_fused(const Geom& g, const StW<R>& w, const RhsC<R>& k, const CkPairArgs<R>& a, int pair,
                           double* partial, hipStream_t s) {
    if (g.nblocks <= 0) return hipSuccess;
    if (pair < 1 || pair > 3) return hipErrorInvalidValue;
    return g.general ? launch_ck_pair<R, ORD, RHS_V>(g, w, k, a, pair, partial, s)
                     : launch_ck_pair<R, ORD, 0>(g, w, k, a, pair, partial, s);
}

template <int ORD, int TERMS>
static hipError_t launch_ckmx_pair(const Geom& g, const StW<double>& w, const RhsC<double>& k, const CkMxArgs& a,
                                   int pair, double* partial, hipStream_t s) {
    const dim3 grid(g.nblocks), block(1024);
    if (pair == 1) hipLaunchKernelGGL((k_ckmx_pair<ORD, TERMS, 1>), grid, block, 0, s, g, w, k, a, partial);
    else if (pair == 2) hipLaunchKernelGGL((k_ckmx_pair<ORD, TERMS, 2>), grid, block, 0, s, g, w, k, a, partial);
    else hipLaunchKernelGGL((k_ckmx_pair<ORD, TERMS, 3>), grid, block, 0, s, g, w, k, a, partial);
    return hipGetLastError();
}

template <int ORD>
hipError_t launch_ckmx_fused(const Geom& g, const StW<double>& w, const RhsC<double>& k, const CkMxArgs& a, int pair,
                             double* partial, hipStream_t s) {
    if (g.nblocks <= 0) return hipSuccess;
    if (pair < 1 || pair > 3) return hipErrorInvalidValue;
    return g.general ? launch_ckmx_pair<ORD, RHS_V>(g, w, k, a, pair, partial, s)
                     : launch_ckmx_pair<ORD, 0>(g, w, k, a, pair, partial, s);
}

template hipError_t launch_ckmx_fused<2>(const Geom&, const StW<double>&, const RhsC<double>&, const CkMxArgs&, int,
                                         double*, hipStream_t);
template hipError_t launch_ckmx_fused<4>(const Geom&, const StW<double>&, const RhsC<double>&, const CkMxArgs&, int,
                                         double*, hipStream_t);

#define GPE_CKF_INST(R, ORD)                                                                                  \
    template hipError_t launch_ck_fused<R, ORD>(const Geom&, const StW<R>&, const RhsC<R>&, const CkPairArgs<R>&, \
                                                int, double*, hipStream_t);
GPE_CKF_INST(double, 2)
GPE_CKF_INST(double, 4)
GPE_CKF_INST(float, 2)
GPE_CKF_INST(float, 4)

}  // namespace gpe
