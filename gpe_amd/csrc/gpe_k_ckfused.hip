// Fused Cash-Karp attempt (three kernels of two stages each, gpe_fused.h):
// launchers and instantiations for complex128 and complex64 storage (mixed
// storage: gpe_k_ckmx.hip).
#include "gpe_fused.h"

namespace gpe {

template <typename R, int ORD, int TERMS, int GM>
static hipError_t launch_ck_pair(const Geom& g, const StW<R>& w, const RhsC<R>& k, const CkPairArgs<R>& a, int pair,
                                 double* partial, hipStream_t s) {
    const dim3 grid(g.nblocks), block(1024);
    if (pair == 1) hipLaunchKernelGGL((k_ck_pair<R, ORD, TERMS, 1, GM>), grid, block, 0, s, g, w, k, a, partial);
    else if (pair == 2) hipLaunchKernelGGL((k_ck_pair<R, ORD, TERMS, 2, GM>), grid, block, 0, s, g, w, k, a, partial);
    else hipLaunchKernelGGL((k_ck_pair<R, ORD, TERMS, 3, GM>), grid, block, 0, s, g, w, k, a, partial);
    return hipGetLastError();
}

template <typename R, int ORD, int TERMS>
static hipError_t launch_ck_pair_gm(const Geom& g, const StW<R>& w, const RhsC<R>& k, const CkPairArgs<R>& a,
                                    int pair, double* partial, hipStream_t s) {
    switch (g_mode((double)a.dt.x, (double)a.dt.y)) {
        case GM_IM: return launch_ck_pair<R, ORD, TERMS, GM_IM>(g, w, k, a, pair, partial, s);
        case GM_RE: return launch_ck_pair<R, ORD, TERMS, GM_RE>(g, w, k, a, pair, partial, s);
        default: return launch_ck_pair<R, ORD, TERMS, GM_GEN>(g, w, k, a, pair, partial, s);
    }
}

template <typename R, int ORD>
hipError_t launch_ck_fused(const Geom& g, const StW<R>& w, const RhsC<R>& k, const CkPairArgs<R>& a, int pair,
                           double* partial, hipStream_t s) {
    if (g.nblocks <= 0) return hipSuccess;
    if (pair < 1 || pair > 3) return hipErrorInvalidValue;
    return g.general ? launch_ck_pair_gm<R, ORD, RHS_V>(g, w, k, a, pair, partial, s)
                     : launch_ck_pair_gm<R, ORD, 0>(g, w, k, a, pair, partial, s);
}

#define GPE_CKF_INST(R, ORD)                                                                                  \
    template hipError_t launch_ck_fused<R, ORD>(const Geom&, const StW<R>&, const RhsC<R>&, const CkPairArgs<R>&, \
                                                int, double*, hipStream_t);
GPE_CKF_INST(double, 2)
GPE_CKF_INST(double, 4)
GPE_CKF_INST(float, 2)
GPE_CKF_INST(float, 4)

}  // namespace gpe
