// Fused Cash-Karp attempt for mixed storage (three kernels of two stages
// each, gpe_fused_mx.h): launchers and instantiations.
#include "gpe_fused_mx.h"

namespace gpe {

template <int ORD, int TERMS, int GM>
static hipError_t launch_ckmx_pair(const Geom& g, const StW<double>& w, const RhsC<double>& k, const CkMxArgs& a,
                                   int pair, double* partial, hipStream_t s) {
    const dim3 grid(g.nblocks), block(1024);
    if (pair == 1) hipLaunchKernelGGL((k_ckmx_pair<ORD, TERMS, 1, GM>), grid, block, 0, s, g, w, k, a, partial);
    else if (pair == 2) hipLaunchKernelGGL((k_ckmx_pair<ORD, TERMS, 2, GM>), grid, block, 0, s, g, w, k, a, partial);
    else hipLaunchKernelGGL((k_ckmx_pair<ORD, TERMS, 3, GM>), grid, block, 0, s, g, w, k, a, partial);
    return hipGetLastError();
}

template <int ORD, int TERMS>
static hipError_t launch_ckmx_pair_gm(const Geom& g, const StW<double>& w, const RhsC<double>& k, const CkMxArgs& a,
                                      int pair, double* partial, hipStream_t s) {
    switch (g_mode(a.dt.x, a.dt.y)) {
        case GM_IM: return launch_ckmx_pair<ORD, TERMS, GM_IM>(g, w, k, a, pair, partial, s);
        case GM_RE: return launch_ckmx_pair<ORD, TERMS, GM_RE>(g, w, k, a, pair, partial, s);
        default: return launch_ckmx_pair<ORD, TERMS, GM_GEN>(g, w, k, a, pair, partial, s);
    }
}

template <int ORD>
hipError_t launch_ckmx_fused(const Geom& g, const StW<double>& w, const RhsC<double>& k, const CkMxArgs& a, int pair,
                             double* partial, hipStream_t s) {
    if (g.nblocks <= 0) return hipSuccess;
    if (pair < 1 || pair > 3) return hipErrorInvalidValue;
    return g.general ? launch_ckmx_pair_gm<ORD, RHS_V>(g, w, k, a, pair, partial, s)
                     : launch_ckmx_pair_gm<ORD, 0>(g, w, k, a, pair, partial, s);
}

template hipError_t launch_ckmx_fused<2>(const Geom&, const StW<double>&, const RhsC<double>&, const CkMxArgs&, int,
                                         double*, hipStream_t);
template hipError_t launch_ckmx_fused<4>(const Geom&, const StW<double>&, const RhsC<double>&, const CkMxArgs&, int,
                                         double*, hipStream_t);

}  // namespace gpe
