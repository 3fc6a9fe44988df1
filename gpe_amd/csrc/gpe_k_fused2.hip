// Fused midpoint RK2 step and fused pair of Euler steps (the stages-1+2
// body of the fused RK4 pair, gpe_fused.h): launchers and instantiations.
#include "gpe_fused.h"

namespace gpe {

template <typename R, int ORD, int GM>
static void launch_pair_fused_gm(const Geom& g, const StW<R>& w, const RhsC<R>& k, const C<R>* psi, C<R>* out,
                                 C<R> dt, int pk, hipStream_t s) {
    constexpr int NT = FusedShape<ORD, kRk4OX, kRk4TYB>::NT;
    if (pk == PK_RK2) {
        const C<R> hdt = mkc<R>(R(0.5) * dt.x, R(0.5) * dt.y);
        if (g.general)
            hipLaunchKernelGGL((k_rk2_fused<R, ORD, RHS_V, GM>), dim3(g.nblocks), dim3(NT), 0, s, g, w, k, psi, out, hdt,
                               dt);
        else
            hipLaunchKernelGGL((k_rk2_fused<R, ORD, 0, GM>), dim3(g.nblocks), dim3(NT), 0, s, g, w, k, psi, out, hdt, dt);
    } else {
        if (g.general)
            hipLaunchKernelGGL((k_euler2_fused<R, ORD, RHS_V, GM>), dim3(g.nblocks), dim3(NT), 0, s, g, w, k, psi, out,
                               dt);
        else
            hipLaunchKernelGGL((k_euler2_fused<R, ORD, 0, GM>), dim3(g.nblocks), dim3(NT), 0, s, g, w, k, psi, out, dt);
    }
}

template <typename R, int ORD>
hipError_t launch_pair_fused(const Geom& g, const StW<R>& w, const RhsC<R>& k, const C<R>* psi, C<R>* out, C<R> dt,
                             int pk, hipStream_t s) {
    if (g.nblocks <= 0) return hipSuccess;
    if (pk != PK_RK2 && pk != PK_EULER2) return hipErrorInvalidValue;
    switch (g_mode((double)dt.x, (double)dt.y)) {
        case GM_IM: launch_pair_fused_gm<R, ORD, GM_IM>(g, w, k, psi, out, dt, pk, s); break;
        case GM_RE: launch_pair_fused_gm<R, ORD, GM_RE>(g, w, k, psi, out, dt, pk, s); break;
        default: launch_pair_fused_gm<R, ORD, GM_GEN>(g, w, k, psi, out, dt, pk, s); break;
    }
    return hipGetLastError();
}

#define GPE_PAIR_INST(R, ORD)                                                                                   \
    template hipError_t launch_pair_fused<R, ORD>(const Geom&, const StW<R>&, const RhsC<R>&, const C<R>*, C<R>*, \
                                                  C<R>, int, hipStream_t);
GPE_PAIR_INST(double, 2)
GPE_PAIR_INST(double, 4)
GPE_PAIR_INST(float, 2)
GPE_PAIR_INST(float, 4)

}  // namespace gpe
