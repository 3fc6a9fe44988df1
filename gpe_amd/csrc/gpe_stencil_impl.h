// Body of the stencil launcher: dispatches on the tile shape.  Included only
// by the per-(precision, order) kernel translation units.
#pragma once

#include "gpe_launch.h"

namespace gpe {

template <typename R, int ORD, int TX, int TY, class Op>
static hipError_t launch_tile(Geom g, const StW<R>& w, const C<R>* u, const Op& op, double* partial, hipStream_t s) {
    if (g.nblocks <= 0) return hipSuccess;
    if (g.general)
        hipLaunchKernelGGL((k_stencil<R, ORD, TX, TY, Op, true>), dim3(g.nblocks), dim3(TX * TY), 0, s, g, w, u, op,
                           partial);
    else
        hipLaunchKernelGGL((k_stencil<R, ORD, TX, TY, Op, false>), dim3(g.nblocks), dim3(TX * TY), 0, s, g, w, u, op,
                           partial);
    return hipGetLastError();
}

template <typename R, int ORD, class Op>
hipError_t launch_stencil(const Tiles& t, Geom g, const StW<R>& w, const C<R>* u, const Op& op, double* partial,
                          hipStream_t s) {
    tile_geom(t, g);
    switch (t.tx * 100 + t.ty) {
        case 6404: return launch_tile<R, ORD, 64, 4, Op>(g, w, u, op, partial, s);
        case 6408: return launch_tile<R, ORD, 64, 8, Op>(g, w, u, op, partial, s);
        case 6416: return launch_tile<R, ORD, 64, 16, Op>(g, w, u, op, partial, s);
        case 3204: return launch_tile<R, ORD, 32, 4, Op>(g, w, u, op, partial, s);
        case 3208: return launch_tile<R, ORD, 32, 8, Op>(g, w, u, op, partial, s);
        case 3216: return launch_tile<R, ORD, 32, 16, Op>(g, w, u, op, partial, s);
        default: return hipErrorInvalidValue;
    }
}

template <int ORD, int TX, int TY, class Op>
static hipError_t launch_tile_sum(Geom g, const StW<double>& w, const SumSrc& src, const Op& op, double* partial,
                                  hipStream_t s) {
    if (g.nblocks <= 0) return hipSuccess;
    if (g.general)
        hipLaunchKernelGGL((k_stencil_sum<ORD, TX, TY, Op, true>), dim3(g.nblocks), dim3(TX * TY), 0, s, g, w, src, op,
                           partial);
    else
        hipLaunchKernelGGL((k_stencil_sum<ORD, TX, TY, Op, false>), dim3(g.nblocks), dim3(TX * TY), 0, s, g, w, src, op,
                           partial);
    return hipGetLastError();
}

template <int ORD, class Op>
hipError_t launch_stencil_sum(const Tiles& t, Geom g, const StW<double>& w, const SumSrc& src, const Op& op,
                              double* partial, hipStream_t s) {
    tile_geom(t, g);
    switch (t.tx * 100 + t.ty) {
        case 6404: return launch_tile_sum<ORD, 64, 4, Op>(g, w, src, op, partial, s);
        case 6408: return launch_tile_sum<ORD, 64, 8, Op>(g, w, src, op, partial, s);
        case 6416: return launch_tile_sum<ORD, 64, 16, Op>(g, w, src, op, partial, s);
        case 3204: return launch_tile_sum<ORD, 32, 4, Op>(g, w, src, op, partial, s);
        case 3208: return launch_tile_sum<ORD, 32, 8, Op>(g, w, src, op, partial, s);
        case 3216: return launch_tile_sum<ORD, 32, 16, Op>(g, w, src, op, partial, s);
        default: return hipErrorInvalidValue;
    }
}

#define GPE_INSTANTIATE_SUM(OP, ORD)                                                                       \
    template hipError_t launch_stencil_sum<ORD, OP>(const Tiles&, Geom, const StW<double>&, const SumSrc&, \
                                                    const OP&, double*, hipStream_t);

#define GPE_INSTANTIATE(R, OP, ORD)                                                                                 \
    template hipError_t launch_stencil<R, ORD, OP>(const Tiles&, Geom, const StW<R>&, const C<R>*, const OP&, double*, \
                                                   hipStream_t);

}  // namespace gpe
