// Body of the stencil launcher: dispatches on the tile shape.  Included only
// by the per-(precision, order) kernel translation units.
#pragma once

#include "gpe_launch.h"

namespace gpe {

template <typename R, int ORD, int TX, int TY, class Op>
static hipError_t launch_tile(Geom g, const StW<R>& w, const C<R>* u, const Op& op, double* partial, hipStream_t s) {
    if (g.nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_stencil<R, ORD, TX, TY, Op>), dim3(g.nblocks), dim3(TX * TY), 0, s, g, w, u, op, partial);
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

#define GPE_INSTANTIATE(R, OP, ORD)                                                                                 \
    template hipError_t launch_stencil<R, ORD, OP>(const Tiles&, Geom, const StW<R>&, const C<R>*, const OP&, double*, \
                                                   hipStream_t);

}  // namespace gpe
