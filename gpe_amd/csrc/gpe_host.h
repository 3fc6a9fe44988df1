// Host-side helpers shared by the C-ABI translation units: grid spacing,
// stencil weights, RHS coefficients and complex dt factors.
#pragma once

#include "../../include/gpe_hip.h"
#include "gpe_launch.h"

namespace gpe {

inline void spacing(const gpe_params& p, double d[3]) {
    // dx = (xr - xl)/nx1 with xl = -xr (src/ic.f90:43-46)
    d[0] = (p.xr + p.xr) / (p.nx - 1);
    d[1] = (p.yr + p.yr) / (p.ny - 1);
    d[2] = (p.zr + p.zr) / (p.nz - 1);
}

template <typename R> inline StW<R> make_w(const gpe_params& p) {
    double d[3];
    spacing(p, d);
    StW<R> w;
    if (p.order == 2) {
        w.w1x = (R)(1.0 / (2.0 * d[0])); w.w1y = (R)(1.0 / (2.0 * d[1])); w.w1z = (R)(1.0 / (2.0 * d[2]));
        w.w2x = (R)(1.0 / (d[0] * d[0])); w.w2y = (R)(1.0 / (d[1] * d[1])); w.w2z = (R)(1.0 / (d[2] * d[2]));
        w.w2c = (R)(2.0 * (1.0 / (d[0] * d[0]) + 1.0 / (d[1] * d[1]) + 1.0 / (d[2] * d[2])));
    } else {
        w.w1x = (R)(1.0 / (12.0 * d[0])); w.w1y = (R)(1.0 / (12.0 * d[1])); w.w1z = (R)(1.0 / (12.0 * d[2]));
        w.w2x = (R)(1.0 / (12.0 * d[0] * d[0])); w.w2y = (R)(1.0 / (12.0 * d[1] * d[1]));
        w.w2z = (R)(1.0 / (12.0 * d[2] * d[2]));
        w.w2c = (R)(30.0 * (1.0 / (12.0 * d[0] * d[0]) + 1.0 / (12.0 * d[1] * d[1]) + 1.0 / (12.0 * d[2] * d[2])));
    }
    return w;
}

template <typename R> inline RhsC<R> make_rhs_p(const gpe_params& p, const void* dtab) {
    double d[3];
    spacing(p, d);
    RhsC<R> k{};
    int flags = 0;
    switch (p.eqn_to_solve) {
        case 1:
        case 3:
            // 0.5*(i + diss)*(lap + (1 - |psi|^2 [- vscale V]) psi) + U dpsi/dx  (src/solve.f90:415-418, 425-428)
            k.a = 1; k.b = 1; k.c = -1; k.d = (R)(-p.vscale); k.pim = (R)0.5;
            if (p.vscale != 0.0) flags |= RHS_V;
            if (p.diss_amp != 0.0) flags |= RHS_DISS;
            if (p.Urhs != 0.0) flags |= RHS_U;
            if (p.eqn_to_solve == 3) flags |= RHS_MASK;
            break;
        case 2:
            // i*(lap - |psi|^2 psi) + U dpsi/dx  (src/solve.f90:420-423)
            k.a = 1; k.b = 0; k.c = -1; k.d = 0; k.pim = 1;
            if (p.Urhs != 0.0) flags |= RHS_U;
            break;
        default:
            // -i*(-0.5 lap + V psi + g|psi|^2 psi - mu psi)  (src/solve.f90:430-436)
            k.a = (R)-0.5; k.b = (R)(-p.mu); k.c = (R)p.g; k.d = 1; k.pim = -1;
            flags |= RHS_V;
            break;
    }
    k.U = (R)p.Urhs;
    k.vx = (R)(0.5 * p.omx * p.omx);
    k.vy = (R)(0.5 * p.omy * p.omy);
    k.vz = (R)(0.5 * p.omz * p.omz);
    k.xl = (R)(-p.xr); k.yl = (R)(-p.yr); k.zl = (R)(-p.zr);
    k.dx = (R)d[0]; k.dy = (R)d[1]; k.dz = (R)d[2];
    k.diss_amp = (R)p.diss_amp;
    k.dtab = (const R*)dtab;
    k.nx = p.nx;
    k.ny = p.ny;
    k.flags = flags;
    return k;
}

template <typename R> inline C<R> cdt(double re, double im, double s) { return mkc<R>((R)(re * s), (R)(im * s)); }


}  // namespace gpe
