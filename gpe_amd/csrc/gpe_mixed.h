// Mixed-storage stage operators: the state psi stays complex128 and every
// stencil and update is evaluated in FP64, but the stage data that only
// carries O(dt) increments (k_j = dt f(u_j) and the partial sums of the
// Butcher rows) is stored as complex64.  A stage stencils psi0 + increment
// through SumSrc, so the stage value u_j is never written.  Rounding an
// increment of size O(dt |f|) to FP32 perturbs psi by ~6e-8 dt |f| per stage,
// against ~6e-8 |psi| per step for complex64 state storage.
//
// Field bytes per cell (state 16 B, increment 8 B):
//   RK4  : stage1 24, stage2 32, stage3 32, stage4 56      -> 144 (c128: 208)
//   RKCK : stage1 24, stage2 40, 56, 64, 72, stage6 72     -> 328 (c128: 576)
#pragma once

#include "gpe_kernels.h"

namespace gpe {

using CD = C<double>;
using CF = C<float>;

__device__ __forceinline__ CD tod(CF v) { return mkc<double>(v.x, v.y); }
__device__ __forceinline__ CF tof(CD v) { return mkc<float>((float)v.x, (float)v.y); }

struct NoPreM {};

// RK4 stages 1-3: K = dt f(u), u = psi0 (stage 1), psi0 + K1/2, psi0 + K2/2.
template <int KIND, int BYTES> struct OpMxRk4K {
    static constexpr int kKind = KIND, kBytesPerCell = BYTES;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    using Pre = NoPreM;
    RhsC<double> k;
    CF* __restrict__ kout;
    CD dt;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre&, double*) const {
        st_stream<float>(kout + i, tof(cmul<double>(dt, rhs_eval<double>(k, st, x, y, z))));
    }
    template <class S>
    __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, float2, double* r) const {
        apply(i, x, y, z, st, p, r);
    }
};
using OpMxRk4S1 = OpMxRk4K<16, 24>;
using OpMxRk4S23 = OpMxRk4K<17, 32>;

// RK4 stage 4 over u4 = psi0 + K3 (aux = K3 of the own cell):
// psi_new = psi0 + (K1 + 2 K2 + 2 K3 + k4)/6 = u4 + (K1 + 2 K2 - 4 K3 + k4)/6.
struct OpMxRk4Last {
    static constexpr int kKind = 18, kBytesPerCell = 56;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    struct Pre { CF k1, k2; };
    RhsC<double> k;
    const CF* __restrict__ k1;
    const CF* __restrict__ k2;
    CD* __restrict__ out;
    CD dt;
    __device__ Pre load(Cell i) const { return {ld_stream<float>(k1 + i), ld_stream<float>(k2 + i)}; }
    template <class S>
    __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, float2 k3f, double*) const {
        const CD k4 = cmul<double>(dt, rhs_eval<double>(k, st, x, y, z));
        const CD a = tod(p.k1), b = tod(p.k2), c = tod(k3f);
        const double sx = (a.x + 2.0 * b.x) - 4.0 * c.x + k4.x, sy = (a.y + 2.0 * b.y) - 4.0 * c.y + k4.y;
        const CD u4 = st.c();
        constexpr double sixth = 1.0 / 6.0;
        st_stream<double>(out + i, mkc<double>(u4.x + sx * sixth, u4.y + sy * sixth));
    }
};

// Cash-Karp stage 1: F0 = f(psi0) (kept for the retries of a rejected step).
struct OpMxCk1 {
    static constexpr int kKind = 19, kBytesPerCell = 24;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    using Pre = NoPreM;
    RhsC<double> k;
    CF* __restrict__ f0;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre&, double*) const {
        st_stream<float>(f0 + i, tof(rhs_eval<double>(k, st, x, y, z)));
    }
};

// Cash-Karp stage 2 over u2 = psi0 + b21 dt F0 (aux = F0 of the own cell):
// K2 = dt f(u2), D3 = b31 k1 + b32 k2.
struct OpMxCk2 {
    static constexpr int kKind = 20, kBytesPerCell = 40;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    using Pre = NoPreM;
    RhsC<double> k;
    CF* __restrict__ k2;
    CF* __restrict__ dnext;
    CD dt;
    CkTab<double> t;
    __device__ Pre load(Cell) const { return {}; }
    template <class S>
    __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre&, float2 f0, double*) const {
        const CD ks = cmul<double>(dt, rhs_eval<double>(k, st, x, y, z));
        const CD k1 = cmul<double>(dt, tod(f0));
        st_stream<float>(k2 + i, tof(ks));
        st_stream<float>(dnext + i, tof(caxpy<double>(t.b32, ks, cscale<double>(t.b31, k1))));
    }
};

// Cash-Karp stages s = 3..5 over u_s = psi0 + D_s: K_s = dt f(u_s),
// D_{s+1} = sum_{j <= s} b_{s+1,j} k_j.
template <int S_> struct OpMxCkMid {
    static constexpr int kKind = 18 + S_, kBytesPerCell = 24 + 8 * (S_ - 1) + 16;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    struct Pre { CF f0, k2, k3, k4; };
    RhsC<double> k;
    const CF* __restrict__ f0;
    const CF* __restrict__ k2;
    const CF* __restrict__ k3;
    const CF* __restrict__ k4;
    CF* __restrict__ kout;
    CF* __restrict__ dnext;
    CD dt;
    CkTab<double> t;
    __device__ Pre load(Cell i) const {
        Pre p;
        p.f0 = ld_stream<float>(f0 + i);
        p.k2 = ld_stream<float>(k2 + i);
        if constexpr (S_ >= 4) p.k3 = ld_stream<float>(k3 + i);
        if constexpr (S_ >= 5) p.k4 = ld_stream<float>(k4 + i);
        return p;
    }
    template <class S>
    __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, float2, double*) const {
        const CD ks = cmul<double>(dt, rhs_eval<double>(k, st, x, y, z));
        const CD k1 = cmul<double>(dt, tod(p.f0));
        const CD k2v = tod(p.k2);
        st_stream<float>(kout + i, tof(ks));
        CD d;
        if constexpr (S_ == 3) {
            d = caxpy<double>(t.b43, ks, caxpy<double>(t.b42, k2v, cscale<double>(t.b41, k1)));
        } else if constexpr (S_ == 4) {
            d = caxpy<double>(t.b54, ks,
                              caxpy<double>(t.b53, tod(p.k3), caxpy<double>(t.b52, k2v, cscale<double>(t.b51, k1))));
        } else {
            d = caxpy<double>(
                t.b65, ks,
                caxpy<double>(t.b64, tod(p.k4),
                              caxpy<double>(t.b63, tod(p.k3), caxpy<double>(t.b62, k2v, cscale<double>(t.b61, k1)))));
        }
        st_stream<float>(dnext + i, tof(d));
    }
};

// Cash-Karp stage 6 over u6 = psi0 + D6 (aux = D6 of the own cell), the
// 5th-order solution and the scaled error max-norm, as OpCkLast:
// red[0] = max |err| / (|psi0| + dt0 |f0| + 1e-30), red[1] = max |psi_new|^2.
struct OpMxCkLast {
    static constexpr int kKind = 24, kBytesPerCell = 72;
    static constexpr bool kPrefetch = false;
    static constexpr int kDepth = 1;
    static constexpr int kNRed = 2, kMode = RED_MAX;
    struct Pre { CF f0, k3, k4, k5; };
    RhsC<double> k;
    const CF* __restrict__ f0;
    const CF* __restrict__ k3;
    const CF* __restrict__ k4;
    const CF* __restrict__ k5;
    CD* __restrict__ out;
    CD dt;
    double dt0_re, dt0_im;
    CkTab<double> t;
    __device__ Pre load(Cell i) const {
        return {ld_stream<float>(f0 + i), ld_stream<float>(k3 + i), ld_stream<float>(k4 + i), ld_stream<float>(k5 + i)};
    }
    template <class S>
    __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, float2 d6, double* red) const {
        const CD k6 = cmul<double>(dt, rhs_eval<double>(k, st, x, y, z));
        const CD f0v = tod(p.f0);
        const CD k1 = cmul<double>(dt, f0v);
        const CD k3v = tod(p.k3), k4v = tod(p.k4), k5v = tod(p.k5);
        const CD p0 = csub<double>(st.c(), tod(d6));
        const CD nw = caxpy<double>(t.c6, k6, caxpy<double>(t.c4, k4v, caxpy<double>(t.c3, k3v, caxpy<double>(t.c1, k1, p0))));
        st_stream<double>(out + i, nw);
        const double er = t.dc1 * k1.x + t.dc3 * k3v.x + t.dc4 * k4v.x + t.dc5 * k5v.x + t.dc6 * k6.x;
        const double ei = t.dc1 * k1.y + t.dc3 * k3v.y + t.dc4 * k4v.y + t.dc5 * k5v.y + t.dc6 * k6.y;
        const double ap = sqrt(p0.x * p0.x + p0.y * p0.y);
        const double af = sqrt(f0v.x * f0v.x + f0v.y * f0v.y);
        const double sr = ap + dt0_re * af + 1e-30, si = dt0_im * af;
        const double e = sqrt((er * er + ei * ei) / (sr * sr + si * si));
        red[0] = e > red[0] ? e : red[0];
        const double a2 = nw.x * nw.x + nw.y * nw.y;
        red[1] = a2 > red[1] ? a2 : red[1];
    }
};

}  // namespace gpe
