// Fused Cash-Karp attempt with mixed storage: the complex128 state psi0 and
// FP64 stencils and updates of the mixed-storage stage kernels (OpMxCk*,
// gpe_mixed.h), with their complex64 increments F0, K2..K4 and partial
// Butcher sums D, in the three stage pairs of the complex128 fused attempt
// (gpe_fused.h).  Every value that the stage kernels round to FP32 before
// storing is rounded at the same point here, and the stage values are
// formed as the stage kernels' stencil source forms them (SumSrc::comb:
// psi0 + s F0 for u2, psi0 + D for u3..u6).
//   pair 1: F0 = f(psi0), u2 = psi0 + b21 dt F0 (whole block);
//           K2 = dt f(u2), D3 = b31 k1 + b32 k2 (inner)
//                                   reads psi0 (16 B)          writes F0, K2, D3 (24 B)
//   pair 2: K3 = dt f(u3), D4 = b41 k1 + b42 k2 + b43 k3, u4 = psi0 + D4;
//           K4 = dt f(u4), D5 = b51 k1 + .. + b54 k4
//                                   reads psi0, D3 (stencil), F0, K2 (40 B)   writes K3, K4, D5 (24 B)
//   pair 3: k5 = dt f(u5), D6 = b61 k1 + .. + b65 k5, u6 = psi0 + D6;
//           k6 = dt f(u6), psi_new, err and the block maxima as OpMxCkLast
//                                   reads psi0, D5 (stencil), F0, K2, K3, K4 (56 B)   writes psi_new (16 B)
// 176 B per cell per attempt (the mixed stage kernels: 328 B).  Pair 2
// keeps psi0 of the own column from its stencil source's load in an LDS ring
// until the first stage needs it, H planes later; pair 3 (its LDS is full)
// reads it again at the first stage's plane, mostly from cache.
#pragma once

#include "gpe_fused.h"
#include "gpe_mixed.h"

namespace gpe {

struct CkMxArgs {
    const CD* p0;  // the state psi0
    const CF* d;   // pairs 2, 3: D3 or D5 of the stencil source psi0 + D
    const CF* f0;
    const CF* k2;
    const CF* k3;
    const CF* k4;
    CF* f0out;  // pair 1
    CF* ka;     // pair 2: K3
    CF* kb;     // pair 1: K2, pair 2: K4
    CF* dnext;  // pair 1: D3, pair 2: D5
    CD* out;    // pair 3: psi_new
    CD dt;
    CD s21;  // b21 dt
    double dt0_re, dt0_im;
};

// GM: the form of G = i pim dt (gmul, hmul).
template <int ORD, int TERMS, int PAIR, int GM>
__device__ __forceinline__ void ckmx_pair_body(const Geom& g, const StW<double>& w0, const RhsC<double>& k,
                                               const CkMxArgs& a, double* __restrict__ partial, int bid, int pidx) {
    using S = FusedShape<ORD>;
    using R = double;
    constexpr int H = S::H, NB = H + 1, Q = 2 * H + 1;
    constexpr bool P3 = PAIR == 3;
    __shared__ CD ldsA[2][S::LHA * S::LWA];           // stencil source plane p with halo
    __shared__ CD ldsB[NB][S::NTB];                   // first-stage u_next, planes z..z+H
    // pair 3: the psi_new and error partials of the inner tile, planes
    // z..z+H, in LDS rings indexed by the inner thread (inner-first map);
    // in registers they spill
    constexpr int NIN = S::OX * S::OY;
    __shared__ CD ldsS[P3 ? NB : 1][P3 ? NIN : 1];
    __shared__ CD ldsE[P3 ? NB : 1][P3 ? NIN : 1];
    // pair 2: psi0 of the own column, planes p..p+2H-1, kept from the stencil
    // source's load (psi0 + D3) until the first stage needs it at plane p;
    // read again from memory two planes later it mostly missed L2 (1.36x
    // the algorithmic bytes).  H+2 slots, so a plane's slot is written
    // before the barrier of the iteration after the one that read it last.
    constexpr bool kP0Ring = PAIR == 2;
    constexpr int NR = H + 2;
    __shared__ CD ldsR[kP0Ring ? NR : 1][kP0Ring ? S::NTB : 1];
    __shared__ double red_sh[16];
    const FusedCtx<ORD, PAIR != 1> c(g, bid);  // the map measured faster per pair
    const StW<double> w = scale_w<double>(w0, k.a);
    const CD G = g_of<double>(k.pim, a.dt);  // k = dt f = G w
    const SumSrc sum1{a.p0, a.d, mkc<double>(1.0, 0.0)};
    const SumSrc sum21{a.p0, a.f0, a.s21};
    constexpr CkTab<double> t = ck_tab<double>();  // constants, not kernel arguments held in registers
    double red0 = 0.0, red1 = 0.0;
    auto src = [&](Cell ci) -> CD {
        if constexpr (PAIR == 1) return *(a.p0 + ci);
        else return sum1.comb(*(a.p0 + ci), *(a.d + ci));
    };
    if (c.z0 < c.z1) {  // uniform per block
        const int start = c.z0 - 2 * H;
        // Queues, slot (plane - start) % Q: the stencil source for planes
        // z..z+2H and the first stage's u for planes z-H..z+H; for planes
        // z..z+H: qs the inner tile's partial sum of D_next (pairs 1, 2), and
        // qd (pair 3) |psi_scal|^2.
        CD qu[Q], qv[Q], qs[NB];
        double qd[NB];
        int ra = 0, rz = 0, rp = H;  // (z - start) & 1, (z - start) % NB, (p - start) % NB
#pragma unroll
        for (int d = 0; d < Q; ++d) qv[d] = mkc<double>(0.0, 0.0);
#pragma unroll
        for (int d = 0; d < NB; ++d) {
            qs[d] = mkc<double>(0.0, 0.0);
            qd[d] = 1.0;
        }
        // ring slot of plane m: (m - start) % NR; iteration z writes plane
        // z + 2H and reads plane p = z + H
        int rw = (2 * H) % NR, rr = H % NR;
        if (c.active) {
#pragma unroll
            for (int d = 0; d < 2 * H; ++d) {
                const Cell ci{plane_off(g, bcmap1(start + d, g.nz, g.bcs)), c.col};
                if constexpr (kP0Ring) {
                    const CD p0 = *(a.p0 + ci);
                    qu[d] = sum1.comb(p0, *(a.d + ci));
                    ldsR[d % NR][c.lb] = p0;
                } else {
                    qu[d] = src(ci);
                }
            }
        }
        struct Pw {
            CD p0;
            CF f0, k2, k3, k4;
        };
        auto load_pw = [&](Pw& v, long long pb) {
            const Cell ci{pb, c.col};
            if constexpr (PAIR >= 2) {
                if constexpr (!kP0Ring) v.p0 = *(a.p0 + ci);
                v.f0 = *(a.f0 + ci);
                v.k2 = *(a.k2 + ci);
            }
            if constexpr (P3) {
                v.k3 = *(a.k3 + ci);
                v.k4 = *(a.k4 + ci);
            }
        };
        // loads of iteration zz: centre of plane zz + 2H, halo and (pair 2)
        // the pointwise inputs of plane zz + H
        struct Pf {
            CD qn;  // pair 2: psi0 of the centre (the source adds D at the step)
            CF qd;  // pair 2: D3 of the centre
            CD hr[S::NSLOT];
            Pw pw;
        };
        auto fetch = [&](Pf& F, int zz) {
            const long long pb = plane_off(g, bcmap1(zz + H, g.nz, g.bcs));
            if (c.active) {
                const Cell cc{plane_off(g, bcmap1(zz + 2 * H, g.nz, g.bcs)), c.col};
                if constexpr (kP0Ring) {
                    F.qn = *(a.p0 + cc);
                    F.qd = *(a.d + cc);
                } else {
                    F.qn = src(cc);
                }
                if constexpr (PAIR == 2) load_pw(F.pw, pb);
            }
#pragma unroll
            for (int s = 0; s < S::NSLOT; ++s)
                if (c.hv[s]) F.hr[s] = src(Cell{pb, c.hoff[s]});
        };
        Pf F{};
        fetch(F, start);
        auto step = [&](int z, auto phase) {
            constexpr int PH = decltype(phase)::value;
            const int p = z + H;
            if constexpr (kP0Ring) {
                qu[(PH + 2 * H) % Q] = sum1.comb(F.qn, F.qd);
                if (c.active) ldsR[rw][c.lb] = F.qn;
            } else {
                qu[(PH + 2 * H) % Q] = F.qn;
            }
            Pw pw = F.pw;
            CD* A = ldsA[ra];
            if (c.active) A[c.la] = qu[(PH + H) % Q];
#pragma unroll
            for (int s = 0; s < S::NSLOT; ++s)
                if (c.hv[s]) A[c.hl[s]] = F.hr[s];
            __syncthreads();
            if (z + 1 < c.z1) fetch(F, z + 1);
            if (c.active) {
                // first stage at plane p, whole block
                FusedSten<R, ORD, S::LWA, TERMS> sa;
                sa.L = A;
                sa.lc = c.la;
#pragma unroll
                for (int d = 0; d <= 2 * H; ++d) sa.q[d] = qu[(PH + d) % Q];
                sa.w = w;
                const CD wa = rhs_w<R>(k, sa, c.mx, c.my, bcmap1(p, g.nz, g.bcs));
                const bool own_p = c.store && p >= c.z0 && p < c.z1;
                const Cell op{plane_off(g, p), c.own};
                CD ub;
                if constexpr (PAIR == 1) {
                    // OpMxCk1, then the source of OpMxCk2 and its D3 partial
                    const CF f0 = tof(rot_pim<double>(k.pim, wa));
                    if (own_p) st_stream<float>(a.f0out + op, f0);
                    ub = sum21.comb(qu[(PH + H) % Q], f0);
                    qs[H] = cscale<double>(t.b31, hmul<double, GM>(a.dt, tod(f0)));
                } else if constexpr (PAIR == 2) {
                    // OpMxCkMid<3>, then the source of OpMxCkMid<4> and its D5 partial
                    const CD k3 = gmul<double, GM>(G, wa);
                    const CF k3f = tof(k3);
                    if (own_p) st_stream<float>(a.ka + op, k3f);
                    const CD k1 = hmul<double, GM>(a.dt, tod(pw.f0));
                    const CD k2v = tod(pw.k2);
                    const CF d4 = tof(caxpy<double>(t.b43, k3, caxpy<double>(t.b42, k2v, cscale<double>(t.b41, k1))));
                    ub = sum1.comb(ldsR[rr][c.lb], d4);
                    qs[H] = caxpy<double>(t.b53, tod(k3f), caxpy<double>(t.b52, k2v, cscale<double>(t.b51, k1)));
                } else {
                    // OpMxCkMid<5>, then the source of OpMxCkLast and its partials;
                    // the pointwise inputs are loaded after the stencil (a set
                    // live across it spills)
                    load_pw(pw, plane_off(g, bcmap1(p, g.nz, g.bcs)));
                    const CD k5 = gmul<double, GM>(G, wa);
                    const CD k1 = hmul<double, GM>(a.dt, tod(pw.f0));
                    const CD k3v = tod(pw.k3), k4v = tod(pw.k4), k5v = tod(tof(k5));
                    const CF d6 = tof(caxpy<double>(
                        t.b65, k5,
                        caxpy<double>(t.b64, k4v,
                                      caxpy<double>(t.b63, k3v, caxpy<double>(t.b62, tod(pw.k2), cscale<double>(t.b61, k1))))));
                    ub = sum1.comb(pw.p0, d6);
                    if (c.winner) {
                        const CD p0 = csub<double>(ub, tod(d6));  // as OpMxCkLast: u6 - D6
                        ldsS[rp][threadIdx.x] =
                            caxpy<double>(t.c4, k4v, caxpy<double>(t.c3, k3v, caxpy<double>(t.c1, k1, p0)));
                        ldsE[rp][threadIdx.x] =
                            mkc<double>(t.dc1 * k1.x + t.dc3 * k3v.x + t.dc4 * k4v.x + t.dc5 * k5v.x,
                                        t.dc1 * k1.y + t.dc3 * k3v.y + t.dc4 * k4v.y + t.dc5 * k5v.y);
                        const CD f0v = tod(pw.f0);
                        const double ap = sqrt(p0.x * p0.x + p0.y * p0.y);
                        const double af = sqrt(f0v.x * f0v.x + f0v.y * f0v.y);
                        const double sr = ap + a.dt0_re * af + 1e-30, si = a.dt0_im * af;
                        qd[H] = sr * sr + si * si;
                    }
                }
                qv[(PH + 2 * H) % Q] = ub;
                ldsB[rp][c.lb] = ub;
            }
            if (z >= c.z0 && c.winner) {
                // second stage at plane z, inner tile (ldsB of plane z was
                // written H iterations ago)
                FusedSten<R, ORD, S::TXB, TERMS> sb;
                sb.L = ldsB[rz];
                sb.lc = c.lb;
#pragma unroll
                for (int d = 0; d <= 2 * H; ++d) sb.q[d] = qv[(PH + d) % Q];
                sb.w = w;
                const CD kb = gmul<double, GM>(G, rhs_w<R>(k, sb, c.mx, c.my, z));
                const Cell oz{plane_off(g, z), c.own};
                if constexpr (PAIR <= 2) {
                    if (c.store) {
                        const double bn = PAIR == 1 ? t.b32 : t.b54;
                        st_stream<float>(a.kb + oz, tof(kb));
                        st_stream<float>(a.dnext + oz, tof(caxpy<double>(bn, kb, qs[0])));
                    }
                } else {
                    if (c.store) {
                        const CD nw = caxpy<double>(t.c6, kb, ldsS[rz][threadIdx.x]);
                        st_stream<double>(a.out + oz, nw);
                        const CD se = ldsE[rz][threadIdx.x];
                        const double er = se.x + t.dc6 * kb.x;
                        const double ei = se.y + t.dc6 * kb.y;
                        const double e = sqrt((er * er + ei * ei) / qd[0]);
                        red0 = e > red0 ? e : red0;
                        const double a2 = nw.x * nw.x + nw.y * nw.y;
                        red1 = a2 > red1 ? a2 : red1;
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < H; ++d) {
                qs[d] = qs[d + 1];
                qd[d] = qd[d + 1];
            }
            ra ^= 1;
            rz = rz == NB - 1 ? 0 : rz + 1;
            rp = rp == NB - 1 ? 0 : rp + 1;
            if constexpr (kP0Ring) {
                rw = rw == NR - 1 ? 0 : rw + 1;
                rr = rr == NR - 1 ? 0 : rr + 1;
            }
        };
        march_rotating<Q>(start, c.z1, step);
    }
    if constexpr (P3) {
        const double v0 = block_reduce<1024, RED_MAX>(red0, red_sh);
        const double v1 = block_reduce<1024, RED_MAX>(red1, red_sh);
        if (threadIdx.x == 0) {
            partial[pidx] = v0;
            partial[(long long)g.nblocks + pidx] = v1;
        }
    }
}

template <int ORD, int TERMS, int PAIR, int GM>
__global__ __launch_bounds__(1024) void k_ckmx_pair(Geom g, StW<double> w, RhsC<double> k, CkMxArgs a,
                                                    double* __restrict__ partial) {
    ckmx_pair_body<ORD, TERMS, PAIR, GM>(g, w, k, a, partial, xcd_remap(blockIdx.x, g.nblocks), blockIdx.x);
}

template <int ORD>
hipError_t launch_ckmx_fused(const Geom& g, const StW<double>& w, const RhsC<double>& k, const CkMxArgs& a, int pair,
                             double* partial, hipStream_t s);  // g.general: the trap term

}  // namespace gpe
