// Fused RK4: the four stages of a classic RK4 step in two kernels that each
// evaluate two stages (temporal blocking in the z-march), for complex128
// (FP64) or complex64 (FP32) fields and a right-hand side whose only optional
// term is the trap V.
//
//   F12: u1 = psi0 + dt/2 f(psi0) on the whole block, then
//        u2 = psi0 + dt/2 f(u1) on the block's inner tile      reads psi0        writes u2
//   F34: u3 = psi0 + dt f(u2) on the whole block, and on the inner tile
//        acc = psi0 + dt f(psi0)/6 + 2(u2 - psi0)/3 + dt f(u2)/3
//        psi_new = acc + dt/6 f(u3)                           reads u2, psi0    writes psi_new
//
// (dt k2 = 2(u2 - psi0) recovers k2 from u2, k1 = f(psi0) is re-evaluated.)
// The step moves 5 fields per cell instead of 13 and evaluates the
// right-hand side about 5.9 times per cell instead of 4.
//
// Geometry: a block of TXB x TYB = (64 + 2H) x 15 threads owns one column
// of cells each; its inner (64 x (15 - 2H)) tile is the output, the H-wide
// rim is recomputed by the neighbouring blocks (overlapped tiling).  A block
// marches its z-range [z0, z1): at iteration z it evaluates the first stage
// at plane p = z + H over the whole block and the second stage at plane z
// over the inner tile.  2H warm-up iterations before z0 fill the queues.
//
// F34's register queues are indexed by (plane - start) % (2H + 1) with the
// z-loop unrolled 2H + 1 times (march_rotating), so they rotate by renaming
// instead of being copied down one slot per plane.
//
// Per thread (own column): register queues of the inputs for planes z..z+2H
// and of the first stage's result for planes z-H..z+H; F34's acc for the
// inner planes z..z+H is a register queue in FP32 and an LDS ring indexed by
// the inner thread in FP64.  LDS: the first stage's input plane p with an H
// halo (double buffered), F34's psi0 plane p (double buffered), and the first
// stage's results in a ring of planes z..z+H, written when computed and read
// H iterations later, so one barrier per iteration orders every LDS reuse.
// F12 keeps the loads of the next GPE_FUSED_DEPTH (3) planes in flight while a
// plane is evaluated; F34 one plane.
//
// Cells outside the domain are evaluated at their unfolded positions with
// bcmap-mapped inputs.  With periodic boundaries that is the wrapped value;
// with even reflection the axis stencils of the Laplacian are symmetric, so
// the unfolded value equals the value at the mirrored index bit for bit.
// The trap V (TERMS = RHS_V), a pointwise position-dependent term, is
// evaluated at the mapped indices, so it keeps that equality.  The other
// optional terms are left to the unfused kernels: Urhs d/dx is antisymmetric
// under reflection, and the dissipation and eqn-3 mask would push these
// kernels past their registers.
#pragma once

#include "gpe_kernels.h"

namespace gpe {

constexpr int kFusedTYB = 15;

// Block shape: an inner tile OX_ wide, blocks of (OX_ + 2H) x TYB_ threads
// in NT-thread workgroups (1024: one block per CU; 512: two).
template <int ORD, int OX_ = 64, int TYB_ = kFusedTYB> struct FusedShape {
    static constexpr int H = ORD / 2;
    static constexpr int TXB = OX_ + 2 * H, TYB = TYB_, NTB = TXB * TYB;
    static constexpr int NT = NTB <= 512 ? 512 : 1024;
    static constexpr int OX = OX_, OY = TYB - 2 * H;       // output tile
    static constexpr int LWA = TXB + 2 * H, LHA = TYB + 2 * H;  // LDS plane with halo
    static constexpr int NHALO = 2 * H * TXB + 2 * H * TYB;
    static constexpr int NSLOT = (NHALO + NT - 1) / NT;
    static_assert(NTB <= NT, "block shape exceeds the workgroup");
};

// The shape of the fused RK4 pair: inner tiles GPE_RK4_OX wide in blocks of
// GPE_RK4_TYB rows (a block of at most 512 threads runs two to a CU).
#ifndef GPE_RK4_OX
#define GPE_RK4_OX 64
#endif
#ifndef GPE_RK4_TYB
#define GPE_RK4_TYB kFusedTYB
#endif
constexpr int kRk4OX = GPE_RK4_OX;
constexpr int kRk4TYB = GPE_RK4_TYB;
template <int ORD> constexpr int kRk4NT = FusedShape<ORD, kRk4OX, kRk4TYB>::NT;
// waves per SIMD the RK4 pair asks the register allocator for: the FP32
// stages 1+2 fit 64 VGPRs, so two 1024-thread blocks share a CU
template <typename R, int ORD, int PAIR> constexpr int kRk4Waves = sizeof(R) == 4 && PAIR == 12 ? 8 : 1;

// Per-thread geometry of a fused block.  The thread-to-column map is a
// per-kernel choice (measured): inner-first saves the rim lanes' share of
// the second stage, but it moves register allocation, and some kernels then
// spill or run slower.
//
//
// REP (a per-kernel choice, measured): the workgroup's threads beyond the
// block's NTB columns (NT - NTB of them) each repeat a rim column, so every
// lane is active and the per-plane branches on `active` fold
// away (a repeated column computes and writes the same values to the same
// LDS cells, and a rim column stores nothing); and with the inner-first map
// and an inner tile of whole wavefronts, `winner` is the inner flag as a
// wave-uniform (scalar) value, so the second-stage branch is a scalar
// branch.  The larger basic blocks move register allocation: some kernels
// spill with it.  Without REP, winner == inner.
template <int ORD, bool INNER_FIRST = true, int OX_ = 64, int TYB_ = kFusedTYB, bool REP = false> struct FusedCtx {
    using S = FusedShape<ORD, OX_, TYB_>;
    static constexpr bool kWaveInner = REP && INNER_FIRST && (S::OX * S::OY) % 64 == 0;
    int z0, z1;
    bool active, inner, winner, store;
    int lx, ly, gx, gy;
    int mx, my;  // the column's in-domain (mapped) indices
    unsigned col, own;
    int la, lb;
    unsigned hoff[S::NSLOT];
    int hl[S::NSLOT];
    bool hv[S::NSLOT];

    __device__ __forceinline__ FusedCtx(const Geom& g, int bid) {
        // with REP, threads beyond the block repeat rim columns: the last
        // rim threads of the map (inner-first: the left and right rim
        // columns; otherwise the last row)
        constexpr int NIN = S::OX * S::OY;
        constexpr int NREP = INNER_FIRST ? S::NTB - NIN : S::TXB;
        const int t0 = (int)threadIdx.x;
        const int tid = !REP || t0 < S::NTB ? t0 : S::NTB - 1 - (t0 - S::NTB) % NREP;
        const int ntxy = g.ntx * g.nty;
        const int txy = bid % ntxy, zc = bid / ntxy;
        // Tiles run in column groups 2 tiles wide (all rows of a group, then
        // the next group), so the blocks one XCD runs at a time form a patch
        // taller than wide: the y rims, 4 of every 15 rows, are re-read from
        // that XCD's L2 more often, at the price of the x rims (4 of 68).
        // Measured at 512^3 against whole rows / 4-wide / 1-wide groups: the
        // Cash-Karp attempt 7.8 against 8.5 / 8.4 / 8.8 ms, RK4 2.64 against
        // 2.70 / 2.69 / 2.71 ms.
        constexpr int kGroupW = 128 / S::OX;
        const int gw = g.ntx % kGroupW == 0 ? kGroupW : g.ntx;
        const int grp = txy / (gw * g.nty), rem = txy % (gw * g.nty);
        const int ox0 = (grp * gw + rem % gw) * S::OX, oy0 = (rem / gw) * S::OY;
        z0 = zc * g.tz;
        z1 = min(z0 + g.tz, g.nz);
        // INNER_FIRST: threads 0 .. OX*OY-1 own the inner tile, OX = 64
        // columns per row, so whole wavefronts are inner and the second stage
        // runs wave-uniform; the others own the H-wide rim: the H rows above
        // and below, then the H columns left and right of each inner row.
        // Otherwise thread t owns column (t % TXB, t / TXB) of the block.
        active = REP || tid < S::NTB;
        if (!INNER_FIRST) {
            lx = tid % S::TXB;
            ly = active ? tid / S::TXB : 0;
        } else if (tid < NIN) {
            lx = S::H + tid % S::OX;
            ly = S::H + tid / S::OX;
        } else {
            int r = tid - NIN;
            if (r < 2 * S::H * S::TXB) {
                const int row = r / S::TXB;
                lx = r % S::TXB;
                ly = row < S::H ? row : S::OY + row;
            } else if (r < 2 * S::H * S::TXB + 2 * S::H * S::OY) {
                r -= 2 * S::H * S::TXB;
                const int cc = r % (2 * S::H);
                ly = S::H + r / (2 * S::H);
                lx = cc < S::H ? cc : S::OX + cc;
            } else {
                lx = 0;
                ly = 0;
            }
        }
        gx = ox0 - S::H + lx;
        gy = oy0 - S::H + ly;
        inner = INNER_FIRST ? tid < NIN : active && lx >= S::H && lx < S::TXB - S::H && ly >= S::H && ly < S::TYB - S::H;
        if constexpr (kWaveInner) winner = __builtin_amdgcn_readfirstlane((int)inner) != 0;
        else winner = inner;
        store = inner && gx < g.nx && gy < g.ny;
        mx = bcmap(gx, g.nx, g.bcs);
        my = bcmap(gy, g.ny, g.bcs);
        col = (unsigned)mx + (unsigned)g.nx * (unsigned)my;
        own = store ? (unsigned)gx + (unsigned)g.nx * (unsigned)gy : 0u;
        la = (ly + S::H) * S::LWA + lx + S::H;
        lb = ly * S::TXB + lx;
        // halo of the block plane: 2H rows above/below, then 2H columns left/right
#pragma unroll
        for (int s = 0; s < S::NSLOT; ++s) {
            const int e = (int)threadIdx.x + s * S::NT;  // every workgroup thread takes halo cells
            hv[s] = e < S::NHALO;
            int hx = ox0 - S::H, hy = oy0 - S::H, lxx = S::H, lyy = S::H;
            if (e < 2 * S::H * S::TXB) {
                const int r = e / S::TXB, c = e % S::TXB;
                lxx = c + S::H;
                lyy = r < S::H ? r : S::TYB + r;
                hx = ox0 - S::H + c;
                hy = oy0 - 2 * S::H + lyy;
            } else if (e < S::NHALO) {
                const int e2 = e - 2 * S::H * S::TXB;
                const int r = e2 / (2 * S::H), c = e2 % (2 * S::H);
                lyy = r + S::H;
                lxx = c < S::H ? c : S::TXB + c;
                hx = ox0 - 2 * S::H + lxx;
                hy = oy0 - S::H + r;
            }
            hl[s] = lyy * S::LWA + lxx;
            hoff[s] = (unsigned)bcmap(hx, g.nx, g.bcs) + (unsigned)g.nx * (unsigned)bcmap(hy, g.ny, g.bcs);
        }
    }
};

// Planes of loads in flight (1, 2 or 3) in the fused kernels.
#ifndef GPE_FUSED_DEPTH
#define GPE_FUSED_DEPTH 3
#endif

// Runs step(z, F) for z = start .. z1-1 with DEPTH sets of in-flight loads:
// fetch(F, z) issues the loads iteration z consumes, step(z, F) consumes
// them and refills F for iteration z + DEPTH.
// pad() issues a store with no lanes in range after each fetch of the
// preamble, so the preamble leaves the same memory operations in flight as
// the loop's iterations (a fetch then a store) and the loop's first wait
// covers one plane's loads, as every later one does.
template <int DEPTH, class Pf, class Fetch, class Step, class Pad>
__device__ __forceinline__ void fused_march(int start, int z1, Fetch fetch, Step step, Pad pad) {
    if constexpr (DEPTH == 1) {
        Pf F0{};
        fetch(F0, start);
        pad();
        for (int z = start; z < z1; ++z) step(z, F0);
    } else if constexpr (DEPTH == 3) {
        // whole triples in the loop, the remainder after it: no step inside
        // the loop is conditional, so the loop's memory operations are the
        // same on every path
        Pf F0{}, F1{}, F2{};
        fetch(F0, start);
        pad();
        fetch(F1, start + 1);
        pad();
        fetch(F2, start + 2);
        pad();
        int z = start;
        for (; z + 2 < z1; z += 3) {
            step(z, F0);
            step(z + 1, F1);
            step(z + 2, F2);
        }
        if (z < z1) step(z, F0);
        if (z + 1 < z1) step(z + 1, F1);
    } else {
        Pf F0{}, F1{};
        fetch(F0, start);
        pad();
        fetch(F1, start + 1);
        pad();
        int z = start;
        for (; z + 1 < z1; z += 2) {
            step(z, F0);
            step(z + 1, F1);
        }
        if (z < z1) step(z, F0);
    }
}

// Calls step(z, std::integral_constant<int, PH>) for z = start .. z1-1 with
// PH = (z - start) % U known at compile time: register queues indexed by
// (PH + d) % U rotate by renaming instead of copying every entry each plane.
template <int U, int PH = 0, class Step>
__device__ __forceinline__ void march_phases(int z, int z1, Step& step) {
    if constexpr (PH < U) {
        if (z < z1) {
            step(z, std::integral_constant<int, PH>{});
            march_phases<U, PH + 1>(z + 1, z1, step);
        }
    }
}

template <int U, int PH = 0, class Step>
__device__ __forceinline__ void march_group(int z, Step& step) {
    if constexpr (PH < U) {
        step(z, std::integral_constant<int, PH>{});
        march_group<U, PH + 1>(z + 1, step);
    }
}

// Whole groups of U planes in the loop (no step inside it is conditional,
// so its memory operations are the same on every path), the remainder after.
template <int U, class Step>
__device__ __forceinline__ void march_rotating(int start, int z1, Step step) {
    int z = start;
    for (; z + U <= z1; z += U) march_group<U>(z, step);
    march_phases<U>(z, z1, step);
}

// Stages 1+2 of the fused block `bid` of a grid (the body of k_rk4_f12 and
// of its batched form).
// TERMS: 0 (no optional right-hand-side term) or RHS_V (the trap only).
template <typename R, int ORD, int LW, int TERMS>
using FusedSten = std::conditional_t<TERMS == 0, Sten<R, ORD, LW, false>, StenV<R, ORD, LW>>;

// What the stages-1+2 body computes from psi0 (h1, h2 the steps of its two
// stages): the first two stages of RK4, u2 = psi0 + dt/2 f(psi0 + dt/2 f(psi0));
// one midpoint RK2 step (src/solve.f90:78-105), psi0 + dt f(psi0 + dt/2 f(psi0));
// or two Euler steps (src/solve.f90:57-74), u1 + dt f(u1) with
// u1 = psi0 + dt f(psi0).
enum { PK_RK4 = 0, PK_RK2 = 1, PK_EULER2 = 2 };

// GM: the form of G = i pim h of the stage updates (gaxpy).
template <typename R, int ORD, int TERMS, int GM, int OX_ = 64, int TYB_ = kFusedTYB, int PK = PK_RK4>
__device__ __forceinline__ void rk4_f12_body(const Geom& g, const StW<R>& w0, const RhsC<R>& k,
                                             const C<R>* __restrict__ psi, C<R>* __restrict__ u2out, C<R> h1,
                                             C<R> h2, int bid) {
    using S = FusedShape<ORD, OX_, TYB_>;
    constexpr int H = S::H, D = GPE_FUSED_DEPTH;
    constexpr unsigned E = sizeof(C<R>);
    __shared__ C<R> ldsA[2][S::LHA * S::LWA];
    __shared__ C<R> ldsB[H + 1][S::NTB];
    const FusedCtx<ORD, true, OX_, TYB_, sizeof(R) == 4> c(g, bid);  // REP measured: FP32 faster, FP64 slower
    if (c.z0 >= c.z1) return;  // uniform per block
    const StW<R> w = scale_w<R>(w0, k.a);
    const C<R> G = g_of<R>(k.pim, h1);  // u = q + G w: the first stage
    const C<R> G2 = PK == PK_RK4 ? G : g_of<R>(k.pim, h2);  // the second
    const int start = c.z0 - 2 * H;
    // byte offsets within a plane; kNoOff where the lane has no access
    const unsigned pbytes = (unsigned)g.plane * E;
    const unsigned ocol = c.active ? c.col * E : kNoOff;
    const unsigned oown = c.store ? c.own * E : kNoOff;
    unsigned ohal[S::NSLOT];
#pragma unroll
    for (int s = 0; s < S::NSLOT; ++s) ohal[s] = c.hv[s] ? c.hoff[s] * E : kNoOff;
    auto pl = [&](const C<R>* f, int zz) { return plane_res_m<R>(f, (unsigned)bcmap1(zz, g.nz, g.bcs), pbytes); };

    C<R> qp[2 * H + 1];  // psi0, planes z..z+2H
    C<R> qv[2 * H + 1];  // u1, planes z-H..z+H
    // LDS buffer and ring slots of the iteration: (z - start) & 1,
    // (z - start) % (H + 1) and (p - start) % (H + 1), advanced by one each
    // iteration instead of divided out
    int ra = 0, rz = 0, rp = H;
#pragma unroll
    for (int d = 0; d <= 2 * H; ++d) qv[d] = mkc<R>(0.0, 0.0);
#pragma unroll
    for (int d = 0; d < 2 * H; ++d) qp[d] = bld<R>(pl(psi, start + d), ocol);
    // loads of iteration zz: centre of plane zz + 2H, halo of plane zz + H.
    // Every iteration issues them (the last D iterations reload the last
    // planes), so the memory instructions in flight are the same on every
    // path and a plane's loads are waited for alone.
    struct Pf {
        C<R> qn;
        C<R> hr[S::NSLOT];
    };
    auto fetch = [&](Pf& F, int zz) {
        zz = zz < c.z1 ? zz : c.z1 - 1;
        F.qn = bld<R>(pl(psi, zz + 2 * H), ocol);
        const BRes rh = pl(psi, zz + H);
#pragma unroll
        for (int s = 0; s < S::NSLOT; ++s) F.hr[s] = bld<R>(rh, ohal[s]);
    };
    auto step = [&](int z, Pf& F) {
        const int p = z + H;
        qp[2 * H] = F.qn;
        C<R>* A = ldsA[ra];
        if (c.active) A[c.la] = qp[H];
#pragma unroll
        for (int s = 0; s < S::NSLOT; ++s)
            if (c.hv[s]) A[c.hl[s]] = F.hr[s];
        __syncthreads();
        fetch(F, z + D);
        if (c.active) {
            // first stage at plane p, whole block
            FusedSten<R, ORD, S::LWA, TERMS> sa;
            sa.L = A;
            sa.lc = c.la;
#pragma unroll
            for (int d = 0; d <= 2 * H; ++d) sa.q[d] = qp[d];
            sa.w = w;
            const C<R> u1 = gaxpy<R, GM>(G, rhs_w<R>(k, sa, c.mx, c.my, bcmap1(p, g.nz, g.bcs)), qp[H]);
            qv[2 * H] = u1;
            ldsB[rp][c.lb] = u1;
        }
        C<R> u2 = mkc<R>(0.0, 0.0);
        const bool second = z >= c.z0;
        if (second && c.winner) {
            // second stage at plane z, inner tile (ldsB of plane z was written H iterations ago)
            FusedSten<R, ORD, S::TXB, TERMS> sb;
            sb.L = ldsB[rz];
            sb.lc = c.lb;
#pragma unroll
            for (int d = 0; d <= 2 * H; ++d) sb.q[d] = qv[d];
            sb.w = w;
            // the second stage's base: psi0, or (two Euler steps) u1 of plane z
            u2 = gaxpy<R, GM>(G2, rhs_w<R>(k, sb, c.mx, c.my, z), PK == PK_EULER2 ? qv[H] : qp[0]);
        }
        // the store is issued on every path (kNoOff where there is no output)
        bst<R>(plane_res_m<R>(u2out, (unsigned)(second ? z : c.z0), pbytes), second ? oown : kNoOff, u2);
#pragma unroll
        for (int d = 0; d < 2 * H; ++d) {
            qp[d] = qp[d + 1];
            qv[d] = qv[d + 1];
        }
        ra ^= 1;
        rz = rz == H ? 0 : rz + 1;
        rp = rp == H ? 0 : rp + 1;
    };
    fused_march<D, Pf>(start, c.z1, fetch, step,
                       [&] { bst<R>(plane_res_m<R>(u2out, 0u, pbytes), kNoOff, mkc<R>(0.0, 0.0)); });
}

template <typename R, int ORD, int TERMS, int GM>
__global__ __launch_bounds__(kRk4NT<ORD>, (kRk4Waves<R, ORD, 12>)) void k_rk4_f12(Geom g, StW<R> w, RhsC<R> k,
                                                                             const C<R>* __restrict__ psi,
                                                                             C<R>* __restrict__ u2out, C<R> hdt) {
    rk4_f12_body<R, ORD, TERMS, GM, kRk4OX, kRk4TYB>(g, w, k, psi, u2out, hdt, hdt,
                                                     xcd_remap(blockIdx.x, g.nblocks));
}

// One midpoint RK2 step (hdt = dt/2) or two Euler steps (hdt = dt) of the
// fused block: psi -> out, 2 field accesses per cell instead of 5 (RK2) or 4
// (two Euler steps) in the stage kernels.
template <typename R, int ORD, int TERMS, int GM>
__global__ __launch_bounds__(kRk4NT<ORD>, (kRk4Waves<R, ORD, 12>)) void k_rk2_fused(Geom g, StW<R> w, RhsC<R> k,
                                                                               const C<R>* __restrict__ psi,
                                                                               C<R>* __restrict__ out, C<R> hdt,
                                                                               C<R> dt) {
    rk4_f12_body<R, ORD, TERMS, GM, kRk4OX, kRk4TYB, PK_RK2>(g, w, k, psi, out, hdt, dt,
                                                             xcd_remap(blockIdx.x, g.nblocks));
}

template <typename R, int ORD, int TERMS, int GM>
__global__ __launch_bounds__(kRk4NT<ORD>, (kRk4Waves<R, ORD, 12>)) void k_euler2_fused(Geom g, StW<R> w, RhsC<R> k,
                                                                                  const C<R>* __restrict__ psi,
                                                                                  C<R>* __restrict__ out, C<R> dt) {
    rk4_f12_body<R, ORD, TERMS, GM, kRk4OX, kRk4TYB, PK_EULER2>(g, w, k, psi, out, dt, dt,
                                                                xcd_remap(blockIdx.x, g.nblocks));
}

// Stages 3+4 of the fused block `bid` of a grid.
// Planes of loads in flight (1 or 2; FP64 keeps one: a second set spills)
// and repeated rim columns (FusedCtx; with the trap term FP64 spills with
// them) of k_rk4_f34.
#ifndef GPE_F34_DEPTH
#define GPE_F34_DEPTH 1
#endif
#ifndef GPE_F34_REP
#define GPE_F34_REP 1
#endif
template <typename R> constexpr int kF34Depth = sizeof(R) == 4 ? GPE_F34_DEPTH : 1;
template <typename R, int TERMS> constexpr bool kF34Rep = GPE_F34_REP != 0 && (sizeof(R) == 4 || TERMS == 0);
template <typename R, int ORD, int TERMS, int GM, int OX_ = 64, int TYB_ = kFusedTYB>
__device__ __forceinline__ void rk4_f34_body(const Geom& g, const StW<R>& w0, const RhsC<R>& k,
                                             const C<R>* __restrict__ u2, const C<R>* __restrict__ psi,
                                             C<R>* __restrict__ out, C<R> dt, C<R> dt3, C<R> dt6, int bid) {
    using S = FusedShape<ORD, OX_, TYB_>;
    // one plane of loads in flight: a second set would not fit beside the
    // three register queues without spilling
    constexpr int H = S::H, NB = H + 1, Q = 2 * H + 1;
    constexpr unsigned E = sizeof(C<R>);
    __shared__ C<R> ldsA[2][S::LHA * S::LWA];  // u2 plane p with halo
    __shared__ C<R> ldsP[2][S::NTB];           // psi0 plane p
    __shared__ C<R> ldsB[NB][S::NTB];          // u3 planes z..z+H
    // FP64: acc of the inner tile for planes z..z+H lives in an LDS ring
    // indexed by the inner thread (with the inner-first map, threads
    // 0 .. OX*OY-1), which leaves room in registers for that map (159 KB of
    // LDS at order 4; 1.53 ms against 1.61 at 512^3).  FP32 keeps it in
    // registers.
    constexpr bool kAccLds = sizeof(R) == 8;
    __shared__ C<R> ldsQ[kAccLds ? NB : 1][kAccLds ? S::OX * S::OY : 1];
    const FusedCtx<ORD, true, OX_, TYB_, kF34Rep<R, TERMS>> c(g, bid);
    constexpr int DEPTH = kF34Depth<R>;
    if (c.z0 >= c.z1) return;
    const StW<R> w = scale_w<R>(w0, k.a);
    // u = q + G w for the steps dt, dt/3 and dt/6
    const C<R> G1 = g_of<R>(k.pim, dt), G3 = g_of<R>(k.pim, dt3), G6 = g_of<R>(k.pim, dt6);
    const int start = c.z0 - 2 * H;
    const unsigned pbytes = (unsigned)g.plane * E;
    const unsigned ocol = c.active ? c.col * E : kNoOff;
    const unsigned oown = c.store ? c.own * E : kNoOff;
    unsigned ohal[S::NSLOT];
#pragma unroll
    for (int s = 0; s < S::NSLOT; ++s) ohal[s] = c.hv[s] ? c.hoff[s] * E : kNoOff;
    auto pl = [&](const C<R>* f, int zz) { return plane_res_m<R>(f, (unsigned)bcmap1(zz, g.nz, g.bcs), pbytes); };

    // Register queues, slot (plane - start) % Q: u2 and psi0 for planes
    // z..z+2H, u3 for planes z-H..z+H, and (FP32) acc for the inner planes
    // z..z+H.
    C<R> qu[Q], qp[Q], qv[Q], qa[H + 1];
    int ra = 0, rz = 0, rp = H;  // (z - start) & 1, (z - start) % NB, (p - start) % NB
#pragma unroll
    for (int d = 0; d < Q; ++d) qv[d] = mkc<R>(0.0, 0.0);
#pragma unroll
    for (int d = 0; d <= H; ++d) qa[d] = mkc<R>(0.0, 0.0);
#pragma unroll
    for (int d = 0; d < 2 * H; ++d) {
        qu[d] = bld<R>(pl(u2, start + d), ocol);
        qp[d] = bld<R>(pl(psi, start + d), ocol);
    }
    constexpr R two3 = R(2.0 / 3.0);
    struct Pf {
        C<R> qnu, qnp;
        C<R> hr[S::NSLOT];
    };
    // loads of iteration zz (every iteration issues them; the last one
    // reloads its own planes)
    auto fetch = [&](Pf& F, int zz) {
        zz = zz < c.z1 ? zz : c.z1 - 1;
        F.qnu = bld<R>(pl(u2, zz + 2 * H), ocol);
        F.qnp = bld<R>(pl(psi, zz + 2 * H), ocol);
        const BRes rh = pl(u2, zz + H);
#pragma unroll
        for (int s = 0; s < S::NSLOT; ++s) F.hr[s] = bld<R>(rh, ohal[s]);
    };
    // DEPTH sets of loads in flight: iteration z consumes set z % depth
    // and refills it for iteration z + depth
    Pf FA{}, FB{};
    fetch(FA, start);
    bst<R>(plane_res_m<R>(out, 0u, pbytes), kNoOff, mkc<R>(0.0, 0.0));  // as the iterations' store (see fused_march)
    if constexpr (DEPTH == 2) {
        fetch(FB, start + 1);
        bst<R>(plane_res_m<R>(out, 0u, pbytes), kNoOff, mkc<R>(0.0, 0.0));
    }
    auto step = [&](int z, auto phase) {
        constexpr int PH = decltype(phase)::value;
        Pf& F = (DEPTH == 2 && PH % 2 == 1) ? FB : FA;
        const int p = z + H;
        qu[(PH + 2 * H) % Q] = F.qnu;
        qp[(PH + 2 * H) % Q] = F.qnp;
        C<R>* A = ldsA[ra];
        C<R>* P = ldsP[ra];
        if (c.active) {
            A[c.la] = qu[(PH + H) % Q];
            P[c.lb] = qp[(PH + H) % Q];
        }
#pragma unroll
        for (int s = 0; s < S::NSLOT; ++s)
            if (c.hv[s]) A[c.hl[s]] = F.hr[s];
        __syncthreads();
        // the next plane's loads issued at a raised wave priority, so they
        // leave before the other waves' stencil arithmetic (k_rk4_f34 1.5-3%
        // faster at 512^3; with the trap term, 2.5% slower on the 256^3
        // trapped vortex, so not there; in k_rk4_f12 9% slower)
        if constexpr (TERMS == 0) __builtin_amdgcn_s_setprio(2);
        fetch(F, z + DEPTH);
        if constexpr (TERMS == 0) __builtin_amdgcn_s_setprio(0);
        if (c.active) {
            // third stage at plane p, whole block: u3 = psi0 + dt f(u2)
            FusedSten<R, ORD, S::LWA, TERMS> sa;
            sa.L = A;
            sa.lc = c.la;
#pragma unroll
            for (int d = 0; d <= 2 * H; ++d) sa.q[d] = qu[(PH + d) % Q];
            sa.w = w;
            const int pm = bcmap1(p, g.nz, g.bcs);
            const C<R> w2 = rhs_w<R>(k, sa, c.mx, c.my, pm);
            const C<R> p0 = qp[(PH + H) % Q], uu = qu[(PH + H) % Q];
            const C<R> u3 = gaxpy<R, GM>(G1, w2, p0);
            qv[(PH + 2 * H) % Q] = u3;
            ldsB[rp][c.lb] = u3;
            if (c.winner) {
                // acc at plane p from k1 = f(psi0) and the recovered dt k2 = 2 (u2 - psi0)
                FusedSten<R, ORD, S::TXB, TERMS> sp;
                sp.L = P;
                sp.lc = c.lb;
#pragma unroll
                for (int d = 0; d <= 2 * H; ++d) sp.q[d] = qp[(PH + d) % Q];
                sp.w = w;
                const C<R> w1 = rhs_w<R>(k, sp, c.mx, c.my, pm);
                const C<R> base = mkc<R>(p0.x + two3 * (uu.x - p0.x), p0.y + two3 * (uu.y - p0.y));
                const C<R> acc = gaxpy<R, GM>(G3, w2, gaxpy<R, GM>(G6, w1, base));
                if constexpr (kAccLds) ldsQ[rp][threadIdx.x] = acc;
                else qa[H] = acc;
            }
        }
        C<R> nw = mkc<R>(0.0, 0.0);
        const bool second = z >= c.z0;
        if (second && c.winner) {
            // fourth stage at plane z, inner tile (ldsB of plane z was written
            // H iterations ago)
            FusedSten<R, ORD, S::TXB, TERMS> sb;
            sb.L = ldsB[rz];
            sb.lc = c.lb;
#pragma unroll
            for (int d = 0; d <= 2 * H; ++d) sb.q[d] = qv[(PH + d) % Q];
            sb.w = w;
            nw = gaxpy<R, GM>(G6, rhs_w<R>(k, sb, c.mx, c.my, z), kAccLds ? ldsQ[rz][threadIdx.x] : qa[0]);
        }
        bst<R>(plane_res_m<R>(out, (unsigned)(second ? z : c.z0), pbytes), second ? oown : kNoOff, nw);
#pragma unroll
        for (int d = 0; d < H; ++d) qa[d] = qa[d + 1];
        ra ^= 1;
        rz = rz == NB - 1 ? 0 : rz + 1;
        rp = rp == NB - 1 ? 0 : rp + 1;
    };
    march_rotating<(DEPTH == 2 && Q % 2 == 1) ? 2 * Q : Q>(start, c.z1, step);
}

template <typename R, int ORD, int TERMS, int GM>
__global__ __launch_bounds__(kRk4NT<ORD>) void k_rk4_f34(
    Geom g, StW<R> w, RhsC<R> k, const C<R>* __restrict__ u2, const C<R>* __restrict__ psi, C<R>* __restrict__ out,
    C<R> dt, C<R> dt3, C<R> dt6) {
    rk4_f34_body<R, ORD, TERMS, GM, kRk4OX, kRk4TYB>(g, w, k, u2, psi, out, dt, dt3, dt6,
                                                     xcd_remap(blockIdx.x, g.nblocks));
}

// Fused Cash-Karp attempt (rkck, src/solve.f90:248-378, with the error
// max-norm of rkqs, src/solve.f90:176-185): the six stages in three kernels
// of two stages each, on the same overlapped tiles, register queues and LDS
// ring as the RK4 pairs.  k_j = dt f(u_j), u_1 = psi0.
//   pair 1: f0 = f(psi0), u2 = psi0 + b21 k1 on the whole block;
//           k2 = dt f(u2), u3 = psi0 + b31 k1 + b32 k2 on the inner tile
//                                          reads psi0                 writes f0, k2, u3
//   pair 2: k3 = dt f(u3), u4 = psi0 + b41 k1 + b42 k2 + b43 k3;
//           k4 = dt f(u4), u5 = psi0 + b51 k1 + .. + b54 k4
//                                          reads u3, psi0, f0, k2     writes k3, k4, u5
//   pair 3: k5 = dt f(u5), u6 = psi0 + b61 k1 + .. + b65 k5;
//           k6 = dt f(u6), psi_new = psi0 + c1 k1 + c3 k3 + c4 k4 + c6 k6,
//           err = dc1 k1 + dc3 k3 + dc4 k4 + dc5 k5 + dc6 k6 and the block
//           maxima of |err| / |psi_scal| and |psi_new|^2
//                                          reads u5, psi0, f0, k2, k3, k4   writes psi_new
// 18 field accesses per attempt instead of 36.  Every sum adds its terms in
// the order of the stage kernels (OpCkMid, OpCkLast).  Pairs 1 and 2 form
// the part of the second stage's sum that does not depend on its own k at
// the first stage and carry it in a register queue for the H planes in
// between.  Pair 3 needs three such partials (psi_new, the error, the
// scale) and keeps part of them in LDS rings (see ck_pair_body).
template <typename R> struct CkPairArgs {
    const C<R>* u;  // stencil input: psi0 (pair 1), u3 (pair 2), u5 (pair 3)
    const C<R>* p0;
    const C<R>* f0;
    const C<R>* k2;
    const C<R>* k3;
    const C<R>* k4;
    C<R>* f0out;  // pair 1
    C<R>* ka;     // pair 2: k3
    C<R>* kb;     // pair 1: k2, pair 2: k4
    C<R>* unext;  // pair 1: u3, pair 2: u5
    C<R>* out;    // pair 3: psi_new
    C<R> dt;
    double dt0_re, dt0_im;  // the step's dt on entry to rkqs (psi_scal)
};

// GM: the form of G = i pim dt (gmul, hmul).
template <typename R, int ORD, int TERMS, int PAIR, int GM>
__device__ __forceinline__ void ck_pair_body(const Geom& g, const StW<R>& w0, const RhsC<R>& k, const CkPairArgs<R>& a,
                                             double* __restrict__ partial, int bid, int pidx) {
    using S = FusedShape<ORD>;
    constexpr int H = S::H, NB = H + 1, Q = 2 * H + 1;
    constexpr bool P3 = PAIR == 3;
    // FP64 pairs 2 and 3 load their pointwise inputs after the first stage's
    // stencil: loaded earlier they are live across the stencil, and pair 3
    // spilled 12 registers (4.8 ms against 3.9 ms at 512^3); pair 2 takes
    // 2.91 against 2.95 ms.  In FP32 there is room, and earlier loads hide
    // their latency (pair 3: 2.2 ms against 2.6 ms).
    constexpr bool kLate = sizeof(R) == 8;
    // Pair 3 carries its second stage's partials from the first stage: the
    // psi_new partial in an LDS ring (ldsK), and the error partial and the
    // scale |psi_scal|^2.  FP32 keeps those two in registers; FP64 keeps the
    // error partial in a second LDS ring (ldsE), both rings indexed by the
    // inner thread (inner-first map), 160 KB of LDS in all at order 4.
    // Carried in registers the FP64 state spills (5.3 ms at 512^3); read
    // again at the second stage, psi0, f0, k3 and k4 miss L2 (1.9x the
    // algorithmic bytes, 3.9 ms).
    constexpr bool kErrLds = sizeof(R) == 8;
    constexpr int NIN = S::OX * S::OY;
    __shared__ C<R> ldsA[2][S::LHA * S::LWA];         // stencil input plane p with halo
    __shared__ C<R> ldsB[NB][S::NTB];                 // first-stage u_next, planes z..z+H
    __shared__ C<R> ldsK[P3 ? NB : 1][P3 ? (kErrLds ? NIN : S::NTB) : 1];
    __shared__ double2 ldsE[P3 && kErrLds ? NB : 1][P3 && kErrLds ? NIN : 1];
    __shared__ double red_sh[16];
    const FusedCtx<ORD, !(P3 && sizeof(R) == 4)> c(g, bid);  // the map measured faster per kernel
    const StW<R> w = scale_w<R>(w0, k.a);
    const C<R> G = g_of<R>(k.pim, a.dt);  // k = dt f = G w
    double red0 = 0.0, red1 = 0.0;
    if (c.z0 < c.z1) {  // uniform per block
        const int start = c.z0 - 2 * H;
        // Queues, slot (plane - start) % Q: the stencil input for planes
        // z..z+2H and the first stage's u for planes z-H..z+H; for planes
        // z..z+H qs (pairs 1, 2) the inner tile's partial sum of u_next, and
        // qe, qd (FP32 pair 3) the error partial and |psi_scal|^2.
        C<R> qu[Q], qv[Q], qs[NB];
        double2 qe[NB];
        double qd[NB];
        int ra = 0, rz = 0, rp = H;  // (z - start) & 1, (z - start) % NB, (p - start) % NB
#pragma unroll
        for (int d = 0; d < Q; ++d) qv[d] = mkc<R>(0.0, 0.0);
#pragma unroll
        for (int d = 0; d < NB; ++d) {
            qs[d] = mkc<R>(0.0, 0.0);
            qe[d] = mkc<double>(0.0, 0.0);
            qd[d] = 1.0;
        }
        if (c.active) {
#pragma unroll
            for (int d = 0; d < 2 * H; ++d) qu[d] = *(a.u + Cell{plane_off(g, bcmap1(start + d, g.nz, g.bcs)), c.col});
        }
        struct Pw {
            C<R> p0, f0, k2, k3, k4;
        };
        auto load_pw = [&](Pw& v, long long pb) {
            const Cell ci{pb, c.col};
            if constexpr (PAIR >= 2) {
                v.p0 = *(a.p0 + ci);
                v.f0 = *(a.f0 + ci);
                v.k2 = *(a.k2 + ci);
            }
            if constexpr (P3) {
                v.k3 = *(a.k3 + ci);
                v.k4 = *(a.k4 + ci);
            }
        };
        // loads of iteration zz: centre of plane zz + 2H, halo of plane zz + H
        // and (FP32 pair 2) the pointwise inputs of plane zz + H
        struct Pf {
            C<R> qn;
            C<R> hr[S::NSLOT];
            Pw pw;
        };
        auto fetch = [&](Pf& F, int zz) {
            const long long pb = plane_off(g, bcmap1(zz + H, g.nz, g.bcs));
            if (c.active) {
                F.qn = *(a.u + Cell{plane_off(g, bcmap1(zz + 2 * H, g.nz, g.bcs)), c.col});
                if constexpr (PAIR == 2 && !kLate) load_pw(F.pw, pb);
            }
#pragma unroll
            for (int s = 0; s < S::NSLOT; ++s)
                if (c.hv[s]) F.hr[s] = *(a.u + Cell{pb, c.hoff[s]});
        };
        Pf F{};
        fetch(F, start);
        constexpr CkTab<R> t = ck_tab<R>();  // constants, not kernel arguments held in registers
        auto step = [&](int z, auto phase) {
            constexpr int PH = decltype(phase)::value;
            const int p = z + H;
            qu[(PH + 2 * H) % Q] = F.qn;
            Pw pw = F.pw;
            if constexpr (P3 && !kLate) {
                // pair 3 loads its pointwise inputs at the top of the plane's
                // iteration: a second set in flight would not fit
                if (c.active) load_pw(pw, plane_off(g, bcmap1(p, g.nz, g.bcs)));
            }
            C<R>* A = ldsA[ra];
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
                const C<R> wa = rhs_w<R>(k, sa, c.mx, c.my, bcmap1(p, g.nz, g.bcs));
                C<R> ub;
                if constexpr (PAIR == 1) {
                    const C<R> p0 = qu[(PH + H) % Q];
                    const C<R> k1 = gmul<R, GM>(G, wa);
                    if (c.store && p >= c.z0 && p < c.z1)
                        st_stream<R>(a.f0out + Cell{plane_off(g, p), c.own}, rot_pim<R>(k.pim, wa));
                    ub = caxpy<R>(t.b21, k1, p0);
                    qs[H] = caxpy<R>(t.b31, k1, p0);
                } else if constexpr (PAIR == 2) {
                    if constexpr (kLate) load_pw(pw, plane_off(g, bcmap1(p, g.nz, g.bcs)));
                    const C<R> k3 = gmul<R, GM>(G, wa);
                    const C<R> k1 = hmul<R, GM>(a.dt, pw.f0);
                    if (c.store && p >= c.z0 && p < c.z1) st_stream<R>(a.ka + Cell{plane_off(g, p), c.own}, k3);
                    ub = caxpy<R>(t.b43, k3, caxpy<R>(t.b42, pw.k2, caxpy<R>(t.b41, k1, pw.p0)));
                    qs[H] = caxpy<R>(t.b53, k3, caxpy<R>(t.b52, pw.k2, caxpy<R>(t.b51, k1, pw.p0)));
                } else {
                    if constexpr (kLate) load_pw(pw, plane_off(g, bcmap1(p, g.nz, g.bcs)));
                    const C<R> k5 = gmul<R, GM>(G, wa);
                    const C<R> k1 = hmul<R, GM>(a.dt, pw.f0);
                    ub = caxpy<R>(t.b65, k5,
                                  caxpy<R>(t.b64, pw.k4,
                                           caxpy<R>(t.b63, pw.k3, caxpy<R>(t.b62, pw.k2, caxpy<R>(t.b61, k1, pw.p0)))));
                    if (c.winner) {
                        const int slot = rp, li = kErrLds ? (int)threadIdx.x : c.lb;
                        ldsK[slot][li] = caxpy<R>(t.c4, pw.k4, caxpy<R>(t.c3, pw.k3, caxpy<R>(t.c1, k1, pw.p0)));
                        const double2 se = mkc<double>(t.dc1 * (double)k1.x + t.dc3 * (double)pw.k3.x +
                                                           t.dc4 * (double)pw.k4.x + t.dc5 * (double)k5.x,
                                                       t.dc1 * (double)k1.y + t.dc3 * (double)pw.k3.y +
                                                           t.dc4 * (double)pw.k4.y + t.dc5 * (double)k5.y);
                        if constexpr (kErrLds) ldsE[slot][li] = se;
                        else qe[H] = se;
                        const double ap = sqrt((double)pw.p0.x * pw.p0.x + (double)pw.p0.y * pw.p0.y);
                        const double af = sqrt((double)pw.f0.x * pw.f0.x + (double)pw.f0.y * pw.f0.y);
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
                const C<R> kb = gmul<R, GM>(G, rhs_w<R>(k, sb, c.mx, c.my, z));
                const Cell oz{plane_off(g, z), c.own};
                if constexpr (PAIR == 1) {
                    if (c.store) {
                        st_stream<R>(a.kb + oz, kb);
                        st_stream<R>(a.unext + oz, caxpy<R>(t.b32, kb, qs[0]));
                    }
                } else if constexpr (PAIR == 2) {
                    if (c.store) {
                        st_stream<R>(a.kb + oz, kb);
                        st_stream<R>(a.unext + oz, caxpy<R>(t.b54, kb, qs[0]));
                    }
                } else {
                    if (c.store) {
                        const int slot = rz, li = kErrLds ? (int)threadIdx.x : c.lb;
                        const C<R> nw = caxpy<R>(t.c6, kb, ldsK[slot][li]);
                        const double2 se = kErrLds ? ldsE[slot][li] : qe[0];
                        const double er = se.x + t.dc6 * (double)kb.x;
                        const double ei = se.y + t.dc6 * (double)kb.y;
                        const double den = qd[0];
                        st_stream<R>(a.out + oz, nw);
                        const double e = sqrt((er * er + ei * ei) / den);
                        red0 = e > red0 ? e : red0;
                        const double a2 = (double)nw.x * nw.x + (double)nw.y * nw.y;
                        red1 = a2 > red1 ? a2 : red1;
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < H; ++d) {
                qs[d] = qs[d + 1];
                qe[d] = qe[d + 1];
                qd[d] = qd[d + 1];
            }
            ra ^= 1;
            rz = rz == NB - 1 ? 0 : rz + 1;
            rp = rp == NB - 1 ? 0 : rp + 1;
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

template <typename R, int ORD, int TERMS, int PAIR, int GM>
__global__ __launch_bounds__(1024) void k_ck_pair(Geom g, StW<R> w, RhsC<R> k, CkPairArgs<R> a,
                                                  double* __restrict__ partial) {
    ck_pair_body<R, ORD, TERMS, PAIR, GM>(g, w, k, a, partial, xcd_remap(blockIdx.x, g.nblocks), blockIdx.x);
}

// Batched form for an ensemble of grids: persistent blocks walk a work list
// of (member, fused block) items, members in descriptor order.  A member's
// state alternates between psi and alt from step to step (flip = 1: the
// current state is in alt), so no copy closes a step.
struct FusedMember {
    Geom g;  // fused_tiles geometry of the member
    StW<double> w;
    RhsC<double> k;
    double2* psi;
    double2* alt;
    double2* u2;
    int begin;  // first work item
};

__device__ __forceinline__ int fused_member_of(const FusedMember* __restrict__ mem, int nmem, int item) {
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mem[mid].begin <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int ORD, int GM>
__global__ __launch_bounds__(1024) void k_rk4_f12_batched(const FusedMember* __restrict__ mem, int nmem, int total,
                                                          double2 hdt, int flip) {
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const FusedMember& M = mem[fused_member_of(mem, nmem, item)];
        rk4_f12_body<double, ORD, 0, GM>(M.g, M.w, M.k, flip ? M.alt : M.psi, M.u2, hdt, hdt, item - M.begin);
        __syncthreads();  // the next item reuses the LDS planes
    }
}

template <int ORD, int GM>
__global__ __launch_bounds__(1024) void k_rk4_f34_batched(const FusedMember* __restrict__ mem, int nmem, int total,
                                                          double2 dt, double2 dt3, double2 dt6, int flip) {
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const FusedMember& M = mem[fused_member_of(mem, nmem, item)];
        rk4_f34_body<double, ORD, 0, GM>(M.g, M.w, M.k, M.u2, flip ? M.alt : M.psi, flip ? M.psi : M.alt, dt, dt3, dt6,
                          item - M.begin);
        __syncthreads();
    }
}

// Host side (gpe_k_fused.hip).
template <int ORD> int fused_tiles(int nx, int ny, int nz, int ncu, Geom& g);      // Cash-Karp pairs, batched
template <int ORD> int rk4_fused_tiles(int nx, int ny, int nz, int ncu, Geom& g);  // k_rk4_f12 / k_rk4_f34
template <typename R, int ORD>
hipError_t launch_rk4_fused(const Geom& g, const StW<R>& w, const RhsC<R>& k, const C<R>* psi, C<R>* u2, C<R>* out,
                            C<R> dt, int stage_pair, hipStream_t s);  // g.general: the trap term
template <int ORD>
hipError_t launch_rk4_fused_batched(const FusedMember* mem, int nmem, int total, int grid, double2 dt, int stage_pair,
                                    int flip, hipStream_t s);
// one RK2 step (pk = PK_RK2) or two Euler steps (PK_EULER2): psi -> out, on
// the tiling of rk4_fused_tiles (gpe_k_fused2.hip)
template <typename R, int ORD>
hipError_t launch_pair_fused(const Geom& g, const StW<R>& w, const RhsC<R>& k, const C<R>* psi, C<R>* out, C<R> dt,
                             int pk, hipStream_t s);
template <typename R, int ORD>
hipError_t launch_ck_fused(const Geom& g, const StW<R>& w, const RhsC<R>& k, const CkPairArgs<R>& a, int pair,
                           double* partial, hipStream_t s);  // g.general: the trap term

}  // namespace gpe
