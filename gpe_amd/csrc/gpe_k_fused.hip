// Fused RK4 stage pairs (complex128, plain right-hand side): launchers and
// the z-range choice.
#include "gpe_fused.h"

namespace gpe {

// Output tiles of 64 x (15 - 2H) cells; the z-range is split into C chunks
// so that the block count fills whole waves of one block per CU, each chunk
// paying 2H warm-up planes.  Returns the block count.
template <int ORD, int OX, int TYB> static int fused_tiles_shape(int nx, int ny, int nz, int ncu, Geom& g) {
    using S = FusedShape<ORD, OX, TYB>;
    g.ntx = (nx + S::OX - 1) / S::OX;
    g.nty = (ny + S::OY - 1) / S::OY;
    const long long tiles = (long long)g.ntx * g.nty;
    double best = -1.0;
    int best_tz = nz;
    for (int cch = 1; cch <= 16 && cch <= nz; ++cch) {
        const int tz = (nz + cch - 1) / cch;
        const int nzc = (nz + tz - 1) / tz;
        const long long b = tiles * nzc;
        const long long waves = (b + ncu - 1) / ncu;
        const double eff = (double)b / (double)(waves * ncu) * (double)tz / (double)(tz + 2 * S::H);
        if (eff > best + 1e-9) {
            best = eff;
            best_tz = tz;
        }
    }
    g.tz = best_tz;
    g.nblocks = (int)(tiles * ((nz + best_tz - 1) / best_tz));
    return g.nblocks;
}

template <int ORD> int fused_tiles(int nx, int ny, int nz, int ncu, Geom& g) {
    return fused_tiles_shape<ORD, 64, kFusedTYB>(nx, ny, nz, ncu, g);
}

// the RK4 pair's shape; a 512-thread block runs two to a CU
template <int ORD> int rk4_fused_tiles(int nx, int ny, int nz, int ncu, Geom& g) {
    constexpr int per_cu = FusedShape<ORD, kRk4OX, kRk4TYB>::NT == 512 ? 2 : 1;
    return fused_tiles_shape<ORD, kRk4OX, kRk4TYB>(nx, ny, nz, ncu * per_cu, g);
}

// The kernels of one G form (gaxpy): stage pair 12 or 34, with or without the trap term.
template <typename R, int ORD, int GM>
static void launch_rk4_fused_gm(const Geom& g, const StW<R>& w, const RhsC<R>& k, const C<R>* psi, C<R>* u2, C<R>* out,
                                C<R> dt, int stage_pair, hipStream_t s) {
    constexpr int NT = FusedShape<ORD, kRk4OX, kRk4TYB>::NT;
    if (stage_pair == 12) {
        const C<R> hdt = mkc<R>(R(0.5) * dt.x, R(0.5) * dt.y);
        if (g.general)
            hipLaunchKernelGGL((k_rk4_f12<R, ORD, RHS_V, GM>), dim3(g.nblocks), dim3(NT), 0, s, g, w, k, psi, u2, hdt);
        else
            hipLaunchKernelGGL((k_rk4_f12<R, ORD, 0, GM>), dim3(g.nblocks), dim3(NT), 0, s, g, w, k, psi, u2, hdt);
    } else {
        const C<R> dt3 = mkc<R>(dt.x / R(3), dt.y / R(3)), dt6 = mkc<R>(dt.x / R(6), dt.y / R(6));
        if (g.general)
            hipLaunchKernelGGL((k_rk4_f34<R, ORD, RHS_V, GM>), dim3(g.nblocks), dim3(NT), 0, s, g, w, k,
                               (const C<R>*)u2, psi, out, dt, dt3, dt6);
        else
            hipLaunchKernelGGL((k_rk4_f34<R, ORD, 0, GM>), dim3(g.nblocks), dim3(NT), 0, s, g, w, k, (const C<R>*)u2,
                               psi, out, dt, dt3, dt6);
    }
}

template <typename R, int ORD>
hipError_t launch_rk4_fused(const Geom& g, const StW<R>& w, const RhsC<R>& k, const C<R>* psi, C<R>* u2, C<R>* out,
                            C<R> dt, int stage_pair, hipStream_t s) {
    if (g.nblocks <= 0) return hipSuccess;
    switch (g_mode((double)dt.x, (double)dt.y)) {
        case GM_IM: launch_rk4_fused_gm<R, ORD, GM_IM>(g, w, k, psi, u2, out, dt, stage_pair, s); break;
        case GM_RE: launch_rk4_fused_gm<R, ORD, GM_RE>(g, w, k, psi, u2, out, dt, stage_pair, s); break;
        default: launch_rk4_fused_gm<R, ORD, GM_GEN>(g, w, k, psi, u2, out, dt, stage_pair, s); break;
    }
    return hipGetLastError();
}

template <int ORD, int GM>
static void launch_rk4_fused_batched_gm(const FusedMember* mem, int nmem, int total, int grid, double2 dt,
                                        int stage_pair, int flip, hipStream_t s) {
    if (stage_pair == 12) {
        const double2 hdt = mkc<double>(0.5 * dt.x, 0.5 * dt.y);
        hipLaunchKernelGGL((k_rk4_f12_batched<ORD, GM>), dim3(grid), dim3(1024), 0, s, mem, nmem, total, hdt, flip);
    } else {
        const double2 dt3 = mkc<double>(dt.x / 3.0, dt.y / 3.0), dt6 = mkc<double>(dt.x / 6.0, dt.y / 6.0);
        hipLaunchKernelGGL((k_rk4_f34_batched<ORD, GM>), dim3(grid), dim3(1024), 0, s, mem, nmem, total, dt, dt3,
                           dt6, flip);
    }
}

template <int ORD>
hipError_t launch_rk4_fused_batched(const FusedMember* mem, int nmem, int total, int grid, double2 dt, int stage_pair,
                                    int flip, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    switch (g_mode(dt.x, dt.y)) {
        case GM_IM: launch_rk4_fused_batched_gm<ORD, GM_IM>(mem, nmem, total, grid, dt, stage_pair, flip, s); break;
        case GM_RE: launch_rk4_fused_batched_gm<ORD, GM_RE>(mem, nmem, total, grid, dt, stage_pair, flip, s); break;
        default: launch_rk4_fused_batched_gm<ORD, GM_GEN>(mem, nmem, total, grid, dt, stage_pair, flip, s); break;
    }
    return hipGetLastError();
}

template int fused_tiles<2>(int, int, int, int, Geom&);
template int rk4_fused_tiles<2>(int, int, int, int, Geom&);
template int rk4_fused_tiles<4>(int, int, int, int, Geom&);
template int fused_tiles<4>(int, int, int, int, Geom&);
#define GPE_FUSED_INST(R, ORD)                                                                                     \
    template hipError_t launch_rk4_fused<R, ORD>(const Geom&, const StW<R>&, const RhsC<R>&, const C<R>*, C<R>*, C<R>*, \
                                                 C<R>, int, hipStream_t);
GPE_FUSED_INST(double, 2)
GPE_FUSED_INST(double, 4)
GPE_FUSED_INST(float, 2)
GPE_FUSED_INST(float, 4)
template hipError_t launch_rk4_fused_batched<2>(const FusedMember*, int, int, int, double2, int, int, hipStream_t);
template hipError_t launch_rk4_fused_batched<4>(const FusedMember*, int, int, int, double2, int, int, hipStream_t);

}  // namespace gpe
