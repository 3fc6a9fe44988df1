// Host-side launch interface of the stencil kernels and the pointwise
// kernels.  The stencil launchers are explicitly instantiated per
// (precision, order) in gpe_k_<p><order>.hip so those translation units
// compile in parallel.
#pragma once

#include "gpe_mixed.h"

namespace gpe {

struct Tiles {
    int tx, ty, tz;
};

// Tile shapes compiled into the library (tx, ty).
inline bool tile_supported(int tx, int ty) {
    return (tx == 64 && (ty == 4 || ty == 8 || ty == 16)) || (tx == 32 && (ty == 4 || ty == 8 || ty == 16));
}

// Fills the tiling fields of g for tile shape t and returns the block count.
inline int tile_geom(const Tiles& t, Geom& g) {
    g.ntx = (g.nx + t.tx - 1) / t.tx;
    g.nty = (g.ny + t.ty - 1) / t.ty;
    g.tz = t.tz;
    const int nzc = (g.nz + t.tz - 1) / t.tz;
    g.nblocks = g.ntx * g.nty * nzc;
    return g.nblocks;
}

template <typename R> using OpCk2 = OpCkMid<R, 2>;
template <typename R> using OpCk3 = OpCkMid<R, 3>;
template <typename R> using OpCk4 = OpCkMid<R, 4>;
template <typename R> using OpCk5 = OpCkMid<R, 5>;

template <typename R, int ORD, class Op>
hipError_t launch_stencil(const Tiles& t, Geom g, const StW<R>& w, const C<R>* u, const Op& op, double* partial,
                          hipStream_t s);

#define GPE_FOR_EACH_OP(X, R)                                                                                      \
    X(R, OpAxpySelf<R>)                                                                                            \
    X(R, OpAxpyP0<R>)                                                                                              \
    X(R, OpRk4First<R>)                                                                                            \
    X(R, OpRk4Second<R>)                                                                                           \
    X(R, OpRk4Third<R>)                                                                                            \
    X(R, OpRk4Last<R>)                                                                                             \
    X(R, OpCk1<R>)                                                                                                 \
    X(R, OpCk2<R>)                                                                                       \
    X(R, OpCk3<R>)                                                                                       \
    X(R, OpCk4<R>)                                                                                       \
    X(R, OpCk5<R>)                                                                                       \
    X(R, OpCkLast<R>)                                                                                              \
    X(R, OpDeriv<R>)                                                                                               \
    X(R, OpRhs<R>)                                                                                                 \
    X(R, OpEnergy<R>)                                                                                              \
    X(R, OpHamiltonian<R>)                                                                                         \
    X(R, OpMomentum<R>)                                                                                            \
    X(R, OpVelocity<R>)                                                                                            \
    X(R, OpVelocityStats<R>)

// Mixed-storage operators: those over a plain complex128 stencil use
// launch_stencil<double>, the others launch_stencil_sum.
template <int ORD, class Op>
hipError_t launch_stencil_sum(const Tiles& t, Geom g, const StW<double>& w, const SumSrc& src, const Op& op,
                              double* partial, hipStream_t s);
using OpMxCk3 = OpMxCkMid<3>;
using OpMxCk4 = OpMxCkMid<4>;
using OpMxCk5 = OpMxCkMid<5>;
#define GPE_FOR_EACH_MIXED_PLAIN(X) X(double, OpMxRk4S1) X(double, OpMxCk1)
#define GPE_FOR_EACH_MIXED_SUM(X) X(OpMxRk4S23) X(OpMxRk4Last) X(OpMxCk2) X(OpMxCk3) X(OpMxCk4) X(OpMxCk5) X(OpMxCkLast)

#define GPE_DECLARE_EXTERN_ORD(R, OP, ORD) \
    extern template hipError_t launch_stencil<R, ORD, OP>(const Tiles&, Geom, const StW<R>&, const C<R>*, const OP&, \
                                                          double*, hipStream_t);
#define GPE_DECLARE_EXTERN(R, OP) GPE_DECLARE_EXTERN_ORD(R, OP, 2) GPE_DECLARE_EXTERN_ORD(R, OP, 4)
GPE_FOR_EACH_OP(GPE_DECLARE_EXTERN, double)
GPE_FOR_EACH_OP(GPE_DECLARE_EXTERN, float)
GPE_FOR_EACH_MIXED_PLAIN(GPE_DECLARE_EXTERN)
#define GPE_DECLARE_EXTERN_SUM_ORD(OP, ORD)                                                                        \
    extern template hipError_t launch_stencil_sum<ORD, OP>(const Tiles&, Geom, const StW<double>&, const SumSrc&, \
                                                           const OP&, double*, hipStream_t);
#define GPE_DECLARE_EXTERN_SUM(OP) GPE_DECLARE_EXTERN_SUM_ORD(OP, 2) GPE_DECLARE_EXTERN_SUM_ORD(OP, 4)
GPE_FOR_EACH_MIXED_SUM(GPE_DECLARE_EXTERN_SUM)

// Pointwise kernels (gpe_pointwise.hip).
template <typename R>
hipError_t launch_ck1_retry(const C<R>* p0, const C<R>* f0, C<R>* unext, C<R> dt, R b21, long long n, hipStream_t s);
template <typename R>
hipError_t launch_norm_partial(const C<R>* u, int nx, int ny, int nz, int weighted, double* partial, int nblocks,
                               hipStream_t s);
template <typename R> hipError_t launch_scale(C<R>* u, long long n, double snn, double sn, hipStream_t s);

// A vortex line for k_imprint_line: line_param of the reference plus scal
// and the zero-wavelength guards safe1/safe2 of get_r / get_theta.
struct LineParam {
    double x0, y0, z0, amp1, amp2, ll, sgn, scal, safe1, safe2;
    int dir;            // 0 x, 1 y, 2 z
    int imprint_phase;  // 1: phase only
};
template <typename R>
hipError_t launch_imprint_line(C<R>* u, int nx, int ny, long long n, const double lo[3], const double d[3],
                               const LineParam& vl, hipStream_t s);
template <typename R> hipError_t launch_from_c128(const double2* src, C<R>* dst, long long n, hipStream_t s);
template <typename R> hipError_t launch_from_c64(const float2* src, C<R>* dst, long long n, hipStream_t s);
template <typename R> hipError_t launch_to_c128(const C<R>* src, double2* dst, long long n, hipStream_t s);
template <typename R> hipError_t launch_to_c64(const C<R>* src, float2* dst, long long n, hipStream_t s);
hipError_t launch_copy16(void* dst, const void* src, long long n16, hipStream_t s);
hipError_t launch_stream5(void* c, void* d, const void* a, const void* b, const void* e, long long n16, hipStream_t s);
hipError_t launch_reduce_final(int mode, const double* partial, int n, int nval, double* out, hipStream_t s);

// Diagnostics (gpe_diag.hip).
hipError_t launch_vel_hist(const double* vel, long long n, int nbins, const double* vmax3, unsigned long long* hist,
                           int nblocks, hipStream_t s);
hipError_t launch_vcf(const double* vx, int nx, long long nlines, double* partial, int nblocks, double* out,
                      hipStream_t s);

// Line length (gpe_linelength.hip): last holds 5, carry 3 doubles per x-line
// ((ny-2)(nz-2) lines), runs 7 doubles per linelength_runs, partial 2
// linelength_blocks doubles; out[0] = the sum of the cell contributions
// (times dx gives the reference's value).
long long linelength_lines(int ny, int nz);
long long linelength_blocks(int ny, int nz);
long long linelength_runs(int ny, int nz);
hipError_t launch_linelength(const double2* psi, int nx, int ny, int nz, double* last, double* carry, double* runs,
                             double* partial, double* out, hipStream_t s);

// Spectral diagnostics (gpe_spectral.hip); out holds 2 kSpectrumShells + nhalf
// values: shell sums, shell counts, F bins.
constexpr int kSpectrumShells = 7;
// entries of the per-k2 table launch_spectrum fills and reads (k2 = 0 .. max)
long long spectrum_table_len(int nx, int ny, int nz);
template <typename R>
hipError_t launch_spectrum(const C<R>* a, int nx, int ny, int nz, int nhalf, double* partial, int nb, double* out,
                           unsigned* tab, hipStream_t s);
hipError_t launch_condensed_tot(int nx, int ny, int nz, const double d[3], double rho0, double* partial, int nb,
                                double* out, hipStream_t s);

constexpr int kNormBlocks = 2048;

}  // namespace gpe
