// Diagnostics of the state: the velocity PDF (get_pdf_velocity + get_pdf,
// src/variables.f90:377-463; OpVelocity<R, true> forms the velocity field and
// its statistics in one pass) and the velocity correlation function along x
// (get_vcf, src/variables.f90:467-507), from a velocity field that
// OpVelocity wrote.  All sums are FP64 in a fixed order (per-block partials,
// then one block); histogram counts are integers.
#include <hip/hip_runtime.h>

#include "gpe_launch.h"

namespace gpe {

constexpr int kDiagThreads = 256;
constexpr int kVelValid = 500;  // the reference keeps values < 0.5e3 (marker 1e3)

// Bin edges of get_pdf (src/variables.f90:448-451): bin i in
// (-nbins/2, nbins/2] holds (thr(i-1), thr(i)], thr(i) = ((2 i) vmax) / nbins,
// evaluated in that order so the edges equal the reference's.
__device__ __forceinline__ double pdf_edge(int i, double vmax, int nbins) {
    return (2.0 * (double)i) * vmax / (double)nbins;
}

// x > pdf_edge(i, vmax, nbins), with the division only where x lies within
// a few ulp of the edge: t * inv_nb is the edge to within 2 ulp, so outside
// 4 ulp of it the comparison is decided the same way.
__device__ __forceinline__ bool above_edge(double x, int i, double vmax, int nbins, double inv_nb) {
    const double t = (2.0 * (double)i) * vmax;
    const double ea = t * inv_nb;
    const double tol = 8.9e-16 * fabs(ea);
    if (x > ea + tol) return true;
    if (x < ea - tol) return false;
    return x > t / (double)nbins;
}

// Histogram of the valid values per component, counted in LDS per block and
// added into hist[a * nbins + (i + nbins/2 - 1)].  Lanes count into ncopy
// interleaved copies of the bins (lane % ncopy; each copy padded by one word
// so equal bins of different copies fall in different LDS banks): the
// values pile up in a few central bins, and atomics of one wavefront to one
// LDS word are serialised.
constexpr int kHistMaxCopies = 8;
constexpr int kHistUnroll = 4;
constexpr int kHistLdsBytes = 64 * 1024;
inline int hist_copies(int nbins) {
    const int c = kHistLdsBytes / ((3 * nbins + 1) * (int)sizeof(unsigned int));
    return c < 1 ? 1 : (c > kHistMaxCopies ? kHistMaxCopies : c);
}
__global__ __launch_bounds__(kDiagThreads) void k_vel_hist(const double* __restrict__ vel, long long n, int nbins,
                                                          int ncopy, const double* __restrict__ vmax3,
                                                          unsigned long long* __restrict__ hist) {
    extern __shared__ unsigned int hsh[];  // ncopy * (3 * nbins + 1)
    const int stride = 3 * nbins + 1;
    for (int j = threadIdx.x; j < ncopy * stride; j += kDiagThreads) hsh[j] = 0u;
    __syncthreads();
    unsigned int* mine = hsh + (threadIdx.x % ncopy) * stride;
    const int lo = -nbins / 2 + 1, hi = nbins / 2;
    const double inv_nb = 1.0 / (double)nbins;
    for (int a = 0; a < 3; ++a) {
        const double* v = vel + a * n;
        const double vmax = vmax3[a];
        if (!(vmax > 0.0)) continue;  // uniform
        const double scale = (double)nbins / (2.0 * vmax);
        // kHistUnroll values of the grid-stride sequence are loaded together,
        // so their loads are in flight at once, then binned one by one
        const long long stride = (long long)gridDim.x * kDiagThreads;
        for (long long i0 = (long long)blockIdx.x * kDiagThreads + threadIdx.x; i0 < n; i0 += kHistUnroll * stride) {
            double xs[kHistUnroll];
#pragma unroll
            for (int u = 0; u < kHistUnroll; ++u) {
                const long long i = i0 + u * stride;
                xs[u] = i < n ? v[i] : 1e3;  // 1e3: the invalid marker, not counted
            }
#pragma unroll
            for (int u = 0; u < kHistUnroll; ++u) {
                const double x = xs[u];
                if (!(x < kVelValid)) continue;
                // a first guess of the bin, then settle on the bin whose edges
                // (as the reference forms them) hold x
                double q = ceil(x * scale);
                q = q < lo - 1 ? lo - 1 : (q > hi + 1 ? hi + 1 : q);
                int b = (int)q;
                while (b <= hi && above_edge(x, b, vmax, nbins, inv_nb)) ++b;
                while (b >= lo && !above_edge(x, b - 1, vmax, nbins, inv_nb)) --b;
                if (b >= lo && b <= hi) atomicAdd(&mine[a * nbins + (b - lo)], 1u);
            }
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < 3 * nbins; j += kDiagThreads) {
        unsigned int c = 0u;
        for (int r = 0; r < ncopy; ++r) c += hsh[r * stride + j];
        if (c) atomicAdd(&hist[j], (unsigned long long)c);
    }
}

// Velocity correlation along x: block b sums, over its lines (y, z), the
// products v(i) v(i+r) for every r into partial[b * nx + r].
__global__ __launch_bounds__(kDiagThreads) void k_vcf_partial(const double* __restrict__ vx, int nx, long long nlines,
                                                             double* __restrict__ partial) {
    extern __shared__ double line[];  // nx
    const long long per = (nlines + gridDim.x - 1) / gridDim.x;
    const long long l0 = (long long)blockIdx.x * per, l1 = l0 + per < nlines ? l0 + per : nlines;
    // each thread owns the lags r = tid, tid + blockDim, ...; at most 32 of them
    constexpr int kMaxLags = 32;
    double acc[kMaxLags];
#pragma unroll
    for (int q = 0; q < kMaxLags; ++q) acc[q] = 0.0;
    for (long long l = l0; l < l1; ++l) {
        __syncthreads();
        for (int i = threadIdx.x; i < nx; i += kDiagThreads) line[i] = vx[l * nx + i];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kMaxLags; ++q) {
            const int r = threadIdx.x + q * kDiagThreads;
            if (r < nx) {
                double s = 0.0;
                for (int i = 0; i + r < nx; ++i) s += line[i] * line[i + r];
                acc[q] += s;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kMaxLags; ++q) {
        const int r = threadIdx.x + q * kDiagThreads;
        if (r < nx) partial[(long long)blockIdx.x * nx + r] = acc[q];
    }
}

// The same sums, lag-blocked (nx <= 64 K): the block's four wavefronts each
// take one line of a group of four; lane t owns the K lags r = tK .. tK+K-1
// and runs i = 0, K, 2K, .. with the K x K products of a step unrolled, a
// broadcast read of v(i + ii) and K + K reads of v(i + r) per step instead of
// two LDS reads per product.  Lines are zero-padded in LDS to nx + 2K, so
// the products past the line's end add exact zeros and need no masks.  The
// four wavefronts' sums are added in a fixed order.  Element e of a line
// sits at e + e / K (one unused slot after every K): the lanes' windows,
// K elements apart, are then K + 1 doubles apart, an odd stride that spreads
// a read of the 64 lanes over all LDS banks (K apart, they fell on a few).
template <int K> __device__ __forceinline__ int vcf_slot(int e) { return e + e / K; }
template <int K> constexpr int vcf_line_len(int nx) { return (nx + 2 * K) + (nx + 2 * K) / K + 1; }

template <int K>
__global__ __launch_bounds__(kDiagThreads) void k_vcf_lags(const double* __restrict__ vx, int nx, long long nlines,
                                                          double* __restrict__ partial) {
    extern __shared__ double lines[];  // 4 lines of nx + 2K, reused for the wave sums
    constexpr int NW = kDiagThreads / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lw = nx + 2 * K;
    double* L = lines + wv * vcf_line_len<K>(nx);
    const long long per = (nlines + gridDim.x - 1) / gridDim.x;
    const long long l0 = (long long)blockIdx.x * per, l1 = l0 + per < nlines ? l0 + per : nlines;
    const int r0 = lane * K;
    double acc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = 0.0;
    for (long long lg = l0; lg < l1; lg += NW) {
        const long long l = lg + wv;
        __syncthreads();
        for (int i = lane; i < lw; i += 64) L[vcf_slot<K>(i)] = (l < l1 && i < nx) ? vx[l * nx + i] : 0.0;
        __syncthreads();
        if (l < l1 && r0 < nx) {
            double b[2 * K];
#pragma unroll
            for (int j = 0; j < K; ++j) b[j] = L[vcf_slot<K>(r0) + j];
            for (int i0 = 0; i0 < nx - r0; i0 += K) {
                const int sb = vcf_slot<K>(i0 + r0 + K), sa = vcf_slot<K>(i0);  // i0, r0: multiples of K
#pragma unroll
                for (int j = 0; j < K; ++j) b[K + j] = L[sb + j];
#pragma unroll
                for (int ii = 0; ii < K; ++ii) {
                    const double a = L[sa + ii];
#pragma unroll
                    for (int q = 0; q < K; ++q) acc[q] += a * b[ii + q];
                }
#pragma unroll
                for (int j = 0; j < K; ++j) b[j] = b[K + j];
            }
        }
    }
    __syncthreads();
    double* red = lines;  // NW x nx
#pragma unroll
    for (int q = 0; q < K; ++q)
        if (r0 + q < nx) red[wv * nx + r0 + q] = acc[q];
    __syncthreads();
    for (int r = threadIdx.x; r < nx; r += kDiagThreads) {
        double sum = red[r];
#pragma unroll
        for (int w = 1; w < NW; ++w) sum += red[w * nx + r];
        partial[(long long)blockIdx.x * nx + r] = sum;
    }
}

// out[r] = sum over blocks (fixed order) of partial[b * nx + r].
__global__ __launch_bounds__(kDiagThreads) void k_vcf_final(const double* __restrict__ partial, int nx, int nb,
                                                           double* __restrict__ out) {
    const int r = blockIdx.x * kDiagThreads + threadIdx.x;
    if (r >= nx) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[(long long)b * nx + r];
    out[r] = s;
}

hipError_t launch_vel_hist(const double* vel, long long n, int nbins, const double* vmax3, unsigned long long* hist,
                           int nblocks, hipStream_t s) {
    const int ncopy = hist_copies(nbins);
    hipLaunchKernelGGL(k_vel_hist, dim3(nblocks), dim3(kDiagThreads),
                       (size_t)ncopy * (3 * nbins + 1) * sizeof(unsigned int), s, vel, n, nbins, ncopy, vmax3, hist);
    return hipGetLastError();
}

template <int K>
static void launch_vcf_lags(const double* vx, int nx, long long nlines, double* partial, int nblocks, hipStream_t s) {
    const size_t lds = (size_t)(kDiagThreads / 64) * vcf_line_len<K>(nx) * sizeof(double);
    hipLaunchKernelGGL((k_vcf_lags<K>), dim3(nblocks), dim3(kDiagThreads), lds, s, vx, nx, nlines, partial);
}

hipError_t launch_vcf(const double* vx, int nx, long long nlines, double* partial, int nblocks, double* out,
                      hipStream_t s) {
    if (nx <= 64) launch_vcf_lags<1>(vx, nx, nlines, partial, nblocks, s);
    else if (nx <= 128) launch_vcf_lags<2>(vx, nx, nlines, partial, nblocks, s);
    else if (nx <= 256) launch_vcf_lags<4>(vx, nx, nlines, partial, nblocks, s);
    else if (nx <= 512) launch_vcf_lags<8>(vx, nx, nlines, partial, nblocks, s);
    else if (nx <= 1024) launch_vcf_lags<16>(vx, nx, nlines, partial, nblocks, s);
    else
        hipLaunchKernelGGL(k_vcf_partial, dim3(nblocks), dim3(kDiagThreads), (size_t)nx * sizeof(double), s, vx, nx,
                           nlines, partial);
    hipLaunchKernelGGL(k_vcf_final, dim3((nx + kDiagThreads - 1) / kDiagThreads), dim3(kDiagThreads), 0, s, partial, nx,
                       nblocks, out);
    return hipGetLastError();
}

}  // namespace gpe
