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
// LDS word are serialised.  EDGES: the block first tabulates the nbins + 1
// edges of each component in LDS with pdf_edge, so a comparison with an edge
// is one LDS read (the same edges, so the same bins); else above_edge.
constexpr int kHistMaxCopies = 8;
constexpr int kHistUnroll = 4;
constexpr int kHistLdsBytes = 64 * 1024;
constexpr int kHistEdgeBytes = 32 * 1024;  // the edge table is used up to this size
inline int hist_copies(int nbins) {
    const int c = kHistLdsBytes / ((3 * nbins + 1) * (int)sizeof(unsigned int));
    return c < 1 ? 1 : (c > kHistMaxCopies ? kHistMaxCopies : c);
}
inline size_t hist_edge_bytes(int nbins) { return (size_t)3 * (nbins + 1) * sizeof(double); }
// 32-bit words of bins, rounded up to an even count so the edges that follow
// are 8-byte aligned
inline int hist_bin_words(int nbins, int ncopy) { return (ncopy * (3 * nbins + 1) + 1) & ~1; }

template <bool EDGES>
__global__ __launch_bounds__(kDiagThreads) void k_vel_hist(const double* __restrict__ vel, long long n, int nbins,
                                                          int ncopy, const double* __restrict__ vmax3,
                                                          unsigned long long* __restrict__ hist) {
    extern __shared__ unsigned int hsh[];  // ncopy * (3 * nbins + 1), then (EDGES) 3 * (nbins + 1) doubles
    const int stride = 3 * nbins + 1;
    for (int j = threadIdx.x; j < ncopy * stride; j += kDiagThreads) hsh[j] = 0u;
    const int lo = -nbins / 2 + 1, hi = nbins / 2;
    double* edges = reinterpret_cast<double*>(hsh + ((ncopy * stride + 1) & ~1));
    if constexpr (EDGES) {
        // edges[a * (nbins + 1) + (i - lo + 1)] = pdf_edge(i) for i = lo - 1 .. hi
        for (int j = threadIdx.x; j < 3 * (nbins + 1); j += kDiagThreads) {
            const int a = j / (nbins + 1), i = j - a * (nbins + 1) + lo - 1;
            edges[j] = pdf_edge(i, vmax3[a], nbins);
        }
    }
    __syncthreads();
    unsigned int* mine = hsh + (threadIdx.x % ncopy) * stride;
    const double inv_nb = 1.0 / (double)nbins;
    for (int a = 0; a < 3; ++a) {
        const double* v = vel + a * n;
        const double vmax = vmax3[a];
        if (!(vmax > 0.0)) continue;  // uniform
        const double scale = (double)nbins / (2.0 * vmax);
        const double* E = edges + a * (nbins + 1) - (lo - 1);  // E[i] = pdf_edge(i), i = lo - 1 .. hi
        auto above = [&](double x, int i) {
            if constexpr (EDGES) return x > E[i];
            else return above_edge(x, i, vmax, nbins, inv_nb);
        };
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
                while (b <= hi && above(x, b)) ++b;
                while (b >= lo && !above(x, b - 1)) --b;
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

// The same sums, lag-blocked (nx <= 64 K).  The lags fall in 64 blocks of
// K, r0 = tK .. tK+K-1 for t = 0 .. 63; a step of block t takes i = i0 ..
// i0+K-1 (i0 a multiple of K) and does its K x K products unrolled, from K
// reads of v(i) and 2K reads of v(i + r) instead of two LDS reads per
// product.  Block t needs ceil((nx - tK) / K) steps, so blocks t and 63 - t
// together need about the same number for every t (65 at nx = 64 K).  Each
// wavefront takes two lines at a time, lanes 0-31 the first and 32-63 the
// second, and lane u of a half runs block u and then block 63 - u in one
// loop of uniform length: no lane idles while the short lags finish.  Lines
// are zero-padded in LDS to nx + 2K, so the products past the line's end
// add exact zeros and need no masks.  The eight (wavefront, half) sums are
// added in a fixed order.  Element e of a line sits at e + e / K (one unused
// slot after every K): windows K elements apart are then K + 1 doubles
// apart, an odd stride that spreads a read over all LDS banks.
template <int K> __device__ __forceinline__ int vcf_slot(int e) { return e + e / K; }
template <int K> constexpr int vcf_line_len(int nx) { return (nx + 2 * K) + (nx + 2 * K) / K + 1; }

template <int K>
__global__ __launch_bounds__(kDiagThreads) void k_vcf_lags(const double* __restrict__ vx, int nx, long long nlines,
                                                          double* __restrict__ partial) {
    extern __shared__ double lines[];  // 2 lines per wavefront, then the lag sums
    constexpr int NW = kDiagThreads / 64, NL = 2 * NW;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, half = lane >> 5, u = lane & 31;
    const int lw = nx + 2 * K, ll = vcf_line_len<K>(nx);
    const double* L = lines + (2 * wv + half) * ll;
    const long long per = (nlines + gridDim.x - 1) / gridDim.x;
    const long long l0 = (long long)blockIdx.x * per, l1 = l0 + per < nlines ? l0 + per : nlines;
    auto nsteps = [&](int t) { const int r = t * K; return r < nx ? (nx - r + K - 1) / K : 0; };
    const int b1 = u, b2 = 63 - u, n1 = nsteps(b1), n2 = nsteps(b2), nt = n1 + n2;
    int tmax = nt;  // the longest lane of the wavefront
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int v = __shfl_xor(tmax, o, 64);
        tmax = v > tmax ? v : tmax;
    }
    double acc1[K], acc2[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc1[q] = acc2[q] = 0.0;
    for (long long lg = l0; lg < l1; lg += NL) {
        __syncthreads();
        // the wavefront stages its two lines
        for (int i = lane; i < 2 * lw; i += 64) {
            const int h = i >= lw, e = i - h * lw;
            const long long l = lg + 2 * wv + h;
            lines[(2 * wv + h) * ll + vcf_slot<K>(e)] = (l < l1 && e < nx) ? vx[l * nx + e] : 0.0;
        }
        __syncthreads();
        if (lg + 2 * wv + half < l1) {
            double cur[K];
#pragma unroll
            for (int q = 0; q < K; ++q) cur[q] = 0.0;
            for (int st = 0; st < tmax; ++st) {
                if (st < nt) {
                    const bool first = st < n1;
                    const int r0 = (first ? b1 : b2) * K, i0 = (first ? st : st - n1) * K;
                    const int sa = vcf_slot<K>(i0), sb = vcf_slot<K>(i0 + r0), sc = vcf_slot<K>(i0 + r0 + K);
                    double b[2 * K];
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        b[j] = L[sb + j];
                        b[K + j] = L[sc + j];
                    }
#pragma unroll
                    for (int ii = 0; ii < K; ++ii) {
                        const double a = L[sa + ii];
#pragma unroll
                        for (int q = 0; q < K; ++q) cur[q] += a * b[ii + q];
                    }
                    if (st == n1 - 1) {  // the first block is done
#pragma unroll
                        for (int q = 0; q < K; ++q) {
                            acc1[q] += cur[q];
                            cur[q] = 0.0;
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < K; ++q) acc2[q] += cur[q];
        }
    }
    __syncthreads();
    double* red = lines;  // NL x nx: the lag sums of each (wavefront, half)
    const int row = 2 * wv + half;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        if (b1 * K + q < nx) red[row * nx + b1 * K + q] = acc1[q];
        if (b2 * K + q < nx) red[row * nx + b2 * K + q] = acc2[q];
    }
    __syncthreads();
    for (int r = threadIdx.x; r < nx; r += kDiagThreads) {
        double sum = red[r];
#pragma unroll
        for (int w = 1; w < NL; ++w) sum += red[w * nx + r];
        partial[(long long)blockIdx.x * nx + r] = sum;
    }
}

// out[r] = sum over blocks (fixed order) of partial[b * nx + r].
__global__ __launch_bounds__(kDiagThreads) void k_vcf_final(const double* __restrict__ partial, int nx, int nb,
                                                           double* __restrict__ out) {
    const int r = blockIdx.x * kDiagThreads + threadIdx.x;
    if (r >= nx) return;
    // groups of 8 partials loaded together, added in block order
    double s = 0.0;
    int b = 0;
    for (; b + 8 <= nb; b += 8) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = partial[(long long)(b + q) * nx + r];
#pragma unroll
        for (int q = 0; q < 8; ++q) s += v[q];
    }
    for (; b < nb; ++b) s += partial[(long long)b * nx + r];
    out[r] = s;
}

hipError_t launch_vel_hist(const double* vel, long long n, int nbins, const double* vmax3, unsigned long long* hist,
                           int nblocks, hipStream_t s) {
    const int ncopy = hist_copies(nbins);
    const size_t bins = (size_t)hist_bin_words(nbins, ncopy) * sizeof(unsigned int);
    if (hist_edge_bytes(nbins) <= (size_t)kHistEdgeBytes)
        hipLaunchKernelGGL(k_vel_hist<true>, dim3(nblocks), dim3(kDiagThreads), bins + hist_edge_bytes(nbins), s, vel,
                           n, nbins, ncopy, vmax3, hist);
    else
        hipLaunchKernelGGL(k_vel_hist<false>, dim3(nblocks), dim3(kDiagThreads), bins, s, vel, n, nbins, ncopy,
                           vmax3, hist);
    return hipGetLastError();
}

template <int K>
static void launch_vcf_lags(const double* vx, int nx, long long nlines, double* partial, int nblocks, hipStream_t s) {
    const size_t lds = (size_t)(2 * (kDiagThreads / 64)) * vcf_line_len<K>(nx) * sizeof(double);
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
