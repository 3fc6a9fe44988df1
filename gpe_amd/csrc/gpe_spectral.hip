// Spectral diagnostics of the state, after its normalised backward FFT
// a = FFT+(psi) / sqrt(nx ny nz) (`fft(in_var, a, 'backward')`,
// src/ic.f90:905-1001, done by hipFFT in the caller):
//   * spectrum (src/io.f90:632-725): per log2 shell m = 1..7 of
//     log2k = log2(sqrt(k2) / (2 pi)) (|log2k| in [m-1, m)) the sum of |a|^2
//     and the mode count, and the integral distribution
//     F(kk) = sum over modes with sqrt(k2) <= kk of |a|^2, kk = 1..nx/2;
//   * condensed_particles (src/io.f90:315-404): the k-grid sum
//     tot = sum (k2 + rho0) / (k2^2 + 2 rho0 k2) over the 4th-order symbol k2
//     of kx = -pi + i dk, dk = 2 pi / (nx - 1) (the same dk on every axis).
// k2 of a mode index follows random_phase / spectrum: i^2 for
// i <= (n-1)/2 + 1, else (n - i)^2.  Sums are FP64: per-thread, then per
// block (fixed order), then over blocks in a fixed order; the F bins
// accumulate in LDS with atomics, so their last bits depend on timing.
#include <hip/hip_runtime.h>

#include "gpe_launch.h"

namespace gpe {

constexpr int kSpThreads = 256;
constexpr double kSpPi = 3.1415926535897932384626433832795;

__device__ __forceinline__ long long mode_sq(int i, int n) {
    const int n1 = n - 1;
    const long long m = i <= n1 / 2 + 1 ? i : n1 - i + 1;
    return m * m;
}

// tab[k2] for k2 = 0 .. max: the F bin s (the smallest kk with
// sqrt(k2) <= kk; 0 when k2 = 0 or s > nhalf) and, in bits 24..31, the log2
// shell m with |log2(sqrt(k2) / (2 pi))| in [m, m + 1) (kSpectrumShells when
// none), with the arithmetic of the spectrum's per-mode sqrt, search and
// log.  k_spectrum_partial reads the shells from it.
__global__ __launch_bounds__(kSpThreads) void k_spectrum_table(long long len, int nhalf, unsigned* __restrict__ tab) {
    const long long k2 = (long long)blockIdx.x * kSpThreads + threadIdx.x;
    if (k2 >= len) return;
    unsigned sbin = 0, m = kSpectrumShells;
    if (k2 > 0) {
        long long s = (long long)sqrt((double)k2);
        while (s * s < k2) ++s;
        while (s > 1 && (s - 1) * (s - 1) >= k2) --s;
        sbin = s <= nhalf ? (unsigned)s : 0u;
        const double lk = fabs(log(0.5 * sqrt((double)k2) / kSpPi) / log(2.0));
        for (int q = kSpectrumShells - 1; q >= 0; --q)
            if (lk < (double)(q + 1) && lk >= (double)q) m = (unsigned)q;
    }
    tab[k2] = sbin | (m << 24);
}

// Per block b, rows of partial[row * nb + b]: 0..6 the shell sums of |a|^2,
// 7..13 the shell counts, 14 + (s - 1) for s = 1..nhalf the sum of |a|^2 over
// the modes whose smallest kk with k2 <= kk^2 is s.
//
// A mode's s comes from a float square root settled with integer compares
// (exact), and its shell from a per-block LDS table over s built from tab,
// so no mode reads tab[k2] (a scattered read per mode).  The k2 of one s
// fill ((s-1)^2, s^2]; the shell boundaries sqrt(k2) = 2 pi 2^m (m = -3 .. 7) are more than one
// apart above 1.6 and the range of s = 1 and 2 holds at most one of them, so
// each range holds at most one change of shell: the table keeps the shells
// at its two ends and the first k2 of the second, found by bisection of tab.
// Entry s: bits 0..31 that k2 (past the range: none), 32..39 and 40..47 the
// shells below and from it.
template <typename R>
__global__ __launch_bounds__(kSpThreads) void k_spectrum_partial(const C<R>* __restrict__ a, int nx, int ny, int nz,
                                                                 int nhalf, double inv_n, long long len, int smax,
                                                                 const unsigned* __restrict__ tab,
                                                                 double* __restrict__ partial) {
    extern __shared__ double hsh[];  // nhalf bins, kSpThreads / 64 reduction slots, then smax + 1 entries
    double* red = hsh + nhalf;
    unsigned long long* stab = reinterpret_cast<unsigned long long*>(red + kSpThreads / 64);
    for (int s = threadIdx.x; s < nhalf; s += kSpThreads) hsh[s] = 0.0;
    for (int s = threadIdx.x; s <= smax; s += kSpThreads) {
        const long long lo = (long long)(s - 1) * (s - 1) + 1;
        const long long hi = (long long)s * s < len - 1 ? (long long)s * s : len - 1;
        unsigned long long e = 0;
        if (s >= 1 && lo <= hi) {
            const unsigned sa = tab[lo] >> 24, sb = tab[hi] >> 24;
            long long thr = hi + 1;
            if (sa != sb) {  // tab[l] has shell sa, tab[h] not
                long long l = lo, h = hi;
                while (h - l > 1) {
                    const long long mid = l + (h - l) / 2;
                    if ((tab[mid] >> 24) == sa) l = mid;
                    else h = mid;
                }
                thr = h;
            }
            e = (unsigned long long)(unsigned)thr | ((unsigned long long)sa << 32) | ((unsigned long long)sb << 40);
        }
        stab[s] = e;
    }
    __syncthreads();
    double e[kSpectrumShells];
    unsigned cnt[kSpectrumShells];  // exact counts (a thread sees under 2^32 modes), summed as doubles per block
#pragma unroll
    for (int m = 0; m < kSpectrumShells; ++m) {
        e[m] = 0.0;
        cnt[m] = 0u;
    }
    // rows (j, k) block by block, x across the threads: the row's indices
    // are formed once per row, not divided out of a linear index per mode
    const long long nrow = (long long)ny * nz;
    for (long long row = blockIdx.x; row < nrow; row += gridDim.x) {
      const int j = (int)(row % ny), k = (int)(row / ny);
      const long long kjk = mode_sq(j, ny) + mode_sq(k, nz);
      const C<R>* __restrict__ ar = a + row * nx;
      for (int i = threadIdx.x; i < nx; i += kSpThreads) {
        const long long k2 = mode_sq(i, nx) + kjk;
        if (k2 == 0) continue;
        const C<R> v = ar[i];
        const double p2 = ((double)v.x * (double)v.x + (double)v.y * (double)v.y) * inv_n;
        // the float root is within a few units of sqrt(k2) (exactly rounded
        // below 2^24); the compares settle the smallest sq with sq^2 >= k2
        int sq = (int)sqrtf((float)k2);
        while ((long long)sq * sq < k2) ++sq;
        while (sq > 1 && (long long)(sq - 1) * (sq - 1) >= k2) --sq;
        const unsigned long long te = stab[sq];
        const unsigned sbin = sq <= nhalf ? (unsigned)sq : 0u;
        const unsigned sh = (unsigned)k2 < (unsigned)te ? (unsigned)(te >> 32) & 0xFFu : (unsigned)(te >> 40) & 0xFFu;
        if (sbin) atomicAdd(&hsh[sbin - 1], p2);
#pragma unroll
        for (int m = 0; m < kSpectrumShells; ++m) {
            if (sh == (unsigned)m) {
                e[m] += p2;
                cnt[m] += 1u;
            }
        }
      }
    }
    const int nb = gridDim.x;
#pragma unroll
    for (int m = 0; m < kSpectrumShells; ++m) {
        const double se = block_reduce<kSpThreads, RED_SUM>(e[m], red);
        const double sc = block_reduce<kSpThreads, RED_SUM>((double)cnt[m], red);
        if (threadIdx.x == 0) {
            partial[(long long)m * nb + blockIdx.x] = se;
            partial[(long long)(kSpectrumShells + m) * nb + blockIdx.x] = sc;
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nhalf; s += kSpThreads)
        partial[(long long)(2 * kSpectrumShells + s) * nb + blockIdx.x] = hsh[s];
}

// -(1/12) (-2 cos(2 k d) + 32 cos(k d) - 30) / d^2 for one axis (the
// 4th-order symbol of condensed_particles, src/io.f90:357-365)
__device__ __forceinline__ double sym4(double kk, double d, double d2) {
    return (-2.0 * cos(2.0 * kk * d) + 32.0 * cos(kk * d) - 30.0) / d2;
}

// sym[0..nx) = the x terms, then ny y terms, then nz z terms
__global__ __launch_bounds__(kSpThreads) void k_tot_axes(int nx, int ny, int nz, double dk, double dx, double dy,
                                                         double dz, double* __restrict__ sym) {
    const int q = blockIdx.x * kSpThreads + threadIdx.x;
    if (q >= nx + ny + nz) return;
    const int i = q < nx ? q : (q < nx + ny ? q - nx : q - nx - ny);
    const double d = q < nx ? dx : (q < nx + ny ? dy : dz);
    sym[q] = sym4(-kSpPi + (double)i * dk, d, d * d);
}

__global__ __launch_bounds__(kSpThreads) void k_tot_partial(int nx, int ny, int nz, const double* __restrict__ sym,
                                                            double rho0, double* __restrict__ partial) {
    __shared__ double red[kSpThreads / 64];
    const double* sx = sym;
    const double* sy = sym + nx;
    const double* sz = sym + nx + ny;
    double acc = 0.0;
    const long long nrow = (long long)ny * nz;
    for (long long row = blockIdx.x; row < nrow; row += gridDim.x) {
        const int j = (int)(row % ny), k = (int)(row / ny);
        const double yj = sy[j], zk = sz[k];
        for (int i = threadIdx.x; i < nx; i += kSpThreads) {
            const double k2 = -(1.0 / 12.0) * (sx[i] + yj + zk);
            if (k2 == 0.0) continue;
            const double k4 = k2 * k2;
            acc += (k2 + rho0) / (k4 + 2.0 * rho0 * k2);
        }
    }
    const double s = block_reduce<kSpThreads, RED_SUM>(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[row] = sum over blocks b (in order) of partial[row * nb + b]
__global__ __launch_bounds__(kSpThreads) void k_sum_rows(const double* __restrict__ partial, int nrows, int nb,
                                                         double* __restrict__ out) {
    const int row = blockIdx.x * kSpThreads + threadIdx.x;
    if (row >= nrows) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[(long long)row * nb + b];
    out[row] = s;
}

long long spectrum_table_len(int nx, int ny, int nz) {
    auto mx = [](int n) {
        const int n1 = n - 1, h = n1 / 2 + 1 < n1 ? n1 / 2 + 1 : n1;
        return (long long)h * h;
    };
    return mx(nx) + mx(ny) + mx(nz) + 1;
}

template <typename R>
hipError_t launch_spectrum(const C<R>* a, int nx, int ny, int nz, int nhalf, double* partial, int nb, double* out,
                           unsigned* tab, hipStream_t s) {
    const double inv_n = 1.0 / ((double)nx * (double)ny * (double)nz);
    const size_t lds = ((size_t)nhalf + kSpThreads / 64) * sizeof(double);
    const long long len = spectrum_table_len(nx, ny, nz);
    hipLaunchKernelGGL(k_spectrum_table, dim3((unsigned)((len + kSpThreads - 1) / kSpThreads)), dim3(kSpThreads), 0, s,
                       len, nhalf, tab);
    // the s table: s = 1 .. smax, smax^2 >= the largest k2 (len - 1)
    int smax = 1;
    while ((long long)smax * smax < len - 1) ++smax;
    // entries hold k2 in 32 bits; 64 KiB of entries reach past 9000^3 grids
    const size_t stab = (size_t)(smax + 1) * sizeof(unsigned long long);
    if (len >= (1LL << 32) || stab > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_spectrum_partial<R>), dim3(nb), dim3(kSpThreads), lds + stab, s, a, nx, ny, nz, nhalf, inv_n,
                       len, smax, (const unsigned*)tab, partial);
    const int nrows = 2 * kSpectrumShells + nhalf;
    hipLaunchKernelGGL(k_sum_rows, dim3((nrows + kSpThreads - 1) / kSpThreads), dim3(kSpThreads), 0, s, partial, nrows,
                       nb, out);
    return hipGetLastError();
}

hipError_t launch_condensed_tot(int nx, int ny, int nz, const double d[3], double rho0, double* partial, int nb,
                                double* out, hipStream_t s) {
    const double dk = 2.0 * kSpPi / (double)(nx - 1);
    double* sym = partial + nb;  // nx + ny + nz values after the block partials
    hipLaunchKernelGGL(k_tot_axes, dim3((nx + ny + nz + kSpThreads - 1) / kSpThreads), dim3(kSpThreads), 0, s, nx, ny,
                       nz, dk, d[0], d[1], d[2], sym);
    hipLaunchKernelGGL(k_tot_partial, dim3(nb), dim3(kSpThreads), 0, s, nx, ny, nz, (const double*)sym, rho0,
                       partial);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(kSpThreads), 0, s, partial, 1, nb, out);
    return hipGetLastError();
}

template hipError_t launch_spectrum<double>(const double2*, int, int, int, int, double*, int, double*, unsigned*,
                                            hipStream_t);
template hipError_t launch_spectrum<float>(const float2*, int, int, int, int, double*, int, double*, unsigned*,
                                           hipStream_t);

}  // namespace gpe
