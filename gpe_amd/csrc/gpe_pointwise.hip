// Pointwise kernels: Cash-Karp retry stage, norm/mass partial sums,
// renormalisation, precision conversion, final partial reduction.
#include "gpe_launch.h"

namespace gpe {

static constexpr int kPwThreads = 256;

static inline int pw_blocks(long long n) {
    long long b = (n + kPwThreads - 1) / kPwThreads;
    if (b > 65536) b = 65536;
    return b < 1 ? 1 : (int)b;
}

// Retry of a rejected Cash-Karp attempt: f(psi0) is unchanged, so stage 1
// reduces to u2 = psi0 + b21 (dt f0).
template <typename R>
__global__ __launch_bounds__(kPwThreads) void k_ck1_retry(const C<R>* __restrict__ p0, const C<R>* __restrict__ f0,
                                                          C<R>* __restrict__ unext, C<R> dt, R b21, long long n) {
    for (long long i = blockIdx.x * (long long)kPwThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kPwThreads) {
        const C<R> k1 = cmul<R>(dt, f0[i]);
        unext[i] = caxpy<R>(b21, k1, p0[i]);
    }
}

template <typename R>
hipError_t launch_ck1_retry(const C<R>* p0, const C<R>* f0, C<R>* unext, C<R> dt, R b21, long long n, hipStream_t s) {
    hipLaunchKernelGGL((k_ck1_retry<R>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, s, p0, f0, unext, dt, b21, n);
    return hipGetLastError();
}

// End-corrected quadrature weight of index i on an axis of n points
// (src/variables.f90:49-51, 713-823): 3/8, 7/6, 23/24 at both ends.
__device__ __forceinline__ double qweight(int i, int n) {
    const int e = i < n - 1 - i ? i : n - 1 - i;
    return e == 0 ? 3.0 / 8.0 : e == 1 ? 7.0 / 6.0 : e == 2 ? 23.0 / 24.0 : 1.0;
}

// Partial sums of w_i w_j w_k |psi|^2 (weighted = 1, get_norm) or |psi|^2
// (weighted = 0, mass) in FP64.  Each thread walks whole x rows so the
// summation order is fixed by the grid.
template <typename R>
__global__ __launch_bounds__(kPwThreads) void k_norm_partial(const C<R>* __restrict__ u, int nx, int ny, int nz,
                                                             int weighted, double* __restrict__ partial) {
    __shared__ double sh[kPwThreads / 64];
    const long long nrow = (long long)ny * nz;
    double acc = 0.0;
    // rows are split over blocks; x is split over the threads of a block
    for (long long r = blockIdx.x; r < nrow; r += gridDim.x) {
        const int j = (int)(r % ny), k = (int)(r / ny);
        const double wjk = weighted ? qweight(j, ny) * qweight(k, nz) : 1.0;
        double rs = 0.0;
        for (int i = threadIdx.x; i < nx; i += kPwThreads) {
            const C<R> v = u[r * nx + i];
            const double a2 = (double)v.x * v.x + (double)v.y * v.y;
            rs += weighted ? qweight(i, nx) * a2 : a2;
        }
        acc += wjk * rs;
    }
    const double v = block_reduce<kPwThreads, RED_SUM>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

template <typename R>
hipError_t launch_norm_partial(const C<R>* u, int nx, int ny, int nz, int weighted, double* partial, int nblocks,
                               hipStream_t s) {
    hipLaunchKernelGGL((k_norm_partial<R>), dim3(nblocks), dim3(kPwThreads), 0, s, u, nx, ny, nz, weighted, partial);
    return hipGetLastError();
}

// psi <- sqrt(nn) psi / sqrt(norm)  (src/variables.f90:698-709)
template <typename R>
__global__ __launch_bounds__(kPwThreads) void k_scale(C<R>* __restrict__ u, long long n, double snn, double sn) {
    for (long long i = blockIdx.x * (long long)kPwThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kPwThreads) {
        const C<R> v = u[i];
        u[i] = mkc<R>((R)((snn * (double)v.x) / sn), (R)((snn * (double)v.y) / sn));
    }
}

template <typename R> hipError_t launch_scale(C<R>* u, long long n, double snn, double sn, hipStream_t s) {
    hipLaunchKernelGGL((k_scale<R>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, s, u, n, snn, sn);
    return hipGetLastError();
}

// psi <- |psi| vortex_line(vl) (imprint_vortex_line, src/variables.f90:1148-1162):
// the line's phase exp(i theta), times amp(r) = 1 - exp(-0.7 r^1.15) unless
// only the phase is imprinted, with theta and r of get_theta / get_r
// (src/ic.f90:730-884) at the grid point x(i) = xl + i dx (get_grid).
template <typename R>
__global__ __launch_bounds__(kPwThreads) void k_imprint_line(C<R>* __restrict__ u, int nx, int ny, long long n,
                                                             double xl, double yl, double zl, double dx, double dy,
                                                             double dz, LineParam vl) {
    constexpr double kPi = 3.1415926535897932384626433832795;
    for (long long q = blockIdx.x * (long long)kPwThreads + threadIdx.x; q < n; q += (long long)gridDim.x * kPwThreads) {
        const int i = (int)(q % nx);
        const long long r = q / nx;
        const int j = (int)(r % ny), k = (int)(r / ny);
        const double x = xl + (double)i * dx, y = yl + (double)j * dy, z = zl + (double)k * dz;
        // (t, t0) along the line; (a, a0) displaced by amp1 sin, (b, b0) by amp2 cos
        double t, t0, a, a0, b, b0;
        if (vl.dir == 0) {
            t = x, t0 = vl.x0, a = y, a0 = vl.y0, b = z, b0 = vl.z0;
        } else if (vl.dir == 1) {
            t = y, t0 = vl.y0, a = z, a0 = vl.z0, b = x, b0 = vl.x0;
        } else {
            t = z, t0 = vl.z0, a = x, a0 = vl.x0, b = y, b0 = vl.y0;
        }
        const double da = a - a0 - vl.amp1 * sin(2.0 * kPi * (t - t0) / (vl.ll + vl.safe1));
        const double db = b - b0 - vl.amp2 * cos(2.0 * kPi * (t - t0) / (vl.ll + vl.safe2));
        const double theta = vl.sgn * atan2(vl.scal * db, vl.scal * da);
        double amp = 1.0;
        if (!vl.imprint_phase) {
            const double rr = sqrt(vl.scal * da * da + vl.scal * db * db);
            amp = 1.0 - exp(-0.7 * pow(rr, 1.15));
        }
        const C<R> v = u[q];
        const double m = hypot((double)v.x, (double)v.y) * amp;
        u[q] = mkc<R>((R)(m * cos(theta)), (R)(m * sin(theta)));
    }
}

template <typename R>
hipError_t launch_imprint_line(C<R>* u, int nx, int ny, long long n, const double lo[3], const double d[3],
                               const LineParam& vl, hipStream_t s) {
    hipLaunchKernelGGL((k_imprint_line<R>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, s, u, nx, ny, n, lo[0], lo[1],
                       lo[2], d[0], d[1], d[2], vl);
    return hipGetLastError();
}

template <typename A, typename B>
__global__ __launch_bounds__(kPwThreads) void k_convert(const A* __restrict__ src, B* __restrict__ dst, long long n) {
    for (long long i = blockIdx.x * (long long)kPwThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kPwThreads) {
        const A v = src[i];
        B o;
        o.x = v.x;
        o.y = v.y;
        dst[i] = o;
    }
}

template <typename R> hipError_t launch_from_c128(const double2* src, C<R>* dst, long long n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<double2, C<R>>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, s, src, dst, n);
    return hipGetLastError();
}
template <typename R> hipError_t launch_from_c64(const float2* src, C<R>* dst, long long n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<float2, C<R>>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, s, src, dst, n);
    return hipGetLastError();
}
template <typename R> hipError_t launch_to_c128(const C<R>* src, double2* dst, long long n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<C<R>, double2>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, s, src, dst, n);
    return hipGetLastError();
}
template <typename R> hipError_t launch_to_c64(const C<R>* src, float2* dst, long long n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<C<R>, float2>), dim3(pw_blocks(n)), dim3(kPwThreads), 0, s, src, dst, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(kPwThreads) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                       long long n) {
    for (long long i = blockIdx.x * (long long)kPwThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kPwThreads)
        dst[i] = src[i];
}

// streaming variant: 3 reads + 2 writes per element (the pointwise traffic
// pattern of an RK4 middle stage without the stencil)
__global__ __launch_bounds__(kPwThreads) void k_stream5(const uint4* __restrict__ a, const uint4* __restrict__ b,
                                                        uint4* __restrict__ c, uint4* __restrict__ d,
                                                        const uint4* __restrict__ e, long long n) {
    for (long long i = blockIdx.x * (long long)kPwThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kPwThreads) {
        const uint4 x = a[i], y = b[i], z = e[i];
        c[i] = make_uint4(x.x ^ y.x, x.y ^ y.y, x.z ^ z.z, x.w ^ z.w);
        d[i] = make_uint4(y.x ^ z.x, y.y ^ z.y, x.z, y.w);
    }
}

hipError_t launch_stream5(void* c, void* d, const void* a, const void* b, const void* e, long long n16, hipStream_t s) {
    hipLaunchKernelGGL(k_stream5, dim3(pw_blocks(n16)), dim3(kPwThreads), 0, s, (const uint4*)a, (const uint4*)b,
                       (uint4*)c, (uint4*)d, (const uint4*)e, n16);
    return hipGetLastError();
}

hipError_t launch_copy16(void* dst, const void* src, long long n16, hipStream_t s) {
    hipLaunchKernelGGL(k_copy16, dim3(pw_blocks(n16)), dim3(kPwThreads), 0, s, (const uint4*)src, (uint4*)dst, n16);
    return hipGetLastError();
}

hipError_t launch_reduce_final(int mode, const double* partial, int n, int nval, double* out, hipStream_t s) {
    if (mode == RED_SUM)
        hipLaunchKernelGGL((k_reduce_final<RED_SUM>), dim3(1), dim3(1024), 0, s, partial, n, nval, out);
    else
        hipLaunchKernelGGL((k_reduce_final<RED_MAX>), dim3(1), dim3(1024), 0, s, partial, n, nval, out);
    return hipGetLastError();
}

#define GPE_PW_INST(R)                                                                                             \
    template hipError_t launch_ck1_retry<R>(const C<R>*, const C<R>*, C<R>*, C<R>, R, long long, hipStream_t);     \
    template hipError_t launch_norm_partial<R>(const C<R>*, int, int, int, int, double*, int, hipStream_t);       \
    template hipError_t launch_scale<R>(C<R>*, long long, double, double, hipStream_t);                           \
    template hipError_t launch_imprint_line<R>(C<R>*, int, int, long long, const double*, const double*,         \
                                               const LineParam&, hipStream_t);                                     \
    template hipError_t launch_from_c128<R>(const double2*, C<R>*, long long, hipStream_t);                       \
    template hipError_t launch_from_c64<R>(const float2*, C<R>*, long long, hipStream_t);                         \
    template hipError_t launch_to_c128<R>(const C<R>*, double2*, long long, hipStream_t);                         \
    template hipError_t launch_to_c64<R>(const C<R>*, float2*, long long, hipStream_t);
GPE_PW_INST(double)
GPE_PW_INST(float)

}  // namespace gpe
