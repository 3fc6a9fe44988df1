// FP32 multilayer perceptron for the learned RKF45 step-size proposer:
// n_in -> 64 -> 64 -> 1 with tanh hidden layers, trained on the GPU with the
// mean squared error and Adam.  No reference counterpart.
//
// Layout of the parameter vector (floats): W1[64][n_in], b1[64],
// W2[64][64], b2[64], W3[64], b3[1].
//
// Kernels map one wavefront to one sample and lane j to hidden unit j of
// both layers (64-wide wavefronts, 64 hidden units).  Gradients are
// accumulated per lane over the samples of a block, reduced across the
// block's waves in LDS, written as per-block partials and summed in a fixed
// order, so training is deterministic.
#include <math.h>

#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/gpe_hip.h"

namespace {

constexpr int H = 64;         // hidden width (= wavefront size)
constexpr int MAX_IN = 16;
constexpr int WAVES = 4;      // samples in flight per block
constexpr int NTHREADS = 64 * WAVES;

__host__ __device__ inline int nparams(int nin) { return H * nin + H + H * H + H + H + 1; }

struct Off {
    int w1, b1, w2, b2, w3, b3, n;
};
__host__ __device__ inline Off offsets(int nin) {
    Off o;
    o.w1 = 0;
    o.b1 = H * nin;
    o.w2 = o.b1 + H;
    o.b2 = o.w2 + H * H;
    o.w3 = o.b2 + H;
    o.b3 = o.w3 + H;
    o.n = o.b3 + 1;
    return o;
}

__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Forward pass of sample s for the calling wave; returns the output (all
// lanes) and leaves h1 (LDS) and h2 (register) for the backward pass.
__device__ inline float forward_one(const float* P, const Off& o, int nin, const float* x, float* h1s, float& h2,
                                    int lane) {
    float a1 = P[o.b1 + lane];
    for (int i = 0; i < nin; ++i) a1 += P[o.w1 + lane * nin + i] * x[i];
    const float h1 = tanhf(a1);
    h1s[lane] = h1;
    __syncthreads();
    float a2 = P[o.b2 + lane];
#pragma unroll 8
    for (int k = 0; k < H; ++k) a2 += P[o.w2 + lane * H + k] * h1s[k];
    h2 = tanhf(a2);
    return wave_sum(P[o.w3 + lane] * h2) + P[o.b3];
}

__global__ __launch_bounds__(NTHREADS) void k_mlp_forward(const float* __restrict__ Pg, int nin,
                                                          const float* __restrict__ X, int n, float* __restrict__ Y) {
    extern __shared__ float sh[];
    const Off o = offsets(nin);
    float* P = sh;                       // parameters
    float* xs = P + o.n;                 // WAVES * MAX_IN
    float* h1s = xs + WAVES * MAX_IN;    // WAVES * H
    for (int i = threadIdx.x; i < o.n; i += NTHREADS) P[i] = Pg[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rounds = (n + gridDim.x * WAVES - 1) / (gridDim.x * WAVES);
    for (int r = 0; r < rounds; ++r) {
        const int s = (r * gridDim.x + blockIdx.x) * WAVES + wave;
        const bool live = s < n;
        if (lane < nin) xs[wave * MAX_IN + lane] = live ? X[(long long)s * nin + lane] : 0.f;
        __syncthreads();
        float h2;
        const float y = forward_one(P, o, nin, xs + wave * MAX_IN, h1s + wave * H, h2, lane);
        if (live && lane == 0) Y[s] = y;
        __syncthreads();
    }
}

// Loss and gradient partials of the samples [0, n) (mean squared error).
__global__ __launch_bounds__(NTHREADS) void k_mlp_grad(const float* __restrict__ Pg, int nin,
                                                       const float* __restrict__ X, const float* __restrict__ Yt,
                                                       int n, float* __restrict__ partial) {
    extern __shared__ float sh[];
    const Off o = offsets(nin);
    float* P = sh;
    float* xs = P + o.n;
    float* h1s = xs + WAVES * MAX_IN;
    float* d2s = h1s + WAVES * H;
    float* red = d2s + WAVES * H;        // WAVES * (H*H + ...) is too large: reduce row by row
    for (int i = threadIdx.x; i < o.n; i += NTHREADS) P[i] = Pg[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float gw1[MAX_IN], gw2[H];
#pragma unroll
    for (int i = 0; i < MAX_IN; ++i) gw1[i] = 0.f;
#pragma unroll
    for (int k = 0; k < H; ++k) gw2[k] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, loss = 0.f;
    const float inv_n = 1.0f / (float)n;
    const int rounds = (n + gridDim.x * WAVES - 1) / (gridDim.x * WAVES);
    for (int r = 0; r < rounds; ++r) {
        const int s = (r * gridDim.x + blockIdx.x) * WAVES + wave;
        const bool live = s < n;
        float* x = xs + wave * MAX_IN;
        float* h1 = h1s + wave * H;
        float* d2 = d2s + wave * H;
        if (lane < nin) x[lane] = live ? X[(long long)s * nin + lane] : 0.f;
        __syncthreads();
        float h2;
        const float out = forward_one(P, o, nin, x, h1, h2, lane);
        const float e = live ? out - Yt[s] : 0.f;
        const float d = 2.f * e * inv_n;          // dL/dout
        loss += e * e * inv_n;
        gw3 += d * h2;
        gb3 += d;
        const float dz2 = d * P[o.w3 + lane] * (1.f - h2 * h2);
        d2[lane] = dz2;
        gb2 += dz2;
#pragma unroll
        for (int k = 0; k < H; ++k) gw2[k] += dz2 * h1[k];
        __syncthreads();
        float back = 0.f;
#pragma unroll 8
        for (int k = 0; k < H; ++k) back += P[o.w2 + k * H + lane] * d2[k];
        const float h1j = h1[lane];
        const float dz1 = back * (1.f - h1j * h1j);
        gb1 += dz1;
#pragma unroll
        for (int i = 0; i < MAX_IN; ++i)
            if (i < nin) gw1[i] += dz1 * x[i];
        __syncthreads();
    }
    // reduce the WAVES waves of the block, one parameter row at a time, in a fixed order
    float* blk = partial + (long long)blockIdx.x * (o.n + 1);
    auto reduce_row = [&](const float* vals, int count, int dst) {
        // vals: per-lane array of `count` values belonging to unit `lane`
        for (int c = 0; c < count; ++c) {
            red[wave * H + lane] = vals[c];
            __syncthreads();
            if (wave == 0) {
                float v = red[lane];
                for (int w = 1; w < WAVES; ++w) v += red[w * H + lane];
                blk[dst + lane * count + c] = v;
            }
            __syncthreads();
        }
    };
    reduce_row(gw1, nin, o.w1);
    {
        float v[1] = {gb1};
        reduce_row(v, 1, o.b1);
    }
    reduce_row(gw2, H, o.w2);
    {
        float v[1] = {gb2};
        reduce_row(v, 1, o.b2);
        v[0] = gw3;
        reduce_row(v, 1, o.w3);
    }
    // b3 and the loss are wave-uniform sums over the wave's samples
    red[wave * H + lane] = lane == 0 ? gb3 : 0.f;
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = 0.f;
        for (int w = 0; w < WAVES; ++w) b += red[w * H];
        blk[o.b3] = b;
    }
    __syncthreads();
    red[wave * H + lane] = lane == 0 ? loss : 0.f;
    __syncthreads();
    if (threadIdx.x == 0) {
        float l = 0.f;
        for (int w = 0; w < WAVES; ++w) l += red[w * H];
        blk[o.n] = l;
    }
}

// Sum per-block partials (fixed order) into g[0..np] (np params + loss).
__global__ void k_sum_partials(const float* __restrict__ partial, int nblocks, int np1, float* __restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np1) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial[(long long)b * np1 + i];
    g[i] = s;
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M, float* __restrict__ V,
                       int np, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.f - b1) * g;
    const float v = b2 * V[i] + (1.f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

struct gpe_mlp {
    int nin = 0, np = 0;
    hipStream_t stream = nullptr;
    float* P = nullptr;   // parameters
    float* G = nullptr;   // gradient + loss
    float* M = nullptr;   // Adam moments
    float* V = nullptr;
    float* partial = nullptr;
    int partial_blocks = 0;
    float* X = nullptr;   // data set / inference inputs
    float* Y = nullptr;
    float* out = nullptr;
    int cap = 0;          // samples the X/Y/out buffers hold
    long long adam_t = 0;
    std::string err;
};

namespace {
thread_local std::string g_mlp_error;

int mfail(gpe_mlp* m, int code, const std::string& s) {
    if (m) m->err = s;
    g_mlp_error = s;
    return code;
}

#define MHIP(m, expr)                                                                                              \
    do {                                                                                                           \
        hipError_t x_ = (expr);                                                                                    \
        if (x_ != hipSuccess) return mfail(m, GPE_EHIP, std::string(#expr) + ": " + hipGetErrorString(x_));      \
    } while (0)

size_t shmem_bytes(int nin) {
    return sizeof(float) * ((size_t)nparams(nin) + WAVES * MAX_IN + 3 * WAVES * H);
}

int ensure_data(gpe_mlp* m, int n) {
    if (n <= m->cap) return 0;
    if (m->X) (void)hipFree(m->X);
    if (m->Y) (void)hipFree(m->Y);
    if (m->out) (void)hipFree(m->out);
    m->X = m->Y = m->out = nullptr;
    m->cap = 0;
    MHIP(m, hipMalloc(&m->X, sizeof(float) * (size_t)n * m->nin));
    MHIP(m, hipMalloc(&m->Y, sizeof(float) * (size_t)n));
    MHIP(m, hipMalloc(&m->out, sizeof(float) * (size_t)n));
    m->cap = n;
    return 0;
}

int grid_for(int n) {
    int g = (n + WAVES * 8 - 1) / (WAVES * 8);  // ~8 samples per wave
    if (g < 1) g = 1;
    if (g > 512) g = 512;
    return g;
}

int ensure_partial(gpe_mlp* m, int blocks) {
    if (blocks <= m->partial_blocks) return 0;
    if (m->partial) (void)hipFree(m->partial);
    m->partial = nullptr;
    MHIP(m, hipMalloc(&m->partial, sizeof(float) * (size_t)blocks * (m->np + 1)));
    m->partial_blocks = blocks;
    return 0;
}

// gradient (and loss) of the samples resident in X/Y[0..n) into G
int grad_dev(gpe_mlp* m, int n) {
    const int g = grid_for(n);
    int rc = ensure_partial(m, g);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mlp_grad, dim3(g), dim3(NTHREADS), shmem_bytes(m->nin), m->stream, m->P, m->nin, m->X, m->Y, n,
                       m->partial);
    MHIP(m, hipGetLastError());
    const int np1 = m->np + 1;
    hipLaunchKernelGGL(k_sum_partials, dim3((np1 + 255) / 256), dim3(256), 0, m->stream, m->partial, g, np1, m->G);
    MHIP(m, hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

const char* gpe_mlp_last_error(const gpe_mlp* m) { return m ? m->err.c_str() : g_mlp_error.c_str(); }

int gpe_mlp_create(int n_in, gpe_mlp** out) {
    if (!out || n_in < 1 || n_in > MAX_IN) return mfail(nullptr, GPE_EINVAL, "gpe_mlp_create: n_in must be 1..16");
    gpe_mlp* m = new gpe_mlp();
    m->nin = n_in;
    m->np = nparams(n_in);
    hipError_t x = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (x == hipSuccess) x = hipMalloc(&m->P, sizeof(float) * m->np);
    if (x == hipSuccess) x = hipMalloc(&m->G, sizeof(float) * (m->np + 1));
    if (x == hipSuccess) x = hipMalloc(&m->M, sizeof(float) * m->np);
    if (x == hipSuccess) x = hipMalloc(&m->V, sizeof(float) * m->np);
    if (x == hipSuccess) x = hipMemsetAsync(m->P, 0, sizeof(float) * m->np, m->stream);
    if (x == hipSuccess) x = hipMemsetAsync(m->M, 0, sizeof(float) * m->np, m->stream);
    if (x == hipSuccess) x = hipMemsetAsync(m->V, 0, sizeof(float) * m->np, m->stream);
    if (x == hipSuccess) x = hipStreamSynchronize(m->stream);
    if (x != hipSuccess) {
        int rc = mfail(nullptr, GPE_EHIP, std::string("gpe_mlp_create: ") + hipGetErrorString(x));
        gpe_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return 0;
}

void gpe_mlp_destroy(gpe_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    float* bufs[] = {m->P, m->G, m->M, m->V, m->partial, m->X, m->Y, m->out};
    for (float* b : bufs)
        if (b) (void)hipFree(b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int gpe_mlp_num_params(const gpe_mlp* m) { return m ? m->np : -1; }

int gpe_mlp_set_params(gpe_mlp* m, const float* p, int64_t n) {
    if (!m || !p || n != m->np) return mfail(m, GPE_EINVAL, "gpe_mlp_set_params: bad arguments");
    MHIP(m, hipMemcpyAsync(m->P, p, sizeof(float) * m->np, hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemsetAsync(m->M, 0, sizeof(float) * m->np, m->stream));
    MHIP(m, hipMemsetAsync(m->V, 0, sizeof(float) * m->np, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    m->adam_t = 0;
    return 0;
}

int gpe_mlp_get_params(gpe_mlp* m, float* p, int64_t n) {
    if (!m || !p || n != m->np) return mfail(m, GPE_EINVAL, "gpe_mlp_get_params: bad arguments");
    MHIP(m, hipMemcpyAsync(p, m->P, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return 0;
}

int gpe_mlp_forward(gpe_mlp* m, const float* x, int n, float* y) {
    if (!m || !x || !y || n < 0) return mfail(m, GPE_EINVAL, "gpe_mlp_forward: bad arguments");
    if (n == 0) return 0;
    int rc = ensure_data(m, n);
    if (rc) return rc;
    MHIP(m, hipMemcpyAsync(m->X, x, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n)), dim3(NTHREADS), shmem_bytes(m->nin), m->stream, m->P, m->nin,
                       m->X, n, m->out);
    MHIP(m, hipGetLastError());
    MHIP(m, hipMemcpyAsync(y, m->out, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return 0;
}

int gpe_mlp_grad(gpe_mlp* m, const float* x, const float* y, int n, float* grad, float* loss) {
    if (!m || !x || !y || !grad || !loss || n <= 0) return mfail(m, GPE_EINVAL, "gpe_mlp_grad: bad arguments");
    int rc = ensure_data(m, n);
    if (rc) return rc;
    MHIP(m, hipMemcpyAsync(m->X, x, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->Y, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, m->stream));
    if ((rc = grad_dev(m, n))) return rc;
    std::vector<float> g(m->np + 1);
    MHIP(m, hipMemcpyAsync(g.data(), m->G, sizeof(float) * (m->np + 1), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    for (int i = 0; i < m->np; ++i) grad[i] = g[i];
    *loss = g[m->np];
    return 0;
}

int gpe_mlp_train(gpe_mlp* m, const float* x, const float* y, int n, int steps, int batch, float lr,
                  float* losses) {
    if (!m || !x || !y || n <= 0 || steps < 0 || batch <= 0 || !(lr > 0))
        return mfail(m, GPE_EINVAL, "gpe_mlp_train: bad arguments");
    int rc = ensure_data(m, n);
    if (rc) return rc;
    MHIP(m, hipMemcpyAsync(m->X, x, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->Y, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, m->stream));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int nb = batch < n ? batch : n;
    const int nbatches = (n + nb - 1) / nb;
    float* Xall = m->X;
    float* Yall = m->Y;
    for (int s = 0; s < steps; ++s) {
        // cyclic mini-batches over the (host-shuffled) data set
        const int b = s % nbatches;
        const int off = b * nb;
        const int cnt = (off + nb <= n) ? nb : n - off;
        m->X = Xall + (size_t)off * m->nin;
        m->Y = Yall + off;
        rc = grad_dev(m, cnt);
        m->X = Xall;
        m->Y = Yall;
        if (rc) return rc;
        ++m->adam_t;
        const float c1 = 1.f - powf(b1, (float)m->adam_t), c2 = 1.f - powf(b2, (float)m->adam_t);
        hipLaunchKernelGGL(k_adam, dim3((m->np + 255) / 256), dim3(256), 0, m->stream, m->P, m->G, m->M, m->V, m->np, lr,
                           b1, b2, eps, c1, c2);
        MHIP(m, hipGetLastError());
        if (losses) MHIP(m, hipMemcpyAsync(losses + s, m->G + m->np, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    }
    MHIP(m, hipStreamSynchronize(m->stream));
    return 0;
}

}  // extern "C"
