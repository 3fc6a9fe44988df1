// The z-marching stencil kernel and the fused stage operators.
#pragma once

#include <type_traits>

#include "gpe_device.h"

namespace gpe {

// ---------------------------------------------------------------------------
// Block reductions (deterministic: fixed butterfly within a wavefront, fixed
// wave order across the block).
// RED_MIXED: rows [0, Op::kNSum) are sums, the rest maxima (starting at -1e300)
enum { RED_NONE = 0, RED_SUM = 1, RED_MAX = 2, RED_MIXED = 3 };

template <int MODE> __device__ __forceinline__ double red_op(double a, double b) {
    if constexpr (MODE == RED_SUM) return a + b;
    else return a > b ? a : b;
}

template <int NT, int MODE> __device__ __forceinline__ double block_reduce(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = red_op<MODE>(v, __shfl_xor(v, o, 64));
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) r = red_op<MODE>(r, sh[i]);
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// The stencil kernel.  Op supplies:
//   kNRed (0..2 reduction values), kMode (RED_SUM / RED_MAX),
//   struct Pre; Pre load(Cell i) const   -- pointwise inputs of cell i
//   void apply(Cell i, int x, int y, int z, const Sten& st, const Pre& p, double* red) const
// GPE_WAVES_PER_EU > 0 asks the compiler for at least that many waves per
// SIMD (caps VGPRs); GPE_PW_PREFETCH = 1 prefetches the pointwise inputs
// together with the stencil planes, 0 loads them at the top of the plane's
// iteration.
#ifndef GPE_WAVES_PER_EU
#define GPE_WAVES_PER_EU 0
#endif
#ifndef GPE_PW_PREFETCH
#define GPE_PW_PREFETCH 1
#endif
// Planes whose loads are in flight while a plane is evaluated (1 or 2).
#ifndef GPE_PF_DEPTH
#define GPE_PF_DEPTH 2
#endif
#if GPE_WAVES_PER_EU > 0
#define GPE_OCC __attribute__((amdgpu_waves_per_eu(GPE_WAVES_PER_EU)))
#else
#define GPE_OCC
#endif

// Index of a cell as a plane base (uniform across the wavefront, kept in
// scalar registers) plus the 32-bit offset within the plane.  `ptr + cell`
// forms the address as (ptr + zb) + own * sizeof(T), a 64-bit scalar base and
// a 32-bit per-lane byte offset, so a kernel keeps one VGPR per distinct
// element size instead of a 64-bit address per field.  Requires
// nx * ny * 16 < 2^32 (checked at context creation).
struct Cell {
    long long zb;
    unsigned own;
};
// readfirstlane pins the (uniform) plane offset in scalar registers and
// hides it from loop strength reduction, which would otherwise turn every
// field into a 64-bit per-lane pointer stepped by one plane per iteration.
// The pointer itself is only offset (never rebuilt from an integer), so the
// compiler keeps the global address space: global_load with a scalar base
// and a 32-bit vector offset.
__device__ __forceinline__ long long uniform_i64(long long v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(unsigned long long)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}
template <typename T> __device__ __forceinline__ T* operator+(T* p, Cell c) {
    return (T*)((char*)(p + uniform_i64(c.zb)) + c.own * (unsigned)sizeof(T));
}
template <typename T> __device__ __forceinline__ const T* operator+(const T* p, Cell c) {
    return (const T*)((const char*)(p + uniform_i64(c.zb)) + c.own * (unsigned)sizeof(T));
}

// Predicated plane access without branches.  A plane of a field becomes a
// buffer resource (scalar base = the plane's first element, range = the
// plane's bytes); lanes pass a 32-bit byte offset, and a lane that must not
// access memory passes kNoOff, which the hardware range check turns into a
// load of 0 or a dropped store with no memory traffic.  With no branch around
// a load or store every wave issues the same memory instructions in the same
// order, so the compiler's wait counts cover exactly the loads a value needs
// (a load under a divergent branch may be skipped by a whole wave, and the
// compiler then waits for every memory operation in flight, prefetches and
// store acknowledgements included).  Needs nx * ny * sizeof(value) <=
// kNoOff, i.e. nx * ny < 2^28 for complex128 (checked at context creation).
constexpr unsigned kNoOff = 0xFFFFFFF0u;
using BRes = __amdgpu_buffer_rsrc_t;
enum { BUF_CACHED = 0, BUF_NT = 2 };  // cache policy bits of the buffer instructions (2: nt)

template <typename R> __device__ __forceinline__ BRes plane_res(const C<R>* f, long long zb, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)(f + uniform_i64(zb)), (short)0, (int)bytes, 0x00020000);
}

// Plane m of a field whose planes are pbytes long (pbytes < 2^32): the byte
// offset is a 32 x 32-bit product, two scalar instructions where a signed
// 64-bit index product takes seven.
template <typename R> __device__ __forceinline__ BRes plane_res_m(const C<R>* f, unsigned m, unsigned pbytes) {
    const char* b = reinterpret_cast<const char*>(f) + (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(m) * pbytes;
    return __builtin_amdgcn_make_buffer_rsrc((void*)b, (short)0, (int)pbytes, 0x00020000);
}

template <typename R, int AUX = BUF_CACHED> __device__ __forceinline__ C<R> bld(BRes r, unsigned off) {
    if constexpr (sizeof(R) == 8) {
        typedef unsigned u4 __attribute__((ext_vector_type(4)));
        const u4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, AUX);
        return __builtin_bit_cast(C<R>, v);
    } else {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        const u2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, AUX);
        return __builtin_bit_cast(C<R>, v);
    }
}

template <typename R, int AUX = BUF_NT> __device__ __forceinline__ void bst(BRes r, unsigned off, C<R> a) {
    if constexpr (sizeof(R) == 8) {
        typedef unsigned u4 __attribute__((ext_vector_type(4)));
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, a), r, (int)off, 0, AUX);
    } else {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, a), r, (int)off, 0, AUX);
    }
}

// Sources of the stencil input.  PlainSrc reads one field.  SumSrc forms
// u = a + s d from an FP64 field a and a complex64 increment d scaled by the
// FP64 complex s, so a stage can stencil psi0 + (increment) without the sum
// ever being stored; it also hands the operator the raw d of the thread's own
// cell (Aux), which the increments-only stages need.
struct NoAux {};
template <typename R> struct PlainSrc {
    using Aux = NoAux;
    const C<R>* __restrict__ u;
    __device__ __forceinline__ C<R> operator()(Cell i) const { return *(u + i); }
    __device__ __forceinline__ C<R> operator()(Cell i, NoAux&) const { return *(u + i); }
};
struct SumSrc {
    using Aux = float2;
    const double2* __restrict__ a;
    const float2* __restrict__ d;
    double2 s;
    __device__ __forceinline__ double2 comb(double2 p, float2 v) const {
        const double vx = v.x, vy = v.y;
        return mkc<double>(p.x + (s.x * vx - s.y * vy), p.y + (s.x * vy + s.y * vx));
    }
    __device__ __forceinline__ double2 operator()(Cell i) const { return comb(*(a + i), *(d + i)); }
    __device__ __forceinline__ double2 operator()(Cell i, float2& dv) const {
        dv = *(d + i);
        return comb(*(a + i), dv);
    }
};

// An operator may set kPrefetch = false: its pointwise inputs are then
// loaded at the top of each plane instead of one plane ahead, which frees the
// second register set where that set would spill (spill reloads wait for
// every outstanding load, the plane prefetch included).
template <class Op, class = void> struct OpPrefetch : std::true_type {};
template <class Op> struct OpPrefetch<Op, std::void_t<decltype(Op::kPrefetch)>> : std::bool_constant<Op::kPrefetch> {};
template <class Op> constexpr bool op_prefetch() { return OpPrefetch<Op>::value; }
// An operator may also cap the prefetch depth (kDepth = 1) where two sets
// of in-flight registers would spill.
template <class Op, class = void> struct OpDepth : std::integral_constant<int, GPE_PF_DEPTH> {};
template <class Op>
struct OpDepth<Op, std::void_t<decltype(Op::kDepth)>>
    : std::integral_constant<int, (Op::kDepth < GPE_PF_DEPTH ? Op::kDepth : GPE_PF_DEPTH)> {};

// Body of the stencil kernel for the tile `bid` of grid g; the block's
// reductions (if any) go to partial[r * g.nblocks + pidx].
template <typename R, int ORD, int TX, int TY, class Op, class Src, bool GEN>
__device__ __forceinline__ void stencil_body_src(const Geom& g, const StW<R>& w, const Src& src, const Op& op,
                                                 double* __restrict__ partial, int bid, int pidx) {
    using Aux = typename Src::Aux;
    constexpr bool kAux = !std::is_same<Aux, NoAux>::value;
    constexpr bool PF = GPE_PW_PREFETCH && op_prefetch<Op>();
    constexpr int H = ORD / 2, NT = TX * TY, LW = TX + 2 * H, LH = TY + 2 * H;
    constexpr int NHALO = 2 * H * TX + 2 * H * TY;
    constexpr int NSLOT = (NHALO + NT - 1) / NT;
    __shared__ C<R> lds[2][LH * LW];
    __shared__ double red_sh[NT / 64];

    const int tid = threadIdx.x;
    const int ntxy = g.ntx * g.nty;
    const int txy = bid % ntxy, zc = bid / ntxy;
    const int x0 = (txy % g.ntx) * TX, y0 = (txy / g.ntx) * TY;
    const int z0 = zc * g.tz;
    const int z1 = min(z0 + g.tz, g.nz);
    const int lx = tid % TX, ly = tid / TX;
    const int x = x0 + lx, y = y0 + ly;
    const bool valid = (x < g.nx) && (y < g.ny);
    const unsigned col = (unsigned)bcmap(x, g.nx, g.bcs) + (unsigned)g.nx * (unsigned)bcmap(y, g.ny, g.bcs);
    const unsigned own = (unsigned)x + (unsigned)g.nx * (unsigned)y;  // cell offset within a plane (valid threads)
    const int lc = (ly + H) * LW + lx + H;

    // Halo slots: first the 2H rows above/below the tile (coalesced rows of
    // TX), then the 2H columns left/right of each tile row.
    unsigned hoff[NSLOT];
    int hl[NSLOT];
    bool hv[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        const int e = tid + s * NT;
        hv[s] = e < NHALO;
        int gx = x0, gy = y0, lxx = H, lyy = H;
        if (e < 2 * H * TX) {
            const int r = e / TX, c = e % TX;
            lxx = c + H;
            lyy = r < H ? r : TY + r;
            gx = x0 + c;
            gy = y0 - H + lyy;
        } else if (e < NHALO) {
            const int e2 = e - 2 * H * TX;
            const int r = e2 / (2 * H), c = e2 % (2 * H);
            lyy = r + H;
            lxx = c < H ? c : TX + c;
            gx = x0 - H + lxx;
            gy = y0 + r;
        }
        hl[s] = lyy * LW + lxx;
        hoff[s] = (unsigned)bcmap(gx, g.nx, g.bcs) + (unsigned)g.nx * (unsigned)bcmap(gy, g.ny, g.bcs);
    }

    // Register z-window q[0..2H] = planes z-H..z+H; qa the matching Aux of
    // the own column (only qa[H] is read, the rest is dead and removed when
    // Aux is empty).
    C<R> q[2 * H + 1];
    Aux qa[2 * H + 1];
#pragma unroll
    for (int d = 0; d < 2 * H; ++d) q[d] = src(Cell{g.plane * bcmap(z0 - H + d, g.nz, g.bcs), col}, qa[d]);
    double red[Op::kNRed > 2 ? Op::kNRed : 2] = {};
    if constexpr (Op::kMode == RED_MIXED) {
#pragma unroll
        for (int r = Op::kNSum; r < Op::kNRed; ++r) red[r] = -1e300;
    }

    // The loads for plane z: centre of plane z+H (the window's new entry),
    // the halo of plane z and the operator's pointwise inputs.
    struct Pf {
        C<R> qn;
        Aux qan;
        C<R> hr[NSLOT];
        typename Op::Pre pw;
    };
    auto fetch = [&](Pf& S, int zz) {
        S.qn = src(Cell{g.plane * bcmap(zz + H, g.nz, g.bcs), col}, S.qan);
#pragma unroll
        for (int s = 0; s < NSLOT; ++s)
            if (hv[s]) S.hr[s] = src(Cell{g.plane * zz, hoff[s]});
        if (PF && valid) S.pw = op.load(Cell{g.plane * zz, own});
    };
    // One plane: stage plane z in LDS (double buffered, one barrier), start
    // the loads of plane z + DEPTH into S, evaluate and store plane z.
    constexpr int DEPTH = OpDepth<Op>::value;
    auto step = [&](int z, Pf& S) {
        q[2 * H] = S.qn;
        qa[2 * H] = S.qan;
        typename Op::Pre pc = S.pw;
        if (!PF && valid) pc = op.load(Cell{g.plane * z, own});
        C<R>* L = lds[(z - z0) & 1];
        L[lc] = q[H];
#pragma unroll
        for (int s = 0; s < NSLOT; ++s)
            if (hv[s]) L[hl[s]] = S.hr[s];
        __syncthreads();
        if (z + DEPTH < z1) fetch(S, z + DEPTH);
        if (valid) {
            Sten<R, ORD, LW, GEN> st;
            st.L = L;
            st.lc = lc;
#pragma unroll
            for (int d = 0; d <= 2 * H; ++d) st.q[d] = q[d];
            st.w = w;
            if constexpr (kAux) op.apply(Cell{g.plane * z, own}, x, y, z, st, pc, qa[H], red);
            else op.apply(Cell{g.plane * z, own}, x, y, z, st, pc, red);
        }
#pragma unroll
        for (int d = 0; d < 2 * H; ++d) {
            q[d] = q[d + 1];
            qa[d] = qa[d + 1];
        }
    };
    if constexpr (DEPTH == 1) {
        Pf S0{};
        if (z0 < z1) fetch(S0, z0);
        for (int z = z0; z < z1; ++z) step(z, S0);
    } else {
        // two planes in flight: S0 carries planes z0, z0+2, ..., S1 the others
        Pf S0{}, S1{};
        if (z0 < z1) fetch(S0, z0);
        if (z0 + 1 < z1) fetch(S1, z0 + 1);
        for (int z = z0; z < z1; z += 2) {
            step(z, S0);
            if (z + 1 < z1) step(z + 1, S1);
        }
    }

    if constexpr (Op::kNRed > 0) {
#pragma unroll
        for (int r = 0; r < Op::kNRed; ++r) {
            double v;
            if constexpr (Op::kMode == RED_MIXED)
                v = r < Op::kNSum ? block_reduce<NT, RED_SUM>(red[r], red_sh) : block_reduce<NT, RED_MAX>(red[r], red_sh);
            else
                v = block_reduce<NT, Op::kMode>(red[r], red_sh);
            if (tid == 0) partial[(long long)r * g.nblocks + pidx] = v;
        }
    }
}

template <typename R, int ORD, int TX, int TY, bool GEN, class Op>
__device__ __forceinline__ void stencil_body(const Geom& g, const StW<R>& w, const C<R>* __restrict__ u, const Op& op,
                                             double* __restrict__ partial, int bid, int pidx) {
    stencil_body_src<R, ORD, TX, TY, Op, PlainSrc<R>, GEN>(g, w, PlainSrc<R>{u}, op, partial, bid, pidx);
}

template <typename R, int ORD, int TX, int TY, class Op, bool GEN>
__global__ __launch_bounds__(TX* TY) GPE_OCC void k_stencil(Geom g, StW<R> w, const C<R>* __restrict__ u, Op op,
                                                             double* __restrict__ partial) {
    stencil_body<R, ORD, TX, TY, GEN, Op>(g, w, u, op, partial, xcd_remap(blockIdx.x, g.nblocks), blockIdx.x);
}

// Stencil of a + s d (SumSrc): FP64 arithmetic over an FP64 state and a
// complex64 increment.
template <int ORD, int TX, int TY, class Op, bool GEN>
__global__ __launch_bounds__(TX* TY) GPE_OCC void k_stencil_sum(Geom g, StW<double> w, SumSrc src, Op op,
                                                                 double* __restrict__ partial) {
    stencil_body_src<double, ORD, TX, TY, Op, SumSrc, GEN>(g, w, src, op, partial, xcd_remap(blockIdx.x, g.nblocks),
                                                            blockIdx.x);
}

// Final reduction of per-block partials (one block, fixed order).
template <int MODE>
__global__ __launch_bounds__(1024) void k_reduce_final(const double* __restrict__ partial, int n, int nval,
                                                        double* __restrict__ out) {
    __shared__ double sh[16];
    for (int r = 0; r < nval; ++r) {
        double v = MODE == RED_SUM ? 0.0 : 0.0;
        for (int i = threadIdx.x; i < n; i += 1024) v = red_op<MODE>(v, partial[(long long)r * n + i]);
        v = block_reduce<1024, MODE>(v, sh);
        if (threadIdx.x == 0) out[r] = v;
    }
}

// ---------------------------------------------------------------------------
// Stage operators.  Complex dt factors are passed premultiplied (e.g. dt/2).

struct NoPre {};

// out = psi + a*f(psi)                              (euler; rk2 stage 1)
template <typename R> struct OpAxpySelf {
    static constexpr int kKind = 0, kAccess = 2;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    using Pre = NoPre;
    RhsC<R> k;
    C<R>* __restrict__ out;
    C<R> a;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre&, double*) const {
        st_stream<R>(out + i, cfma<R>(a, rhs_eval<R>(k, st, x, y, z), st.c()));
    }
};

// out = psi0 + a*f(u)                               (rk2 stage 2)
template <typename R> struct OpAxpyP0 {
    static constexpr int kKind = 1, kAccess = 3;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    struct Pre { C<R> p0; };
    RhsC<R> k;
    const C<R>* __restrict__ p0;
    C<R>* __restrict__ out;
    C<R> a;
    __device__ Pre load(Cell i) const { return {ld_stream<R>(p0 + i)}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, double*) const {
        st_stream<R>(out + i, cfma<R>(a, rhs_eval<R>(k, st, x, y, z), p.p0));
    }
};

// Classic RK4 in 13 field accesses per cell (instead of 16): the
// accumulator psi0 + dt (k1/6 + k2/3 + k3/3) is formed once, in stage 3,
// with dt k1/6 = (u1 - psi0)/3 and dt k2/3 = 2 (u2 - psi0)/3 recovered from
// the stage fields that stage reads anyway.
//   stage 1: u1 = psi0 + (dt/2) f(psi0)                       reads psi0        writes u1
//   stage 2: u2 = psi0 + (dt/2) f(u1)                         reads u1, psi0    writes u2
//   stage 3: u3 = psi0 + dt f(u2),
//            acc = psi0 + ((u1 - psi0) + 2 (u2 - psi0))/3 + (dt/3) f(u2)
//                                                             reads u2, psi0, u1 writes u3 (over u1), acc
//   stage 4: psi_new = acc + (dt/6) f(u3)                     reads u3, acc     writes psi_new

// RK4 stage 1: u1 = psi0 + (dt/2) f
template <typename R> struct OpRk4First {
    static constexpr int kKind = 2, kAccess = 2;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    using Pre = NoPre;
    RhsC<R> k;
    C<R>* __restrict__ unext;
    C<R> a1;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre&, double*) const {
        st_stream<R>(unext + i, cfma<R>(a1, rhs_eval<R>(k, st, x, y, z), st.c()));
    }
};

// RK4 stage 2: u2 = psi0 + (dt/2) f(u1)
template <typename R> struct OpRk4Second {
    static constexpr int kKind = 3, kAccess = 3;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    struct Pre { C<R> p0; };
    RhsC<R> k;
    const C<R>* __restrict__ p0;
    C<R>* __restrict__ unext;
    C<R> a1;
    __device__ Pre load(Cell i) const { return {ld_stream<R>(p0 + i)}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, double*) const {
        st_stream<R>(unext + i, cfma<R>(a1, rhs_eval<R>(k, st, x, y, z), p.p0));
    }
};

// RK4 stage 3 (see above).  u1 is read and u3 written at the same index by
// the same thread through one pointer.
template <typename R> struct OpRk4Third {
    static constexpr int kKind = 15, kAccess = 5;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    struct Pre { C<R> p0, u1; };
    RhsC<R> k;
    const C<R>* __restrict__ p0;
    C<R>* u13;
    C<R>* __restrict__ acc;
    C<R> a1, a3;  // dt, dt/3
    __device__ Pre load(Cell i) const { return {ld_stream<R>(p0 + i), ld_stream<R>(u13 + i)}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, double*) const {
        const C<R> f = rhs_eval<R>(k, st, x, y, z);
        const C<R> u2 = st.c();
        const R third = R(1) / R(3);
        const C<R> s = mkc<R>((p.u1.x - p.p0.x) + R(2) * (u2.x - p.p0.x), (p.u1.y - p.p0.y) + R(2) * (u2.y - p.p0.y));
        st_stream<R>(u13 + i, cfma<R>(a1, f, p.p0));
        st_stream<R>(acc + i, cfma<R>(a3, f, caxpy<R>(third, s, p.p0)));
    }
};

// RK4 stage 4: out = acc + (dt/6) f
template <typename R> struct OpRk4Last {
    static constexpr int kKind = 4, kAccess = 3;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    struct Pre { C<R> acc; };
    RhsC<R> k;
    const C<R>* __restrict__ acc;
    C<R>* __restrict__ out;
    C<R> a;
    __device__ Pre load(Cell i) const { return {ld_stream<R>(acc + i)}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, double*) const {
        st_stream<R>(out + i, cfma<R>(a, rhs_eval<R>(k, st, x, y, z), p.acc));
    }
};

// Cash-Karp tableau (src/solve.f90:255-281), stored in the field precision.
template <typename R> struct CkTab {
    R b21, b31, b32, b41, b42, b43, b51, b52, b53, b54, b61, b62, b63, b64, b65;
    R c1, c3, c4, c6;
    double dc1, dc3, dc4, dc5, dc6;
};

// Cash-Karp coefficients (src/solve.f90:255-281), dc_i = c_i - c*_i.  A
// compile-time constant, so kernels can use it without holding it in
// registers (the compiler materialises each coefficient where it is used).
template <typename R> constexpr CkTab<R> ck_tab() {
    constexpr double c1 = 37.0 / 378.0, c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c5 = 0.0, c6 = 512.0 / 1771.0;
    CkTab<R> t{};
    t.b21 = (R)0.2; t.b31 = (R)0.075; t.b32 = (R)0.225; t.b41 = (R)0.3; t.b42 = (R)-0.9; t.b43 = (R)1.2;
    t.b51 = (R)(-11.0 / 54.0); t.b52 = (R)2.5; t.b53 = (R)(-70.0 / 27.0); t.b54 = (R)(35.0 / 27.0);
    t.b61 = (R)(1631.0 / 55296.0); t.b62 = (R)(175.0 / 512.0); t.b63 = (R)(575.0 / 13824.0);
    t.b64 = (R)(44275.0 / 110592.0); t.b65 = (R)(253.0 / 4096.0);
    t.c1 = (R)c1; t.c3 = (R)c3; t.c4 = (R)c4; t.c6 = (R)c6;
    t.dc1 = c1 - 2825.0 / 27648.0;
    t.dc3 = c3 - 18575.0 / 48384.0;
    t.dc4 = c4 - 13525.0 / 55296.0;
    t.dc5 = c5 - 277.0 / 14336.0;
    t.dc6 = c6 - 0.25;
    return t;
}

// Cash-Karp stage 1 (first attempt): f0 = f(psi0) ; u2 = psi0 + b21 (dt f0)
template <typename R> struct OpCk1 {
    static constexpr int kKind = 5, kAccess = 3;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    using Pre = NoPre;
    RhsC<R> k;
    C<R>* __restrict__ f0;
    C<R>* __restrict__ unext;
    C<R> dt;
    R b21;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre&, double*) const {
        const C<R> f = rhs_eval<R>(k, st, x, y, z);
        st_stream<R>(f0 + i, f);
        const C<R> k1 = cmul<R>(dt, f);
        st_stream<R>(unext + i, caxpy<R>(b21, k1, st.c()));
    }
};

// Cash-Karp stage s = 2..5: k_s = dt f(u_s) ; u_{s+1} = psi0 + sum_j b_{s+1,j} k_j
template <typename R, int S_> struct OpCkMid {
    // kernel s reads u_s, psi0, f0, k2..k_{s-1} and writes k_s, u_{s+1}
    static constexpr int kKind = 4 + S_, kAccess = S_ + 3;
    static constexpr int kDepth = S_ >= 4 ? 1 : 2;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    struct Pre { C<R> p0, f0, k2, k3, k4; };
    RhsC<R> k;
    const C<R>* __restrict__ p0;
    const C<R>* __restrict__ f0;
    const C<R>* __restrict__ k2;
    const C<R>* __restrict__ k3;
    const C<R>* __restrict__ k4;
    C<R>* __restrict__ kout;
    C<R>* __restrict__ unext;
    C<R> dt;
    CkTab<R> t;
    __device__ Pre load(Cell i) const {
        Pre p;
        p.p0 = ld_stream<R>(p0 + i);
        p.f0 = ld_stream<R>(f0 + i);
        if constexpr (S_ >= 3) p.k2 = ld_stream<R>(k2 + i);
        if constexpr (S_ >= 4) p.k3 = ld_stream<R>(k3 + i);
        if constexpr (S_ >= 5) p.k4 = ld_stream<R>(k4 + i);
        return p;
    }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, double*) const {
        const C<R> ks = cmul<R>(dt, rhs_eval<R>(k, st, x, y, z));
        const C<R> k1 = cmul<R>(dt, p.f0);
        st_stream<R>(kout + i, ks);
        C<R> u;
        if constexpr (S_ == 2) {
            u = caxpy<R>(t.b32, ks, caxpy<R>(t.b31, k1, p.p0));
        } else if constexpr (S_ == 3) {
            u = caxpy<R>(t.b43, ks, caxpy<R>(t.b42, p.k2, caxpy<R>(t.b41, k1, p.p0)));
        } else if constexpr (S_ == 4) {
            u = caxpy<R>(t.b54, ks, caxpy<R>(t.b53, p.k3, caxpy<R>(t.b52, p.k2, caxpy<R>(t.b51, k1, p.p0))));
        } else {
            u = caxpy<R>(t.b65, ks,
                         caxpy<R>(t.b64, p.k4, caxpy<R>(t.b63, p.k3, caxpy<R>(t.b62, p.k2, caxpy<R>(t.b61, k1, p.p0)))));
        }
        st_stream<R>(unext + i, u);
    }
};

// Cash-Karp stage 6 + 5th-order solution + embedded error + scaled max-norm
// (src/solve.f90:176-185, 365-375).  red[0] = max |err| / |psi_scal| with
// psi_scal = |psi0| + dt0 |f0| + 1e-30 (complex), red[1] = max |psi_new|^2.
// The error combination and scale are evaluated in FP64.
template <typename R> struct OpCkLast {
    static constexpr int kKind = 10, kAccess = 7;
    static constexpr int kDepth = 1;
    static constexpr int kNRed = 2, kMode = RED_MAX;
    struct Pre { C<R> p0, f0, k3, k4, k5; };
    RhsC<R> k;
    const C<R>* __restrict__ p0;
    const C<R>* __restrict__ f0;
    const C<R>* __restrict__ k3;
    const C<R>* __restrict__ k4;
    const C<R>* __restrict__ k5;
    C<R>* __restrict__ out;
    C<R> dt;
    double dt0_re, dt0_im;
    CkTab<R> t;
    __device__ Pre load(Cell i) const {
        return {ld_stream<R>(p0 + i), ld_stream<R>(f0 + i), ld_stream<R>(k3 + i), ld_stream<R>(k4 + i),
                ld_stream<R>(k5 + i)};
    }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre& p, double* red) const {
        const C<R> k6 = cmul<R>(dt, rhs_eval<R>(k, st, x, y, z));
        const C<R> k1 = cmul<R>(dt, p.f0);
        const C<R> nw = caxpy<R>(t.c6, k6, caxpy<R>(t.c4, p.k4, caxpy<R>(t.c3, p.k3, caxpy<R>(t.c1, k1, p.p0))));
        st_stream<R>(out + i, nw);
        const double er = t.dc1 * (double)k1.x + t.dc3 * (double)p.k3.x + t.dc4 * (double)p.k4.x +
                          t.dc5 * (double)p.k5.x + t.dc6 * (double)k6.x;
        const double ei = t.dc1 * (double)k1.y + t.dc3 * (double)p.k3.y + t.dc4 * (double)p.k4.y +
                          t.dc5 * (double)p.k5.y + t.dc6 * (double)k6.y;
        const double ap = sqrt((double)p.p0.x * p.p0.x + (double)p.p0.y * p.p0.y);
        const double af = sqrt((double)p.f0.x * p.f0.x + (double)p.f0.y * p.f0.y);
        const double sr = ap + dt0_re * af + 1e-30, si = dt0_im * af;
        const double e = sqrt((er * er + ei * ei) / (sr * sr + si * si));
        red[0] = e > red[0] ? e : red[0];
        const double a2 = (double)nw.x * nw.x + (double)nw.y * nw.y;
        red[1] = a2 > red[1] ? a2 : red[1];
    }
};

// Standalone derivative (derivs module drop-in): out = D_op(u)
template <typename R> struct OpDeriv {
    static constexpr int kKind = 11, kAccess = 2;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    using Pre = NoPre;
    C<R>* __restrict__ out;
    int op;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell i, int, int, int, const S& st, const Pre&, double*) const {
        C<R> v;
        switch (op) {
            case 0: v = st.d1x(); break;
            case 1: v = st.d1y(); break;
            case 2: v = st.d1z(); break;
            case 3: v = st.d2x(); break;
            case 4: v = st.d2y(); break;
            case 5: v = st.d2z(); break;
            default: v = st.lap(); break;
        }
        st_stream<R>(out + i, v);
    }
};

// Standalone get_rhs: out = f(u)
template <typename R> struct OpRhs {
    static constexpr int kKind = 12, kAccess = 2;
    static constexpr int kNRed = 0, kMode = RED_NONE;
    using Pre = NoPre;
    RhsC<R> k;
    C<R>* __restrict__ out;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell i, int x, int y, int z, const S& st, const Pre&, double*) const {
        st_stream<R>(out + i, rhs_eval<R>(k, st, x, y, z));
    }
};

// Energy density of the reference (src/variables.f90:563-587):
// Re(-lap(psi) conj(psi)) + 0.5 |psi|^4, summed in FP64.
template <typename R> struct OpEnergy {
    static constexpr int kKind = 13, kAccess = 1;
    static constexpr int kNRed = 1, kMode = RED_SUM;
    using Pre = NoPre;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell, int, int, int, const S& st, const Pre&, double* red) const {
        const C<R> L = st.lap();
        const C<R> p = st.c();
        const double pr = p.x, pi = p.y;
        const double a2 = pr * pr + pi * pi;
        red[0] += -((double)L.x * pr + (double)L.y * pi) + 0.5 * (a2 * a2);
    }
};

// Conserved energy of the equation being solved (no reference counterpart:
// the reference's energy() omits V and mu).  For f = i*pim*(a lap + (b + c|psi|^2
// + dV) psi) the density is -sgn(pim) (a Re(conj(psi) lap psi) + b|psi|^2
// + c|psi|^4/2 + dV|psi|^2): eqn 1 gives |grad psi|^2 - |psi|^2 + |psi|^4/2,
// eqn 4 gives |grad psi|^2/2 + V|psi|^2 + g|psi|^4/2 - mu|psi|^2 (after
// summation by parts of the discrete Laplacian).  Summed in FP64.
template <typename R> struct OpHamiltonian {
    static constexpr int kKind = 14, kAccess = 1;
    static constexpr int kNRed = 1, kMode = RED_SUM;
    using Pre = NoPre;
    RhsC<R> k;
    double sgn;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell, int x, int y, int z, const S& st, const Pre&, double* red) const {
        const C<R> L = st.lap();
        const C<R> p = st.c();
        const double pr = p.x, pi = p.y;
        const double a2 = pr * pr + pi * pi;
        double v = 0.0;
        if (k.flags & RHS_V) {
            const double xc = (double)k.xl + x * (double)k.dx, yc = (double)k.yl + y * (double)k.dy,
                         zc = (double)k.zl + z * (double)k.dz;
            v = (double)k.d * ((double)k.vx * xc * xc + (double)k.vy * yc * yc + (double)k.vz * zc * zc);
        }
        const double re_pl = (double)L.x * pr + (double)L.y * pi;
        red[0] += -sgn * ((double)k.a * re_pl + ((double)k.b + v) * a2 + 0.5 * (double)k.c * a2 * a2);
    }
};

// End-corrected quadrature weight of index i on an axis of n points
// (src/variables.f90:49-51, 713-823): 3/8, 7/6, 23/24 at both ends.
__device__ __forceinline__ double qweight_axis(int i, int n) {
    const int e = i < n - 1 - i ? i : n - 1 - i;
    return e == 0 ? 3.0 / 8.0 : e == 1 ? 7.0 / 6.0 : e == 2 ? 23.0 / 24.0 : 1.0;
}

// Momentum (src/variables.f90:615-694): per axis a, the end-corrected
// quadrature of (Re psi - 1) Im(d_a psi*) - Im psi Re(d_a psi), with
// d_a psi* = conj(d_a psi) (the derivative weights are real).  FP64 sums.
template <typename R> struct OpMomentum {
    static constexpr int kKind = 27, kAccess = 1;
    static constexpr int kNRed = 3, kMode = RED_SUM;
    using Pre = NoPre;
    int nx, ny, nz;
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell, int x, int y, int z, const S& st, const Pre&, double* red) const {
        const C<R> p = st.c();
        const double w = qweight_axis(x, nx) * qweight_axis(y, ny) * qweight_axis(z, nz);
        const double pr = (double)p.x - 1.0, pi = p.y;
        const C<R> d[3] = {st.d1x(), st.d1y(), st.d1z()};
#pragma unroll
        for (int a = 0; a < 3; ++a) red[a] += w * (pr * (-(double)d[a].y) - pi * (double)d[a].x);
    }
};

// First derivative in the reference's operation order (src/derivs.f90:25-142):
// order 2: (f+1 - f-1) / (2 d); order 4: (((-f+2 + 8 f+1) - 8 f-1) + f-2) / (12 d),
// in FP64 without contraction, so that equal neighbours cancel exactly where
// they cancel in the reference (the sign of a near-zero velocity decides its
// PDF bin).  axis 0..2; den = 2 d or 12 d.
template <typename R, int H, class S>
__device__ __forceinline__ double2 d1_ref(const S& st, int axis, double den) {
#pragma clang fp contract(off)
    auto nb = [&](int o) -> double2 {
        const C<R> v = axis == 0 ? st.ax(o) : axis == 1 ? st.ay(o) : st.q[H + o];
        return mkc<double>(v.x, v.y);
    };
    const double2 p1 = nb(1), m1 = nb(-1);
    if constexpr (H == 1) {
        return mkc<double>((p1.x - m1.x) / den, (p1.y - m1.y) / den);
    } else {
        const double2 p2 = nb(2), m2 = nb(-2);
        return mkc<double>((((-p2.x + 8.0 * p1.x) - 8.0 * m1.x) + m2.x) / den,
                           (((-p2.y + 8.0 * p1.y) - 8.0 * m1.y) + m2.y) / den);
    }
}

// The velocity of get_pdf_velocity and get_vcf (src/variables.f90:393-400,
// 478-483): -i/2 (psi* g - psi g*) / |psi|^2 where the reference's gradient
// arrays g, g* are declared real (lines 387, 474) and so hold Re(d_a psi);
// the expression reduces exactly to -Im(psi) Re(d_a psi) / |psi|^2.
// masked: cells with |psi|^2 <= 0.01 get the reference's marker 1e3 (as
// get_pdf_velocity); otherwise every cell is divided (get_vcf).  ncomp = 3
// writes v_x, v_y, v_z, 1 only v_x.
// STATS: also the statistics of get_pdf_velocity (src/variables.f90:377-463)
// per component a of the values below the validity marker: rows 3a + 0..2
// count, sum, sum of squares; rows 9 + 2a, 10 + 2a max and -min.
template <typename R, bool STATS = false> struct OpVelocity {
    static constexpr int kKind = 28, kBytesPerCell = (int)sizeof(C<R>) + 24;
    static constexpr int kNRed = STATS ? 15 : 0, kMode = STATS ? RED_MIXED : RED_NONE, kNSum = 9;
    static constexpr double kValid = 500.0;  // the reference keeps values < 0.5e3 (marker 1e3)
    using Real = R;
    using Pre = NoPre;
    double* __restrict__ vel;  // component a at vel + a * n
    long long n;
    int masked, ncomp;
    double den[3];             // 2 d or 12 d per axis
    __device__ Pre load(Cell) const { return {}; }
    template <class S> __device__ void apply(Cell i, int, int, int, const S& st, const Pre&, double* red) const {
#pragma clang fp contract(off)
        const C<R> p = st.c();
        const double pr = p.x, pi = p.y;
        const double ab = hypot(pr, pi);  // abs(in_var)**2 as the reference forms it
        const double a2 = ab * ab;
        const bool ok = !masked || a2 > 0.01;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (a < ncomp) {
                const double g = d1_ref<R, S::H>(st, a, den[a]).x;  // the real gradient array
                const double v = ok ? -(pi * g) / a2 : 1e3;
                double* dst = vel + a * n;
                dst[i.zb + i.own] = v;
                if constexpr (STATS) {
                    if (v < kValid) {
                        red[3 * a + 0] += 1.0;
                        red[3 * a + 1] += v;
                        red[3 * a + 2] += v * v;
                        red[9 + 2 * a] = v > red[9 + 2 * a] ? v : red[9 + 2 * a];
                        red[10 + 2 * a] = -v > red[10 + 2 * a] ? -v : red[10 + 2 * a];
                    }
                }
            }
        }
    }
};

template <typename R> using OpVelocityStats = OpVelocity<R, true>;

}  // namespace gpe
