// Device-side building blocks of the gfx950 GPE library: complex helpers,
// boundary index maps, the z-marching LDS stencil kernel and the fused
// per-stage operators.  Included by the per-precision kernel translation
// units and (for the operator structs only) by the host C-ABI code.
//
// Stencil kernel design (one kernel per RK stage):
//   * a workgroup owns a TX x TY tile of an x-y plane and marches over a
//     z-range of TZ planes;
//   * each thread keeps its own column's z-neighbours in a register queue
//     (2H+1 planes, H = order/2) plus one plane prefetched ahead;
//   * the x-y neighbours come from an LDS copy of the current plane with an
//     H-wide halo on the four sides (no corners: the operators are axis
//     stencils).  The centre cells come from the register queue, the halo
//     cells from global memory, prefetched one plane ahead;
//   * the LDS tile is double buffered, so one barrier per plane suffices;
//   * periodic or reflective boundaries are applied by mapping every global
//     index into the domain (the reference's ghost planes and index maps,
//     src/derivs.f90:276-398 and src/variables.f90:186-358, become index
//     arithmetic); tiles that straddle the domain edge load mapped values, so
//     any grid size works with any tile shape;
//   * the stage update (RK combination, error estimate, reductions) is fused
//     into the same pass, so each field is read and written once per stage.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpe {

template <typename R> struct cxt;
template <> struct cxt<double> { using T = double2; };
template <> struct cxt<float> { using T = float2; };
template <typename R> using C = typename cxt<R>::T;

template <typename R> __host__ __device__ __forceinline__ C<R> mkc(R a, R b) {
    C<R> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename R> __device__ __forceinline__ C<R> cadd(C<R> a, C<R> b) { return mkc<R>(a.x + b.x, a.y + b.y); }
template <typename R> __device__ __forceinline__ C<R> csub(C<R> a, C<R> b) { return mkc<R>(a.x - b.x, a.y - b.y); }
template <typename R> __device__ __forceinline__ C<R> cscale(R s, C<R> a) { return mkc<R>(s * a.x, s * a.y); }
template <typename R> __device__ __forceinline__ C<R> cmul(C<R> a, C<R> b) {
    return mkc<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// c + a*b (complex)
template <typename R> __device__ __forceinline__ C<R> cfma(C<R> a, C<R> b, C<R> c) {
    return mkc<R>(c.x + (a.x * b.x - a.y * b.y), c.y + (a.x * b.y + a.y * b.x));
}
// c + s*a (real s)
template <typename R> __device__ __forceinline__ C<R> caxpy(R s, C<R> a, C<R> c) {
    return mkc<R>(c.x + s * a.x, c.y + s * a.y);
}

// Streaming (non-temporal) access for the pointwise fields of a stage: they
// are touched once per kernel, so they bypass cache retention and leave L2 to
// the stencil field's halo re-reads.
#ifndef GPE_STREAM_NT
#define GPE_STREAM_NT 1
#endif
template <typename R> struct vec2;
template <> struct vec2<double> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct vec2<float> { typedef float T __attribute__((ext_vector_type(2))); };

template <typename R> __device__ __forceinline__ C<R> ld_stream(const C<R>* p) {
#if GPE_STREAM_NT
    const typename vec2<R>::T v = __builtin_nontemporal_load((const typename vec2<R>::T*)p);
    return mkc<R>(v.x, v.y);
#else
    return *p;
#endif
}
template <typename R> __device__ __forceinline__ void st_stream(C<R>* p, C<R> a) {
#if GPE_STREAM_NT
    typename vec2<R>::T v = {a.x, a.y};
    __builtin_nontemporal_store(v, (typename vec2<R>::T*)p);
#else
    *p = a;
#endif
}

// Map a global index into [0, n): periodic wrap (bcs = 1) or even reflection
// about the end nodes (bcs = 2): -1 -> 1, -2 -> 2, n -> n-2, n+1 -> n-3.
// Reflection is periodic with period 2(n-1), which makes the map total.
__device__ __forceinline__ int bcmap(int i, int n, int bcs) {
    if (bcs == 1) {
        i %= n;
        return i < 0 ? i + n : i;
    }
    const int p = 2 * (n - 1);
    i %= p;
    if (i < 0) i += p;
    return i >= n ? p - i : i;
}

// bcmap for an index at most n - 2 outside [0, n) (one fold), without the
// integer division: the z-marching kernels map a few planes around their
// range every plane, in scalar registers (n >= 6 and H <= 2 keep them in
// that band).
__device__ __forceinline__ int bcmap1(int i, int n, int bcs) {
    // both rules formed and selected (no branch on bcs)
    const int lo = bcs == 1 ? i + n : -i;
    const int hi = bcs == 1 ? i - n : 2 * (n - 1) - i;
    return i < 0 ? lo : (i >= n ? hi : i);
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD receives a contiguous range of tiles and neighbouring tiles share that
// XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    constexpr int NX = 8;
    const int xcd = bid % NX, k = bid / NX;
    const int q = nb / NX, r = nb % NX;
    return (xcd < r) ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

struct Geom {
    int nx, ny, nz;
    int bcs;
    int ntx, nty, tz, nblocks;
    long long plane;  // nx*ny
    int general;      // 1: the right-hand side has optional terms (RhsC::flags != 0)
};

// Offset in cells of plane m (0 <= m < nz): a 32 x 32-bit product, since a
// plane holds fewer than 2^28 cells (a signed 64-bit product takes seven
// scalar instructions).
__device__ __forceinline__ long long plane_off(const Geom& g, int m) {
    return (long long)((unsigned long long)(unsigned)m * (unsigned)g.plane);
}

// Stencil weights.  2nd order: d1 = (f+1 - f-1) w1, d2 = (f+1 - 2f + f-1) w2
// with w1 = 1/(2d), w2 = 1/d^2.  4th order: d1 = (8(f+1 - f-1) - (f+2 - f-2)) w1,
// d2 = (16(f+1 + f-1) - (f+2 + f-2) - 30 f) w2 with w1 = 1/(12d), w2 = 1/(12d^2)
// (src/derivs.f90:25-272).
template <typename R> struct StW {
    R w1x, w1y, w1z;
    R w2x, w2y, w2z;
    R w2c;  // (2 or 30) * (w2x + w2y + w2z)
};

// Right-hand side in coefficient form (src/solve.f90:414-437):
//   w = a*lap + (b + c|psi|^2 + d*V) psi
//   f = (0.5*diss + i*pim) * w + U dpsi/dx, times sphere() for eqn 3
// eqn 1: a=1, b=1, c=-1, d=-vscale, pim=1/2, diss term
// eqn 2: a=1, b=0, c=-1, d=0, pim=1
// eqn 4: a=-1/2, b=-mu, c=g, d=1, pim=-1
enum { RHS_V = 1, RHS_U = 2, RHS_DISS = 4, RHS_MASK = 8 };
template <typename R> struct RhsC {
    R a, b, c, d, pim, U;
    R vx, vy, vz;        // V = vx x^2 + vy y^2 + vz z^2 with vx = omx^2/2 (src/ic.f90:708-726)
    R xl, yl, zl, dx, dy, dz;
    R diss_amp;
    const R* dtab;       // nx + ny + nz entries: 1 + tanh(s - a) tanh(s + a) per axis (src/solve.f90:397-409)
    int nx, ny;
    int flags;
};

// Accessor for the stencil of one cell: LDS tile of the current plane plus
// the register z-window.
// GEN = false compiles the right-hand side without its optional terms
// (RhsC::flags == 0), which keeps their constants and table addresses out of
// the registers of the common kernels.
template <typename R, int ORD, int LW, bool GEN = true> struct Sten {
    static constexpr int H = ORD / 2;
    static constexpr bool kGeneral = GEN;
    // optional right-hand-side terms compiled in (each still gated by RhsC::flags)
    static constexpr int kTerms = GEN ? (RHS_V | RHS_U | RHS_DISS | RHS_MASK) : 0;
    const C<R>* L;
    int lc;
    C<R> q[2 * H + 1];
    StW<R> w;

    __device__ __forceinline__ C<R> c() const { return q[H]; }
    __device__ __forceinline__ C<R> ax(int o) const { return L[lc + o]; }
    __device__ __forceinline__ C<R> ay(int o) const { return L[lc + o * LW]; }

    __device__ __forceinline__ C<R> d1(C<R> p1, C<R> m1, C<R> p2, C<R> m2, R w1) const {
        if constexpr (ORD == 2) {
            return mkc<R>((p1.x - m1.x) * w1, (p1.y - m1.y) * w1);
        } else {
            return mkc<R>((R(8) * (p1.x - m1.x) - (p2.x - m2.x)) * w1, (R(8) * (p1.y - m1.y) - (p2.y - m2.y)) * w1);
        }
    }
    __device__ __forceinline__ C<R> d2(C<R> p1, C<R> m1, C<R> p2, C<R> m2, C<R> f0, R w2) const {
        if constexpr (ORD == 2) {
            return mkc<R>((p1.x + m1.x - R(2) * f0.x) * w2, (p1.y + m1.y - R(2) * f0.y) * w2);
        } else {
            return mkc<R>((R(16) * (p1.x + m1.x) - (p2.x + m2.x) - R(30) * f0.x) * w2,
                          (R(16) * (p1.y + m1.y) - (p2.y + m2.y) - R(30) * f0.y) * w2);
        }
    }
    // outer neighbours (+-2 for 4th order; for 2nd order d1/d2 ignore them)
    __device__ __forceinline__ C<R> p2x() const { return ax(H); }
    __device__ __forceinline__ C<R> m2x() const { return ax(-H); }
    __device__ __forceinline__ C<R> p2y() const { return ay(H); }
    __device__ __forceinline__ C<R> m2y() const { return ay(-H); }
    __device__ __forceinline__ C<R> p2z() const { return q[2 * H]; }
    __device__ __forceinline__ C<R> m2z() const { return q[0]; }

    __device__ __forceinline__ C<R> d1x() const { return d1(ax(1), ax(-1), p2x(), m2x(), w.w1x); }
    __device__ __forceinline__ C<R> d1y() const { return d1(ay(1), ay(-1), p2y(), m2y(), w.w1y); }
    __device__ __forceinline__ C<R> d1z() const { return d1(q[H + 1], q[H - 1], p2z(), m2z(), w.w1z); }
    __device__ __forceinline__ C<R> d2x() const { return d2(ax(1), ax(-1), p2x(), m2x(), c(), w.w2x); }
    __device__ __forceinline__ C<R> d2y() const { return d2(ay(1), ay(-1), p2y(), m2y(), c(), w.w2y); }
    __device__ __forceinline__ C<R> d2z() const { return d2(q[H + 1], q[H - 1], p2z(), m2z(), c(), w.w2z); }

    // Laplacian with the centre term folded: sum_axis w2 * (neighbour sums) - w2c * f.
    __device__ __forceinline__ C<R> lap() const {
        const C<R> f0 = c();
        const C<R> xp = ax(1), xm = ax(-1), yp = ay(1), ym = ay(-1), zp = q[H + 1], zm = q[H - 1];
        R re, im;
        if constexpr (ORD == 2) {
            re = w.w2x * (xp.x + xm.x) + w.w2y * (yp.x + ym.x) + w.w2z * (zp.x + zm.x) - w.w2c * f0.x;
            im = w.w2x * (xp.y + xm.y) + w.w2y * (yp.y + ym.y) + w.w2z * (zp.y + zm.y) - w.w2c * f0.y;
        } else {
            const C<R> xpp = ax(2), xmm = ax(-2), ypp = ay(2), ymm = ay(-2), zpp = q[4], zmm = q[0];
            re = w.w2x * (R(16) * (xp.x + xm.x) - (xpp.x + xmm.x)) + w.w2y * (R(16) * (yp.x + ym.x) - (ypp.x + ymm.x)) +
                 w.w2z * (R(16) * (zp.x + zm.x) - (zpp.x + zmm.x)) - w.w2c * f0.x;
            im = w.w2x * (R(16) * (xp.y + xm.y) - (xpp.y + xmm.y)) + w.w2y * (R(16) * (yp.y + ym.y) - (ypp.y + ymm.y)) +
                 w.w2z * (R(16) * (zp.y + zm.y) - (zpp.y + zmm.y)) - w.w2c * f0.y;
        }
        return mkc<R>(re, im);
    }
};

// Sten with the trap term V as its only optional term: keeps the registers of
// the other terms out of kernels that run the V-only right-hand side.
template <typename R, int ORD, int LW> struct StenV : Sten<R, ORD, LW, true> {
    static constexpr int kTerms = RHS_V;
};

// Folded form of a stage update u = q + h f for a right-hand side whose only
// optional term is the trap V (the fused kernels): f = i pim w with
// w = a lap + (b + c|psi|^2 + d V) psi.  pim (+-1/2, +-1) and a (1, -1/2) are
// powers of two, so u = q + G w with G = i pim h, and a folded into the
// Laplacian weights, gives the same w and the same product terms as the
// unfolded form; only the roundings of the complex product differ.  With a
// real h (real time) G is imaginary and the product takes two FMAs instead of
// a complex FMA's four and the rotation's two multiplies; with an imaginary h
// (imaginary time) G is real.
enum { GM_GEN = 0, GM_IM = 1, GM_RE = 2 };

// which of the three forms G = i pim h takes (h = 0 counts as real)
__host__ __device__ inline int g_mode(double hre, double him) { return him == 0.0 ? GM_IM : (hre == 0.0 ? GM_RE : GM_GEN); }

// A wave-uniform value moved to scalar registers (FP64 and FP32 arithmetic
// runs on the vector ALU, so a product of two kernel arguments would
// otherwise occupy vector registers for the whole kernel).
__device__ __forceinline__ float uni(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uni(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename R> __device__ __forceinline__ C<R> g_of(R pim, C<R> h) {
    return mkc<R>(uni(-pim * h.y), uni(pim * h.x));
}

// q + G w in the form GM
template <typename R, int GM> __device__ __forceinline__ C<R> gaxpy(C<R> G, C<R> w, C<R> q) {
    if constexpr (GM == GM_IM) {
        return mkc<R>(q.x - G.y * w.y, q.y + G.y * w.x);
    } else if constexpr (GM == GM_RE) {
        return mkc<R>(q.x + G.x * w.x, q.y + G.x * w.y);
    } else {
        return cfma<R>(G, w, q);
    }
}

// G w in the form GM (a right-hand side's increment dt f = G w).
template <typename R, int GM> __device__ __forceinline__ C<R> gmul(C<R> G, C<R> w) {
    if constexpr (GM == GM_IM) {
        return mkc<R>(-G.y * w.y, G.y * w.x);
    } else if constexpr (GM == GM_RE) {
        return mkc<R>(G.x * w.x, G.x * w.y);
    } else {
        return cmul<R>(G, w);
    }
}

// h f for a step h of the form GM (GM_IM: h real, GM_RE: h imaginary): the
// complex product without its zero terms, so the same rounding.
template <typename R, int GM> __device__ __forceinline__ C<R> hmul(C<R> h, C<R> f) {
    if constexpr (GM == GM_IM) {
        return mkc<R>(h.x * f.x, h.x * f.y);
    } else if constexpr (GM == GM_RE) {
        return mkc<R>(-h.y * f.y, h.y * f.x);
    } else {
        return cmul<R>(h, f);
    }
}

// f = i pim w
template <typename R> __device__ __forceinline__ C<R> rot_pim(R pim, C<R> w) { return mkc<R>(-pim * w.y, pim * w.x); }

// The Laplacian weights times a (exact: a is a power of two).
template <typename R> __device__ __forceinline__ StW<R> scale_w(StW<R> w, R a) {
    w.w2x = uni(w.w2x * a);
    w.w2y = uni(w.w2y * a);
    w.w2z = uni(w.w2z * a);
    w.w2c = uni(w.w2c * a);
    return w;
}

// w = a lap + (b + c|psi|^2 [+ d V]) psi with a already in the stencil's
// weights (scale_w); S is Sten<.., false> or StenV.
template <typename R, class S>
__device__ __forceinline__ C<R> rhs_w(const RhsC<R>& k, const S& st, int x, int y, int z) {
    static_assert((S::kTerms & ~RHS_V) == 0, "rhs_w: the trap V is the only optional term");
    const C<R> p = st.c();
    const C<R> L = st.lap();
    R s = k.b + k.c * (p.x * p.x + p.y * p.y);
    if constexpr ((S::kTerms & RHS_V) != 0) {
        if (k.flags & RHS_V) {
            const R xc = k.xl + R(x) * k.dx, yc = k.yl + R(y) * k.dy, zc = k.zl + R(z) * k.dz;
            s += k.d * (k.vx * (xc * xc) + k.vy * (yc * yc) + k.vz * (zc * zc));
        }
    }
    return mkc<R>(L.x + s * p.x, L.y + s * p.y);
}

template <typename R, class S>
__device__ __forceinline__ C<R> rhs_eval(const RhsC<R>& k, const S& st, int x, int y, int z) {
    const C<R> p = st.c();
    const C<R> L = st.lap();
    R s = k.b + k.c * (p.x * p.x + p.y * p.y);
    if constexpr (!S::kGeneral) {
        const C<R> w = mkc<R>(k.a * L.x + s * p.x, k.a * L.y + s * p.y);
        return mkc<R>(-k.pim * w.y, k.pim * w.x);
    } else {
        if (k.flags & RHS_V) {
            const R xc = k.xl + R(x) * k.dx, yc = k.yl + R(y) * k.dy, zc = k.zl + R(z) * k.dz;
            s += k.d * (k.vx * (xc * xc) + k.vy * (yc * yc) + k.vz * (zc * zc));
        }
        const C<R> w = mkc<R>(k.a * L.x + s * p.x, k.a * L.y + s * p.y);
        C<R> f = mkc<R>(-k.pim * w.y, k.pim * w.x);
        if constexpr ((S::kTerms & RHS_DISS) != 0) {
            if (k.flags & RHS_DISS) {
                // diss = diss_amp*(tx(x) + ty(y) + tz(z)); (0.5*(i + diss)) * w adds 0.5*diss*w
                const R diss = k.diss_amp * ((k.dtab[x] + k.dtab[k.nx + y]) + k.dtab[k.nx + k.ny + z]);
                f = caxpy<R>(R(0.5) * diss, w, f);
            }
        }
        if constexpr ((S::kTerms & RHS_U) != 0) {
            if (k.flags & RHS_U) f = caxpy<R>(k.U, st.d1x(), f);
        }
        if constexpr ((S::kTerms & RHS_MASK) == 0) return f;
        if (k.flags & RHS_MASK) {
            // sphere() (src/ic.f90:635-657): 0.5*(1 + tanh(r - 10 - 2))
            const R xc = k.xl + R(x) * k.dx, yc = k.yl + R(y) * k.dy, zc = k.zl + R(z) * k.dz;
            const R m = R(0.5) * (R(1) + tanh(sqrt(xc * xc + yc * yc + zc * zc) - R(12)));
            f = cscale<R>(m, f);
        }
        return f;
    }
}

}  // namespace gpe
