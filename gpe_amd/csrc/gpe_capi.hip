// C ABI of libgpehip.so (declared in include/gpe_hip.h): context, device
// memory, the stage sequences of euler/rk2/rk4/rkck, the rkqs controller,
// reductions and the host-array drop-ins of the reference's module routines.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <utility>
#include <type_traits>
#include <mutex>
#include <unordered_map>
#include <vector>

#include <hipfft/hipfft.h>

#include "../../include/gpe_hip.h"
#include "gpe_fused_mx.h"
#include "gpe_host.h"

using namespace gpe;

namespace {

thread_local std::string g_thread_error;

}  // namespace

// device and pinned host result slots (reductions' final values)
constexpr int kNResults = 16;

struct gpe_ctx {
    gpe_params p;
    int device = 0;
    int ncu = 256;        // compute units of the device (fused RK4 wave sizing)
    bool fused = true;    // fused RK4 where it applies (gpe_set_fused)
    hipStream_t stream = nullptr;
    long long n = 0;   // cells
    size_t esz = 0;    // bytes per complex value
    void* psi = nullptr;
    // RK temporaries
    void* uA = nullptr;
    void* uB = nullptr;
    void* acc = nullptr;
    // Cash-Karp
    void* f0 = nullptr;
    void* k2 = nullptr;
    void* k3 = nullptr;
    void* k4 = nullptr;
    void* k5 = nullptr;
    void* pnew = nullptr;
    // mixed storage (GPE_MIXED): complex64 increments K1..K3 (RK4) or
    // F0, K2..K5, D_A, D_B (RKCK)
    void* mx[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // diagnostics: velocity components (3 n doubles), histogram counts
    void* vel = nullptr;
    void* hist = nullptr;
    // spectral diagnostics: FFT of the state, its hipFFT plan, row sums
    void* fbuf = nullptr;
    hipfftHandle fplan = 0;
    bool fplan_ok = false;
    double* sres = nullptr;
    size_t sres_cap = 0;
    // line length: per-line records, carries, block partials (+ a complex128
    // copy of a complex64 state)
    double* llbuf = nullptr;
    void* llpsi = nullptr;
    size_t hist_cap = 0;
    bool f0_valid = false;
    bool cand_valid = false;
    // auxiliary field (gpe_field_upload): host arrays handed to the
    // diagnostics drop-ins land here, never in the state
    void* aux = nullptr;
    bool use_aux = false;
    // drop-in scratch
    void* sin = nullptr;
    void* sout = nullptr;
    void* stage = nullptr;  // host-dtype staging for conversions
    size_t stage_bytes = 0;
    // reductions
    double* partial = nullptr;
    size_t partial_cap = 0;  // doubles
    double* dres = nullptr;
    double* hres = nullptr;  // pinned, kNResults doubles
    void* dtab = nullptr;    // dissipation tables (storage precision)
    Tiles tiles{64, 8, 0};
    std::string err;
    // per-launch HIP event timing (gpe_prof_enable)
    struct ProfRec {
        int kind;
        double bytes;
        hipEvent_t a, b;
    };
    bool prof = false;
    std::vector<ProfRec> recs;
    // allocation bases of offset fields (see dalloc)
    std::unordered_map<void*, void*> bases;
    int nalloc = 0;
};

static const char* const kKindNames[GPE_NKINDS] = {
    "euler_or_rk2_stage1", "rk2_stage2", "rk4_stage1", "rk4_stage2", "rk4_stage4", "rkck_stage1",
    "rkck_stage2", "rkck_stage3", "rkck_stage4", "rkck_stage5", "rkck_stage6_error", "deriv", "get_rhs", "energy",
    "hamiltonian", "rk4_stage3", "mx_rk4_stage1", "mx_rk4_stage2_3", "mx_rk4_stage4", "mx_rkck_stage1",
    "mx_rkck_stage2", "mx_rkck_stage3", "mx_rkck_stage4", "mx_rkck_stage5", "mx_rkck_stage6_error",
    "rk4_fused12", "rk4_fused34", "momentum", "velocity", "rkck_fused12", "rkck_fused34", "rkck_fused56_error",
    "mx_rkck_fused12", "mx_rkck_fused34", "mx_rkck_fused56_error", "rk2_fused", "euler2_fused"};

static int fail(gpe_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    g_thread_error = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                                          \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess)                                                                                      \
            return fail(ctx, GPE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));                         \
    } while (0)

// Field allocations.  With GPE_PAD_FIELDS each allocation starts at a
// different offset (a multiple of 64 KiB + 256 B) so that equal indices of
// different fields do not share cache sets.
#ifndef GPE_PAD_FIELDS
#define GPE_PAD_FIELDS 1
#endif
static int dalloc(gpe_ctx* c, void** p, size_t bytes, const char* what) {
    if (*p) return 0;
    const size_t pad = GPE_PAD_FIELDS ? (size_t)(c->nalloc % 16) * (65536 + 256) : 0;
    void* base = nullptr;
    hipError_t e = hipMalloc(&base, bytes + (GPE_PAD_FIELDS ? (1u << 20) : 0));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, GPE_ENOMEM, std::string("hipMalloc(") + what + ", " + std::to_string(bytes) + " B): " +
                                       hipGetErrorString(e));
    }
    *p = (char*)base + pad;
    c->bases[*p] = base;
    c->nalloc++;
    return 0;
}

static void dfree(gpe_ctx* c, void* p) {
    if (!p) return;
    auto it = c->bases.find(p);
    (void)hipFree(it != c->bases.end() ? it->second : p);
    if (it != c->bases.end()) c->bases.erase(it);
}

static size_t field_bytes(const gpe_ctx* c) { return (size_t)c->n * c->esz; }

// The field the reductions, diagnostics, gpe_renormalise, gpe_imprint_line,
// gpe_peek and gpe_download act on (gpe_select_field); stepping always acts
// on the state.
static void* sel_field(const gpe_ctx* c) { return c->use_aux ? c->aux : c->psi; }

static int validate(const gpe_params* p, std::string& why) {
    if (!p) { why = "params is NULL"; return GPE_EINVAL; }
    const int mn = p->order == 4 ? 6 : 6;
    if (p->nx < mn || p->ny < mn || p->nz < mn) { why = "nx, ny, nz must be >= 6"; return GPE_EINVAL; }
    // the kernels address a plane with 32-bit byte offsets (16 B per value)
    // and mark lanes without access with the out-of-range offset 0xFFFFFFF0
    if ((long long)p->nx * p->ny >= (1LL << 28)) { why = "nx*ny must be < 2^28"; return GPE_EINVAL; }
    if (p->order != 2 && p->order != 4) { why = "order must be 2 or 4"; return GPE_EINVAL; }
    if (p->bcs != 1 && p->bcs != 2) { why = "bcs must be 1 (periodic) or 2 (reflective)"; return GPE_EINVAL; }
    if (p->eqn_to_solve < 1 || p->eqn_to_solve > 4) { why = "eqn_to_solve must be 1..4"; return GPE_EINVAL; }
    if (p->storage != GPE_C128 && p->storage != GPE_C64 && p->storage != GPE_MIXED) {
        why = "storage must be GPE_C128, GPE_C64 or GPE_MIXED";
        return GPE_EINVAL;
    }
    if (!(p->xr > 0) || !(p->yr > 0) || !(p->zr > 0)) { why = "xr, yr, zr must be > 0"; return GPE_EINVAL; }
    if (!(p->eps > 0) || !(p->safety > 0)) { why = "eps and safety must be > 0"; return GPE_EINVAL; }
    return 0;
}

// Default tiling: 64x16 tiles for wide planes (one 1024-thread block per CU
// marching long z-ranges, so neighbouring tiles advance together and share
// halo rows through L2), 32x8 tiles for small planes; tz is chosen so the
// launch has at least one block per CU (256), or two for complex64 storage,
// whose kernels are small enough for two 1024-thread blocks per CU (the
// second block's loads overlap the first one's barrier and compute).
static int tz_for(const gpe_params& p, int tx, int ty) {
    const long long ntxy = (long long)((p.nx + tx - 1) / tx) * ((p.ny + ty - 1) / ty);
    const long long want = p.storage == GPE_C64 ? 512 : 256;
    long long tz = (p.nz * ntxy) / want;
    if (tz < 4) tz = 4;
    if (tz > p.nz) tz = p.nz;
    return (int)tz;
}

static Tiles default_tiles(const gpe_params& p) {
    // mixed storage: two input streams per stencil load; 64x8 measured best
    Tiles t{64, p.storage == GPE_MIXED ? 8 : 16, 0};
    if ((long long)p.nx * p.ny < 128LL * 128LL) t = Tiles{32, 8, 0};
    t.tz = tz_for(p, t.tx, t.ty);
    return t;
}

template <typename R> static Geom make_geom(const gpe_ctx* c) {
    Geom g{};
    g.nx = c->p.nx;
    g.ny = c->p.ny;
    g.nz = c->p.nz;
    g.bcs = c->p.bcs;
    g.plane = (long long)c->p.nx * c->p.ny;
    g.general = make_rhs_p<R>(c->p, nullptr).flags != 0;
    return g;
}

template <typename R> static RhsC<R> make_rhs(const gpe_ctx* c) { return make_rhs_p<R>(c->p, c->dtab); }

template <typename R> static CkTab<R> make_tab() { return ck_tab<R>(); }

static int ensure_partial(gpe_ctx* c, size_t ndoubles) {
    if (c->partial_cap >= ndoubles) return 0;
    dfree(c, c->partial);
    c->partial = nullptr;
    c->partial_cap = 0;
    int rc = dalloc(c, (void**)&c->partial, ndoubles * sizeof(double), "partials");
    if (rc) return rc;
    c->partial_cap = ndoubles;
    return 0;
}

// Algorithmic field bytes per cell of an operator: kAccess values of the
// storage precision, or the operator's own kBytesPerCell (mixed storage).
template <class Op, class = void> struct OpBytes {
    static double get(size_t esz) { return (double)Op::kAccess * (double)esz; }
};
template <class Op> struct OpBytes<Op, std::void_t<decltype(Op::kBytesPerCell)>> {
    static double get(size_t) { return (double)Op::kBytesPerCell; }
};

// Launches one stencil pass (launch(g) enqueues the kernel) on the tiled
// geometry g, with the optional per-launch event timing and the final
// reduction of its partials.
template <class Op, class Launch> static int run_launch(gpe_ctx* c, Geom g, Launch launch) {
    const int nb = g.nblocks;
    if (Op::kNRed > 0) {
        int rc = ensure_partial(c, (size_t)nb * Op::kNRed);
        if (rc) return rc;
    }
    gpe_ctx::ProfRec rec{Op::kKind, OpBytes<Op>::get(c->esz) * (double)c->n, nullptr, nullptr};
    if (c->prof) {
        HIPCHK(c, hipEventCreate(&rec.a));
        HIPCHK(c, hipEventCreate(&rec.b));
        HIPCHK(c, hipEventRecord(rec.a, c->stream));
    }
    HIPCHK(c, launch(g));
    if (c->prof) {
        HIPCHK(c, hipEventRecord(rec.b, c->stream));
        c->recs.push_back(rec);
    }
    if constexpr (Op::kMode == RED_MIXED) {
        HIPCHK(c, launch_reduce_final(RED_SUM, c->partial, nb, Op::kNSum, c->dres, c->stream));
        HIPCHK(c, launch_reduce_final(RED_MAX, c->partial + (size_t)Op::kNSum * nb, nb, Op::kNRed - Op::kNSum,
                                      c->dres + Op::kNSum, c->stream));
    } else if (Op::kNRed > 0) {
        HIPCHK(c, launch_reduce_final(Op::kMode, c->partial, nb, Op::kNRed, c->dres, c->stream));
    }
    return 0;
}

template <typename R, int ORD, class Op> static int run_op(gpe_ctx* c, const void* u, const Op& op) {
    Geom g0 = make_geom<R>(c);
    tile_geom(c->tiles, g0);
    return run_launch<Op>(c, g0, [&](Geom g) {
        return launch_stencil<R, ORD, Op>(c->tiles, g, make_w<R>(c->p), (const C<R>*)u, op, c->partial, c->stream);
    });
}

// Stencil over psi0 + s d (mixed storage).
template <int ORD, class Op>
static int run_op_sum(gpe_ctx* c, const void* a, const void* d, double2 s, const Op& op) {
    const SumSrc src{(const double2*)a, (const float2*)d, s};
    Geom g0 = make_geom<double>(c);
    tile_geom(c->tiles, g0);
    return run_launch<Op>(c, g0, [&](Geom g) {
        return launch_stencil_sum<ORD, Op>(c->tiles, g, make_w<double>(c->p), src, op, c->partial, c->stream);
    });
}

// ---------------------------------------------------------------------------
// Steppers (src/solve.f90:57-154)

template <typename R, int ORD> static int step_fixed(gpe_ctx* c, int scheme, double dre, double dim) {
    const RhsC<R> k = make_rhs<R>(c);
    C<R>* psi = (C<R>*)c->psi;
    C<R>* uA = (C<R>*)c->uA;
    C<R>* uB = (C<R>*)c->uB;
    int rc;
    if (scheme == GPE_EULER) {
        // new = old + dt*rhs
        OpAxpySelf<R> op{k, uA, cdt<R>(dre, dim, 1.0)};
        if ((rc = run_op<R, ORD>(c, psi, op))) return rc;
        std::swap(c->psi, c->uA);
    } else if (scheme == GPE_RK2) {
        // tmp = old + 0.5 dt kk1 ; new = old + dt kk2
        OpAxpySelf<R> s1{k, uA, cdt<R>(dre, dim, 0.5)};
        if ((rc = run_op<R, ORD>(c, psi, s1))) return rc;
        OpAxpyP0<R> s2{k, psi, uB, cdt<R>(dre, dim, 1.0)};
        if ((rc = run_op<R, ORD>(c, uA, s2))) return rc;
        std::swap(c->psi, c->uB);
    } else {
        // classic RK4 in 13 field accesses (see OpRk4Third): u1 -> uA, u2 -> uB,
        // u3 over u1 in uA, accumulator in acc, psi_new over psi0
        C<R>* acc = (C<R>*)c->acc;
        OpRk4First<R> s1{k, uA, cdt<R>(dre, dim, 0.5)};
        if ((rc = run_op<R, ORD>(c, psi, s1))) return rc;
        OpRk4Second<R> s2{k, psi, uB, cdt<R>(dre, dim, 0.5)};
        if ((rc = run_op<R, ORD>(c, uA, s2))) return rc;
        OpRk4Third<R> s3{k, psi, uA, acc, cdt<R>(dre, dim, 1.0), cdt<R>(dre, dim, 1.0 / 3.0)};
        if ((rc = run_op<R, ORD>(c, uB, s3))) return rc;
        OpRk4Last<R> s4{k, acc, psi, cdt<R>(dre, dim, 1.0 / 6.0)};
        if ((rc = run_op<R, ORD>(c, uA, s4))) return rc;
    }
    return 0;
}

// Fused RK4 applies to complex128 storage with a plain right-hand side.
// Fused RK4 applies to complex128 and complex64 fields and a right-hand side
// whose only optional term is the trap V.
static bool fused_rk4(const gpe_ctx* c, int scheme) {
    const int f = make_rhs_p<double>(c->p, nullptr).flags;
    return c->fused && scheme == GPE_RK4 && (c->p.storage == GPE_C128 || c->p.storage == GPE_C64) &&
           (f == 0 || f == RHS_V);
}

// The fused RK2 step and the fused pair of Euler steps have the same
// conditions as fused RK4.
static bool fused_pair(const gpe_ctx* c, int scheme) {
    const int f = make_rhs_p<double>(c->p, nullptr).flags;
    return c->fused && (scheme == GPE_RK2 || scheme == GPE_EULER) &&
           (c->p.storage == GPE_C128 || c->p.storage == GPE_C64) && (f == 0 || f == RHS_V);
}

// The fused Cash-Karp attempt has the same conditions, and also applies to
// mixed storage.
static bool fused_ck(const gpe_ctx* c) {
    const int f = make_rhs_p<double>(c->p, nullptr).flags;
    return c->fused && (f == 0 || f == RHS_V);
}

static int need_rk_buffers(gpe_ctx* c, int scheme) {
    int rc;
    const size_t fb = field_bytes(c);
    if (fused_rk4(c, scheme)) {
        if ((rc = dalloc(c, &c->uA, fb, "uA"))) return rc;
        return dalloc(c, &c->pnew, fb, "pnew");
    }
    if (fused_pair(c, scheme)) return dalloc(c, &c->uA, fb, "uA");
    if (c->p.storage == GPE_MIXED && (scheme == GPE_RK4 || scheme == GPE_RK45)) {
        const int nmx = scheme == GPE_RK4 ? 3 : 7;
        for (int j = 0; j < nmx; ++j)
            if ((rc = dalloc(c, &c->mx[j], (size_t)c->n * sizeof(float2), "mixed increment"))) return rc;
        return dalloc(c, &c->pnew, fb, "pnew");
    }
    if ((rc = dalloc(c, &c->uA, fb, "uA"))) return rc;
    if (scheme == GPE_EULER) return 0;
    if ((rc = dalloc(c, &c->uB, fb, "uB"))) return rc;
    if (scheme == GPE_RK4 && (rc = dalloc(c, &c->acc, fb, "acc"))) return rc;
    if (scheme == GPE_RK45) {
        if ((rc = dalloc(c, &c->f0, fb, "f0"))) return rc;
        if ((rc = dalloc(c, &c->k2, fb, "k2"))) return rc;
        if ((rc = dalloc(c, &c->k3, fb, "k3"))) return rc;
        if ((rc = dalloc(c, &c->k4, fb, "k4"))) return rc;
        if (!fused_ck(c) && (rc = dalloc(c, &c->k5, fb, "k5"))) return rc;
        if ((rc = dalloc(c, &c->pnew, fb, "pnew"))) return rc;
    }
    return 0;
}

// One Cash-Karp attempt (src/solve.f90:248-378) with the fused error
// max-norm of rkqs (src/solve.f90:176-185).
template <typename R, int ORD>
static int ck_attempt(gpe_ctx* c, double dre, double dim, double d0re, double d0im, double out[2]) {
    const RhsC<R> k = make_rhs<R>(c);
    const CkTab<R> t = make_tab<R>();
    const C<R> dt = cdt<R>(dre, dim, 1.0);
    C<R>* psi = (C<R>*)c->psi;
    C<R>* uA = (C<R>*)c->uA;
    C<R>* uB = (C<R>*)c->uB;
    C<R>* f0 = (C<R>*)c->f0;
    C<R>* k2 = (C<R>*)c->k2;
    C<R>* k3 = (C<R>*)c->k3;
    C<R>* k4 = (C<R>*)c->k4;
    C<R>* k5 = (C<R>*)c->k5;
    int rc;
    if (!c->f0_valid) {
        OpCk1<R> s1{k, f0, uA, dt, t.b21};
        if ((rc = run_op<R, ORD>(c, psi, s1))) return rc;
        c->f0_valid = true;
    } else {
        HIPCHK(c, launch_ck1_retry<R>(psi, f0, uA, dt, t.b21, c->n, c->stream));
    }
    OpCkMid<R, 2> s2{k, psi, f0, nullptr, nullptr, nullptr, k2, uB, dt, t};
    if ((rc = run_op<R, ORD>(c, uA, s2))) return rc;
    OpCkMid<R, 3> s3{k, psi, f0, k2, nullptr, nullptr, k3, uA, dt, t};
    if ((rc = run_op<R, ORD>(c, uB, s3))) return rc;
    OpCkMid<R, 4> s4{k, psi, f0, k2, k3, nullptr, k4, uB, dt, t};
    if ((rc = run_op<R, ORD>(c, uA, s4))) return rc;
    OpCkMid<R, 5> s5{k, psi, f0, k2, k3, k4, k5, uA, dt, t};
    if ((rc = run_op<R, ORD>(c, uB, s5))) return rc;
    OpCkLast<R> s6{k, psi, f0, k3, k4, k5, (C<R>*)c->pnew, dt, d0re, d0im, t};
    if ((rc = run_op<R, ORD>(c, uA, s6))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->hres, c->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[0] = c->hres[0] / c->p.eps;
    out[1] = c->hres[1];
    c->cand_valid = true;
    return 0;
}

// Classic RK4 with complex64 increments (GPE_MIXED): K1 = dt f(psi0),
// K2 = dt f(psi0 + K1/2), K3 = dt f(psi0 + K2/2), then
// psi_new = psi0 + (K1 + 2 K2 + 2 K3 + dt f(psi0 + K3))/6 into pnew (swapped
// in).  144 field bytes per cell.
template <int ORD> static int step_rk4_mixed(gpe_ctx* c, double dre, double dim) {
    const RhsC<double> k = make_rhs<double>(c);
    const CD dt = cdt<double>(dre, dim, 1.0);
    CF* K1 = (CF*)c->mx[0];
    CF* K2 = (CF*)c->mx[1];
    CF* K3 = (CF*)c->mx[2];
    int rc;
    if ((rc = run_op<double, ORD>(c, c->psi, OpMxRk4S1{k, K1, dt}))) return rc;
    if ((rc = run_op_sum<ORD>(c, c->psi, K1, mkc<double>(0.5, 0.0), OpMxRk4S23{k, K2, dt}))) return rc;
    if ((rc = run_op_sum<ORD>(c, c->psi, K2, mkc<double>(0.5, 0.0), OpMxRk4S23{k, K3, dt}))) return rc;
    if ((rc = run_op_sum<ORD>(c, c->psi, K3, mkc<double>(1.0, 0.0), OpMxRk4Last{k, K1, K2, (CD*)c->pnew, dt})))
        return rc;
    std::swap(c->psi, c->pnew);
    return 0;
}

// One Cash-Karp attempt with complex64 increments (GPE_MIXED); same error
// norm and outputs as ck_attempt.  F0 = f(psi0) is kept across retries.
template <int ORD>
static int ck_attempt_mixed(gpe_ctx* c, double dre, double dim, double d0re, double d0im, double out[2]) {
    const RhsC<double> k = make_rhs<double>(c);
    const CkTab<double> t = make_tab<double>();
    const CD dt = cdt<double>(dre, dim, 1.0);
    CF* F0 = (CF*)c->mx[0];
    CF* K2 = (CF*)c->mx[1];
    CF* K3 = (CF*)c->mx[2];
    CF* K4 = (CF*)c->mx[3];
    CF* K5 = (CF*)c->mx[4];
    CF* DA = (CF*)c->mx[5];
    CF* DB = (CF*)c->mx[6];
    const double2 one = mkc<double>(1.0, 0.0);
    int rc;
    if (!c->f0_valid) {
        if ((rc = run_op<double, ORD>(c, c->psi, OpMxCk1{k, F0}))) return rc;
        c->f0_valid = true;
    }
    const double2 s21 = mkc<double>(t.b21 * dt.x, t.b21 * dt.y);
    if ((rc = run_op_sum<ORD>(c, c->psi, F0, s21, OpMxCk2{k, K2, DA, dt, t}))) return rc;
    if ((rc = run_op_sum<ORD>(c, c->psi, DA, one, OpMxCk3{k, F0, K2, nullptr, nullptr, K3, DB, dt, t}))) return rc;
    if ((rc = run_op_sum<ORD>(c, c->psi, DB, one, OpMxCk4{k, F0, K2, K3, nullptr, K4, DA, dt, t}))) return rc;
    if ((rc = run_op_sum<ORD>(c, c->psi, DA, one, OpMxCk5{k, F0, K2, K3, K4, K5, DB, dt, t}))) return rc;
    if ((rc = run_op_sum<ORD>(c, c->psi, DB, one, OpMxCkLast{k, F0, K3, K4, K5, (CD*)c->pnew, dt, d0re, d0im, t})))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->hres, c->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[0] = c->hres[0] / c->p.eps;
    out[1] = c->hres[1];
    c->cand_valid = true;
    return 0;
}

// Launch-accounting tags of the fused kernels (run_launch reads these).
struct FusedTag12 {
    static constexpr int kKind = 25, kAccess = 2, kNRed = 0, kMode = RED_NONE;
};
struct FusedTag34 {
    static constexpr int kKind = 26, kAccess = 3, kNRed = 0, kMode = RED_NONE;
};

// One classic RK4 step as two fused kernels (gpe_fused.h): psi -> u2 (uA),
// then (u2, psi) -> psi_new (pnew, swapped in).
template <typename R, int ORD> static int step_rk4_fused(gpe_ctx* c, double dre, double dim) {
    const RhsC<R> k = make_rhs<R>(c);
    const StW<R> w = make_w<R>(c->p);
    const C<R> dt = cdt<R>(dre, dim, 1.0);
    Geom g = make_geom<R>(c);
    rk4_fused_tiles<ORD>(g.nx, g.ny, g.nz, c->ncu, g);
    const C<R>* psi = (const C<R>*)c->psi;
    C<R>* u2 = (C<R>*)c->uA;
    C<R>* out = (C<R>*)c->pnew;
    int rc;
    if ((rc = run_launch<FusedTag12>(c, g, [&](Geom gg) {
             return launch_rk4_fused<R, ORD>(gg, w, k, psi, u2, nullptr, dt, 12, c->stream);
         })))
        return rc;
    if ((rc = run_launch<FusedTag34>(c, g, [&](Geom gg) {
             return launch_rk4_fused<R, ORD>(gg, w, k, psi, u2, out, dt, 34, c->stream);
         })))
        return rc;
    std::swap(c->psi, c->pnew);
    return 0;
}

struct FusedTagRk2 {
    static constexpr int kKind = 35, kAccess = 2, kNRed = 0, kMode = RED_NONE;
};
struct FusedTagEuler2 {
    static constexpr int kKind = 36, kAccess = 2, kNRed = 0, kMode = RED_NONE;
};

// One midpoint RK2 step (pk = PK_RK2) or two Euler steps (PK_EULER2) as one
// fused kernel (gpe_fused.h): psi -> uA, swapped in.
template <typename R, int ORD> static int step_pair_fused(gpe_ctx* c, double dre, double dim, int pk) {
    const RhsC<R> k = make_rhs<R>(c);
    const StW<R> w = make_w<R>(c->p);
    const C<R> dt = cdt<R>(dre, dim, 1.0);
    Geom g = make_geom<R>(c);
    rk4_fused_tiles<ORD>(g.nx, g.ny, g.nz, c->ncu, g);
    auto launch = [&](Geom gg) {
        return launch_pair_fused<R, ORD>(gg, w, k, (const C<R>*)c->psi, (C<R>*)c->uA, dt, pk, c->stream);
    };
    const int rc = pk == PK_RK2 ? run_launch<FusedTagRk2>(c, g, launch) : run_launch<FusedTagEuler2>(c, g, launch);
    if (rc) return rc;
    std::swap(c->psi, c->uA);
    return 0;
}

template <typename R> static int dispatch_pair(gpe_ctx* c, double dre, double dim, int pk) {
    return c->p.order == 2 ? step_pair_fused<R, 2>(c, dre, dim, pk) : step_pair_fused<R, 4>(c, dre, dim, pk);
}

struct FusedCkTag1 {
    static constexpr int kKind = 29, kAccess = 4, kNRed = 0, kMode = RED_NONE;
};
struct FusedCkTag2 {
    static constexpr int kKind = 30, kAccess = 7, kNRed = 0, kMode = RED_NONE;
};
struct FusedCkTag3 {
    static constexpr int kKind = 31, kAccess = 7, kNRed = 2, kMode = RED_MAX;
};
struct FusedMxCkTag1 {
    static constexpr int kKind = 32, kBytesPerCell = 40, kNRed = 0, kMode = RED_NONE;
};
struct FusedMxCkTag2 {
    static constexpr int kKind = 33, kBytesPerCell = 64, kNRed = 0, kMode = RED_NONE;
};
struct FusedMxCkTag3 {
    static constexpr int kKind = 34, kBytesPerCell = 72, kNRed = 2, kMode = RED_MAX;
};

// One mixed-storage Cash-Karp attempt as three fused kernels
// (gpe_fused_mx.h): psi -> F0, K2, D3 (DA); psi + D3 -> K3, K4, D5 (DB);
// psi + D5 -> psi_new (pnew) with the error max-norm, as ck_attempt_mixed.
template <int ORD>
static int ck_attempt_mixed_fused(gpe_ctx* c, double dre, double dim, double d0re, double d0im, double out[2]) {
    const RhsC<double> k = make_rhs<double>(c);
    const StW<double> w = make_w<double>(c->p);
    Geom g = make_geom<double>(c);
    fused_tiles<ORD>(g.nx, g.ny, g.nz, c->ncu, g);
    CkMxArgs a{};
    a.p0 = (const CD*)c->psi;
    a.f0 = (const CF*)c->mx[0];
    a.k2 = (const CF*)c->mx[1];
    a.k3 = (const CF*)c->mx[2];
    a.k4 = (const CF*)c->mx[3];
    a.dt = cdt<double>(dre, dim, 1.0);
    const CkTab<double> t = make_tab<double>();
    a.s21 = mkc<double>(t.b21 * a.dt.x, t.b21 * a.dt.y);
    a.dt0_re = d0re;
    a.dt0_im = d0im;
    CF* DA = (CF*)c->mx[5];
    CF* DB = (CF*)c->mx[6];
    int rc;
    CkMxArgs a1 = a;
    a1.f0out = (CF*)c->mx[0];
    a1.kb = (CF*)c->mx[1];
    a1.dnext = DA;
    if ((rc = run_launch<FusedMxCkTag1>(c, g, [&](Geom gg) {
             return launch_ckmx_fused<ORD>(gg, w, k, a1, 1, c->partial, c->stream);
         })))
        return rc;
    CkMxArgs a2 = a;
    a2.d = DA;
    a2.ka = (CF*)c->mx[2];
    a2.kb = (CF*)c->mx[3];
    a2.dnext = DB;
    if ((rc = run_launch<FusedMxCkTag2>(c, g, [&](Geom gg) {
             return launch_ckmx_fused<ORD>(gg, w, k, a2, 2, c->partial, c->stream);
         })))
        return rc;
    CkMxArgs a3 = a;
    a3.d = DB;
    a3.out = (CD*)c->pnew;
    if ((rc = run_launch<FusedMxCkTag3>(c, g, [&](Geom gg) {
             return launch_ckmx_fused<ORD>(gg, w, k, a3, 3, c->partial, c->stream);
         })))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->hres, c->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[0] = c->hres[0] / c->p.eps;
    out[1] = c->hres[1];
    c->f0_valid = true;
    c->cand_valid = true;
    return 0;
}

// One Cash-Karp attempt as three fused kernels (gpe_fused.h): psi -> f0, k2,
// u3 (uA); u3 -> k3, k4, u5 (uB); u5 -> psi_new (pnew) with the error
// max-norm, as ck_attempt.  A retry evaluates f(psi0) again inside the first
// kernel (bit for bit the stored f0), so the retry costs one field read more
// than reusing f0 would and no extra pass.
template <typename R, int ORD>
static int ck_attempt_fused(gpe_ctx* c, double dre, double dim, double d0re, double d0im, double out[2]) {
    const RhsC<R> k = make_rhs<R>(c);
    const StW<R> w = make_w<R>(c->p);
    Geom g = make_geom<R>(c);
    fused_tiles<ORD>(g.nx, g.ny, g.nz, c->ncu, g);
    CkPairArgs<R> a{};
    a.p0 = (const C<R>*)c->psi;
    a.f0 = (const C<R>*)c->f0;
    a.k2 = (const C<R>*)c->k2;
    a.k3 = (const C<R>*)c->k3;
    a.k4 = (const C<R>*)c->k4;
    a.dt = cdt<R>(dre, dim, 1.0);
    a.dt0_re = d0re;
    a.dt0_im = d0im;
    int rc;
    CkPairArgs<R> a1 = a;
    a1.u = (const C<R>*)c->psi;
    a1.f0out = (C<R>*)c->f0;
    a1.kb = (C<R>*)c->k2;
    a1.unext = (C<R>*)c->uA;
    if ((rc = run_launch<FusedCkTag1>(c, g, [&](Geom gg) {
             return launch_ck_fused<R, ORD>(gg, w, k, a1, 1, c->partial, c->stream);
         })))
        return rc;
    CkPairArgs<R> a2 = a;
    a2.u = (const C<R>*)c->uA;
    a2.ka = (C<R>*)c->k3;
    a2.kb = (C<R>*)c->k4;
    a2.unext = (C<R>*)c->uB;
    if ((rc = run_launch<FusedCkTag2>(c, g, [&](Geom gg) {
             return launch_ck_fused<R, ORD>(gg, w, k, a2, 2, c->partial, c->stream);
         })))
        return rc;
    CkPairArgs<R> a3 = a;
    a3.u = (const C<R>*)c->uB;
    a3.out = (C<R>*)c->pnew;
    if ((rc = run_launch<FusedCkTag3>(c, g, [&](Geom gg) {
             return launch_ck_fused<R, ORD>(gg, w, k, a3, 3, c->partial, c->stream);
         })))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->hres, c->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[0] = c->hres[0] / c->p.eps;
    out[1] = c->hres[1];
    c->f0_valid = true;
    c->cand_valid = true;
    return 0;
}

template <typename R> static int dispatch_fixed(gpe_ctx* c, int scheme, double dre, double dim) {
    return c->p.order == 2 ? step_fixed<R, 2>(c, scheme, dre, dim) : step_fixed<R, 4>(c, scheme, dre, dim);
}
template <typename R>
static int dispatch_attempt(gpe_ctx* c, double dre, double dim, double d0re, double d0im, double out[2]) {
    if (fused_ck(c))
        return c->p.order == 2 ? ck_attempt_fused<R, 2>(c, dre, dim, d0re, d0im, out)
                               : ck_attempt_fused<R, 4>(c, dre, dim, d0re, d0im, out);
    return c->p.order == 2 ? ck_attempt<R, 2>(c, dre, dim, d0re, d0im, out)
                           : ck_attempt<R, 4>(c, dre, dim, d0re, d0im, out);
}

static int attempt(gpe_ctx* c, double dre, double dim, double d0re, double d0im, double out[2]) {
    int rc = need_rk_buffers(c, GPE_RK45);
    if (rc) return rc;
    if (c->p.storage == GPE_MIXED) {
        if (fused_ck(c))
            return c->p.order == 2 ? ck_attempt_mixed_fused<2>(c, dre, dim, d0re, d0im, out)
                                   : ck_attempt_mixed_fused<4>(c, dre, dim, d0re, d0im, out);
        return c->p.order == 2 ? ck_attempt_mixed<2>(c, dre, dim, d0re, d0im, out)
                               : ck_attempt_mixed<4>(c, dre, dim, d0re, d0im, out);
    }
    return c->p.storage == GPE_C128 ? dispatch_attempt<double>(c, dre, dim, d0re, d0im, out)
                                    : dispatch_attempt<float>(c, dre, dim, d0re, d0im, out);
}

static void accept(gpe_ctx* c) {
    std::swap(c->psi, c->pnew);
    c->f0_valid = false;
    c->cand_valid = false;
}

static double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// rkqs (src/solve.f90:158-244): attempts until errmax <= 1, then the
// elementary step-size controller of the reference.
static int rkqs(gpe_ctx* c, gpe_step_info* io) {
    const gpe_params& p = c->p;
    const double errcon = pow(5.0 / p.safety, 1.0 / p.dt_increase);
    double dre = io->dt_re, dim = io->dt_im;
    const double d0re = dre, d0im = dim;  // psi_scal uses the dt held on entry
    double out[2] = {0, 0};
    int rc;
    for (;;) {
        const bool had_f0 = c->f0_valid;
        if ((rc = attempt(c, dre, dim, d0re, d0im, out))) return rc;
        io->n_rhs += had_f0 ? 5 : 6;
        if (!(out[0] > 1.0)) break;
        io->n_reject += 1;
        // dt_temp = safety*dt*errmax**dt_decrease; no more than a factor 10 decrease
        const double f = pow(out[0], p.dt_decrease);
        const double mt = hypot((p.safety * dre) * f, (p.safety * dim) * f);
        const double ad = hypot(dre, dim);
        const double mag = fsign(fmax(mt, 0.1 * ad), ad);
        if (!p.real_time) { dre = 0.0; dim = -mag; }
        else { dre = mag; dim = 0.0; }
        if (!isfinite(out[0]) || mag == 0.0) return fail(c, GPE_EINVAL, "rkqs: step size underflow or non-finite error");
    }
    double nre, nim;
    if (out[0] > errcon) {
        const double f = pow(out[0], p.dt_increase);
        nre = (p.safety * dre) * f;
        nim = (p.safety * dim) * f;
    } else {
        nre = 5.0 * dre;
        nim = 5.0 * dim;
    }
    io->t += dre;              // t = t + real(dt)
    io->dt_re = nre;           // dt = dt_next
    io->dt_im = nim;
    io->im_t += fabs(nim);     // im_t = im_t + abs(aimag(dt)) with the new dt (src/solve.f90:229)
    io->errmax = out[0];
    io->psi2max = out[1];
    io->n_accept += 1;
    accept(c);
    return 0;
}

// ---------------------------------------------------------------------------
// Host <-> device transfers

static size_t host_esz(int host_dtype) { return host_dtype == GPE_HOST_C64 ? 8 : 16; }

static int ensure_stage(gpe_ctx* c, size_t bytes) {
    if (c->stage_bytes >= bytes) return 0;
    dfree(c, c->stage);
    c->stage = nullptr;
    c->stage_bytes = 0;
    int rc = dalloc(c, &c->stage, bytes, "staging");
    if (rc) return rc;
    c->stage_bytes = bytes;
    return 0;
}

static int upload_to(gpe_ctx* c, void* dst, const void* host, int host_dtype, int64_t ny_alloc, int64_t y_off,
                     int64_t z_off) {
    if (!host) return fail(c, GPE_EINVAL, "upload: host pointer is NULL");
    if (host_dtype != GPE_HOST_C128 && host_dtype != GPE_HOST_C64) return fail(c, GPE_EINVAL, "upload: bad host dtype");
    if (ny_alloc < c->p.ny || y_off < 0 || y_off + c->p.ny > ny_alloc || z_off < 0)
        return fail(c, GPE_EINVAL, "upload: bad halo layout (ny_alloc, y_off, z_off)");
    const size_t he = host_esz(host_dtype);
    const bool same = (he == c->esz);
    void* target = dst;
    if (!same) {
        int rc = ensure_stage(c, (size_t)c->n * he);
        if (rc) return rc;
        target = c->stage;
    }
    const char* src = (const char*)host + ((size_t)z_off * ny_alloc + y_off) * c->p.nx * he;
    const size_t width = (size_t)c->p.nx * c->p.ny * he;
    const size_t spitch = (size_t)c->p.nx * ny_alloc * he;
    HIPCHK(c, hipMemcpy2DAsync(target, width, src, spitch, width, c->p.nz, hipMemcpyHostToDevice, c->stream));
    if (!same) {
        if (c->p.storage != GPE_C64) {
            HIPCHK(c, launch_from_c64<double>((const float2*)c->stage, (double2*)dst, c->n, c->stream));
        } else {
            HIPCHK(c, launch_from_c128<float>((const double2*)c->stage, (float2*)dst, c->n, c->stream));
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int download_from(gpe_ctx* c, const void* src, void* host, int host_dtype) {
    if (!host) return fail(c, GPE_EINVAL, "download: host pointer is NULL");
    if (host_dtype != GPE_HOST_C128 && host_dtype != GPE_HOST_C64) return fail(c, GPE_EINVAL, "download: bad host dtype");
    const size_t he = host_esz(host_dtype);
    const void* from = src;
    if (he != c->esz) {
        int rc = ensure_stage(c, (size_t)c->n * he);
        if (rc) return rc;
        if (c->p.storage != GPE_C64) {
            HIPCHK(c, launch_to_c64<double>((const double2*)src, (float2*)c->stage, c->n, c->stream));
        } else {
            HIPCHK(c, launch_to_c128<float>((const float2*)src, (double2*)c->stage, c->n, c->stream));
        }
        from = c->stage;
    }
    HIPCHK(c, hipMemcpyAsync(host, from, (size_t)c->n * he, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

template <typename R> static int build_dtab(gpe_ctx* c) {
    // 1 + tanh(s - a) tanh(s + a) with a = x(nx-3) etc. (src/solve.f90:397-409)
    const gpe_params& p = c->p;
    double d[3];
    spacing(p, d);
    const int ns[3] = {p.nx, p.ny, p.nz};
    const double lo[3] = {-p.xr, -p.yr, -p.zr};
    std::string tmp;
    R* h = new R[(size_t)p.nx + p.ny + p.nz];
    size_t o = 0;
    for (int ax = 0; ax < 3; ++ax) {
        const double a = lo[ax] + (double)(ns[ax] - 3) * d[ax];
        for (int i = 0; i < ns[ax]; ++i) {
            const double s = lo[ax] + (double)i * d[ax];
            h[o++] = (R)(1.0 + tanh(s - a) * tanh(s + a));
        }
    }
    int rc = dalloc(c, &c->dtab, o * sizeof(R), "diss tables");
    if (!rc) {
        hipError_t e = hipMemcpy(c->dtab, h, o * sizeof(R), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(c, GPE_EHIP, std::string("hipMemcpy(dtab): ") + hipGetErrorString(e));
    }
    delete[] h;
    return rc;
}

// ---------------------------------------------------------------------------
// Public entry points

template <typename R> static int deriv_impl(gpe_ctx* c, int op, const void* u, void* out, bool rhs) {
    if (rhs) {
        OpRhs<R> o{make_rhs<R>(c), (C<R>*)out};
        return c->p.order == 2 ? run_op<R, 2>(c, u, o) : run_op<R, 4>(c, u, o);
    }
    OpDeriv<R> o{(C<R>*)out, op};
    return c->p.order == 2 ? run_op<R, 2>(c, u, o) : run_op<R, 4>(c, u, o);
}

static int deriv_common(gpe_ctx* c, int op, const void* in, int64_t ny_alloc, int64_t y_off, int64_t z_off,
                        void* out, int host_dtype, bool rhs) {
    if (!out) return fail(c, GPE_EINVAL, "output pointer is NULL");
    int rc;
    if ((rc = dalloc(c, &c->sout, field_bytes(c), "deriv out"))) return rc;
    const void* u = c->psi;
    if (in) {
        if ((rc = dalloc(c, &c->sin, field_bytes(c), "deriv in"))) return rc;
        if ((rc = upload_to(c, c->sin, in, host_dtype, ny_alloc, y_off, z_off))) return rc;
        u = c->sin;
    }
    rc = c->p.storage != GPE_C64 ? deriv_impl<double>(c, op, u, c->sout, rhs) : deriv_impl<float>(c, op, u, c->sout, rhs);
    if (rc) return rc;
    return download_from(c, c->sout, out, host_dtype);
}

static int read_result(gpe_ctx* c, double* v) {
    HIPCHK(c, hipMemcpyAsync(c->hres, c->dres, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *v = c->hres[0];
    return 0;
}

static int norm_mass(gpe_ctx* c, double* v, int weighted) {
    if (!c) return fail(nullptr, GPE_EINVAL, "norm/mass: ctx is NULL");
    if (!v) return fail(c, GPE_EINVAL, "norm/mass: output is NULL");
    int rc = ensure_partial(c, kNormBlocks);
    if (rc) return rc;
    if (c->p.storage != GPE_C64) {
        HIPCHK(c, launch_norm_partial<double>((const double2*)sel_field(c), c->p.nx, c->p.ny, c->p.nz, weighted, c->partial,
                                              kNormBlocks, c->stream));
    } else {
        HIPCHK(c, launch_norm_partial<float>((const float2*)sel_field(c), c->p.nx, c->p.ny, c->p.nz, weighted, c->partial,
                                             kNormBlocks, c->stream));
    }
    HIPCHK(c, launch_reduce_final(RED_SUM, c->partial, kNormBlocks, 1, c->dres, c->stream));
    double s;
    if ((rc = read_result(c, &s))) return rc;
    double d[3];
    spacing(c->p, d);
    *v = s * d[0] * d[1] * d[2];
    return 0;
}

static void prof_clear(gpe_ctx* c) {
    for (auto& r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    c->recs.clear();
}

// ---------------------------------------------------------------------------
// gpe_energy without allocation.  The stencil pass of the energy writes one
// FP64 partial per tile block (the tile-block count grows with the grid);
// here those partials live in `scratch`, a static device array of
// kNoAllocPartials doubles (512 KiB, held by each device the library runs
// on until the process ends).  Host code never uses the array's name as an
// address: the first call on a device asks hipGetSymbolAddress for that
// device's copy and keeps the pointer per device.  Every block overwrites
// its own partial, so nothing needs clearing; calls that overlap on one
// device share the array and race.
constexpr int kNoAllocPartials = 65536;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocPartials];
static double* g_scratch_dev[kNoAllocDevices];
static std::mutex g_scratch_mu;

static int noalloc_scratch(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, GPE_EHIP, std::string("gpe_energy_noalloc: hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kNoAllocDevices) return fail(nullptr, GPE_EINVAL, "gpe_energy_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(nullptr, GPE_EHIP, std::string("gpe_energy_noalloc: hipGetSymbolAddress(scratch): ") +
                                               hipGetErrorString(e));
        g_scratch_dev[dev] = (double*)ptr;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

// out = out * a * b * c, in that order (gpe_energy's E * dx * dy * dz)
__global__ void k_scale_result(double* out, double a, double b, double c) {
    if (threadIdx.x == 0) out[0] = out[0] * a * b * c;
}

// The default tiling of the context (gpe_energy) for p and its block count.
template <typename R> static Geom noalloc_geom(const gpe_params& p, Tiles& t) {
    Geom g{};
    g.nx = p.nx;
    g.ny = p.ny;
    g.nz = p.nz;
    g.bcs = p.bcs;
    g.plane = (long long)p.nx * p.ny;
    g.general = make_rhs_p<R>(p, nullptr).flags != 0;
    t = default_tiles(p);
    tile_geom(t, g);
    return g;
}

template <typename R, int ORD>
static int energy_noalloc(const gpe_params& p, const void* psi, double* scr, double* out, hipStream_t s) {
    Tiles t;
    const Geom g = noalloc_geom<R>(p, t);
    const int nb = g.nblocks;
    OpEnergy<R> o;
    hipError_t e = launch_stencil<R, ORD, OpEnergy<R>>(t, g, make_w<R>(p), (const C<R>*)psi, o, scr, s);
    if (e == hipSuccess) e = launch_reduce_final(RED_SUM, scr, nb, 1, out, s);
    if (e == hipSuccess) {
        double d[3];
        spacing(p, d);
        hipLaunchKernelGGL(k_scale_result, dim3(1), dim3(64), 0, s, out, d[0], d[1], d[2]);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(nullptr, GPE_EHIP, std::string("gpe_energy_noalloc: ") + hipGetErrorString(e));
    return 0;
}

extern "C" {

int gpe_energy_noalloc(const gpe_params* p, const void* psi, double* energy, void* stream) {
    std::string why;
    if (validate(p, why)) return fail(nullptr, GPE_EINVAL, "gpe_energy_noalloc: " + why);
    if (!psi || !energy) return fail(nullptr, GPE_EINVAL, "gpe_energy_noalloc: psi and energy must be device pointers");
    Tiles t;
    const int nb = p->storage != GPE_C64 ? noalloc_geom<double>(*p, t).nblocks : noalloc_geom<float>(*p, t).nblocks;
    if (nb > kNoAllocPartials)
        return fail(nullptr, GPE_EINVAL, "gpe_energy_noalloc: the grid needs " + std::to_string(nb) +
                                             " block partials; the static scratch holds " +
                                             std::to_string(kNoAllocPartials));
    double* scr = nullptr;
    int rc = noalloc_scratch(&scr);
    if (rc) return rc;
    const hipStream_t s = (hipStream_t)stream;
    if (p->storage != GPE_C64)
        return p->order == 2 ? energy_noalloc<double, 2>(*p, psi, scr, energy, s)
                             : energy_noalloc<double, 4>(*p, psi, scr, energy, s);
    return p->order == 2 ? energy_noalloc<float, 2>(*p, psi, scr, energy, s)
                         : energy_noalloc<float, 4>(*p, psi, scr, energy, s);
}

int gpe_abi_version(void) { return GPE_ABI_VERSION; }

const char* gpe_last_error(const gpe_ctx* ctx) { return ctx ? ctx->err.c_str() : g_thread_error.c_str(); }

int gpe_ctx_create(const gpe_params* p, gpe_ctx** out) {
    if (!out) return fail(nullptr, GPE_EINVAL, "gpe_ctx_create: out is NULL");
    *out = nullptr;
    std::string why;
    if (validate(p, why)) return fail(nullptr, GPE_EINVAL, "gpe_ctx_create: " + why);
    gpe_ctx* c = new gpe_ctx();
    c->p = *p;
    c->n = (long long)p->nx * p->ny * p->nz;
    c->esz = p->storage == GPE_C64 ? 8 : 16;
    c->tiles = default_tiles(*p);
    int rc = 0;
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        rc = fail(nullptr, GPE_EHIP, std::string("gpe_ctx_create: ") + hipGetErrorString(e));
        delete c;
        return rc;
    }
    if (!rc) rc = dalloc(c, &c->psi, field_bytes(c), "psi");
    if (!rc) rc = dalloc(c, (void**)&c->dres, kNResults * sizeof(double), "results");
    if (!rc) {
        e = hipHostMalloc((void**)&c->hres, kNResults * sizeof(double), hipHostMallocDefault);
        if (e != hipSuccess) rc = fail(c, GPE_ENOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    if (!rc && p->diss_amp != 0.0 && (p->eqn_to_solve == 1 || p->eqn_to_solve == 3))
        rc = p->storage != GPE_C64 ? build_dtab<double>(c) : build_dtab<float>(c);
    if (!rc) {
        e = hipMemsetAsync(c->psi, 0, field_bytes(c), c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(c, GPE_EHIP, std::string("init: ") + hipGetErrorString(e));
    }
    if (rc) {
        g_thread_error = c->err;
        gpe_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

void gpe_ctx_destroy(gpe_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    prof_clear(c);
    void* bufs[] = {c->aux,   c->psi,   c->uA,   c->uB,   c->acc,   c->f0,   c->k2,   c->k3,   c->k4,    c->k5,
                    c->pnew,  c->sin,  c->sout, c->stage, c->partial, c->dres, c->dtab, c->mx[0], c->mx[1],
                    c->mx[2], c->mx[3], c->mx[4], c->mx[5], c->mx[6], c->vel, c->hist, c->fbuf, c->sres, c->llbuf, c->llpsi};
    for (void* b : bufs) dfree(c, b);
    if (c->fplan_ok) (void)hipfftDestroy(c->fplan);
    if (c->hres) (void)hipHostFree(c->hres);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int gpe_set_tiles(gpe_ctx* c, int tx, int ty, int tz) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_set_tiles: ctx is NULL");
    if (!tile_supported(tx, ty)) return fail(c, GPE_EINVAL, "gpe_set_tiles: unsupported tile shape");
    if (tz < 0) return fail(c, GPE_EINVAL, "gpe_set_tiles: tz must be >= 0");
    Tiles t{tx, ty, tz};
    if (tz == 0) t.tz = tz_for(c->p, tx, ty);
    if (t.tz > c->p.nz) t.tz = c->p.nz;
    c->tiles = t;
    return 0;
}

int gpe_set_fused(gpe_ctx* c, int on) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_set_fused: ctx is NULL");
    c->fused = on != 0;
    return 0;
}

int gpe_get_tiles(const gpe_ctx* c, int* tx, int* ty, int* tz) {
    if (!c || !tx || !ty || !tz) return fail(nullptr, GPE_EINVAL, "gpe_get_tiles: NULL argument");
    *tx = c->tiles.tx;
    *ty = c->tiles.ty;
    *tz = c->tiles.tz;
    return 0;
}

int gpe_upload(gpe_ctx* c, const void* host, int host_dtype, int64_t ny_alloc, int64_t y_off, int64_t z_off) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_upload: ctx is NULL");
    int rc = upload_to(c, c->psi, host, host_dtype, ny_alloc, y_off, z_off);
    if (!rc) {
        c->f0_valid = false;
        c->cand_valid = false;
    }
    return rc;
}

int gpe_field_upload(gpe_ctx* c, int field, const void* host, int host_dtype, int64_t ny_alloc, int64_t y_off,
                     int64_t z_off) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_field_upload: ctx is NULL");
    if (field == GPE_FIELD_STATE) return gpe_upload(c, host, host_dtype, ny_alloc, y_off, z_off);
    if (field != GPE_FIELD_AUX) return fail(c, GPE_EINVAL, "gpe_field_upload: field must be GPE_FIELD_STATE or GPE_FIELD_AUX");
    int rc = dalloc(c, &c->aux, field_bytes(c), "auxiliary field");
    if (rc) return rc;
    return upload_to(c, c->aux, host, host_dtype, ny_alloc, y_off, z_off);
}

int gpe_select_field(gpe_ctx* c, int field) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_select_field: ctx is NULL");
    if (field != GPE_FIELD_STATE && field != GPE_FIELD_AUX)
        return fail(c, GPE_EINVAL, "gpe_select_field: field must be GPE_FIELD_STATE or GPE_FIELD_AUX");
    if (field == GPE_FIELD_AUX && !c->aux)
        return fail(c, GPE_ESTATE, "gpe_select_field: no auxiliary field has been uploaded");
    c->use_aux = field == GPE_FIELD_AUX;
    return 0;
}

int gpe_set_real_time(gpe_ctx* c, int real_time) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_set_real_time: ctx is NULL");
    c->p.real_time = real_time != 0;
    return 0;
}

int gpe_download(gpe_ctx* c, void* host, int host_dtype) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_download: ctx is NULL");
    return download_from(c, sel_field(c), host, host_dtype);
}

int gpe_step(gpe_ctx* c, int scheme, int nsteps, gpe_step_info* io) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_step: ctx is NULL");
    if (!io) return fail(c, GPE_EINVAL, "gpe_step: io is NULL");
    if (scheme < GPE_EULER || scheme > GPE_RK45)
        return fail(c, GPE_EINVAL, "ERROR: Unrecognised time stepping scheme.");  // src/solve.f90:50-51
    if (nsteps < 0) return fail(c, GPE_EINVAL, "gpe_step: nsteps < 0");
    int rc = need_rk_buffers(c, scheme);
    if (rc) return rc;
    if (scheme == GPE_RK45) {
        for (int s = 0; s < nsteps; ++s)
            if ((rc = rkqs(c, io))) return rc;
        return 0;
    }
    const int nrhs = scheme == GPE_EULER ? 1 : scheme == GPE_RK2 ? 2 : 4;
    for (int s = 0; s < nsteps; ++s) {
        if (fused_pair(c, scheme) && (scheme == GPE_RK2 || s + 1 < nsteps)) {
            // one fused RK2 step, or two Euler steps in one kernel
            const int pk = scheme == GPE_RK2 ? PK_RK2 : PK_EULER2;
            rc = c->p.storage == GPE_C64 ? dispatch_pair<float>(c, io->dt_re, io->dt_im, pk)
                                         : dispatch_pair<double>(c, io->dt_re, io->dt_im, pk);
            if (rc) return rc;
            const int nst = scheme == GPE_RK2 ? 1 : 2;
            for (int q = 0; q < nst; ++q) {
                io->t += io->dt_re;
                io->im_t += fabs(io->dt_im);
                io->n_accept += 1;
                io->n_rhs += nrhs;
            }
            s += nst - 1;
            continue;
        }
        if (fused_rk4(c, scheme))
            rc = c->p.storage == GPE_C64
                     ? (c->p.order == 2 ? step_rk4_fused<float, 2>(c, io->dt_re, io->dt_im)
                                        : step_rk4_fused<float, 4>(c, io->dt_re, io->dt_im))
                     : (c->p.order == 2 ? step_rk4_fused<double, 2>(c, io->dt_re, io->dt_im)
                                        : step_rk4_fused<double, 4>(c, io->dt_re, io->dt_im));
        else if (c->p.storage == GPE_MIXED && scheme == GPE_RK4)
            rc = c->p.order == 2 ? step_rk4_mixed<2>(c, io->dt_re, io->dt_im)
                                 : step_rk4_mixed<4>(c, io->dt_re, io->dt_im);
        else
            rc = c->p.storage != GPE_C64 ? dispatch_fixed<double>(c, scheme, io->dt_re, io->dt_im)
                                         : dispatch_fixed<float>(c, scheme, io->dt_re, io->dt_im);
        if (rc) return rc;
        io->t += io->dt_re;             // t = t + real(dt)
        io->im_t += fabs(io->dt_im);    // im_t = im_t + abs(aimag(dt))
        io->n_accept += 1;
        io->n_rhs += nrhs;
    }
    c->f0_valid = false;
    c->cand_valid = false;
    HIPCHK(c, hipGetLastError());
    return 0;
}

int gpe_solver(gpe_ctx* c, int scheme, const void* var_in, int64_t ny_alloc, int64_t y_off, int64_t z_off,
               void* var_out, int host_dtype, gpe_step_info* io) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_solver: ctx is NULL");
    int rc = gpe_upload(c, var_in, host_dtype, ny_alloc, y_off, z_off);
    if (!rc) rc = gpe_step(c, scheme, 1, io);
    if (!rc) rc = gpe_download(c, var_out, host_dtype);
    return rc;
}

int gpe_rk45_attempt(gpe_ctx* c, double dt_re, double dt_im, double dt0_re, double dt0_im, double out[2]) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_rk45_attempt: ctx is NULL");
    if (!out) return fail(c, GPE_EINVAL, "gpe_rk45_attempt: out is NULL");
    if (!isfinite(dt_re) || !isfinite(dt_im) || (dt_re == 0.0 && dt_im == 0.0))
        return fail(c, GPE_EINVAL, "gpe_rk45_attempt: dt must be finite and non-zero");
    return attempt(c, dt_re, dt_im, dt0_re, dt0_im, out);
}

int gpe_rk45_accept(gpe_ctx* c) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_rk45_accept: ctx is NULL");
    if (!c->cand_valid) return fail(c, GPE_ESTATE, "gpe_rk45_accept: no attempt to accept");
    accept(c);
    return 0;
}

int gpe_deriv(gpe_ctx* c, int op, const void* in, int64_t ny_alloc, int64_t y_off, int64_t z_off, void* out,
              int host_dtype) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_deriv: ctx is NULL");
    if (op < GPE_D_X || op > GPE_D_LAP) return fail(c, GPE_EINVAL, "gpe_deriv: unknown operator");
    return deriv_common(c, op, in, ny_alloc, y_off, z_off, out, host_dtype, false);
}

int gpe_rhs(gpe_ctx* c, const void* in, int64_t ny_alloc, int64_t y_off, int64_t z_off, void* out, int host_dtype) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_rhs: ctx is NULL");
    return deriv_common(c, 0, in, ny_alloc, y_off, z_off, out, host_dtype, true);
}

int gpe_norm(gpe_ctx* c, double* norm) { return norm_mass(c, norm, 1); }
int gpe_mass(gpe_ctx* c, double* mass) { return norm_mass(c, mass, 0); }

int gpe_energy(gpe_ctx* c, double* energy) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_energy: ctx is NULL");
    if (!energy) return fail(c, GPE_EINVAL, "gpe_energy: output is NULL");
    int rc;
    if (c->p.storage != GPE_C64) {
        OpEnergy<double> o;
        rc = c->p.order == 2 ? run_op<double, 2>(c, sel_field(c), o) : run_op<double, 4>(c, sel_field(c), o);
    } else {
        OpEnergy<float> o;
        rc = c->p.order == 2 ? run_op<float, 2>(c, sel_field(c), o) : run_op<float, 4>(c, sel_field(c), o);
    }
    if (rc) return rc;
    double s;
    if ((rc = read_result(c, &s))) return rc;
    double d[3];
    spacing(c->p, d);
    *energy = s * d[0] * d[1] * d[2];  // E = E*dx*dy*dz (src/variables.f90:584)
    return 0;
}

int gpe_momentum(gpe_ctx* c, double mom[3]) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_momentum: ctx is NULL");
    if (!mom) return fail(c, GPE_EINVAL, "gpe_momentum: output is NULL");
    int rc;
    if (c->p.storage != GPE_C64) {
        OpMomentum<double> o{c->p.nx, c->p.ny, c->p.nz};
        rc = c->p.order == 2 ? run_op<double, 2>(c, sel_field(c), o) : run_op<double, 4>(c, sel_field(c), o);
    } else {
        OpMomentum<float> o{c->p.nx, c->p.ny, c->p.nz};
        rc = c->p.order == 2 ? run_op<float, 2>(c, sel_field(c), o) : run_op<float, 4>(c, sel_field(c), o);
    }
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->hres, c->dres, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double d[3];
    spacing(c->p, d);
    for (int a = 0; a < 3; ++a) mom[a] = c->hres[a] * d[0] * d[1] * d[2];  // integrate_x/y/z multiply by dx, dy, dz
    return 0;
}

// Velocity field into c->vel (OpVelocity), masked as get_pdf_velocity or not.
// stats: the same pass also reduces the statistics of get_pdf_velocity
// into c->dres (OpVelocity<R, true>)
static int velocity_field(gpe_ctx* c, int masked, int ncomp, bool stats = false) {
    int rc;
    if ((rc = dalloc(c, &c->vel, (size_t)c->n * 3 * sizeof(double), "velocity"))) return rc;
    double d[3];
    spacing(c->p, d);
    const double f = c->p.order == 2 ? 2.0 : 12.0;  // (2.0_pr*dx) or (12.0_pr*dx), src/derivs.f90:41-56
    const double den[3] = {f * d[0], f * d[1], f * d[2]};
    auto go = [&](auto o) {
        using R = typename decltype(o)::Real;
        return c->p.order == 2 ? run_op<R, 2>(c, sel_field(c), o) : run_op<R, 4>(c, sel_field(c), o);
    };
    if (c->p.storage != GPE_C64) {
        if (stats) return go(OpVelocity<double, true>{(double*)c->vel, c->n, masked, ncomp, {den[0], den[1], den[2]}});
        return go(OpVelocity<double, false>{(double*)c->vel, c->n, masked, ncomp, {den[0], den[1], den[2]}});
    }
    if (stats) return go(OpVelocity<float, true>{(double*)c->vel, c->n, masked, ncomp, {den[0], den[1], den[2]}});
    return go(OpVelocity<float, false>{(double*)c->vel, c->n, masked, ncomp, {den[0], den[1], den[2]}});
}

int gpe_velocity_pdf(gpe_ctx* c, int nbins, int64_t count[3], double mean[3], double stdev[3], double vmax[3],
                     double* pdf) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_velocity_pdf: ctx is NULL");
    if (!count || !mean || !stdev || !vmax || !pdf) return fail(c, GPE_EINVAL, "gpe_velocity_pdf: output is NULL");
    if (nbins < 2 || nbins > 4096 || (nbins & 1)) return fail(c, GPE_EINVAL, "gpe_velocity_pdf: nbins must be even, 2..4096");
    int rc;
    // the velocity field and, in the same pass, its statistics: dres rows
    // 0..8 count, sum, sum of squares per component, 9..14 max and -min
    if ((rc = velocity_field(c, 1, 3, true))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->hres, c->dres, 15 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double vm[3];
    for (int a = 0; a < 3; ++a) {
        const double n = c->hres[3 * a], s = c->hres[3 * a + 1], s2 = c->hres[3 * a + 2];
        count[a] = (int64_t)n;
        mean[a] = s / n;                                  // mean (src/variables.f90:827-848)
        stdev[a] = sqrt(s2 / n - mean[a] * mean[a]);      // stdev (src/variables.f90:852-875)
        const double mx = c->hres[9 + 2 * a], mn = -c->hres[10 + 2 * a];
        vm[a] = fabs(mx) > fabs(mn) ? fabs(mx) : fabs(mn);  // max_vel (src/variables.f90:441-442)
        vmax[a] = vm[a];
    }
    // histogram with the reference's bin edges
    const size_t hb = (size_t)3 * nbins * sizeof(unsigned long long);
    if (c->hist_cap < hb) {
        dfree(c, c->hist);
        c->hist = nullptr;
        if ((rc = dalloc(c, &c->hist, hb, "histogram"))) return rc;
        c->hist_cap = hb;
    }
    for (int a = 0; a < 3; ++a) c->hres[a] = vm[a];  // pinned staging
    HIPCHK(c, hipMemcpyAsync(c->dres, c->hres, 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->hist, 0, hb, c->stream));
    HIPCHK(c, launch_vel_hist((const double*)c->vel, c->n, nbins, c->dres, (unsigned long long*)c->hist, kNormBlocks,
                              c->stream));
    std::vector<unsigned long long> h((size_t)3 * nbins);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->hist, hb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; ++a) {
        unsigned long long total = 0;
        for (int j = 0; j < nbins; ++j) total += h[(size_t)a * nbins + j];
        for (int j = 0; j < nbins; ++j) pdf[(size_t)a * nbins + j] = (double)h[(size_t)a * nbins + j] / (double)total;
    }
    return 0;
}

int gpe_vcf(gpe_ctx* c, double* f) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_vcf: ctx is NULL");
    if (!f) return fail(c, GPE_EINVAL, "gpe_vcf: output is NULL");
    const int nx = c->p.nx;
    if (nx > 8192) return fail(c, GPE_EINVAL, "gpe_vcf: nx must be <= 8192");
    int rc;
    if ((rc = velocity_field(c, 0, 1))) return rc;
    const long long nlines = (long long)c->p.ny * c->p.nz;
    const int nb = (int)(nlines < 1024 ? nlines : 1024);
    if ((rc = ensure_partial(c, (size_t)nb * nx + nx))) return rc;
    double* out = c->partial + (size_t)nb * nx;
    HIPCHK(c, launch_vcf((const double*)c->vel, nx, nlines, c->partial, nb, out, c->stream));
    std::vector<double> q(nx);
    HIPCHK(c, hipMemcpyAsync(q.data(), out, nx * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // Qxx(r) / (num(r) ny nz), num(r) = nx - r, then f = Qxx / Qxx(0) (src/variables.f90:501-502)
    const double yz = (double)c->p.ny * (double)c->p.nz;
    std::vector<double> qn(nx);
    for (int r = 0; r < nx; ++r) qn[r] = q[r] / ((double)(nx - r) * yz);
    for (int r = 0; r < nx; ++r) f[r] = qn[r] / qn[0];
    return 0;
}

int gpe_hamiltonian(gpe_ctx* c, double* h) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_hamiltonian: ctx is NULL");
    if (!h) return fail(c, GPE_EINVAL, "gpe_hamiltonian: output is NULL");
    int rc;
    if (c->p.storage != GPE_C64) {
        RhsC<double> k = make_rhs<double>(c);
        OpHamiltonian<double> o{k, k.pim >= 0 ? 1.0 : -1.0};
        rc = c->p.order == 2 ? run_op<double, 2>(c, sel_field(c), o) : run_op<double, 4>(c, sel_field(c), o);
    } else {
        RhsC<float> k = make_rhs<float>(c);
        OpHamiltonian<float> o{k, k.pim >= 0 ? 1.0 : -1.0};
        rc = c->p.order == 2 ? run_op<float, 2>(c, sel_field(c), o) : run_op<float, 4>(c, sel_field(c), o);
    }
    if (rc) return rc;
    double s;
    if ((rc = read_result(c, &s))) return rc;
    double d[3];
    spacing(c->p, d);
    *h = s * d[0] * d[1] * d[2];
    return 0;
}

int gpe_renormalise(gpe_ctx* c, double norm) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_renormalise: ctx is NULL");
    if (!(norm > 0.0) || !isfinite(norm)) return fail(c, GPE_EINVAL, "gpe_renormalise: norm must be positive");
    const double snn = sqrt(c->p.nn), sn = sqrt(norm);
    if (c->p.storage != GPE_C64) {
        HIPCHK(c, launch_scale<double>((double2*)sel_field(c), c->n, snn, sn, c->stream));
    } else {
        HIPCHK(c, launch_scale<float>((float2*)sel_field(c), c->n, snn, sn, c->stream));
    }
    if (!c->use_aux) {
        c->f0_valid = false;
        c->cand_valid = false;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int gpe_imprint_line(gpe_ctx* c, const gpe_line* vl, double scal) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_imprint_line: ctx is NULL");
    if (!vl) return fail(c, GPE_EINVAL, "gpe_imprint_line: vl is NULL");
    if (vl->dir < 0 || vl->dir > 2) return fail(c, GPE_EINVAL, "gpe_imprint_line: dir must be 0 (x), 1 (y) or 2 (z)");
    LineParam lp{vl->x0, vl->y0, vl->z0, vl->amp1, vl->amp2, vl->ll, vl->sgn, scal, 0.0, 0.0, vl->dir,
                 vl->imprint_phase != 0};
    // zero-wavelength guard of get_r / get_theta (src/ic.f90:741-755)
    const double eps = 2.220446049250313e-16;
    if (fabs(vl->ll) < eps) {
        if (fabs(vl->amp1) >= eps || fabs(vl->amp2) >= eps)
            return fail(c, GPE_EINVAL, "ERROR: vortex line disturbance wavelength is zero, but amplitude is non-zero.");
        lp.safe1 = lp.safe2 = 1.0;
    }
    double d[3];
    spacing(c->p, d);
    const double lo[3] = {-c->p.xr, -c->p.yr, -c->p.zr};
    if (c->p.storage != GPE_C64) {
        HIPCHK(c, launch_imprint_line<double>((double2*)sel_field(c), c->p.nx, c->p.ny, c->n, lo, d, lp, c->stream));
    } else {
        HIPCHK(c, launch_imprint_line<float>((float2*)sel_field(c), c->p.nx, c->p.ny, c->n, lo, d, lp, c->stream));
    }
    if (!c->use_aux) {
        c->f0_valid = false;
        c->cand_valid = false;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int gpe_peek(gpe_ctx* c, int32_t i, int32_t j, int32_t k, double out[2]) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_peek: ctx is NULL");
    if (!out) return fail(c, GPE_EINVAL, "gpe_peek: output is NULL");
    if (i < 0 || i >= c->p.nx || j < 0 || j >= c->p.ny || k < 0 || k >= c->p.nz)
        return fail(c, GPE_EINVAL, "gpe_peek: cell outside the grid");
    const long long q = ((long long)k * c->p.ny + j) * c->p.nx + i;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->esz == 16) {
        HIPCHK(c, hipMemcpy(out, (const char*)sel_field(c) + q * 16, 16, hipMemcpyDeviceToHost));
    } else {
        float v[2];
        HIPCHK(c, hipMemcpy(v, (const char*)sel_field(c) + q * 8, 8, hipMemcpyDeviceToHost));
        out[0] = v[0];
        out[1] = v[1];
    }
    return 0;
}

int gpe_linelength(gpe_ctx* c, double* length) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_linelength: ctx is NULL");
    if (!length) return fail(c, GPE_EINVAL, "gpe_linelength: output is NULL");
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const long long nl = linelength_lines(ny, nz), nb = linelength_blocks(ny, nz);
    const long long nr = linelength_runs(ny, nz);
    int rc = dalloc(c, (void**)&c->llbuf, (size_t)(8 * nl + 7 * nr + 2 * nb + 1) * sizeof(double), "line length");
    if (rc) return rc;
    const double2* src = (const double2*)sel_field(c);
    if (c->p.storage == GPE_C64) {
        if ((rc = dalloc(c, &c->llpsi, (size_t)c->n * 16, "line length c128 copy"))) return rc;
        HIPCHK(c, launch_to_c128<float>((const float2*)sel_field(c), (double2*)c->llpsi, c->n, c->stream));
        src = (const double2*)c->llpsi;
    }
    double* last = c->llbuf;
    double* carry = last + 5 * nl;
    double* runs = carry + 3 * nl;
    double* partial = runs + 7 * nr;
    HIPCHK(c, launch_linelength(src, nx, ny, nz, last, carry, runs, partial, c->dres, c->stream));
    double s;
    if ((rc = read_result(c, &s))) return rc;
    double d[3];
    spacing(c->p, d);
    *length = s * d[0];  // linelength = linelength*dx (src/variables.f90:1141)
    return 0;
}

int gpe_condensed_particles(gpe_ctx* c, double row[7], double eta[7], int64_t nharm[7], double* F) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_condensed_particles: ctx is NULL");
    if (!row) return fail(c, GPE_EINVAL, "gpe_condensed_particles: row is NULL");
    const bool spec = eta || nharm || F;
    if (spec && !(eta && nharm && F))
        return fail(c, GPE_EINVAL, "gpe_condensed_particles: eta, nharm and F are all given or all NULL");
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const bool dbl = c->esz == 16;
    int rc = dalloc(c, &c->fbuf, (size_t)c->n * c->esz, "fft buffer");
    if (rc) return rc;
    if (!c->fplan_ok) {
        // row-major plan: z slowest, x fastest
        if (hipfftPlan3d(&c->fplan, nz, ny, nx, dbl ? HIPFFT_Z2Z : HIPFFT_C2C) != HIPFFT_SUCCESS)
            return fail(c, GPE_EHIP, "gpe_condensed_particles: hipfftPlan3d failed");
        c->fplan_ok = true;
    }
    if (hipfftSetStream(c->fplan, c->stream) != HIPFFT_SUCCESS)
        return fail(c, GPE_EHIP, "gpe_condensed_particles: hipfftSetStream failed");
    // a = FFT+(psi) (FFTW_BACKWARD), normalised by 1/sqrt(N) in the sums below
    HIPCHK(c, hipMemcpyAsync(c->fbuf, sel_field(c), (size_t)c->n * c->esz, hipMemcpyDeviceToDevice, c->stream));
    hipfftResult fr = dbl ? hipfftExecZ2Z(c->fplan, (hipfftDoubleComplex*)c->fbuf, (hipfftDoubleComplex*)c->fbuf,
                                          HIPFFT_BACKWARD)
                          : hipfftExecC2C(c->fplan, (hipfftComplex*)c->fbuf, (hipfftComplex*)c->fbuf, HIPFFT_BACKWARD);
    if (fr != HIPFFT_SUCCESS) return fail(c, GPE_EHIP, "gpe_condensed_particles: hipfftExec failed");
    const double V = (double)nx * (double)ny * (double)nz;
    double a0[2];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (dbl) {
        HIPCHK(c, hipMemcpy(a0, c->fbuf, 16, hipMemcpyDeviceToHost));
    } else {
        float f[2];
        HIPCHK(c, hipMemcpy(f, c->fbuf, 8, hipMemcpyDeviceToHost));
        a0[0] = f[0];
        a0[1] = f[1];
    }
    // n0 = |a(0,0,0)|^2 (src/io.f90:335-343)
    const double n0 = (a0[0] * a0[0] + a0[1] * a0[1]) / V;
    double M;
    if ((rc = norm_mass(c, &M, 0))) return rc;
    const int nb = 512;
    const int nhalf = nx / 2;
    // block partials, the k-grid symbols, then the per-k2 spectrum table (4-byte entries)
    const size_t tab_at = (size_t)nb * (2 * kSpectrumShells + nhalf + 1) + nx + ny + nz;
    if ((rc = ensure_partial(c, tab_at + (size_t)(spectrum_table_len(nx, ny, nz) + 1) / 2))) return rc;
    double d[3];
    spacing(c->p, d);
    const double rho0 = n0 / V;
    HIPCHK(c, launch_condensed_tot(nx, ny, nz, d, rho0, c->partial, nb, c->dres, c->stream));
    double tot;
    if ((rc = read_result(c, &tot))) return rc;
    // the mass.dat row after (time, im_t) (src/io.f90:375-386)
    const double temp = (M - n0) / tot;
    const double E0 = (1.0 / (2.0 * V)) * (M * M + (M - n0) * (M - n0));
    const double H = E0 + temp * (V - 1.0);
    const double vol = 8.0 * c->p.xr * c->p.yr * c->p.zr;
    const double temp2 = ((M / vol) - (n0 / V)) / tot;
    const double r7[7] = {M / vol, n0 / V, M, n0, temp, temp2, H / V};
    for (int q = 0; q < 7; ++q) row[q] = r7[q];
    if (!spec) return 0;
    const size_t nrows = 2 * kSpectrumShells + nhalf;
    if (c->sres_cap < nrows) {
        dfree(c, c->sres);
        c->sres = nullptr;
        c->sres_cap = 0;
        if ((rc = dalloc(c, (void**)&c->sres, nrows * sizeof(double), "spectrum sums"))) return rc;
        c->sres_cap = nrows;
    }
    if (dbl) {
        HIPCHK(c, launch_spectrum<double>((const double2*)c->fbuf, nx, ny, nz, nhalf, c->partial, nb, c->sres,
                                          (unsigned*)(c->partial + tab_at), c->stream));
    } else {
        HIPCHK(c, launch_spectrum<float>((const float2*)c->fbuf, nx, ny, nz, nhalf, c->partial, nb, c->sres,
                                         (unsigned*)(c->partial + tab_at), c->stream));
    }
    std::vector<double> h(nrows);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(h.data(), c->sres, nrows * sizeof(double), hipMemcpyDeviceToHost));
    // spectrum (src/io.f90:700-716): mean occupation per shell, integral distribution
    for (int m = 0; m < kSpectrumShells; ++m) {
        nharm[m] = (int64_t)h[kSpectrumShells + m];
        eta[m] = h[m] == 0.0 ? 0.0 : h[m] / (double)nharm[m];
    }
    double acc = 0.0;
    for (int kk = 0; kk < nhalf; ++kk) {
        acc += h[2 * kSpectrumShells + kk];
        F[kk] = acc;
    }
    return 0;
}

int gpe_prof_enable(gpe_ctx* c, int on) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_prof_enable: ctx is NULL");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_clear(c);
    c->prof = on != 0;
    return 0;
}

int gpe_prof_read(gpe_ctx* c, int kind, double* total_ms, int64_t* launches, double* bytes_per_launch) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_prof_read: ctx is NULL");
    if (kind < 0 || kind >= GPE_NKINDS || !total_ms || !launches || !bytes_per_launch)
        return fail(c, GPE_EINVAL, "gpe_prof_read: bad argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double ms = 0.0, bytes = 0.0;
    int64_t n = 0;
    for (auto& r : c->recs) {
        if (r.kind != kind) continue;
        float e = 0.f;
        HIPCHK(c, hipEventElapsedTime(&e, r.a, r.b));
        ms += e;
        bytes = r.bytes;
        ++n;
    }
    *total_ms = ms;
    *launches = n;
    *bytes_per_launch = bytes;
    return 0;
}

const char* gpe_kernel_kind_name(int kind) { return (kind >= 0 && kind < GPE_NKINDS) ? kKindNames[kind] : ""; }

int gpe_bench_copy(void* dst, const void* src, int64_t nbytes, void* stream) {
    if (!dst || !src || nbytes < 0 || (nbytes % 16) != 0)
        return fail(nullptr, GPE_EINVAL, "gpe_bench_copy: pointers must be non-NULL and nbytes a multiple of 16");
    hipError_t e = launch_copy16(dst, src, nbytes / 16, (hipStream_t)stream);
    if (e != hipSuccess) return fail(nullptr, GPE_EHIP, std::string("gpe_bench_copy: ") + hipGetErrorString(e));
    return 0;
}

int gpe_bench_stream5(void* c_out, void* d_out, const void* a, const void* b, const void* e, int64_t nbytes,
                      void* stream) {
    if (!c_out || !d_out || !a || !b || !e || nbytes < 0 || (nbytes % 16) != 0)
        return fail(nullptr, GPE_EINVAL, "gpe_bench_stream5: pointers must be non-NULL and nbytes a multiple of 16");
    hipError_t err = launch_stream5(c_out, d_out, a, b, e, nbytes / 16, (hipStream_t)stream);
    if (err != hipSuccess) return fail(nullptr, GPE_EHIP, std::string("gpe_bench_stream5: ") + hipGetErrorString(err));
    return 0;
}

void* gpe_state_ptr(gpe_ctx* c) { return c ? c->psi : nullptr; }
void* gpe_stream(gpe_ctx* c) { return c ? (void*)c->stream : nullptr; }

int gpe_sync(gpe_ctx* c) {
    if (!c) return fail(nullptr, GPE_EINVAL, "gpe_sync: ctx is NULL");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"
