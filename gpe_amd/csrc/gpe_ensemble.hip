// Batched ensemble: many independent members of different grid sizes
// advanced with RK4.  All members share the equation, order, boundaries and
// storage precision; each has its own grid, spacing and fields.
//
// Four ways to run a step (gpe_ens_step):
//   GPE_ENS_PER_MEMBER  one launch per member and stage (4 x members launches)
//   GPE_ENS_BATCHED     one launch per stage for all members: persistent
//                       blocks walk a work list of (member, tile) items
//   GPE_ENS_GRAPH       GPE_ENS_BATCHED captured once into a hipGraph of a
//                       chunk of steps and replayed
//   GPE_ENS_GROUPED     members packed into groups whose fields fit the
//                       256 MiB Infinity Cache; each group advances a chunk
//                       of steps back to back (members are independent, so
//                       the order of their steps is free) while its fields
//                       stay cache resident; the chunk is captured in a
//                       hipGraph and replayed
//   GPE_ENS_FUSED       the fused RK4 of a single context (two kernels of two
//                       stages each, 5 field accesses per cell) for all
//                       members: one persistent launch per stage pair over a
//                       work list of (member, fused block) items, largest
//                       members first; each member's state alternates between
//                       two buffers from step to step
//   GPE_ENS_FUSED_GRAPH GPE_ENS_FUSED captured in a hipGraph (an even number of
//                       steps) and replayed
//   GPE_ENS_PER_MEMBER_FUSED  the single-context fused RK4 launched once per
//                       member (2 x members launches): the baseline of the
//                       fused batched modes
// The fused modes need complex128 members whose right-hand side has no
// optional terms, as the single-context fused path does.
// Each member's arithmetic is the single-context kernel's (same operators,
// same per-member tiling), so results do not depend on the mode.
#include <math.h>

#include <algorithm>
#include <string>
#include <vector>

#include "gpe_fused.h"
#include "gpe_host.h"

using namespace gpe;

namespace gpe {

template <typename R> struct EnsMember {
    Geom g;
    StW<R> w;
    RhsC<R> k;
    C<R>* psi;
    C<R>* ua;
    C<R>* ub;
    C<R>* acc;
    int begin;  // first work item of the member
};

// One RK4 stage for every member.  STAGE 1..4 selects the operator and the
// fields exactly as step_fixed does for one context.
template <typename R, int ORD, int TX, int TY, int STAGE, bool GEN>
__global__ __launch_bounds__(TX* TY) GPE_OCC void k_rk4_batched(const EnsMember<R>* __restrict__ mem, int nmem,
                                                                 int total, C<R> a1, C<R> a2) {
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = nmem - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (mem[mid].begin <= item) lo = mid;
            else hi = mid - 1;
        }
        const EnsMember<R>& M = mem[lo];
        const int local = item - M.begin;
        if constexpr (STAGE == 1) {
            const OpRk4First<R> op{M.k, M.ua, a1};
            stencil_body<R, ORD, TX, TY, GEN>(M.g, M.w, M.psi, op, nullptr, local, 0);
        } else if constexpr (STAGE == 2) {
            const OpRk4Second<R> op{M.k, M.psi, M.ub, a1};
            stencil_body<R, ORD, TX, TY, GEN>(M.g, M.w, M.ua, op, nullptr, local, 0);
        } else if constexpr (STAGE == 3) {
            const OpRk4Third<R> op{M.k, M.psi, M.ua, M.acc, a1, a2};
            stencil_body<R, ORD, TX, TY, GEN>(M.g, M.w, M.ub, op, nullptr, local, 0);
        } else {
            const OpRk4Last<R> op{M.k, M.acc, M.psi, a1};
            stencil_body<R, ORD, TX, TY, GEN>(M.g, M.w, M.ua, op, nullptr, local, 0);
        }
        __syncthreads();  // the next item reuses the LDS tile
    }
}

template <typename R, int ORD, int TX, int TY, bool GEN>
static void launch_batched_stage(int stage, const EnsMember<R>* mem, int nmem, int total, int grid, C<R> a1, C<R> a2,
                                 hipStream_t s) {
    switch (stage) {
        case 1: hipLaunchKernelGGL((k_rk4_batched<R, ORD, TX, TY, 1, GEN>), dim3(grid), dim3(TX * TY), 0, s, mem, nmem, total, a1, a2); break;
        case 2: hipLaunchKernelGGL((k_rk4_batched<R, ORD, TX, TY, 2, GEN>), dim3(grid), dim3(TX * TY), 0, s, mem, nmem, total, a1, a2); break;
        case 3: hipLaunchKernelGGL((k_rk4_batched<R, ORD, TX, TY, 3, GEN>), dim3(grid), dim3(TX * TY), 0, s, mem, nmem, total, a1, a2); break;
        default: hipLaunchKernelGGL((k_rk4_batched<R, ORD, TX, TY, 4, GEN>), dim3(grid), dim3(TX * TY), 0, s, mem, nmem, total, a1, a2); break;
    }
}

// general: some member's right-hand side has optional terms (RhsC::flags != 0)
template <typename R, int ORD, int TX, int TY>
static hipError_t launch_batched_tile(int stage, bool general, const EnsMember<R>* mem, int nmem, int total, int grid,
                                      C<R> a1, C<R> a2, hipStream_t s) {
    if (general) launch_batched_stage<R, ORD, TX, TY, true>(stage, mem, nmem, total, grid, a1, a2, s);
    else launch_batched_stage<R, ORD, TX, TY, false>(stage, mem, nmem, total, grid, a1, a2, s);
    return hipGetLastError();
}

template <typename R, int ORD>
static hipError_t launch_batched(int tx, int ty, int stage, bool general, const EnsMember<R>* mem, int nmem, int total,
                                 int grid, C<R> a1, C<R> a2, hipStream_t s) {
    switch (tx * 100 + ty) {
        case 3208: return launch_batched_tile<R, ORD, 32, 8>(stage, general, mem, nmem, total, grid, a1, a2, s);
        case 3204: return launch_batched_tile<R, ORD, 32, 4>(stage, general, mem, nmem, total, grid, a1, a2, s);
        case 6404: return launch_batched_tile<R, ORD, 64, 4>(stage, general, mem, nmem, total, grid, a1, a2, s);
        case 6408: return launch_batched_tile<R, ORD, 64, 8>(stage, general, mem, nmem, total, grid, a1, a2, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace gpe

struct gpe_ens {
    int n = 0;
    std::vector<gpe_params> mp;
    std::vector<long long> cells;
    hipStream_t stream = nullptr;
    std::vector<void*> psi, ua, ub, acc;
    void* ddesc = nullptr;
    int tx = 32, ty = 8;
    std::vector<int> tz;
    int total = 0;
    int grid = 0;
    // grouped mode: members in descriptor order, groups of that order
    size_t group_budget = (size_t)192 << 20;
    int group_steps = 16;
    std::vector<int> order;               // descriptor slot -> member
    struct Group { int first, count, items, grid; };
    std::vector<Group> groups;
    void* gdesc = nullptr;                // descriptors in group order, begin relative to the group
    bool desc_dirty = true;
    bool general = false;  // some member's right-hand side has optional terms
    // graph of `gsteps` batched steps with dt (gre, gim)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gsteps = 0;
    int gmode = -1;
    double gre = 0, gim = 0;
    // fused modes: descriptors (largest member first), work items, grid, the
    // buffer holding each member's state (0: psi, 1: ub) and their own graph
    void* fdesc = nullptr;
    int ftotal = 0, fgrid = 0;
    bool fused_ok = false;
    int ncu = 256;  // compute units of the device (fused tiling)
    int fflip = 0;
    hipGraph_t fgraph = nullptr;
    hipGraphExec_t fgexec = nullptr;
    int fgsteps = 0;
    double fgre = 0, fgim = 0;
    std::string err;
};

namespace {
thread_local std::string g_ens_error;

int efail(gpe_ens* e, int code, const std::string& m) {
    if (e) e->err = m;
    g_ens_error = m;
    return code;
}

#define EHIP(e, expr)                                                                                              \
    do {                                                                                                           \
        hipError_t x_ = (expr);                                                                                    \
        if (x_ != hipSuccess) return efail(e, GPE_EHIP, std::string(#expr) + ": " + hipGetErrorString(x_));      \
    } while (0)

bool batched_tile_supported(int tx, int ty) {
    return (tx == 32 && (ty == 4 || ty == 8)) || (tx == 64 && (ty == 4 || ty == 8));
}

void drop_graph(gpe_ens* e) {
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
    e->gsteps = 0;
}

void drop_fused_graph(gpe_ens* e) {
    if (e->fgexec) (void)hipGraphExecDestroy(e->fgexec);
    if (e->fgraph) (void)hipGraphDestroy(e->fgraph);
    e->fgexec = nullptr;
    e->fgraph = nullptr;
    e->fgsteps = 0;
}

// Fused-mode descriptors, written in place (a captured graph keeps reading
// the same buffer).  z-chunks of 32 planes keep items of similar size; the
// members go largest first so the persistent blocks end together.
int build_fused_desc(gpe_ens* e) {
    e->fused_ok = e->mp[0].storage == GPE_C128 && !e->general;
    if (!e->fused_ok) return 0;
    int dev = 0, ncu = 0;
    EHIP(e, hipGetDevice(&dev));
    EHIP(e, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    e->ncu = ncu;
    std::vector<int> idx(e->n);
    for (int m = 0; m < e->n; ++m) idx[m] = m;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return e->cells[a] > e->cells[b]; });
    std::vector<FusedMember> h(e->n);
    int begin = 0;
    for (int s = 0; s < e->n; ++s) {
        const int m = idx[s];
        const gpe_params& p = e->mp[m];
        FusedMember& d = h[s];
        d.g = Geom{};
        d.g.nx = p.nx;
        d.g.ny = p.ny;
        d.g.nz = p.nz;
        d.g.bcs = p.bcs;
        d.g.plane = (long long)p.nx * p.ny;
        if (p.order == 2) fused_tiles<2>(p.nx, p.ny, p.nz, ncu, d.g);
        else fused_tiles<4>(p.nx, p.ny, p.nz, ncu, d.g);
        d.g.tz = p.nz < 32 ? p.nz : 32;
        d.g.nblocks = d.g.ntx * d.g.nty * ((p.nz + d.g.tz - 1) / d.g.tz);
        d.w = make_w<double>(p);
        d.k = make_rhs_p<double>(p, nullptr);
        d.psi = (double2*)e->psi[m];
        d.alt = (double2*)e->ub[m];
        d.u2 = (double2*)e->ua[m];
        d.begin = begin;
        begin += d.g.nblocks;
    }
    e->ftotal = begin;
    e->fgrid = ncu < begin ? ncu : begin;  // one 1024-thread block per CU (LDS-limited)
    if (!e->fdesc) EHIP(e, hipMalloc(&e->fdesc, sizeof(FusedMember) * e->n));
    EHIP(e, hipStreamSynchronize(e->stream));
    EHIP(e, hipMemcpy(e->fdesc, h.data(), sizeof(FusedMember) * e->n, hipMemcpyHostToDevice));
    return 0;
}

int default_tz(const gpe_params& p, int tx, int ty) {
    // about 8 planes per block keeps small members spread over many blocks
    (void)tx;
    (void)ty;
    return p.nz < 8 ? p.nz : 8;
}

template <typename R> int build_desc(gpe_ens* e) {
    std::vector<EnsMember<R>> h(e->n);
    int begin = 0;
    e->general = false;
    for (int m = 0; m < e->n; ++m) {
        const gpe_params& p = e->mp[m];
        EnsMember<R>& d = h[m];
        d.g = Geom{};
        d.g.nx = p.nx;
        d.g.ny = p.ny;
        d.g.nz = p.nz;
        d.g.bcs = p.bcs;
        d.g.plane = (long long)p.nx * p.ny;
        tile_geom(Tiles{e->tx, e->ty, e->tz[m]}, d.g);
        d.w = make_w<R>(p);
        d.k = make_rhs_p<R>(p, nullptr);
        d.g.general = d.k.flags != 0;
        e->general = e->general || d.g.general;
        d.psi = (C<R>*)e->psi[m];
        d.ua = (C<R>*)e->ua[m];
        d.ub = (C<R>*)e->ub[m];
        d.acc = (C<R>*)e->acc[m];
        d.begin = begin;
        begin += d.g.nblocks;
    }
    e->total = begin;
    if (!e->ddesc) EHIP(e, hipMalloc(&e->ddesc, sizeof(EnsMember<R>) * e->n));
    EHIP(e, hipMemcpy(e->ddesc, h.data(), sizeof(EnsMember<R>) * e->n, hipMemcpyHostToDevice));
    // persistent grid: enough resident blocks to fill the device, never more than the work
    int dev = 0, ncu = 0, per_cu = 0;
    EHIP(e, hipGetDevice(&dev));
    EHIP(e, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    per_cu = (1024 / (e->tx * e->ty)) * 2;
    e->grid = ncu * per_cu < e->total ? ncu * per_cu : e->total;

    // groups for GPE_ENS_GROUPED: members by increasing size, packed while the
    // group's four fields fit the budget (a larger member is a group alone)
    std::vector<int> idx(e->n);
    for (int m = 0; m < e->n; ++m) idx[m] = m;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return e->cells[a] < e->cells[b]; });
    const size_t esz = e->mp[0].storage == GPE_C128 ? 16 : 8;
    e->order = idx;
    e->groups.clear();
    std::vector<EnsMember<R>> g(e->n);
    size_t used = 0;
    for (int s = 0; s < e->n; ++s) {
        const int m = idx[s];
        const size_t ws = 4 * (size_t)e->cells[m] * esz;
        if (e->groups.empty() || used + ws > e->group_budget) {
            e->groups.push_back({s, 0, 0, 0});
            used = 0;
        }
        auto& G = e->groups.back();
        g[s] = h[m];
        g[s].begin = G.items;
        G.items += h[m].g.nblocks;
        G.count += 1;
        used += ws;
    }
    for (auto& G : e->groups) G.grid = ncu * per_cu < G.items ? ncu * per_cu : G.items;
    if (!e->gdesc) EHIP(e, hipMalloc(&e->gdesc, sizeof(EnsMember<R>) * e->n));
    EHIP(e, hipMemcpy(e->gdesc, g.data(), sizeof(EnsMember<R>) * e->n, hipMemcpyHostToDevice));
    e->desc_dirty = false;
    drop_graph(e);
    return build_fused_desc(e);
}

// One fused RK4 step of every member (two launches); toggles the state buffer.
int step_fused(gpe_ens* e, double dre, double dim) {
    const double2 dt = mkc<double>(dre, dim);
    const FusedMember* mem = (const FusedMember*)e->fdesc;
    const int ord = e->mp[0].order;
    for (int pair : {12, 34}) {
        hipError_t x = ord == 2 ? launch_rk4_fused_batched<2>(mem, e->n, e->ftotal, e->fgrid, dt, pair, e->fflip,
                                                              e->stream)
                                : launch_rk4_fused_batched<4>(mem, e->n, e->ftotal, e->fgrid, dt, pair, e->fflip,
                                                              e->stream);
        EHIP(e, x);
    }
    e->fflip ^= 1;
    return 0;
}

// After an odd number of fused steps the states sit in the alternate
// buffers: exchange the roles (host pointers and every descriptor).
template <typename R> int settle_fused(gpe_ens* e) {
    if (!e->fflip) return 0;
    for (int m = 0; m < e->n; ++m) std::swap(e->psi[m], e->ub[m]);
    e->fflip = 0;
    return build_desc<R>(e);
}

template <typename R> int ens_step_fused(gpe_ens* e, int mode, double dre, double dim, int nsteps) {
    if (!e->fused_ok)
        return efail(e, GPE_EINVAL, "gpe_ens_step: the fused modes need complex128 members whose right-hand side "
                                    "has no optional terms (trap, Urhs, dissipation, eqn 3)");
    int done = 0;
    if (mode == GPE_ENS_FUSED_GRAPH) {
        int chunk = nsteps < 50 ? nsteps : 50;
        chunk -= chunk % 2;  // a replay starts and ends with the states in psi
        if (chunk >= 2) {
            if (!e->fgexec || e->fgsteps != chunk || e->fgre != dre || e->fgim != dim) {
                drop_fused_graph(e);
                EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                int rc = 0;
                for (int s = 0; s < chunk && !rc; ++s) rc = step_fused(e, dre, dim);
                hipGraph_t g = nullptr;
                hipError_t x = hipStreamEndCapture(e->stream, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                EHIP(e, x);
                e->fgraph = g;
                EHIP(e, hipGraphInstantiate(&e->fgexec, e->fgraph, nullptr, nullptr, 0));
                e->fgsteps = chunk;
                e->fgre = dre;
                e->fgim = dim;
            }
            while (done + e->fgsteps <= nsteps) {
                EHIP(e, hipGraphLaunch(e->fgexec, e->stream));
                done += e->fgsteps;
            }
        }
    }
    for (; done < nsteps; ++done) {
        int rc = step_fused(e, dre, dim);
        if (rc) return rc;
    }
    return settle_fused<R>(e);
}

template <typename R>
int step_launches(gpe_ens* e, const EnsMember<R>* mem, int nmem, int total, int grid, double dre, double dim) {
    const int ord = e->mp[0].order;
    struct St { int stage; double s1, s2; } st[4] = {{1, 0.5, 0.0}, {2, 0.5, 0.0}, {3, 1.0, 1.0 / 3.0}, {4, 1.0 / 6.0, 0.0}};
    for (auto& s : st) {
        const C<R> a1 = cdt<R>(dre, dim, s.s1), a2 = cdt<R>(dre, dim, s.s2);
        hipError_t x = ord == 2 ? launch_batched<R, 2>(e->tx, e->ty, s.stage, e->general, mem, nmem, total, grid, a1, a2,
                                                       e->stream)
                                : launch_batched<R, 4>(e->tx, e->ty, s.stage, e->general, mem, nmem, total, grid, a1, a2,
                                                       e->stream);
        EHIP(e, x);
    }
    return 0;
}

template <typename R> int step_batched(gpe_ens* e, double dre, double dim) {
    return step_launches<R>(e, (const EnsMember<R>*)e->ddesc, e->n, e->total, e->grid, dre, dim);
}

// `steps` steps of every member, group after group
template <typename R> int steps_grouped(gpe_ens* e, int steps, double dre, double dim) {
    const EnsMember<R>* base = (const EnsMember<R>*)e->gdesc;
    for (const auto& G : e->groups)
        for (int s = 0; s < steps; ++s) {
            int rc = step_launches<R>(e, base + G.first, G.count, G.items, G.grid, dre, dim);
            if (rc) return rc;
        }
    return 0;
}

template <typename R, int ORD> int step_member(gpe_ens* e, int m, double dre, double dim) {
    const gpe_params& p = e->mp[m];
    const RhsC<R> k = make_rhs_p<R>(p, nullptr);
    const StW<R> w = make_w<R>(p);
    Geom g{};
    g.nx = p.nx;
    g.ny = p.ny;
    g.nz = p.nz;
    g.bcs = p.bcs;
    g.plane = (long long)p.nx * p.ny;
    g.general = k.flags != 0;
    const Tiles t{e->tx, e->ty, e->tz[m]};
    C<R>* psi = (C<R>*)e->psi[m];
    C<R>* ua = (C<R>*)e->ua[m];
    C<R>* ub = (C<R>*)e->ub[m];
    C<R>* acc = (C<R>*)e->acc[m];
    OpRk4First<R> s1{k, ua, cdt<R>(dre, dim, 0.5)};
    EHIP(e, (launch_stencil<R, ORD, OpRk4First<R>>(t, g, w, psi, s1, nullptr, e->stream)));
    OpRk4Second<R> s2{k, psi, ub, cdt<R>(dre, dim, 0.5)};
    EHIP(e, (launch_stencil<R, ORD, OpRk4Second<R>>(t, g, w, ua, s2, nullptr, e->stream)));
    OpRk4Third<R> s3{k, psi, ua, acc, cdt<R>(dre, dim, 1.0), cdt<R>(dre, dim, 1.0 / 3.0)};
    EHIP(e, (launch_stencil<R, ORD, OpRk4Third<R>>(t, g, w, ub, s3, nullptr, e->stream)));
    OpRk4Last<R> s4{k, acc, psi, cdt<R>(dre, dim, 1.0 / 6.0)};
    EHIP(e, (launch_stencil<R, ORD, OpRk4Last<R>>(t, g, w, ua, s4, nullptr, e->stream)));
    return 0;
}

// One fused RK4 step of member m with the single-context launch (the two
// fused kernels on the member's own grid, its default z-chunks); the new
// state lands in ub, which becomes psi.
template <int ORD> int step_member_fused(gpe_ens* e, int m, double dre, double dim) {
    const gpe_params& p = e->mp[m];
    Geom g{};
    g.nx = p.nx;
    g.ny = p.ny;
    g.nz = p.nz;
    g.bcs = p.bcs;
    g.plane = (long long)p.nx * p.ny;
    rk4_fused_tiles<ORD>(p.nx, p.ny, p.nz, e->ncu, g);
    const StW<double> w = make_w<double>(p);
    const RhsC<double> k = make_rhs_p<double>(p, nullptr);
    const double2 dt = mkc<double>(dre, dim);
    const double2* psi = (const double2*)e->psi[m];
    double2* u2 = (double2*)e->ua[m];
    EHIP(e, (launch_rk4_fused<double, ORD>(g, w, k, psi, u2, nullptr, dt, 12, e->stream)));
    EHIP(e, (launch_rk4_fused<double, ORD>(g, w, k, psi, u2, (double2*)e->ub[m], dt, 34, e->stream)));
    std::swap(e->psi[m], e->ub[m]);
    return 0;
}

template <typename R> int ens_step(gpe_ens* e, int mode, double dre, double dim, int nsteps) {
    if (e->desc_dirty) {
        int rc = build_desc<R>(e);
        if (rc) return rc;
    }
    if (mode == GPE_ENS_FUSED || mode == GPE_ENS_FUSED_GRAPH) return ens_step_fused<R>(e, mode, dre, dim, nsteps);
    if (mode == GPE_ENS_PER_MEMBER_FUSED) {
        if (!e->fused_ok)
            return efail(e, GPE_EINVAL, "gpe_ens_step: the fused modes need complex128 members whose right-hand "
                                        "side has no optional terms (trap, Urhs, dissipation, eqn 3)");
        for (int s = 0; s < nsteps; ++s)
            for (int m = 0; m < e->n; ++m) {
                int rc = e->mp[0].order == 2 ? step_member_fused<2>(e, m, dre, dim) : step_member_fused<4>(e, m, dre, dim);
                if (rc) return rc;
            }
        if (nsteps > 0) e->desc_dirty = true;  // the state buffers changed roles
        return 0;
    }
    if (mode == GPE_ENS_PER_MEMBER) {
        for (int s = 0; s < nsteps; ++s)
            for (int m = 0; m < e->n; ++m) {
                int rc = e->mp[0].order == 2 ? step_member<R, 2>(e, m, dre, dim) : step_member<R, 4>(e, m, dre, dim);
                if (rc) return rc;
            }
        return 0;
    }
    if (mode == GPE_ENS_BATCHED) {
        for (int s = 0; s < nsteps; ++s) {
            int rc = step_batched<R>(e, dre, dim);
            if (rc) return rc;
        }
        return 0;
    }
    // GPE_ENS_GRAPH / GPE_ENS_GROUPED: capture a chunk once per (mode, dt), replay it
    const bool grouped = mode == GPE_ENS_GROUPED;
    const int chunk = grouped ? (nsteps < e->group_steps ? nsteps : e->group_steps) : (nsteps < 50 ? nsteps : 50);
    if (chunk <= 0) return 0;
    if (!e->gexec || e->gsteps != chunk || e->gre != dre || e->gim != dim || e->gmode != mode) {
        drop_graph(e);
        EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = 0;
        if (grouped) rc = steps_grouped<R>(e, chunk, dre, dim);
        else
            for (int s = 0; s < chunk && !rc; ++s) rc = step_batched<R>(e, dre, dim);
        hipGraph_t g = nullptr;
        hipError_t x = hipStreamEndCapture(e->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        EHIP(e, x);
        e->graph = g;
        EHIP(e, hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
        e->gsteps = chunk;
        e->gre = dre;
        e->gim = dim;
        e->gmode = mode;
    }
    int done = 0;
    while (done + e->gsteps <= nsteps) {
        EHIP(e, hipGraphLaunch(e->gexec, e->stream));
        done += e->gsteps;
    }
    if (done < nsteps) {
        const int rest = nsteps - done;
        int rc = grouped ? steps_grouped<R>(e, rest, dre, dim) : 0;
        for (int s = 0; !grouped && s < rest && !rc; ++s) rc = step_batched<R>(e, dre, dim);
        if (rc) return rc;
    }
    return 0;
}

}  // namespace

extern "C" {

const char* gpe_ens_last_error(const gpe_ens* e) { return e ? e->err.c_str() : g_ens_error.c_str(); }

int gpe_ens_create(const gpe_params* members, int n, gpe_ens** out) {
    if (!out || !members || n <= 0) return efail(nullptr, GPE_EINVAL, "gpe_ens_create: bad arguments");
    *out = nullptr;
    const gpe_params& p0 = members[0];
    for (int m = 0; m < n; ++m) {
        const gpe_params& p = members[m];
        if (p.order != p0.order || p.bcs != p0.bcs || p.eqn_to_solve != p0.eqn_to_solve || p.storage != p0.storage)
            return efail(nullptr, GPE_EINVAL, "gpe_ens_create: members must share order, bcs, eqn_to_solve, storage");
        if (p.order != 2 && p.order != 4) return efail(nullptr, GPE_EINVAL, "gpe_ens_create: order must be 2 or 4");
        if (p.bcs != 1 && p.bcs != 2) return efail(nullptr, GPE_EINVAL, "gpe_ens_create: bcs must be 1 or 2");
        if (p.eqn_to_solve < 1 || p.eqn_to_solve > 4)
            return efail(nullptr, GPE_EINVAL, "gpe_ens_create: eqn_to_solve must be 1..4");
        if (p.storage != GPE_C128 && p.storage != GPE_C64)
            return efail(nullptr, GPE_EINVAL, "gpe_ens_create: storage must be GPE_C128 or GPE_C64");
        if (p.nx < 6 || p.ny < 6 || p.nz < 6) return efail(nullptr, GPE_EINVAL, "gpe_ens_create: nx, ny, nz must be >= 6");
        if ((long long)p.nx * p.ny >= (1LL << 28)) return efail(nullptr, GPE_EINVAL, "gpe_ens_create: nx*ny must be < 2^28");
        if (!(p.xr > 0) || !(p.yr > 0) || !(p.zr > 0)) return efail(nullptr, GPE_EINVAL, "gpe_ens_create: xr, yr, zr must be > 0");
        if (p.diss_amp != 0.0) return efail(nullptr, GPE_EINVAL, "gpe_ens_create: diss_amp is not supported in an ensemble");
    }
    gpe_ens* e = new gpe_ens();
    e->n = n;
    e->mp.assign(members, members + n);
    const size_t esz = p0.storage == GPE_C128 ? 16 : 8;
    e->psi.assign(n, nullptr);
    e->ua.assign(n, nullptr);
    e->ub.assign(n, nullptr);
    e->acc.assign(n, nullptr);
    e->tz.resize(n);
    hipError_t x = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    for (int m = 0; m < n && x == hipSuccess; ++m) {
        e->cells.push_back((long long)members[m].nx * members[m].ny * members[m].nz);
        const size_t b = (size_t)e->cells[m] * esz;
        e->tz[m] = default_tz(members[m], e->tx, e->ty);
        x = hipMalloc(&e->psi[m], b);
        if (x == hipSuccess) x = hipMalloc(&e->ua[m], b);
        if (x == hipSuccess) x = hipMalloc(&e->ub[m], b);
        if (x == hipSuccess) x = hipMalloc(&e->acc[m], b);
        if (x == hipSuccess) x = hipMemsetAsync(e->psi[m], 0, b, e->stream);
    }
    if (x == hipSuccess) x = hipStreamSynchronize(e->stream);
    if (x != hipSuccess) {
        int rc = efail(nullptr, x == hipErrorOutOfMemory ? GPE_ENOMEM : GPE_EHIP,
                       std::string("gpe_ens_create: ") + hipGetErrorString(x));
        gpe_ens_destroy(e);
        return rc;
    }
    *out = e;
    return 0;
}

void gpe_ens_destroy(gpe_ens* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    drop_fused_graph(e);
    if (e->fdesc) (void)hipFree(e->fdesc);
    for (int m = 0; m < e->n; ++m) {
        void* b[] = {e->psi[m], e->ua[m], e->ub[m], e->acc[m]};
        for (void* p : b)
            if (p) (void)hipFree(p);
    }
    if (e->ddesc) (void)hipFree(e->ddesc);
    if (e->gdesc) (void)hipFree(e->gdesc);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int gpe_ens_set_tiles(gpe_ens* e, int tx, int ty, const int* tz) {
    if (!e) return efail(nullptr, GPE_EINVAL, "gpe_ens_set_tiles: ens is NULL");
    if (!batched_tile_supported(tx, ty)) return efail(e, GPE_EINVAL, "gpe_ens_set_tiles: unsupported tile shape");
    for (int m = 0; m < e->n; ++m) {
        const int z = tz ? tz[m] : default_tz(e->mp[m], tx, ty);
        if (z < 1) return efail(e, GPE_EINVAL, "gpe_ens_set_tiles: tz must be >= 1");
        e->tz[m] = z > e->mp[m].nz ? e->mp[m].nz : z;
    }
    e->tx = tx;
    e->ty = ty;
    e->desc_dirty = true;
    return 0;
}

int gpe_ens_upload(gpe_ens* e, int m, const void* host, int host_dtype) {
    if (!e || !host || m < 0 || m >= e->n) return efail(e, GPE_EINVAL, "gpe_ens_upload: bad arguments");
    const bool c128 = e->mp[0].storage == GPE_C128;
    if ((host_dtype == GPE_HOST_C128) != c128)
        return efail(e, GPE_EINVAL, "gpe_ens_upload: host dtype must match the ensemble storage");
    EHIP(e, hipMemcpyAsync(e->psi[m], host, (size_t)e->cells[m] * (c128 ? 16 : 8), hipMemcpyHostToDevice, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    return 0;
}

int gpe_ens_download(gpe_ens* e, int m, void* host, int host_dtype) {
    if (!e || !host || m < 0 || m >= e->n) return efail(e, GPE_EINVAL, "gpe_ens_download: bad arguments");
    const bool c128 = e->mp[0].storage == GPE_C128;
    if ((host_dtype == GPE_HOST_C128) != c128)
        return efail(e, GPE_EINVAL, "gpe_ens_download: host dtype must match the ensemble storage");
    EHIP(e, hipMemcpyAsync(host, e->psi[m], (size_t)e->cells[m] * (c128 ? 16 : 8), hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    return 0;
}

int gpe_ens_step(gpe_ens* e, int mode, double dt_re, double dt_im, int nsteps) {
    if (!e) return efail(nullptr, GPE_EINVAL, "gpe_ens_step: ens is NULL");
    if (mode < GPE_ENS_PER_MEMBER || mode > GPE_ENS_PER_MEMBER_FUSED || nsteps < 0)
        return efail(e, GPE_EINVAL, "gpe_ens_step: bad mode or nsteps");
    return e->mp[0].storage == GPE_C128 ? ens_step<double>(e, mode, dt_re, dt_im, nsteps)
                                        : ens_step<float>(e, mode, dt_re, dt_im, nsteps);
}

int gpe_ens_work_items(const gpe_ens* e) { return e ? e->total : -1; }

int gpe_ens_set_grouping(gpe_ens* e, int64_t budget_bytes, int steps_per_visit) {
    if (!e) return efail(nullptr, GPE_EINVAL, "gpe_ens_set_grouping: ens is NULL");
    if (budget_bytes <= 0 || steps_per_visit <= 0) return efail(e, GPE_EINVAL, "gpe_ens_set_grouping: bad arguments");
    e->group_budget = (size_t)budget_bytes;
    e->group_steps = steps_per_visit;
    e->desc_dirty = true;
    return 0;
}

int gpe_ens_num_groups(gpe_ens* e) {
    if (!e) return efail(nullptr, GPE_EINVAL, "gpe_ens_num_groups: ens is NULL");
    if (e->desc_dirty) {
        int rc = e->mp[0].storage == GPE_C128 ? build_desc<double>(e) : build_desc<float>(e);
        if (rc) return rc;
    }
    return (int)e->groups.size();
}

void* gpe_ens_stream(gpe_ens* e) { return e ? (void*)e->stream : nullptr; }

int gpe_ens_sync(gpe_ens* e) {
    if (!e) return efail(nullptr, GPE_EINVAL, "gpe_ens_sync: ens is NULL");
    EHIP(e, hipStreamSynchronize(e->stream));
    return 0;
}

}  // extern "C"
