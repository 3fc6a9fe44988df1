// Line length of the zero lines of psi (linelength, src/variables.f90:879-1145).
//
// Per interior cell (i, j, k), each axis in 1..n-2, the reference looks on
// the cell's six faces for the crossing point of the zero lines of Re psi
// and Im psi (a "hit"), ends the cell with 1/2 on a face where one part
// vanishes at all four corners while the other changes sign, or with 1/4 on
// an edge where both vanish, and otherwise adds a |h1 - h2| for the first
// two hits (a = 1, 1/2 or 1/4 for hits on shared faces).  Its hit slots are
// local arrays that persist from cell to cell, so a cell with a single hit
// pairs it with the second hit of the last earlier cell (in k, j, i order)
// that recorded two.  That one serial dependency is carried here in three
// passes:
//   1. one wavefront per x-line (j, k) walks its cells in chunks of 64 and
//      sums their contributions, resolving a single-hit cell against the
//      last two-hit lane before it in the chunk (ballot + shuffle) or the
//      line's last two-hit cell so far; the single-hit cells before the
//      line's first two-hit cell ("leading" cells) need the carry and are
//      only counted.  It also records the line's last two-hit cell;
//   2. a scan over the lines in runs (run ends, one block over the runs, run
//      fills): the carry into each line;
//   3. one wavefront per line that has leading cells walks its start again,
//      chunk by chunk up to the chunk of its first two-hit cell, and adds
//      them against the carry.
// In pass 1 the face rules run only for the cells that are not quiet
// (ll_quiet), pooled over the block's four lines and evaluated on all of its
// threads.
// The cell arithmetic keeps the reference's operation order with
// contraction off, so every cell's contribution and hits are bit-identical
// to the reference's; only the order of the final sum differs.
#include <hip/hip_runtime.h>

#include "gpe_launch.h"

namespace gpe {

constexpr int kLlThreads = 256;  // 4 wavefronts = 4 lines per block

// corners of face n (u(1)..u(4)) as (di, dj, dk)
constexpr signed char kLlFace[6][4][3] = {
    {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}},  // 1 front  (k)
    {{0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}},  // 2 back   (k+1)
    {{1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}},  // 3 right  (i+1)
    {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}},  // 4 left   (i)
    {{0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}},  // 5 top    (j+1)
    {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}},  // 6 bottom (j)
};

struct LlCell {
    int nhits;     // hits recorded (before an ending face or edge)
    int ended;     // 0, 1 (face: 1/2), 2 (edge: 1/4)
    double h1[3];  // first hit
    double h2[3];  // second hit (nhits >= 2)
};

// Sign-change crossings of the zero line of w along the edges 1-2, 2-3,
// 3-4, 4-1 of a face: the first two distinct ones in (xs, ys)[0..1];
// returns how many were kept (the reference's lu - 1).
__device__ __forceinline__ int ll_cross(const double* w, double* xs, double* ys) {
#pragma clang fp contract(off)
    int n = 0;
    double px = 0.0, py = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double a = w[q], b = w[q + 1];
        if (a != 0.0 && a * b < 0.0) {
            double x, y;
            if (q == 0) {
                x = fabs(a / (a - b));
                y = 0.0;
            } else if (q == 1) {
                x = 1.0;
                y = fabs(a / (a - b));
            } else if (q == 2) {
                x = fabs(b / (a - b));
                y = 1.0;
            } else {
                x = 0.0;
                y = fabs(b / (a - b));
            }
            if (n == 0 || x != px || y != py) {
                if (n == 0) {
                    xs[0] = x;
                    ys[0] = y;
                } else if (n == 1) {
                    xs[1] = x;
                    ys[1] = y;
                }
                ++n;
            }
            px = x;
            py = y;
        }
    }
    return n;
}

// Re psi (or Im psi) of one strict sign at all eight corners of cell
// (i, j, k): ll_cell returns at once with no hit and no ending.
__device__ __forceinline__ bool ll_quiet(const double2* __restrict__ psi, int nx, int ny, int i, int j, int k) {
    bool upos = true, uneg = true, vpos = true, vneg = true;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const double2 z = psi[((long long)(k + (q >> 2)) * ny + (j + ((q >> 1) & 1))) * nx + (i + (q & 1))];
        upos = upos && z.x > 0.0;
        uneg = uneg && z.x < 0.0;
        vpos = vpos && z.y > 0.0;
        vneg = vneg && z.y < 0.0;
    }
    return upos || uneg || vpos || vneg;
}

__device__ LlCell ll_cell(const double2* __restrict__ psi, int nx, int ny, int i, int j, int k) {
#pragma clang fp contract(off)
    LlCell r;
    r.nhits = 0;
    r.ended = 0;
    for (int c = 0; c < 3; ++c) r.h1[c] = r.h2[c] = 0.0;
    double2 cr[2][2][2];
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di)
                cr[dk][dj][di] = psi[((long long)(k + dk) * ny + (j + dj)) * nx + (i + di)];
    // Re psi (or Im psi) of one strict sign at all eight corners: no face
    // has a zero, a crossing or a sign change of it, so no rule applies
    // (most cells of a field with sparse vortex lines end here)
    bool upos = true, uneg = true, vpos = true, vneg = true;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const double2 z = cr[q >> 2][(q >> 1) & 1][q & 1];
        upos = upos && z.x > 0.0;
        uneg = uneg && z.x < 0.0;
        vpos = vpos && z.y > 0.0;
        vneg = vneg && z.y < 0.0;
    }
    if (upos || uneg || vpos || vneg) return r;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
        double u[5], v[5];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double2 z = cr[kLlFace[n][c][2]][kLlFace[n][c][1]][kLlFace[n][c][0]];
            u[c] = z.x;
            v[c] = z.y;
        }
        u[4] = u[0];
        v[4] = v[0];
        const bool uzero = u[0] == 0.0 && u[1] == 0.0 && u[2] == 0.0 && u[3] == 0.0;
        const bool vzero = v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0 && v[3] == 0.0;
        const bool vchange = v[0] * v[1] < 0.0 || v[1] * v[2] < 0.0 || v[2] * v[3] < 0.0 || v[3] * v[0] < 0.0;
        const bool uchange = u[0] * u[1] < 0.0 || u[1] * u[2] < 0.0 || u[2] * u[3] < 0.0 || u[3] * u[0] < 0.0;
        if ((uzero && vchange) || (vzero && uchange)) {
            r.ended = 1;
            return r;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (u[q] == 0.0 && u[q + 1] == 0.0 && v[q] == 0.0 && v[q + 1] == 0.0) {
                r.ended = 2;
                return r;
            }
        // ll_cross keeps at most one crossing per edge with a sign change, so
        // a face with fewer than two such edges in u or in v has no hit:
        // skipped before any division
        int nu = 0, nv = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            nu += u[q] != 0.0 && u[q] * u[q + 1] < 0.0;
            nv += v[q] != 0.0 && v[q] * v[q + 1] < 0.0;
        }
        if (nu < 2 || nv < 2) continue;
        double xu[2], yu[2], xv[2], yv[2];
        if (ll_cross(u, xu, yu) < 2 || ll_cross(v, xv, yv) < 2) continue;
        double den = xv[1] * (yu[0] - yu[1]) + xv[0] * (yu[1] - yu[0]) + (xu[0] - xu[1]) * (yv[0] - yv[1]);
        if (den == 0.0) den = den + 0.0000001;
        const double xx = (xu[0] * (xv[1] * (yv[0] - yu[1]) + xv[0] * (yu[1] - yv[1])) +
                           xu[1] * (xv[1] * (yu[0] - yv[0]) + xv[0] * (yv[1] - yu[0]))) /
                          den;
        const double yy = (xv[1] * (yu[0] - yu[1]) * yv[0] + xu[0] * yu[1] * yv[0] - xv[0] * yu[0] * yv[1] -
                           xu[0] * yu[1] * yv[1] + xv[0] * yu[1] * yv[1] + xu[1] * yu[0] * (yv[1] - yv[0])) /
                          den;
        if (xx >= 0.0 && xx <= 1.0 && yy >= 0.0 && yy <= 1.0) {
            double h[3];
            switch (n) {
                case 0: h[0] = xx, h[1] = yy, h[2] = 0.0; break;
                case 1: h[0] = xx, h[1] = yy, h[2] = 1.0; break;
                case 2: h[0] = 1.0, h[1] = yy, h[2] = xx; break;
                case 3: h[0] = 0.0, h[1] = yy, h[2] = xx; break;
                case 4: h[0] = xx, h[1] = 1.0, h[2] = yy; break;
                default: h[0] = xx, h[1] = 0.0, h[2] = yy; break;
            }
            if (r.nhits == 0)
                for (int c = 0; c < 3; ++c) r.h1[c] = h[c];
            else if (r.nhits == 1)
                for (int c = 0; c < 3; ++c) r.h2[c] = h[c];
            ++r.nhits;
        }
    }
    return r;
}

// a |h1 - h2| of a cell that ended its face loop with hits
__device__ __forceinline__ double ll_pair(const double* h1, const double* h2) {
#pragma clang fp contract(off)
    const bool xs = h1[0] == h2[0] && (h1[0] == 0.0 || h1[0] == 1.0);
    const bool ys = h1[1] == h2[1] && (h1[1] == 0.0 || h1[1] == 1.0);
    const bool zs = h1[2] == h2[2] && (h1[2] == 0.0 || h1[2] == 1.0);
    double a = 1.0;
    if (xs || ys) {  // the reference tests y twice where z is meant
        a = a * 0.5;
        if ((xs && ys) || (ys && zs)) a = a * 0.5;
    }
    const double dx = h1[0] - h2[0], dy = h1[1] - h2[1], dz = h1[2] - h2[2];
    return a * sqrt(dx * dx + dy * dy + dz * dz);
}

struct LlGeom {
    int nx, ny, nz;
    long long nlines;  // (ny - 2) (nz - 2)
};

__device__ __forceinline__ void ll_line(const LlGeom& g, long long line, int& j, int& k) {
    j = 1 + (int)(line % (g.ny - 2));
    k = 1 + (int)(line / (g.ny - 2));
}

__device__ __forceinline__ int ll_last_lane(unsigned long long mask) { return 63 - __clzll(mask); }

constexpr int kLlWaves = kLlThreads / 64;
constexpr int kLlWaveBits = 28;  // pooled cell: i | wavefront << 28 (i < nx < 2^28)

// last[] entries per line: has, h2 (3), leading single-hit cells
constexpr int kLlLast = 5;

// Pass 2: carry[line] = the second hit of the last two-hit cell of the
// lines before it, (0, 0, 0) before any.  An exclusive "last entry with
// has != 0" scan over the lines, in runs of kLlRun lines: (a) each thread
// finds its run's last such line, (b) one block scans the runs, (c) each
// thread fills its run's lines.
constexpr int kLlRun = 16;
constexpr int kLlScan = 1024;
constexpr int kLlRunThreads = 256;

// the last entry in [l0, l1) of e (entries of STRIDE doubles: has, x, y, z,
// ...) with has != 0: groups of 8 from the end, their flags loaded together
template <int STRIDE>
__device__ __forceinline__ void ll_scan_last(const double* __restrict__ e, long long l0, long long l1, double& has,
                                             double& x, double& y, double& z) {
    has = 0.0, x = 0.0, y = 0.0, z = 0.0;
    for (long long g1 = l1; g1 > l0 && has == 0.0; g1 -= 8) {
        const long long g0 = g1 - 8 > l0 ? g1 - 8 : l0;
        double f[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) f[q] = g0 + q < g1 ? e[STRIDE * (g0 + q)] : 0.0;
#pragma unroll
        for (int q = 7; q >= 0; --q)
            if (has == 0.0 && f[q] != 0.0) {
                const long long lq = g0 + q;
                has = 1.0, x = e[STRIDE * lq + 1], y = e[STRIDE * lq + 2], z = e[STRIDE * lq + 3];
            }
    }
}

// out[3 l .. 3 l + 2] for l in [l0, l1): the value carried into entry l,
// starting from (x, y, z); groups of 8 entries loaded before the serial
// update, so their loads are in flight together
template <int STRIDE>
__device__ __forceinline__ void ll_scan_fill(const double* __restrict__ e, long long l0, long long l1, double x,
                                             double y, double z, double* __restrict__ out) {
    constexpr int kG = 8;
    long long l = l0;
    for (; l + kG <= l1; l += kG) {
        double v[kG][4];
#pragma unroll
        for (int q = 0; q < kG; ++q)
#pragma unroll
            for (int c = 0; c < 4; ++c) v[q][c] = e[STRIDE * (l + q) + c];
#pragma unroll
        for (int q = 0; q < kG; ++q) {
            out[3 * (l + q) + 0] = x, out[3 * (l + q) + 1] = y, out[3 * (l + q) + 2] = z;
            if (v[q][0] != 0.0) x = v[q][1], y = v[q][2], z = v[q][3];
        }
    }
    for (; l < l1; ++l) {
        out[3 * l + 0] = x, out[3 * l + 1] = y, out[3 * l + 2] = z;
        if (e[STRIDE * l] != 0.0) x = e[STRIDE * l + 1], y = e[STRIDE * l + 2], z = e[STRIDE * l + 3];
    }
}

// (a) rl[run] = (has, h2) of the run's last line with a two-hit cell
__global__ __launch_bounds__(kLlRunThreads) void k_ll_run_last(const double* __restrict__ last, long long nlines,
                                                               double* __restrict__ rl) {
    const long long run = (long long)blockIdx.x * kLlRunThreads + threadIdx.x;
    const long long l0 = run * kLlRun;
    if (l0 >= nlines) return;
    const long long l1 = l0 + kLlRun < nlines ? l0 + kLlRun : nlines;
    double has, x, y, z;
    ll_scan_last<kLlLast>(last, l0, l1, has, x, y, z);
    rl[4 * run + 0] = has, rl[4 * run + 1] = x, rl[4 * run + 2] = y, rl[4 * run + 3] = z;
}

// (b) one block: rin[run] = the value carried into each run
__global__ __launch_bounds__(kLlScan) void k_ll_run_scan(const double* __restrict__ rl, long long nruns,
                                                         double* __restrict__ rin) {
    __shared__ double part_last[kLlScan][4];
    __shared__ double part_in[kLlScan][3];
    const long long per = (nruns + kLlScan - 1) / kLlScan;
    const long long r0 = (long long)threadIdx.x * per, r1 = r0 + per < nruns ? r0 + per : nruns;
    double has, x, y, z;
    ll_scan_last<4>(rl, r0, r1 > r0 ? r1 : r0, has, x, y, z);
    part_last[threadIdx.x][0] = has;
    part_last[threadIdx.x][1] = x;
    part_last[threadIdx.x][2] = y;
    part_last[threadIdx.x][3] = z;
    __syncthreads();
    if (threadIdx.x == 0) {
        double cx = 0.0, cy = 0.0, cz = 0.0;
        for (int t = 0; t < kLlScan; ++t) {
            part_in[t][0] = cx, part_in[t][1] = cy, part_in[t][2] = cz;
            if (part_last[t][0] != 0.0) cx = part_last[t][1], cy = part_last[t][2], cz = part_last[t][3];
        }
    }
    __syncthreads();
    if (r1 > r0) ll_scan_fill<4>(rl, r0, r1, part_in[threadIdx.x][0], part_in[threadIdx.x][1], part_in[threadIdx.x][2], rin);
}

// (c) carry of every line of the run
__global__ __launch_bounds__(kLlRunThreads) void k_ll_run_fill(const double* __restrict__ last, long long nlines,
                                                               const double* __restrict__ rin,
                                                               double* __restrict__ carry) {
    const long long run = (long long)blockIdx.x * kLlRunThreads + threadIdx.x;
    const long long l0 = run * kLlRun;
    if (l0 >= nlines) return;
    const long long l1 = l0 + kLlRun < nlines ? l0 + kLlRun : nlines;
    ll_scan_fill<kLlLast>(last, l0, l1, rin[3 * run + 0], rin[3 * run + 1], rin[3 * run + 2], carry);
}

// Segments of kLlSeg chunks: per segment each wavefront ranks the
// non-quiet cells of its line (most cells are quiet, but nearly every chunk
// of 64 holds a few that are not, and a chunk evaluated lane by lane runs
// the branchy face rules for all of its lanes), the block lays the four
// lists end to end in LDS (entry: i and the wavefront) and evaluates them on
// all threads, adding the contributions that need no
// neighbour (ended cells, two-hit cells) and keeping in LDS what the ordered
// pass needs: whether a cell has two hits (its h2 feeds later single-hit
// cells) or one (its h1).  Each wavefront then resolves its line's
// single-hit cells chunk by chunk in cell order.
constexpr int kLlSeg = 8;

// Pass 1: partial[block] = the block's sum of the contributions of all but
// the leading single-hit cells; last[line] = (1, h2) of the line's last
// two-hit cell, or (0, ...), then the number of leading cells.
__global__ __launch_bounds__(kLlThreads) void k_ll_sum(const double2* __restrict__ psi, LlGeom g,
                                                       double* __restrict__ last, double* __restrict__ partial) {
    constexpr int NP = kLlWaves * kLlSeg * 64;
    __shared__ double red[kLlWaves];
    __shared__ int pool_i[NP];          // pooled cells (i | wavefront << 28), then their flags
    __shared__ double pool_h[3][NP];    // h2 of a two-hit cell, h1 of a single-hit one
    __shared__ int cnt[kLlWaves];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long long line = (long long)blockIdx.x * kLlWaves + wv;
    const bool valid = line < g.nlines;  // uniform per wavefront
    const unsigned long long below = (1ull << lane) - 1ull;
    int j = 1, k = 1;
    if (valid) ll_line(g, line, j, k);
    // the line's last two-hit cell so far (seen: uniform per wavefront)
    double cx = 0.0, cy = 0.0, cz = 0.0;
    bool seen = false;
    int nlead = 0;
    double acc = 0.0;
    for (int s0 = 1; s0 <= g.nx - 2; s0 += 64 * kLlSeg) {  // uniform per block
        unsigned long long m[kLlSeg];
        int base[kLlSeg];
        int nw = 0;
#pragma unroll
        for (int c = 0; c < kLlSeg; ++c) {
            const int i = s0 + 64 * c + lane;
            m[c] = __ballot(valid && i <= g.nx - 2 && !ll_quiet(psi, g.nx, g.ny, i, j, k));
            base[c] = nw;
            nw += __popcll(m[c]);
        }
        if (lane == 0) cnt[wv] = nw;
        __syncthreads();
        int off = 0, nh = 0;
#pragma unroll
        for (int w = 0; w < kLlWaves; ++w) {
            off += w < wv ? cnt[w] : 0;
            nh += cnt[w];
        }
#pragma unroll
        for (int c = 0; c < kLlSeg; ++c)
            if ((m[c] >> lane) & 1ull)
                pool_i[off + base[c] + __popcll(m[c] & below)] = (s0 + 64 * c + lane) | (wv << kLlWaveBits);
        __syncthreads();
        for (int q = threadIdx.x; q < nh; q += kLlThreads) {
            const int e = pool_i[q], w = e >> kLlWaveBits;
            int jq, kq;
            ll_line(g, (long long)blockIdx.x * kLlWaves + w, jq, kq);
            const LlCell c = ll_cell(psi, g.nx, g.ny, e & ((1 << kLlWaveBits) - 1), jq, kq);
            const bool two = c.nhits >= 2, single = c.ended == 0 && c.nhits == 1;
            if (c.ended == 1) acc += 0.5;
            else if (c.ended == 2) acc += 0.25;
            else if (two) acc += ll_pair(c.h1, c.h2);
            pool_i[q] = (two ? 1 : 0) | (single ? 2 : 0);
#pragma unroll
            for (int d = 0; d < 3; ++d) pool_h[d][q] = two ? c.h2[d] : c.h1[d];
        }
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int c = 0; c < kLlSeg; ++c) {
                int fl = 0;
                double hv[3] = {0.0, 0.0, 0.0};
                if ((m[c] >> lane) & 1ull) {
                    const int q = off + base[c] + __popcll(m[c] & below);
                    fl = pool_i[q];
#pragma unroll
                    for (int d = 0; d < 3; ++d) hv[d] = pool_h[d][q];
                }
                const unsigned long long w = __ballot(fl & 1);
                // the last two-hit lane before this one, else the line's last
                // two-hit cell so far, else (a leading cell) the carry
                const unsigned long long bw = w & below;
                const int src = bw ? ll_last_lane(bw) : lane;
                const double sx = __shfl(hv[0], src, 64), sy = __shfl(hv[1], src, 64), sz = __shfl(hv[2], src, 64);
                const bool defer = (fl & 2) && !bw && !seen;
                if ((fl & 2) && !defer) {
                    const double h2[3] = {bw ? sx : cx, bw ? sy : cy, bw ? sz : cz};
                    acc += ll_pair(hv, h2);
                }
                nlead += __popcll(__ballot(defer));
                if (w) {
                    const int src2 = ll_last_lane(w);
                    cx = __shfl(hv[0], src2, 64);
                    cy = __shfl(hv[1], src2, 64);
                    cz = __shfl(hv[2], src2, 64);
                    seen = true;
                }
            }
        }
        __syncthreads();  // the next segment rewrites cnt and the pool
    }
    if (valid && lane == 0) {
        last[kLlLast * line + 0] = seen ? 1.0 : 0.0;
        last[kLlLast * line + 1] = cx;
        last[kLlLast * line + 2] = cy;
        last[kLlLast * line + 3] = cz;
        last[kLlLast * line + 4] = (double)nlead;
    }
    const double s = block_reduce<kLlThreads, RED_SUM>(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Pass 3: partial[block] = the block's sum over its lines' leading
// single-hit cells, each against the carry into its line: a line that has
// them evaluates its cells again from the start, chunk by chunk, up to the
// chunk of its first two-hit cell.
__global__ __launch_bounds__(kLlThreads) void k_ll_lead(const double2* __restrict__ psi, LlGeom g,
                                                        const double* __restrict__ last,
                                                        const double* __restrict__ carry, double* __restrict__ partial) {
    __shared__ double red[kLlWaves];
    const int lane = threadIdx.x & 63;
    const long long line = (long long)blockIdx.x * kLlWaves + (threadIdx.x >> 6);
    double acc = 0.0;
    if (line < g.nlines && last[kLlLast * line + 4] != 0.0) {  // uniform per wavefront
        const double cv[3] = {carry[3 * line + 0], carry[3 * line + 1], carry[3 * line + 2]};
        int j, k;
        ll_line(g, line, j, k);
        for (int i0 = 1; i0 <= g.nx - 2; i0 += 64) {
            const int i = i0 + lane;
            LlCell c{};
            if (i <= g.nx - 2) c = ll_cell(psi, g.nx, g.ny, i, j, k);
            const unsigned long long w = __ballot(c.nhits >= 2);
            if (c.ended == 0 && c.nhits == 1 && !(w & ((1ull << lane) - 1ull))) acc += ll_pair(c.h1, cv);
            if (w) break;  // uniform per wavefront
        }
    }
    const double s = block_reduce<kLlThreads, RED_SUM>(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kLlScan) void k_ll_final(const double* __restrict__ partial, long long nb,
                                                      double* __restrict__ out) {
    __shared__ double red[kLlScan / 64];
    double s = 0.0;
    for (long long b = threadIdx.x; b < nb; b += kLlScan) s += partial[b];
    s = block_reduce<kLlScan, RED_SUM>(s, red);
    if (threadIdx.x == 0) out[0] = s;
}

long long linelength_lines(int ny, int nz) { return (long long)(ny - 2) * (nz - 2); }
long long linelength_blocks(int ny, int nz) {
    return (linelength_lines(ny, nz) + kLlThreads / 64 - 1) / (kLlThreads / 64);
}

long long linelength_runs(int ny, int nz) { return (linelength_lines(ny, nz) + kLlRun - 1) / kLlRun; }

hipError_t launch_linelength(const double2* psi, int nx, int ny, int nz, double* last, double* carry, double* runs,
                             double* partial, double* out, hipStream_t s) {
    const LlGeom g{nx, ny, nz, linelength_lines(ny, nz)};
    const long long nb = linelength_blocks(ny, nz);
    hipLaunchKernelGGL(k_ll_sum, dim3((unsigned)nb), dim3(kLlThreads), 0, s, psi, g, last, partial);
    const long long nruns = linelength_runs(ny, nz);
    const unsigned rb = (unsigned)((nruns + kLlRunThreads - 1) / kLlRunThreads);
    hipLaunchKernelGGL(k_ll_run_last, dim3(rb), dim3(kLlRunThreads), 0, s, (const double*)last, g.nlines, runs);
    hipLaunchKernelGGL(k_ll_run_scan, dim3(1), dim3(kLlScan), 0, s, (const double*)runs, nruns, runs + 4 * nruns);
    hipLaunchKernelGGL(k_ll_run_fill, dim3(rb), dim3(kLlRunThreads), 0, s, (const double*)last, g.nlines,
                       (const double*)(runs + 4 * nruns), carry);
    hipLaunchKernelGGL(k_ll_lead, dim3((unsigned)nb), dim3(kLlThreads), 0, s, psi, g, (const double*)last,
                       (const double*)carry, partial + nb);
    hipLaunchKernelGGL(k_ll_final, dim3(1), dim3(kLlScan), 0, s, partial, 2 * nb, out);
    return hipGetLastError();
}

}  // namespace gpe
