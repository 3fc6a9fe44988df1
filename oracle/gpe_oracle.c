/*
 * CPU oracle: plain complex128 restatement of the reference's time-stepping
 * path (liangjj/gpe, src/derivs.f90, src/solve.f90, src/variables.f90,
 * src/ic.f90).  TEST INFRASTRUCTURE ONLY -- it is the checker for the HIP
 * library, never the thing measured or shipped.
 *
 * Design: the reference evaluates stencils on "halo" arrays
 * (0:nx1, js-2:je+2, ks-2:ke+2) whose y/z ghost planes are filled by MPI
 * (src/variables.f90:186-358), while x boundaries use index maps
 * (src/derivs.f90:276-398).  This restatement keeps that structure for ONE
 * rank: it builds the halo array, fills the ghost planes the way the
 * one-rank MPI exchange does (periodic: wrap; reflective: untouched), then
 * applies the index-mapped stencils with the reference's operation order.
 * The HIP library instead indexes the whole domain with in-kernel wrap or
 * reflection, so agreement between the two is a real check of the boundary
 * handling, not a restatement of the same code.
 *
 * Arithmetic follows Fortran semantics for complex(8): a*b for complex is
 * (ar*br - ai*bi, ar*bi + ai*br), abs() is hypot, real*complex scales both
 * parts, and expressions are evaluated left to right.  Compile with
 * -ffp-contract=off so no FMA is formed.
 */
#include "gpe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cx;

static inline cx mk(double re, double im) { cx r = {re, im}; return r; }
static inline cx cadd(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
static inline cx cneg(cx a) { return mk(-a.re, -a.im); }
static inline cx cmul(cx a, cx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx rmul(double r, cx a) { return mk(r * a.re, r * a.im); }
static inline cx cdivr(cx a, double r) { return mk(a.re / r, a.im / r); }
static inline double cabs_(cx a) { return hypot(a.re, a.im); }
/* Fortran complex division (Smith's algorithm is what compilers use; for the
 * quantities divided here -- err/psi_scal -- only the modulus is used, so the
 * modulus of the quotient is taken as |a|/|b|, see orc_rkqs). */

int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Index maps: src/derivs.f90:276-317 (second_index), 321-398 (fourth_index) */

void orc_second_index(int bcs, int indx, int n, int x_deriv, int minus[2], int plus[2]) {
    minus[0] = indx - 1;
    plus[0] = indx + 1;
    if (bcs == 1) {
        if (x_deriv) {
            if (indx == 0) { minus[0] = n; plus[0] = 1; }
            if (indx == n) { minus[0] = n - 1; plus[0] = 0; }
        }
    } else if (bcs == 2) {
        if (indx == 0) { minus[0] = 1; plus[0] = 1; }
        if (indx == n) { minus[0] = n - 1; plus[0] = n - 1; }
    }
}

void orc_fourth_index(int bcs, int indx, int n, int x_deriv, int minus[2], int plus[2]) {
    minus[0] = indx - 1; minus[1] = indx - 2;
    plus[0] = indx + 1; plus[1] = indx + 2;
    if (bcs == 1) {
        if (x_deriv) {
            if (indx == 0) { minus[0] = n; minus[1] = n - 1; plus[0] = 1; plus[1] = 2; }
            if (indx == 1) { minus[0] = 0; minus[1] = n; plus[0] = 2; plus[1] = 3; }
            if (indx == n - 1) { minus[0] = n - 2; minus[1] = n - 3; plus[0] = n; plus[1] = 0; }
            if (indx == n) { minus[0] = n - 1; minus[1] = n - 2; plus[0] = 0; plus[1] = 1; }
        }
    } else if (bcs == 2) {
        if (indx == 0) { minus[0] = 1; minus[1] = 2; plus[0] = 1; plus[1] = 2; }
        if (indx == 1) { minus[0] = 0; minus[1] = 1; plus[0] = 2; plus[1] = 3; }
        if (indx == n - 1) { minus[0] = n - 2; minus[1] = n - 3; plus[0] = n; plus[1] = n - 1; }
        if (indx == n) { minus[0] = n - 1; minus[1] = n - 2; plus[0] = n - 1; plus[1] = n - 2; }
    }
}

/* ------------------------------------------------------------------------ */
/* Grid: src/ic.f90:32-69 */

void orc_grid(const orc_params* p, double* x, double* y, double* z, double d[3]) {
    double xl = -p->xr, yl = -p->yr, zl = -p->zr;
    double dx = (p->xr - xl) / (p->nx - 1);
    double dy = (p->yr - yl) / (p->ny - 1);
    double dz = (p->zr - zl) / (p->nz - 1);
    for (int i = 0; i < p->nx; i++) x[i] = xl + (double)i * dx;
    for (int j = 0; j < p->ny; j++) y[j] = yl + (double)j * dy;
    for (int k = 0; k < p->nz; k++) z[k] = zl + (double)k * dz;
    d[0] = dx; d[1] = dy; d[2] = dz;
}

/* ------------------------------------------------------------------------ */
/* Halo array (0:nx1, -2:ny+1, -2:nz+1) for one rank. */

typedef struct {
    int nx, ny, nz;
    cx* a;
} halo_t;

static inline size_t hidx(const halo_t* h, int i, int j, int k) {
    return (size_t)i + (size_t)h->nx * ((size_t)(j + 2) + (size_t)(h->ny + 4) * (size_t)(k + 2));
}
static inline size_t iidx(const orc_params* p, int i, int j, int k) {
    return (size_t)i + (size_t)p->nx * ((size_t)j + (size_t)p->ny * (size_t)k);
}

static int halo_alloc(const orc_params* p, halo_t* h) {
    h->nx = p->nx; h->ny = p->ny; h->nz = p->nz;
    size_t n = (size_t)p->nx * (p->ny + 4) * (p->nz + 4);
    h->a = (cx*)calloc(n, sizeof(cx));
    return h->a ? 0 : -1;
}

/* Ghost fill the one-rank MPI exchange performs: send_recv_z (planes
 * ke-1, ke -> ks-2, ks-1 and ks+1, ks -> ke+2, ke+1, src/variables.f90:
 * 186-233) then pack_y/send_recv_y/unpack_y (rows je-1:je -> js-2:js-1 and
 * js:js+1 -> je+1:je+2 for k in kks..kke, 236-358).  With reflective BCs the
 * neighbours are MPI_PROC_NULL and nothing is received. */
static void halo_fill_ghosts(const orc_params* p, halo_t* h) {
    if (p->bcs != 1) return;
    int nx = p->nx, ny = p->ny, nz = p->nz;
    for (int g = 0; g < 2; g++) {
        int kdst_lo = -2 + g, ksrc_lo = nz - 2 + g;   /* ks-2 <- ke-1, ks-1 <- ke */
        int kdst_hi = nz + 1 - g, ksrc_hi = 1 - g;    /* ke+2 <- ks+1, ke+1 <- ks */
        for (int j = 0; j < ny; j++) {
            memcpy(&h->a[hidx(h, 0, j, kdst_lo)], &h->a[hidx(h, 0, j, ksrc_lo)], sizeof(cx) * nx);
            memcpy(&h->a[hidx(h, 0, j, kdst_hi)], &h->a[hidx(h, 0, j, ksrc_hi)], sizeof(cx) * nx);
        }
    }
    for (int k = 0; k < nz; k++) {
        for (int g = 0; g < 2; g++) {
            memcpy(&h->a[hidx(h, 0, -2 + g, k)], &h->a[hidx(h, 0, ny - 2 + g, k)], sizeof(cx) * nx);
            memcpy(&h->a[hidx(h, 0, ny + g, k)], &h->a[hidx(h, 0, g, k)], sizeof(cx) * nx);
        }
    }
}

static void halo_set_interior(const orc_params* p, halo_t* h, const cx* in) {
    for (int k = 0; k < p->nz; k++)
        for (int j = 0; j < p->ny; j++)
            memcpy(&h->a[hidx(h, 0, j, k)], &in[iidx(p, 0, j, k)], sizeof(cx) * p->nx);
    halo_fill_ghosts(p, h);
}

/* ------------------------------------------------------------------------ */
/* Derivatives: src/derivs.f90:25-272. */

static void deriv_x(const orc_params* p, const halo_t* f, cx* fx, double dx) {
    int nx1 = p->nx - 1;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < p->nz; k++) {
        int mi[2], pl[2];
        for (int j = 0; j < p->ny; j++) {
            for (int i = 0; i <= nx1; i++) {
                size_t o = iidx(p, i, j, k);
                if (p->order == 2) {
                    orc_second_index(p->bcs, i, nx1, 1, mi, pl);
                    fx[o] = cdivr(csub(f->a[hidx(f, pl[0], j, k)], f->a[hidx(f, mi[0], j, k)]), 2.0 * dx);
                } else {
                    orc_fourth_index(p->bcs, i, nx1, 1, mi, pl);
                    cx s = cneg(f->a[hidx(f, pl[1], j, k)]);
                    s = cadd(s, rmul(8.0, f->a[hidx(f, pl[0], j, k)]));
                    s = csub(s, rmul(8.0, f->a[hidx(f, mi[0], j, k)]));
                    s = cadd(s, f->a[hidx(f, mi[1], j, k)]);
                    fx[o] = cdivr(s, 12.0 * dx);
                }
            }
        }
    }
}

/* axis: 1 = y, 2 = z.  First derivative on the halo array. */
static void deriv_yz(const orc_params* p, const halo_t* f, cx* out, int axis, double d) {
    int n1 = (axis == 1 ? p->ny : p->nz) - 1;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < p->nz; k++) {
        int mi[2], pl[2];
        for (int j = 0; j < p->ny; j++) {
            int idx = axis == 1 ? j : k;
            if (p->order == 2) orc_second_index(p->bcs, idx, n1, 0, mi, pl);
            else orc_fourth_index(p->bcs, idx, n1, 0, mi, pl);
            for (int i = 0; i <= p->nx - 1; i++) {
                cx fp1, fm1, fp2 = mk(0, 0), fm2 = mk(0, 0);
                if (axis == 1) {
                    fp1 = f->a[hidx(f, i, pl[0], k)]; fm1 = f->a[hidx(f, i, mi[0], k)];
                    if (p->order == 4) { fp2 = f->a[hidx(f, i, pl[1], k)]; fm2 = f->a[hidx(f, i, mi[1], k)]; }
                } else {
                    fp1 = f->a[hidx(f, i, j, pl[0])]; fm1 = f->a[hidx(f, i, j, mi[0])];
                    if (p->order == 4) { fp2 = f->a[hidx(f, i, j, pl[1])]; fm2 = f->a[hidx(f, i, j, mi[1])]; }
                }
                size_t o = iidx(p, i, j, k);
                if (p->order == 2) {
                    out[o] = cdivr(csub(fp1, fm1), 2.0 * d);
                } else {
                    cx s = cneg(fp2);
                    s = cadd(s, rmul(8.0, fp1));
                    s = csub(s, rmul(8.0, fm1));
                    s = cadd(s, fm2);
                    out[o] = cdivr(s, 12.0 * d);
                }
            }
        }
    }
}

/* Second derivatives: axis 0 = x, 1 = y, 2 = z (src/derivs.f90:146-272). */
static void deriv_2(const orc_params* p, const halo_t* f, cx* out, int axis, double d2) {
    int n1 = (axis == 0 ? p->nx : axis == 1 ? p->ny : p->nz) - 1;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < p->nz; k++) {
        int mi[2], pl[2];
        for (int j = 0; j < p->ny; j++) {
            if (axis == 1 || axis == 2) {
                int idx = axis == 1 ? j : k;
                if (p->order == 2) orc_second_index(p->bcs, idx, n1, 0, mi, pl);
                else orc_fourth_index(p->bcs, idx, n1, 0, mi, pl);
            }
            for (int i = 0; i < p->nx; i++) {
                cx fp1, fm1, fp2 = mk(0, 0), fm2 = mk(0, 0);
                cx f0 = f->a[hidx(f, i, j, k)];
                if (axis == 0) {
                    if (p->order == 2) orc_second_index(p->bcs, i, n1, 1, mi, pl);
                    else orc_fourth_index(p->bcs, i, n1, 1, mi, pl);
                    fp1 = f->a[hidx(f, pl[0], j, k)]; fm1 = f->a[hidx(f, mi[0], j, k)];
                    if (p->order == 4) { fp2 = f->a[hidx(f, pl[1], j, k)]; fm2 = f->a[hidx(f, mi[1], j, k)]; }
                } else if (axis == 1) {
                    fp1 = f->a[hidx(f, i, pl[0], k)]; fm1 = f->a[hidx(f, i, mi[0], k)];
                    if (p->order == 4) { fp2 = f->a[hidx(f, i, pl[1], k)]; fm2 = f->a[hidx(f, i, mi[1], k)]; }
                } else {
                    fp1 = f->a[hidx(f, i, j, pl[0])]; fm1 = f->a[hidx(f, i, j, mi[0])];
                    if (p->order == 4) { fp2 = f->a[hidx(f, i, j, pl[1])]; fm2 = f->a[hidx(f, i, j, mi[1])]; }
                }
                size_t o = iidx(p, i, j, k);
                if (p->order == 2) {
                    cx s = csub(fp1, rmul(2.0, f0));
                    s = cadd(s, fm1);
                    out[o] = cdivr(s, d2);
                } else {
                    cx s = cneg(fp2);
                    s = cadd(s, rmul(16.0, fp1));
                    s = csub(s, rmul(30.0, f0));
                    s = cadd(s, rmul(16.0, fm1));
                    s = csub(s, fm2);
                    out[o] = cdivr(s, 12.0 * d2);
                }
            }
        }
    }
}

static size_t ncell(const orc_params* p) { return (size_t)p->nx * p->ny * p->nz; }

static int check_params(const orc_params* p) {
    if (p->nx < 6 || p->ny < 6 || p->nz < 6) return -1;
    if (p->order != 2 && p->order != 4) return -2;
    if (p->bcs != 1 && p->bcs != 2) return -3;
    if (p->eqn_to_solve < 1 || p->eqn_to_solve > 4) return -4;
    return 0;
}

/* laplacian = dxx + dyy + dzz (src/variables.f90:55-71) on a filled halo array. */
static int laplacian_h(const orc_params* p, const halo_t* h, cx* lap) {
    double x0[1], d[3];
    (void)x0;
    double dx = (p->xr + p->xr) / (p->nx - 1), dy = (p->yr + p->yr) / (p->ny - 1), dz = (p->zr + p->zr) / (p->nz - 1);
    d[0] = dx; d[1] = dy; d[2] = dz;
    size_t n = ncell(p);
    cx* dyy = (cx*)malloc(n * sizeof(cx));
    cx* dzz = (cx*)malloc(n * sizeof(cx));
    if (!dyy || !dzz) { free(dyy); free(dzz); return -10; }
    deriv_2(p, h, lap, 0, d[0] * d[0]);
    deriv_2(p, h, dyy, 1, d[1] * d[1]);
    deriv_2(p, h, dzz, 2, d[2] * d[2]);
#pragma omp parallel for schedule(static)
    for (size_t o = 0; o < n; o++) lap[o] = cadd(cadd(lap[o], dyy[o]), dzz[o]);
    free(dyy); free(dzz);
    return 0;
}

int orc_deriv(const orc_params* p, int op, const double* f, double* out) {
    int rc = check_params(p);
    if (rc) return rc;
    halo_t h;
    if (halo_alloc(p, &h)) return -10;
    halo_set_interior(p, &h, (const cx*)f);
    double dx = (p->xr + p->xr) / (p->nx - 1), dy = (p->yr + p->yr) / (p->ny - 1), dz = (p->zr + p->zr) / (p->nz - 1);
    cx* o = (cx*)out;
    switch (op) {
        case 0: deriv_x(p, &h, o, dx); break;
        case 1: deriv_yz(p, &h, o, 1, dy); break;
        case 2: deriv_yz(p, &h, o, 2, dz); break;
        case 3: deriv_2(p, &h, o, 0, dx * dx); break;
        case 4: deriv_2(p, &h, o, 1, dy * dy); break;
        case 5: deriv_2(p, &h, o, 2, dz * dz); break;
        case 6: rc = laplacian_h(p, &h, o); break;
        default: rc = -5;
    }
    free(h.a);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* get_rhs: src/solve.f90:382-440 on a filled halo array. */

static int rhs_h(const orc_params* p, const halo_t* h, cx* rhs) {
    int nx = p->nx, ny = p->ny, nz = p->nz;
    size_t n = ncell(p);
    double* x = (double*)malloc(sizeof(double) * (nx + ny + nz));
    if (!x) return -10;
    double *y = x + nx, *z = y + ny, d[3];
    orc_grid(p, x, y, z, d);
    cx* lap = (cx*)malloc(n * sizeof(cx));
    cx* dpsidx = NULL;
    if (!lap) { free(x); return -10; }
    int rc = laplacian_h(p, h, lap);
    if (rc) { free(x); free(lap); return rc; }
    /* deriv_x is evaluated unconditionally by the reference (solve.f90:411)
     * but only enters through Urhs*dpsidx; with Urhs == 0 the term adds an
     * exact zero, so it is only formed when Urhs != 0. */
    if (p->Urhs != 0.0) {
        dpsidx = (cx*)malloc(n * sizeof(cx));
        if (!dpsidx) { free(x); free(lap); return -10; }
        deriv_x(p, h, dpsidx, d[0]);
    }
    /* diss (solve.f90:397-409): a = x(nx-3), b = y(ny-3), c = z(nz-3). */
    double a = x[nx - 3], b = y[ny - 3], c = z[nz - 3];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++) {
        for (int j = 0; j < ny; j++) {
            for (int i = 0; i < nx; i++) {
                size_t o = iidx(p, i, j, k);
                cx psi = h->a[hidx(h, i, j, k)];
                cx L = lap[o];
                cx r;
                double ab = cabs_(psi);
                double ab2 = ab * ab;
                if (p->eqn_to_solve == 1 || p->eqn_to_solve == 3) {
                    double diss = 0.0;
                    if (p->diss_amp != 0.0)
                        diss = p->diss_amp * ((1.0 + tanh(x[i] - a) * tanh(x[i] + a)) +
                                              (1.0 + tanh(y[j] - b) * tanh(y[j] + b)) +
                                              (1.0 + tanh(z[k] - c) * tanh(z[k] + c)));
                    /* 0.5*(eye+diss) * (laplacian + (1-|psi|^2)*psi) + Urhs*dpsidx */
                    double s = 1.0 - ab2;
                    if (p->vscale != 0.0) {
                        /* extension: i/2 (lap + (1 - |psi|^2 - vscale*V) psi) */
                        double V = 0.5 * ((p->omx * p->omx) * (x[i] * x[i]) + (p->omy * p->omy) * (y[j] * y[j]) +
                                          (p->omz * p->omz) * (z[k] * z[k]));
                        s = s - p->vscale * V;
                    }
                    cx t1 = cadd(L, rmul(s, psi));
                    r = cmul(rmul(0.5, mk(diss, 1.0)), t1);
                    if (dpsidx) r = cadd(r, rmul(p->Urhs, dpsidx[o]));
                    if (p->eqn_to_solve == 3) {
                        /* sphere() (src/ic.f90:635-657): 0.5*(1+tanh(r - rad - eps1)), rad=10, eps1=2 */
                        double rr = sqrt(x[i] * x[i] + y[j] * y[j] + z[k] * z[k]);
                        double m = 0.5 * (1.0 + tanh(rr - 10.0 - 2.0));
                        r = cmul(r, mk(m, 0.0));
                    }
                } else if (p->eqn_to_solve == 2) {
                    /* eye * (laplacian - |psi|^2 psi) + Urhs*dpsidx */
                    cx t1 = csub(L, rmul(ab2, psi));
                    r = cmul(mk(0.0, 1.0), t1);
                    if (dpsidx) r = cadd(r, rmul(p->Urhs, dpsidx[o]));
                } else {
                    /* -eye * (-0.5*lap + Vtrap*psi + g*(abs(psi**2)*psi) - mu*psi) */
                    double V = 0.5 * ((p->omx * p->omx) * (x[i] * x[i]) + (p->omy * p->omy) * (y[j] * y[j]) +
                                      (p->omz * p->omz) * (z[k] * z[k]));
                    cx psi2 = cmul(psi, psi);
                    double ap2 = cabs_(psi2);
                    cx w = rmul(-0.5, L);
                    w = cadd(w, rmul(V, psi));
                    w = cadd(w, rmul(p->g, rmul(ap2, psi)));
                    w = csub(w, rmul(p->mu, psi));
                    r = cmul(mk(-0.0, -1.0), w);
                }
                rhs[o] = r;
            }
        }
    }
    free(x); free(lap); free(dpsidx);
    return 0;
}

int orc_rhs(const orc_params* p, const double* in, double* out) {
    int rc = check_params(p);
    if (rc) return rc;
    halo_t h;
    if (halo_alloc(p, &h)) return -10;
    halo_set_interior(p, &h, (const cx*)in);
    rc = rhs_h(p, &h, (cx*)out);
    free(h.a);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Steppers: src/solve.f90:57-154 */

static void advance_time(double dt_re, double dt_im, double t[2]) {
    t[0] = t[0] + dt_re;       /* t = t + real(dt) */
    t[1] = t[1] + fabs(dt_im); /* im_t = im_t + abs(aimag(dt)) */
}

int orc_euler(const orc_params* p, double dt_re, double dt_im, const double* in, double* out, double t[2]) {
    int rc = check_params(p);
    if (rc) return rc;
    size_t n = ncell(p);
    halo_t h;
    if (halo_alloc(p, &h)) return -10;
    halo_set_interior(p, &h, (const cx*)in);
    cx* rhs = (cx*)malloc(n * sizeof(cx));
    if (!rhs) { free(h.a); return -10; }
    rc = rhs_h(p, &h, rhs);
    cx dt = mk(dt_re, dt_im);
    const cx* ci = (const cx*)in;
    cx* co = (cx*)out;
    for (size_t o = 0; o < n; o++) co[o] = cadd(ci[o], cmul(dt, rhs[o]));
    free(rhs); free(h.a);
    if (!rc) advance_time(dt_re, dt_im, t);
    return rc;
}

/* Stage update tmp = old + c*k into the halo array interior, then ghost fill
 * (the send_recv_z/pack_y/send_recv_y/unpack_y sequence). */
static void stage_into(const orc_params* p, halo_t* tmp, const cx* old, cx cdt, const cx* k) {
#pragma omp parallel for schedule(static)
    for (int kk = 0; kk < p->nz; kk++)
        for (int j = 0; j < p->ny; j++)
            for (int i = 0; i < p->nx; i++) {
                size_t o = iidx(p, i, j, kk);
                tmp->a[hidx(tmp, i, j, kk)] = cadd(old[o], cmul(cdt, k[o]));
            }
    halo_fill_ghosts(p, tmp);
}

int orc_rk2(const orc_params* p, double dt_re, double dt_im, const double* in, double* out, double t[2]) {
    int rc = check_params(p);
    if (rc) return rc;
    size_t n = ncell(p);
    halo_t h, tmp;
    if (halo_alloc(p, &h)) return -10;
    if (halo_alloc(p, &tmp)) { free(h.a); return -10; }
    halo_set_interior(p, &h, (const cx*)in);
    cx* k1 = (cx*)malloc(n * sizeof(cx));
    cx* k2 = (cx*)malloc(n * sizeof(cx));
    const cx* old = (const cx*)in;
    cx dt = mk(dt_re, dt_im);
    if (!k1 || !k2) { rc = -10; goto done; }
    if ((rc = rhs_h(p, &h, k1))) goto done;
    stage_into(p, &tmp, old, rmul(0.5, dt), k1); /* tmp = old + 0.5*dt*kk1 */
    if ((rc = rhs_h(p, &tmp, k2))) goto done;
    for (size_t o = 0; o < n; o++) ((cx*)out)[o] = cadd(old[o], cmul(dt, k2[o]));
    advance_time(dt_re, dt_im, t);
done:
    free(k1); free(k2); free(h.a); free(tmp.a);
    return rc;
}

typedef struct {
    halo_t h, tmp;
    cx *k1, *k2, *k3, *k4;
} rk4_work;

static int rk4_work_alloc(const orc_params* p, rk4_work* w) {
    memset(w, 0, sizeof(*w));
    size_t n = ncell(p);
    if (halo_alloc(p, &w->h) || halo_alloc(p, &w->tmp)) return -10;
    w->k1 = (cx*)malloc(n * sizeof(cx)); w->k2 = (cx*)malloc(n * sizeof(cx));
    w->k3 = (cx*)malloc(n * sizeof(cx)); w->k4 = (cx*)malloc(n * sizeof(cx));
    return (w->k1 && w->k2 && w->k3 && w->k4) ? 0 : -10;
}
static void rk4_work_free(rk4_work* w) {
    free(w->h.a); free(w->tmp.a); free(w->k1); free(w->k2); free(w->k3); free(w->k4);
}

static int rk4_core(const orc_params* p, rk4_work* w, double dt_re, double dt_im, const cx* old, cx* out, double t[2]) {
    size_t n = ncell(p);
    cx dt = mk(dt_re, dt_im);
    int rc;
    halo_set_interior(p, &w->h, old);
    if ((rc = rhs_h(p, &w->h, w->k1))) return rc;
    stage_into(p, &w->tmp, old, rmul(0.5, dt), w->k1);
    if ((rc = rhs_h(p, &w->tmp, w->k2))) return rc;
    stage_into(p, &w->tmp, old, rmul(0.5, dt), w->k2);
    if ((rc = rhs_h(p, &w->tmp, w->k3))) return rc;
    stage_into(p, &w->tmp, old, dt, w->k3);
    if ((rc = rhs_h(p, &w->tmp, w->k4))) return rc;
    /* new = old + dt*(kk1/6 + kk2/3 + kk3/3 + kk4/6)  (solve.f90:144-148) */
#pragma omp parallel for schedule(static)
    for (size_t o = 0; o < n; o++) {
        cx s = cadd(cadd(cadd(cdivr(w->k1[o], 6.0), cdivr(w->k2[o], 3.0)), cdivr(w->k3[o], 3.0)), cdivr(w->k4[o], 6.0));
        out[o] = cadd(old[o], cmul(dt, s));
    }
    advance_time(dt_re, dt_im, t);
    return 0;
}

int orc_rk4(const orc_params* p, double dt_re, double dt_im, const double* in, double* out, double t[2]) {
    int rc = check_params(p);
    if (rc) return rc;
    rk4_work w;
    if ((rc = rk4_work_alloc(p, &w)) == 0) rc = rk4_core(p, &w, dt_re, dt_im, (const cx*)in, (cx*)out, t);
    rk4_work_free(&w);
    return rc;
}

int orc_rk4_n(const orc_params* p, double dt_re, double dt_im, double* psi, int nsteps, double t[2]) {
    int rc = check_params(p);
    if (rc) return rc;
    rk4_work w;
    size_t n = ncell(p);
    cx* nxt = (cx*)malloc(n * sizeof(cx));
    if ((rc = rk4_work_alloc(p, &w)) == 0 && nxt) {
        for (int s = 0; s < nsteps && !rc; s++) {
            rc = rk4_core(p, &w, dt_re, dt_im, (const cx*)psi, nxt, t);
            memcpy(psi, nxt, n * sizeof(cx));
        }
    } else if (!rc) {
        rc = -10;
    }
    rk4_work_free(&w);
    free(nxt);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Cash-Karp: src/solve.f90:248-378 */

static const double b21 = 0.2, b31 = 0.075, b32 = 0.225, b41 = 0.3, b42 = -0.9, b43 = 1.2;
#define B51 (-11.0 / 54.0)
#define B52 2.5
#define B53 (-70.0 / 27.0)
#define B54 (35.0 / 27.0)
#define B61 (1631.0 / 55296.0)
#define B62 (175.0 / 512.0)
#define B63 (575.0 / 13824.0)
#define B64 (44275.0 / 110592.0)
#define B65 (253.0 / 4096.0)
#define C1 (37.0 / 378.0)
#define C2 0.0
#define C3 (250.0 / 621.0)
#define C4 (125.0 / 594.0)
#define C5 0.0
#define C6 (512.0 / 1771.0)

static int rkck_core(const orc_params* p, cx dt, const cx* old, halo_t* hold, halo_t* tmp, cx* kk[6], cx* out, cx* err,
                     int have_k1) {
    const double dc1 = C1 - 2825.0 / 27648.0, dc2 = C2 - 0.0, dc3 = C3 - 18575.0 / 48384.0;
    const double dc4 = C4 - 13525.0 / 55296.0, dc5 = C5 - 277.0 / 14336.0, dc6 = C6 - 0.25;
    size_t n = ncell(p);
    int nx = p->nx, ny = p->ny, nz = p->nz;
    int rc;
    cx *k1 = kk[0], *k2 = kk[1], *k3 = kk[2], *k4 = kk[3], *k5 = kk[4], *k6 = kk[5];
    /* get_rhs(old, kk1) -- the reference recomputes f(old) on every attempt;
     * the values are identical, so a cached f(old) may be passed in k1. */
    if (!have_k1) {
        if ((rc = rhs_h(p, hold, k1))) return rc;
    }
    /* kk1 = dt*kk1; tmp = old + b21*kk1 */
    cx* K1 = (cx*)malloc(n * sizeof(cx));
    if (!K1) return -10;
#define LOOP for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++)
#define O iidx(p, i, j, k)
#define T tmp->a[hidx(tmp, i, j, k)]
    LOOP { size_t o = O; K1[o] = cmul(dt, k1[o]); T = cadd(old[o], rmul(b21, K1[o])); }
    halo_fill_ghosts(p, tmp);
    if ((rc = rhs_h(p, tmp, k2))) goto fail;
    LOOP { size_t o = O; k2[o] = cmul(dt, k2[o]); T = cadd(cadd(old[o], rmul(b31, K1[o])), rmul(b32, k2[o])); }
    halo_fill_ghosts(p, tmp);
    if ((rc = rhs_h(p, tmp, k3))) goto fail;
    LOOP {
        size_t o = O; k3[o] = cmul(dt, k3[o]);
        T = cadd(cadd(cadd(old[o], rmul(b41, K1[o])), rmul(b42, k2[o])), rmul(b43, k3[o]));
    }
    halo_fill_ghosts(p, tmp);
    if ((rc = rhs_h(p, tmp, k4))) goto fail;
    LOOP {
        size_t o = O; k4[o] = cmul(dt, k4[o]);
        T = cadd(cadd(cadd(cadd(old[o], rmul(B51, K1[o])), rmul(B52, k2[o])), rmul(B53, k3[o])), rmul(B54, k4[o]));
    }
    halo_fill_ghosts(p, tmp);
    if ((rc = rhs_h(p, tmp, k5))) goto fail;
    LOOP {
        size_t o = O; k5[o] = cmul(dt, k5[o]);
        T = cadd(cadd(cadd(cadd(cadd(old[o], rmul(B61, K1[o])), rmul(B62, k2[o])), rmul(B63, k3[o])), rmul(B64, k4[o])),
                 rmul(B65, k5[o]));
    }
    halo_fill_ghosts(p, tmp);
    if ((rc = rhs_h(p, tmp, k6))) goto fail;
    LOOP {
        size_t o = O; k6[o] = cmul(dt, k6[o]);
        out[o] = cadd(cadd(cadd(cadd(cadd(cadd(old[o], rmul(C1, K1[o])), rmul(C2, k2[o])), rmul(C3, k3[o])),
                                rmul(C4, k4[o])), rmul(C5, k5[o])), rmul(C6, k6[o]));
        err[o] = cadd(cadd(cadd(cadd(cadd(rmul(dc1, K1[o]), rmul(dc2, k2[o])), rmul(dc3, k3[o])), rmul(dc4, k4[o])),
                           rmul(dc5, k5[o])), rmul(dc6, k6[o]));
    }
#undef LOOP
#undef O
#undef T
    free(K1);
    return 0;
fail:
    free(K1);
    return rc;
}

typedef struct {
    halo_t h, tmp;
    cx* kk[6];
    cx* err;
} rkck_work;

static int rkck_alloc(const orc_params* p, rkck_work* w) {
    memset(w, 0, sizeof(*w));
    size_t n = ncell(p);
    if (halo_alloc(p, &w->h) || halo_alloc(p, &w->tmp)) return -10;
    for (int s = 0; s < 6; s++) if (!(w->kk[s] = (cx*)malloc(n * sizeof(cx)))) return -10;
    if (!(w->err = (cx*)malloc(n * sizeof(cx)))) return -10;
    return 0;
}
static void rkck_free(rkck_work* w) {
    free(w->h.a); free(w->tmp.a); free(w->err);
    for (int s = 0; s < 6; s++) free(w->kk[s]);
}

int orc_rkck(const orc_params* p, double dt_re, double dt_im, const double* in, double* out, double* err) {
    int rc = check_params(p);
    if (rc) return rc;
    rkck_work w;
    if ((rc = rkck_alloc(p, &w)) == 0) {
        halo_set_interior(p, &w.h, (const cx*)in);
        rc = rkck_core(p, mk(dt_re, dt_im), (const cx*)in, &w.h, &w.tmp, w.kk, (cx*)out, (cx*)err, 0);
    }
    rkck_free(&w);
    return rc;
}

static double sign_(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

/* rkqs: src/solve.f90:158-244.  The controller constants come from p
 * (eps = 1e-6, safety = 0.9, dt_decrease = -0.25, dt_increase = -0.2 in the
 * reference, src/parameters.f90:89-92). */
int orc_rkqs(const orc_params* p, double dtc[2], double t[2], const double* in, double* out, double info[3]) {
    int rc = check_params(p);
    if (rc) return rc;
    size_t n = ncell(p);
    rkck_work w;
    cx* scal = (cx*)malloc(n * sizeof(cx));
    cx* f0 = (cx*)malloc(n * sizeof(cx));
    if ((rc = rkck_alloc(p, &w)) != 0 || !scal || !f0) { rc = rc ? rc : -10; goto done; }
    {
        const cx* old = (const cx*)in;
        cx dt = mk(dtc[0], dtc[1]);
        double errcon = pow(5.0 / p->safety, 1.0 / p->dt_increase);
        halo_set_interior(p, &w.h, old);
        /* psi_scal = abs(in) + dt*abs(f(in)) + 1e-30 (complex) */
        if ((rc = rhs_h(p, &w.h, f0))) goto done;
        for (size_t o = 0; o < n; o++) {
            cx s = cadd(mk(cabs_(old[o]), 0.0), cmul(dt, mk(cabs_(f0[o]), 0.0)));
            scal[o] = cadd(s, mk(1e-30, 0.0));
        }
        double errmax = 0.0;
        int nrej = 0, nrhs = 1;
        for (;;) {
            memcpy(w.kk[0], f0, n * sizeof(cx)); /* f(old) is unchanged between attempts */
            rc = rkck_core(p, dt, old, &w.h, &w.tmp, w.kk, (cx*)out, w.err, 1);
            nrhs += 6;
            if (rc) goto done;
            /* errmax = maxval(abs(err/psi_scal))/eps ; |a/b| = |a|/|b| */
            double m = 0.0;
            for (size_t o = 0; o < n; o++) {
                double e = cabs_(w.err[o]) / cabs_(scal[o]);
                if (e > m) m = e;
            }
            errmax = m / p->eps;
            if (errmax <= 1.0) break;
            nrej++;
            cx dt_temp = cmul(rmul(p->safety, dt), mk(pow(errmax, p->dt_decrease), 0.0));
            double mag = fmax(cabs_(dt_temp), 0.1 * cabs_(dt));
            if (!p->real_time) dt = mk(0.0, -sign_(mag, cabs_(dt)));
            else dt = mk(sign_(mag, cabs_(dt)), 0.0);
        }
        cx dt_next;
        if (errmax > errcon) dt_next = cmul(rmul(p->safety, dt), mk(pow(errmax, p->dt_increase), 0.0));
        else dt_next = rmul(5.0, dt);
        t[0] = t[0] + dt.re;           /* t = t + real(dt) */
        dt = dt_next;                  /* dt = dt_next */
        t[1] = t[1] + fabs(dt.im);     /* im_t = im_t + abs(aimag(dt)) -- the NEW dt */
        dtc[0] = dt.re; dtc[1] = dt.im;
        info[0] = errmax; info[1] = nrej; info[2] = nrhs;
    }
done:
    rkck_free(&w);
    free(scal); free(f0);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Reductions: src/variables.f90:537-611, 698-709, 713-823 */

static const double q1 = 3.0 / 8.0, q2 = 7.0 / 6.0, q3 = 23.0 / 24.0;

double orc_norm(const orc_params* p, const double* in) {
    int nx = p->nx, ny = p->ny, nz = p->nz;
    const cx* a = (const cx*)in;
    double dx = (p->xr + p->xr) / (nx - 1), dy = (p->yr + p->yr) / (ny - 1), dz = (p->zr + p->zr) / (nz - 1);
    double* xint = (double*)malloc(sizeof(double) * ny * nz);
    double* yint = (double*)malloc(sizeof(double) * nz);
    double* v = (double*)malloc(sizeof(double) * nx);
    /* integrate_x */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            for (int i = 0; i < nx; i++) { double h = cabs_(a[iidx(p, i, j, k)]); v[i] = h * h; }
            double s = 0.0;
            for (int i = 3; i <= nx - 4; i++) s += v[i];
            xint[j + ny * k] = (q1 * v[0] + q2 * v[1] + q3 * v[2] + s + q3 * v[nx - 3] + q2 * v[nx - 2] + q1 * v[nx - 1]) * dx;
        }
    /* integrate_y */
    for (int k = 0; k < nz; k++) {
        double s = 0.0;
        for (int j = 0; j < ny; j++) {
            double tv = xint[j + ny * k];
            if (j == 0) tv = q1 * xint[j + ny * k];
            if (j == 1) tv = q2 * xint[j + ny * k];
            if (j == 2) tv = q3 * xint[j + ny * k];
            if (j == ny - 3) tv = q3 * xint[j + ny * k];
            if (j == ny - 2) tv = q2 * xint[j + ny * k];
            if (j == ny - 1) tv = q1 * xint[j + ny * k];
            s += tv;
        }
        yint[k] = s * dy;
    }
    /* integrate_z */
    double s = 0.0;
    for (int k = 0; k < nz; k++) {
        double tv = yint[k];
        if (k == 0) tv = q1 * yint[k];
        if (k == 1) tv = q2 * yint[k];
        if (k == 2) tv = q3 * yint[k];
        if (k == nz - 3) tv = q3 * yint[k];
        if (k == nz - 2) tv = q2 * yint[k];
        if (k == nz - 1) tv = q1 * yint[k];
        s += tv;
    }
    free(xint); free(yint); free(v);
    return s * dz;
}

double orc_energy(const orc_params* p, const double* in) {
    size_t n = ncell(p);
    halo_t h;
    cx* lap = (cx*)malloc(n * sizeof(cx));
    if (halo_alloc(p, &h) || !lap) { free(h.a); free(lap); return NAN; }
    halo_set_interior(p, &h, (const cx*)in);
    laplacian_h(p, &h, lap);
    const cx* a = (const cx*)in;
    double s = 0.0;
    /* real(-laplacian*conjg(psi) + 0.5*abs(psi)**4) summed, times dx*dy*dz */
    for (size_t o = 0; o < n; o++) {
        double ab = cabs_(a[o]);
        cx v = cmul(cneg(lap[o]), mk(a[o].re, -a[o].im));
        v = cadd(v, mk(0.5 * ((ab * ab) * (ab * ab)), 0.0)); /* abs()**4 by squaring */
        s += v.re;
    }
    free(lap); free(h.a);
    double dx = (p->xr + p->xr) / (p->nx - 1), dy = (p->yr + p->yr) / (p->ny - 1), dz = (p->zr + p->zr) / (p->nz - 1);
    return s * dx * dy * dz;
}

double orc_mass(const orc_params* p, const double* in) {
    size_t n = ncell(p);
    const cx* a = (const cx*)in;
    double s = 0.0;
    for (size_t o = 0; o < n; o++) { double h = cabs_(a[o]); s += h * h; }
    double dx = (p->xr + p->xr) / (p->nx - 1), dy = (p->yr + p->yr) / (p->ny - 1), dz = (p->zr + p->zr) / (p->nz - 1);
    return s * dx * dy * dz;
}

void orc_renormalise(const orc_params* p, double* var, double norm) {
    size_t n = ncell(p);
    cx* a = (cx*)var;
    double snn = sqrt(p->nn), sn = sqrt(norm);
    for (size_t o = 0; o < n; o++) a[o] = cdivr(rmul(snn, a[o]), sn); /* sqrt(nn)*var/sqrt(norm) */
}


/* ------------------------------------------------------------------------
 * Line length of the zero lines of psi (linelength, src/variables.f90:879-1145).
 * Each interior cell (i, j, k) in 1..n-2 on every axis, visited in the
 * reference's order (k, then j, then i fastest), looks for the point where
 * the zero lines of Re psi and Im psi cross on each of its six faces: per
 * face the first two sign-change crossings of u (and of v) along the edges
 * 1-2, 2-3, 3-4, 4-1 define two straight lines whose intersection, if it lies
 * in the unit face, is a hit.  A face where u (or v) vanishes at all four
 * corners while the other changes sign adds 1/2 and ends the cell; a face
 * with an edge where both vanish adds 1/4 and ends the cell.  Otherwise a
 * cell with hits adds a sqrt(|h1 - h2|^2) with a = 1/2 or 1/4 for hits on
 * shared faces (the reference's test compares y twice where z is meant).
 * The hit slots persist from cell to cell, as the reference's local arrays
 * do: a cell with a single hit pairs it with the second hit of the last
 * cell that recorded one (0 before any, where the reference reads an
 * undefined value).  The sum is multiplied by dx.
 * counts: [0] cells with two or more hits, [1] cells with exactly one,
 * [2] cells ended by a face, [3] cells ended by an edge.
 * ---------------------------------------------------------------------- */

/* corners of face n (u(1), u(2), u(3), u(4)) as (di, dj, dk) */
static const int ll_face[6][4][3] = {
    {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}},  /* 1 front  (k)   */
    {{0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}},  /* 2 back   (k+1) */
    {{1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}},  /* 3 right  (i+1) */
    {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}},  /* 4 left   (i)   */
    {{0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}},  /* 5 top    (j+1) */
    {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}},  /* 6 bottom (j)   */
};

/* the crossings of the zero line of w (w[0..4], w[4] = w[0]) with the face
 * edges: returns lu - 1, the number of distinct sign-change crossings kept,
 * with their positions in (xs, ys) */
static int ll_crossings(const double* w, double* xs, double* ys) {
    int lu = 1;  /* 1-based slot, as the reference */
    for (int q = 1; q <= 4; q++) {
        const double a = w[q - 1], b = w[q];
        if (a == 0.0) {
            /* a corner on the line: stored in slot lu, not counted */
            static const double cx_[4] = {0.0, 1.0, 1.0, 0.0}, cy_[4] = {0.0, 0.0, 1.0, 1.0};
            xs[lu] = cx_[q - 1];
            ys[lu] = cy_[q - 1];
        } else if (a * b < 0.0) {
            switch (q) {
                case 1: xs[lu] = fabs(a / (a - b)); ys[lu] = 0.0; break;
                case 2: xs[lu] = 1.0; ys[lu] = fabs(a / (a - b)); break;
                case 3: xs[lu] = fabs(b / (a - b)); ys[lu] = 1.0; break;
                default: xs[lu] = 0.0; ys[lu] = fabs(b / (a - b)); break;
            }
            if (lu == 1 || xs[lu] != xs[lu - 1] || ys[lu] != ys[lu - 1]) lu++;
        }
    }
    return lu;
}

double orc_linelength(const orc_params* p, const double* in, long long counts[4]) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    const cx* a = (const cx*)in;
    const double dx = (p->xr + p->xr) / (nx - 1);
    double hx[7] = {0}, hy[7] = {0}, hz[7] = {0};  /* hit slots 1..6, kept across cells */
    double total = 0.0;
    for (int q = 0; q < 4; q++) counts[q] = 0;
    for (int k = 1; k <= nz - 2; k++)
        for (int j = 1; j <= ny - 2; j++)
            for (int i = 1; i <= nx - 2; i++) {
                int m = 1, ended = 0;
                for (int n = 0; n < 6 && !ended; n++) {
                    double u[5], v[5];
                    for (int c = 0; c < 4; c++) {
                        const int *o = ll_face[n][c];
                        const cx z = a[((size_t)(k + o[2]) * ny + (j + o[1])) * nx + (i + o[0])];
                        u[c] = z.re;
                        v[c] = z.im;
                    }
                    u[4] = u[0];
                    v[4] = v[0];
                    const int uzero = u[0] == 0.0 && u[1] == 0.0 && u[2] == 0.0 && u[3] == 0.0;
                    const int vzero = v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0 && v[3] == 0.0;
                    const int vchange = v[0] * v[1] < 0.0 || v[1] * v[2] < 0.0 || v[2] * v[3] < 0.0 || v[3] * v[0] < 0.0;
                    const int uchange = u[0] * u[1] < 0.0 || u[1] * u[2] < 0.0 || u[2] * u[3] < 0.0 || u[3] * u[0] < 0.0;
                    if ((uzero && vchange) || (vzero && uchange)) {
                        total += 0.5;
                        counts[2]++;
                        ended = 1;
                        break;
                    }
                    for (int q = 0; q < 4; q++)
                        if (u[q] == 0.0 && u[q + 1] == 0.0 && v[q] == 0.0 && v[q + 1] == 0.0) {
                            ended = 1;
                            break;
                        }
                    if (ended) {
                        total += 0.25;
                        counts[3]++;
                        break;
                    }
                    double xu[7], yu[7], xv[7], yv[7];
                    const int lu = ll_crossings(u, xu, yu);
                    const int lv = ll_crossings(v, xv, yv);
                    if (lu > 2 && lv > 2) {
                        double den = xv[2] * (yu[1] - yu[2]) + xv[1] * (yu[2] - yu[1]) + (xu[1] - xu[2]) * (yv[1] - yv[2]);
                        if (den == 0.0) den = den + 0.0000001;
                        const double xx = (xu[1] * (xv[2] * (yv[1] - yu[2]) + xv[1] * (yu[2] - yv[2])) +
                                           xu[2] * (xv[2] * (yu[1] - yv[1]) + xv[1] * (yv[2] - yu[1]))) / den;
                        const double yy = (xv[2] * (yu[1] - yu[2]) * yv[1] + xu[1] * yu[2] * yv[1] - xv[1] * yu[1] * yv[2] -
                                           xu[1] * yu[2] * yv[2] + xv[1] * yu[2] * yv[2] + xu[2] * yu[1] * (yv[2] - yv[1])) / den;
                        if (xx >= 0.0 && xx <= 1.0 && yy >= 0.0 && yy <= 1.0) {
                            double h[3];
                            switch (n + 1) {
                                case 1: h[0] = xx; h[1] = yy; h[2] = 0.0; break;
                                case 2: h[0] = xx; h[1] = yy; h[2] = 1.0; break;
                                case 3: h[0] = 1.0; h[1] = yy; h[2] = xx; break;
                                case 4: h[0] = 0.0; h[1] = yy; h[2] = xx; break;
                                case 5: h[0] = xx; h[1] = 1.0; h[2] = yy; break;
                                default: h[0] = xx; h[1] = 0.0; h[2] = yy; break;
                            }
                            hx[m] = h[0];
                            hy[m] = h[1];
                            hz[m] = h[2];
                            m++;
                        }
                    }
                }
                if (ended || m == 1) continue;
                counts[m == 2 ? 1 : 0]++;
                const int xs = hx[1] == hx[2] && (hx[1] == 0.0 || hx[1] == 1.0);
                const int ys = hy[1] == hy[2] && (hy[1] == 0.0 || hy[1] == 1.0);
                const int zs = hz[1] == hz[2] && (hz[1] == 0.0 || hz[1] == 1.0);
                double w = 1.0;
                if (xs || ys || ys) {
                    w = w * 0.5;
                    if ((xs && (ys || ys)) || (ys && zs)) w = w * 0.5;
                }
                const double ddx = hx[1] - hx[2], ddy = hy[1] - hy[2], ddz = hz[1] - hz[2];
                total = total + w * sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
            }
    return total * dx;
}
