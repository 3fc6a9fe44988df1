/*
 * CPU oracle for the GPE time-stepping path: a plain complex128 restatement
 * of the reference Fortran (liangjj/gpe) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * Fortran host, the Python host mirror) links or calls this code; only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark do, and
 * there only as the checker.
 *
 * Layout: every field is nx*ny*nz complex values, x fastest, interleaved
 * (re, im) doubles -- the layout of the reference's interior arrays
 * complex(pr), dimension(0:nx1, js:je, ks:ke) with one rank (src/solve.f90:31).
 */
#ifndef GPE_ORACLE_H
#define GPE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int32_t nx, ny, nz;
    int32_t order;          /* 2 or 4                       (src/parameters.f90:32) */
    int32_t bcs;            /* 1 periodic, 2 reflective      (src/parameters.f90:31) */
    int32_t eqn_to_solve;   /* 1..4                          (src/solve.f90:414-437) */
    int32_t real_time;      /* 1 real time, 0 imaginary time (src/parameters.f90:38) */
    int32_t pad_;
    double xr, yr, zr;      /* box half widths               (src/parameters.f90:28) */
    double Urhs, diss_amp, g, mu, nn, omx, omy, omz; /* eqn_params (src/parameters.f90:44-55) */
    double vscale;          /* coefficient of V in eqn 1 (0 reproduces the reference) */
    double eps, safety, dt_decrease, dt_increase; /* src/parameters.f90:89-92 */
} orc_params;

/* Index maps second_index / fourth_index (src/derivs.f90:276-398). */
void orc_second_index(int bcs, int indx, int n, int x_deriv, int minus[2], int plus[2]);
void orc_fourth_index(int bcs, int indx, int n, int x_deriv, int minus[2], int plus[2]);

/* Grid (src/ic.f90:32-69): x[0..nx-1] etc. and dx, dx2. */
void orc_grid(const orc_params* p, double* x, double* y, double* z, double d[3]);

/* Derivatives of the interior field f (ghost fill for one rank is done
 * internally).  op: 0 x, 1 y, 2 z, 3 xx, 4 yy, 5 zz, 6 laplacian.
 * (src/derivs.f90:25-272, src/variables.f90:55-71) */
int orc_deriv(const orc_params* p, int op, const double* f, double* out);

/* get_rhs (src/solve.f90:382-440). */
int orc_rhs(const orc_params* p, const double* in, double* out);

/* One step of euler / rk2 / rk4 (src/solve.f90:57-154).  t[0] = t, t[1] = im_t
 * are advanced exactly as the reference does. */
int orc_euler(const orc_params* p, double dt_re, double dt_im, const double* in, double* out, double t[2]);
int orc_rk2(const orc_params* p, double dt_re, double dt_im, const double* in, double* out, double t[2]);
int orc_rk4(const orc_params* p, double dt_re, double dt_im, const double* in, double* out, double t[2]);
/* nsteps RK4 steps in place. */
int orc_rk4_n(const orc_params* p, double dt_re, double dt_im, double* psi, int nsteps, double t[2]);

/* Cash-Karp attempt rkck (src/solve.f90:248-378): new and err. */
int orc_rkck(const orc_params* p, double dt_re, double dt_im, const double* in, double* out, double* err);

/* One accepted rkqs step (src/solve.f90:158-244).  dt[2] is the complex dt,
 * updated in place; t[2] = (t, im_t); info[0] = errmax of the accepted
 * attempt, info[1] = number of rejected attempts, info[2] = RHS evaluations. */
int orc_rkqs(const orc_params* p, double dt[2], double t[2], const double* in, double* out, double info[3]);

/* Reductions (src/variables.f90:537-611, 698-709, 713-823). */
double orc_norm(const orc_params* p, const double* in);
double orc_energy(const orc_params* p, const double* in);
double orc_mass(const orc_params* p, const double* in);
void   orc_renormalise(const orc_params* p, double* var, double norm);

/* Line length of the zero lines of psi (linelength, src/variables.f90:879-1145),
 * times dx; counts: cells with >= 2 hits, with 1 hit, ended by a face, by an edge. */
double orc_linelength(const orc_params* p, const double* in, long long counts[4]);

/* Number of OpenMP threads the stencil loops use (1 without OpenMP). */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
