/*
 * gpe_hip.h -- C ABI of the MI355X (gfx950) GPE time-stepping library
 * (libgpehip.so).
 *
 * The library replaces the time-stepping hot path of the reference Fortran
 * solver (liangjj/gpe): the derivs module (src/derivs.f90), get_rhs and the
 * euler/rk2/rk4/rkqs/rkck steppers of the solve module (src/solve.f90), the
 * field reductions of the variables module (src/variables.f90) and the MPI
 * ghost-plane exchange (src/variables.f90:186-358), which becomes in-kernel
 * periodic or reflective indexing over one whole-domain grid on one GPU.
 *
 * Conventions (all entry points):
 *  - return 0 on success, a negative GPE_E* code on failure; the message is
 *    available from gpe_last_error(ctx) (or gpe_last_error(NULL) when no
 *    context exists).  The Fortran host turns a failure into
 *    emergency_stop(message), the reference's error convention
 *    (src/error.f90:24-46).
 *  - host arrays are borrowed for the duration of the call only; the context
 *    owns all device memory.  Fields are x-fastest (Fortran column-major),
 *    complex values interleaved (re, im): complex(c_double_complex) <->
 *    double _Complex, complex(c_float_complex) <-> float _Complex.
 *  - one host thread per context; every kernel of a context runs on that
 *    context's HIP stream (gpe_stream).  Calls that return host values
 *    synchronise the stream.
 */
#ifndef GPE_HIP_H
#define GPE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPE_ABI_VERSION 2

/* error codes */
#define GPE_OK 0
#define GPE_EINVAL (-1)   /* invalid argument */
#define GPE_EHIP (-2)     /* HIP runtime error */
#define GPE_ENOMEM (-3)   /* device allocation failed */
#define GPE_ESTATE (-4)   /* call out of order (e.g. accept without attempt) */

/* storage precision of device fields */
#define GPE_C128 0        /* complex128 fields, FP64 stencil arithmetic */
#define GPE_C64 1         /* complex64 fields, FP32 stencil, FP64 reductions and time */
#define GPE_MIXED 2       /* complex128 state and FP64 stencil; the RK4/RKF45 stage increments
                             (k_j = dt f(u_j) and partial Butcher sums) stored as complex64 */

/* host array element type */
#define GPE_HOST_C128 0
#define GPE_HOST_C64 1

/* time-stepping schemes: the reference's `scheme` values (src/solve.f90:34-50) */
#define GPE_EULER 0       /* 'euler' */
#define GPE_RK2 1         /* 'rk2'   */
#define GPE_RK4 2         /* 'rk4'   */
#define GPE_RK45 3        /* 'rk45'  (rkqs controller with Cash-Karp rkck) */

/* derivative operators (src/derivs.f90:20-21, src/variables.f90:55-71) */
#define GPE_D_X 0
#define GPE_D_Y 1
#define GPE_D_Z 2
#define GPE_D_XX 3
#define GPE_D_YY 4
#define GPE_D_ZZ 5
#define GPE_D_LAP 6

/* Run-time parameters: the subset of the reference's module parameters and
 * its run_params / eqn_params namelists the path reads
 * (src/parameters.f90:25-55, 89-93; src/parameters.in:22-29). */
typedef struct gpe_params {
    int32_t nx, ny, nz;     /* grid points per axis (compile-time nx,ny,nz in the reference) */
    int32_t order;          /* 2 or 4 */
    int32_t bcs;            /* 1 periodic, 2 reflective */
    int32_t eqn_to_solve;   /* 1..4 (src/solve.f90:414-437) */
    int32_t real_time;      /* 1: dt real, 0: imaginary time */
    int32_t storage;        /* GPE_C128, GPE_C64 or GPE_MIXED */
    double xr, yr, zr;      /* half widths: dx = 2 xr/(nx-1), x(i) = -xr + i dx (src/ic.f90:43-56) */
    double Urhs, diss_amp, g, mu, nn, omx, omy, omz;
    double vscale;          /* coefficient of the trap V in eqn 1: i/2(lap + (1-|psi|^2 - vscale V)psi);
                               0 reproduces the reference's eqn 1 */
    double eps, safety, dt_decrease, dt_increase; /* rkqs constants; reference: 1e-6, 0.9, -0.25, -0.2 */
} gpe_params;

/* In/out record of gpe_step.  t, im_t, dt are the reference's module
 * variables of the same names (src/parameters.f90:115-117). */
typedef struct gpe_step_info {
    double t, im_t;         /* in/out: advanced as the reference does */
    double dt_re, dt_im;    /* in/out: complex dt; for RK45 the dt to try next */
    double errmax;          /* RK45: errmax of the last accepted attempt */
    double psi2max;         /* RK45: max |psi|^2 of the last accepted state */
    int64_t n_accept;       /* accumulated accepted steps */
    int64_t n_reject;       /* accumulated rejected RK45 attempts */
    int64_t n_rhs;          /* accumulated RHS evaluations */
} gpe_step_info;

typedef struct gpe_ctx gpe_ctx;

int gpe_abi_version(void);

/* Last error message of ctx, or of the calling thread when ctx == NULL. */
const char* gpe_last_error(const gpe_ctx* ctx);

/* Create a context on the current HIP device: validates p, allocates the
 * state field, creates the context's stream.  Replaces the reference's
 * start-up of the path (get_grid src/ic.f90:32-69; array allocation
 * src/gpe.f90:57-62; decomposition set-up src/variables.f90:98-182). */
int gpe_ctx_create(const gpe_params* p, gpe_ctx** out);
void gpe_ctx_destroy(gpe_ctx* ctx);

/* Tile shape of the stencil kernels (tx in {32, 64}, ty in {4, 8, 16},
 * tz >= 1 planes per block, or 0 for the default).  Used by the autotuner. */
int gpe_set_tiles(gpe_ctx* ctx, int tx, int ty, int tz);
int gpe_get_tiles(const gpe_ctx* ctx, int* tx, int* ty, int* tz);

/* Fused stages (on by default) for complex128 or complex64 storage and a
 * right-hand side whose only optional term is the trap V (no Urhs,
 * dissipation or eqn-3 mask): gpe_step with GPE_RK4 runs the four stages as
 * two kernels that each evaluate two stages (5 field accesses per cell
 * instead of 13); with GPE_RK2 one kernel evaluates both stages of a step
 * (2 accesses instead of 5); with GPE_EULER one kernel evaluates two steps
 * (2 accesses instead of 4; an odd last step runs alone); each Cash-Karp
 * attempt of GPE_RK45 (and gpe_rk45_attempt) runs its six stages as three
 * such kernels (18 accesses instead of 36).  on = 0 selects the
 * stage-by-stage kernels.  No counterpart in the reference: a performance
 * switch of the same euler, rk2, rk4 and rkck (src/solve.f90:57-154,
 * 248-378). */
int gpe_set_fused(gpe_ctx* ctx, int on);

/* Copy a host field into the device state.  The host array has extent
 * nx * ny_alloc * (any) and the interior starts at (0, y_off, z_off): pass
 * (ny, 0, 0) for an interior array and (ny+4, 2, 2) for the reference's
 * halo array var(0:nx1, js-2:je+2, ks-2:ke+2) with one rank.  Replaces the
 * main-loop copy psi%old = psi%new and ghost fill (src/gpe.f90:165-174). */
int gpe_upload(gpe_ctx* ctx, const void* host, int host_dtype, int64_t ny_alloc, int64_t y_off, int64_t z_off);
/* Copy the selected field (the state unless gpe_select_field chose the
 * auxiliary field) into an interior host array (nx*ny*nz values). */
int gpe_download(gpe_ctx* ctx, void* host, int host_dtype);

/* Fields of a context: the state, which every stepper advances, and an
 * auxiliary field that receives host arrays handed to the reference's
 * diagnostics (get_norm(in_var, norm), energy(in_var, E), mass, momentum,
 * get_vcf, linelength, renormalise(var, norm), imprint_vortex_line(in_var):
 * src/variables.f90:537-611, 615-694, 467-507, 698-709, 879-1162), so that
 * passing a stale or auxiliary array (the reference's save_energy(t,
 * psi%old)) never replaces the device-resident state. */
#define GPE_FIELD_STATE 0
#define GPE_FIELD_AUX 1
/* Upload into `field` (layout as gpe_upload); GPE_FIELD_STATE is gpe_upload. */
int gpe_field_upload(gpe_ctx* ctx, int field, const void* host, int host_dtype, int64_t ny_alloc, int64_t y_off,
                     int64_t z_off);
/* The field that gpe_norm, gpe_mass, gpe_energy, gpe_hamiltonian,
 * gpe_momentum, gpe_velocity_pdf, gpe_vcf, gpe_linelength,
 * gpe_condensed_particles, gpe_peek, gpe_renormalise, gpe_imprint_line and
 * gpe_download act on until the next call (GPE_FIELD_STATE at creation).
 * gpe_step, gpe_solver, gpe_rk45_*, gpe_upload and gpe_deriv/gpe_rhs with
 * in == NULL always use the state.  GPE_FIELD_AUX needs a prior
 * gpe_field_upload of it (GPE_ESTATE otherwise). */
int gpe_select_field(gpe_ctx* ctx, int field);

/* The reference's module variable real_time (src/parameters.f90:26), which
 * the driver sets to .true. when imaginary-time relaxation converges
 * (src/gpe.f90:251-271): rkqs picks the direction of a retried step from it
 * (dt = (0, -|dt_temp|) in imaginary time, (|dt_temp|, 0) in real time,
 * src/solve.f90:195-201).  gpe_ctx_create takes it from gpe_params; a host
 * whose real_time changes during the run passes the current value before
 * each gpe_step or gpe_solver call. */
int gpe_set_real_time(gpe_ctx* ctx, int real_time);

/* Advance the device state.  scheme GPE_EULER/RK2/RK4: nsteps fixed steps
 * of dt = (io->dt_re, io->dt_im) (src/solve.f90:57-154).  GPE_RK45: nsteps
 * accepted steps of rkqs (src/solve.f90:158-244) starting from io->dt; on
 * return io->dt holds dt_next.  t and im_t advance as in the reference. */
int gpe_step(gpe_ctx* ctx, int scheme, int nsteps, gpe_step_info* io);

/* Host-array drop-ins of solver(var_in, var_out) (src/solve.f90:25-53):
 * upload var_in (halo layout as in gpe_upload), one step, download var_out. */
int gpe_solver(gpe_ctx* ctx, int scheme, const void* var_in, int64_t ny_alloc, int64_t y_off, int64_t z_off,
               void* var_out, int host_dtype, gpe_step_info* io);

/* One Cash-Karp attempt from the current state with step dt; dt0 is the
 * step held on entry to rkqs, used for the error scale
 * psi_scal = |psi| + dt0 |f(psi)| + 1e-30 (src/solve.f90:176-178).  f(psi)
 * is evaluated on the first attempt after each accept (or upload) and reused
 * on retries.  out[0] = errmax = max|err/psi_scal|/eps (src/solve.f90:184),
 * out[1] = max |psi_new|^2.  The candidate is kept until gpe_rk45_accept. */
int gpe_rk45_attempt(gpe_ctx* ctx, double dt_re, double dt_im, double dt0_re, double dt0_im, double out[2]);
/* Make the last attempt's candidate the state (src/solve.f90:232). */
int gpe_rk45_accept(gpe_ctx* ctx);

/* derivs module drop-in (src/derivs.f90:25-272) and laplacian
 * (src/variables.f90:55-71): out = D_op(in).  in == NULL uses the device
 * state; otherwise in has the gpe_upload layout.  out is an interior array. */
int gpe_deriv(gpe_ctx* ctx, int op, const void* in, int64_t ny_alloc, int64_t y_off, int64_t z_off, void* out,
              int host_dtype);
/* get_rhs (src/solve.f90:382-440), same argument conventions as gpe_deriv. */
int gpe_rhs(gpe_ctx* ctx, const void* in, int64_t ny_alloc, int64_t y_off, int64_t z_off, void* out,
            int host_dtype);

/* gpe_energy without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  psi is a device field of the grid and storage that
 * *p describes (nx*ny*nz values, x fastest); the energy of gpe_energy is
 * written to the device double *energy, all on `stream` (a hipStream_t, or
 * NULL), with no allocation and no synchronisation.  The kernels' block
 * partials live in a static device array of the library (65536 doubles,
 * 512 KiB, held by each device until the process ends); a grid whose tile
 * blocks outnumber it is refused (GPE_EINVAL).  Calls that overlap on one
 * device share that array and race.  Same kernels, tiling and summation
 * order as gpe_energy: the same result bit for bit. */
int gpe_energy_noalloc(const gpe_params* p, const void* psi, double* energy, void* stream);

/* Reductions of the device state, accumulated in FP64 in a fixed order:
 * get_norm with the end-corrected weights (src/variables.f90:537-559,
 * 713-823), energy (563-587), mass (591-611). */
int gpe_norm(gpe_ctx* ctx, double* norm);
int gpe_energy(gpe_ctx* ctx, double* energy);
int gpe_mass(gpe_ctx* ctx, double* mass);
/* A vortex line: line_param of the reference's parameters.in (a point on
 * the line, displacement amplitudes amp1 (sin) and amp2 (cos) of
 * wavelength ll, sign of the winding, direction, phase-only imprint). */
typedef struct gpe_line {
    double x0, y0, z0, amp1, amp2, ll, sgn;
    int32_t dir;            /* 0 'x', 1 'y', 2 'z' */
    int32_t imprint_phase;  /* 1: exp(i theta) only, no core */
} gpe_line;

/* psi <- |psi| vortex_line(vl) on the device state, with scal of
 * eqn_params: imprint_vortex_line (src/variables.f90:1148-1162) over
 * vortex_line, get_r and get_theta (src/ic.f90:327-348, 730-884).  Fails as
 * get_r does when ll = 0 with a non-zero amplitude. */
int gpe_imprint_line(gpe_ctx* ctx, const gpe_line* vl, double scal);

/* The state's value at cell (i, j, k) (x fastest).  No reference
 * counterpart: save_time (src/io.f90:147-195) reads psi at the box centre,
 * which a device-resident host fetches with this instead of the field. */
int gpe_peek(gpe_ctx* ctx, int32_t i, int32_t j, int32_t k, double out[2]);

/* Total length of the vortex lines (zero lines of psi) of the state:
 * linelength (src/variables.f90:879-1145), including its face and edge
 * rules and its reuse of the previous two-hit cell's second hit by a cell
 * with one hit (zero before the first such cell, where the reference reads
 * an undefined value).  Assumes dx = dy = dz, as the reference does. */
int gpe_linelength(gpe_ctx* ctx, double* length);

/* Condensed particles and spectra of the state: condensed_particles
 * (src/io.f90:315-404) and, when eta, nharm and F are given, spectrum
 * (src/io.f90:632-725), from the normalised backward FFT
 * a = FFT+(psi)/sqrt(nx ny nz) (fft 'backward', src/ic.f90:905-1001; hipFFT
 * here).  row = the mass.dat values after (time, im_t): M/(8 xr yr zr),
 * n0/V, M, n0, temp, temp2, H/V, with n0 = |a(0,0,0)|^2 and M the mass.
 * eta[m], nharm[m]: mean |a|^2 and mode count of log2 shell m+1 (eta 0 where
 * the shell holds nothing); F[kk-1], kk = 1..nx/2: sum of |a|^2 over the
 * modes with sqrt(k2) <= kk.  The first call allocates a field-sized FFT
 * buffer and a hipFFT plan, kept until gpe_ctx_destroy. */
int gpe_condensed_particles(gpe_ctx* ctx, double row[7], double eta[7], int64_t nharm[7], double* F);

/* Momentum (variables::momentum, src/variables.f90:615-694): mom[a] is the
 * end-corrected quadrature (weights as get_norm) of
 * (Re psi - 1) Im(d_a psi*) - Im psi Re(d_a psi), a = x, y, z. */
int gpe_momentum(gpe_ctx* ctx, double mom[3]);

/* Velocity statistics and PDF (get_pdf_velocity + get_pdf,
 * src/variables.f90:377-463, as save_velocity_pdf uses them,
 * src/io.f90:222-285): the reference's velocity -i/2 (psi* g - psi g*)/|psi|^2
 * with real gradient arrays g = g* = Re(d_a psi), i.e. -Im(psi) Re(d_a psi)/|psi|^2,
 * where |psi|^2 > 0.01.
 * Per component a: count[a] valid cells, mean[a], stdev[a] =
 * sqrt(<v^2> - mean^2), vmax[a] = max |v|, and pdf[a*nbins + j], the fraction
 * of the binned values in bin i = j - nbins/2 + 1, (2(i-1) vmax/nbins,
 * 2 i vmax/nbins].  nbins even, 2..4096.  Uses 24 bytes per cell of device
 * scratch. */
int gpe_velocity_pdf(gpe_ctx* ctx, int nbins, int64_t count[3], double mean[3], double stdev[3], double vmax[3],
                     double* pdf);

/* Velocity correlation function along x (get_vcf, src/variables.f90:467-507):
 * f[r] = Q(r)/Q(0), Q(r) = mean over (i, y, z), i + r < nx, of v_x(i) v_x(i+r),
 * with the same velocity, unmasked.  nx <= 8192. */
int gpe_vcf(gpe_ctx* ctx, double* f);

/* Conserved energy of the equation solved (eqn 1: sum |grad psi|^2 - |psi|^2
 * + |psi|^4/2; eqn 4: sum |grad psi|^2/2 + V|psi|^2 + g|psi|^4/2 - mu|psi|^2,
 * times dx dy dz), for drift reports.  No reference counterpart (the
 * reference's energy() omits V and mu). */
int gpe_hamiltonian(gpe_ctx* ctx, double* h);
/* psi <- sqrt(nn) psi / sqrt(norm) (src/variables.f90:698-709). */
int gpe_renormalise(gpe_ctx* ctx, double norm);

/* Device pointer of the state field (nx*ny*nz complex of the storage
 * precision) and the context's stream (a hipStream_t). */
/* Per-launch timing: with profiling on, every stencil launch of ctx is
 * bracketed by HIP events on the context's stream.  gpe_prof_read returns,
 * for one kernel kind, the summed duration, the launch count and the
 * algorithmic HBM bytes of one launch (fields read + written per cell times
 * the cell count).  No reference counterpart: measurement support. */
#define GPE_NKINDS 37
int gpe_prof_enable(gpe_ctx* ctx, int on);
int gpe_prof_read(gpe_ctx* ctx, int kind, double* total_ms, int64_t* launches, double* bytes_per_launch);
const char* gpe_kernel_kind_name(int kind);

/* Plain device copy of nbytes (a multiple of 16) on a HIP stream: the
 * calibration kernel for HBM byte counters.  No reference counterpart. */
int gpe_bench_copy(void* dst, const void* src, int64_t nbytes, void* stream);
/* Streaming kernel with three reads (a, b, e) and two writes (c, d) of
 * nbytes each: the achievable-bandwidth reference for the pointwise traffic
 * of a stage kernel.  No reference counterpart. */
int gpe_bench_stream5(void* c, void* d, const void* a, const void* b, const void* e, int64_t nbytes, void* stream);

/* ------------------------------------------------------------------------
 * Batched ensemble (no reference counterpart: the reference runs one field
 * per process group).  Members share order, bcs, eqn_to_solve and storage;
 * grid sizes and spacings differ.  RK4 only. */
#define GPE_ENS_PER_MEMBER 0  /* one launch per member and stage */
#define GPE_ENS_BATCHED 1     /* one persistent launch per stage for all members */
#define GPE_ENS_GRAPH 2       /* GPE_ENS_BATCHED captured in a hipGraph and replayed */
#define GPE_ENS_GROUPED 3     /* cache-sized groups advanced a chunk of steps each, in a hipGraph */
#define GPE_ENS_FUSED 4       /* fused RK4 (two persistent launches per step) for all members;
                                 complex128 members with the plain right-hand side */
#define GPE_ENS_FUSED_GRAPH 5 /* GPE_ENS_FUSED captured in a hipGraph and replayed */
#define GPE_ENS_PER_MEMBER_FUSED 6 /* the single-context fused RK4 (two launches) once per member:
                                      the baseline of the fused batched modes */
typedef struct gpe_ens gpe_ens;
int gpe_ens_create(const gpe_params* members, int n, gpe_ens** out);
void gpe_ens_destroy(gpe_ens* ens);
const char* gpe_ens_last_error(const gpe_ens* ens);
/* tile shape of the batch (tx, ty) in {32, 64} x {4, 8}; tz per member (NULL: default) */
int gpe_ens_set_tiles(gpe_ens* ens, int tx, int ty, const int* tz);
/* member field in the storage precision, nx*ny*nz values, x fastest */
int gpe_ens_upload(gpe_ens* ens, int member, const void* host, int host_dtype);
int gpe_ens_download(gpe_ens* ens, int member, void* host, int host_dtype);
int gpe_ens_step(gpe_ens* ens, int mode, double dt_re, double dt_im, int nsteps);
int gpe_ens_work_items(const gpe_ens* ens);
/* GPE_ENS_GROUPED: byte budget of a group's four fields (default 192 MiB)
 * and the steps a group advances per visit (default 16) */
int gpe_ens_set_grouping(gpe_ens* ens, int64_t budget_bytes, int steps_per_visit);
int gpe_ens_num_groups(gpe_ens* ens);
void* gpe_ens_stream(gpe_ens* ens);
int gpe_ens_sync(gpe_ens* ens);

/* ------------------------------------------------------------------------
 * FP32 MLP n_in -> 64 -> 64 -> 1 (tanh) of the learned RKF45 step-size
 * proposer: inference, gradients and Adam training on the GPU.  No
 * reference counterpart.  Parameters are packed W1[64][n_in], b1[64],
 * W2[64][64], b2[64], W3[64], b3; the loss is the mean squared error. */
typedef struct gpe_mlp gpe_mlp;
int gpe_mlp_create(int n_in, gpe_mlp** out);
void gpe_mlp_destroy(gpe_mlp* mlp);
const char* gpe_mlp_last_error(const gpe_mlp* mlp);
int gpe_mlp_num_params(const gpe_mlp* mlp);
int gpe_mlp_set_params(gpe_mlp* mlp, const float* params, int64_t n);
int gpe_mlp_get_params(gpe_mlp* mlp, float* params, int64_t n);
/* y[s] = mlp(x[s, :]) for s < n (x row-major n x n_in) */
int gpe_mlp_forward(gpe_mlp* mlp, const float* x, int n, float* y);
/* gradient of mean((mlp(x) - y)^2) over the n samples, and the loss */
int gpe_mlp_grad(gpe_mlp* mlp, const float* x, const float* y, int n, float* grad, float* loss);
/* `steps` Adam steps over cyclic mini-batches of `batch` samples (the data
 * set is taken in the given order); losses[s] receives each step's loss */
int gpe_mlp_train(gpe_mlp* mlp, const float* x, const float* y, int n, int steps, int batch, float lr, float* losses);

void* gpe_state_ptr(gpe_ctx* ctx);
void* gpe_stream(gpe_ctx* ctx);
int gpe_sync(gpe_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
