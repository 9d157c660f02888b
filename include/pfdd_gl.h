/*
 * pfdd_gl.h -- C ABI of the MI355X (gfx950) Ginzburg-Landau phase-field solver.
 *
 * One handle owns one periodic grid of `ns` complex order parameters (one per
 * slip system) resident on one GPU, the packed Green's-function matrix B(k)
 * and the work buffers of the GL step.  Plain pointers and sizes only; every
 * function returns 0 (PFDD_OK) or an error code, with a message available from
 * pfdd_gl_last_error() on the calling thread.
 *
 * Each entry point names the reference interface it replaces
 * (lanl/Phase-Field-Dislocation-Dynamics-PFDD-2.x, src/ directory).
 *
 * Host layout of a field is the reference's `fft->xi[0]`:
 *   xi[2*(k + j*nz + i*ny*nz + a*nx*ny*nz) + {0: re, 1: im}]  (float64)
 *
 * Threading: a handle is used from one host thread at a time; all its work is
 * ordered on one HIP stream.  Different handles are independent.
 */
#ifndef PFDD_GL_H
#define PFDD_GL_H

#ifdef __cplusplus
extern "C" {
#endif

#define PFDD_GL_MAX_SYSTEMS 12
#define PFDD_GL_MAX_COEF 8

/* pfdd_gl_desc.flags */
#define PFDD_GL_FLAG_STORED_B 1 /* build packed B(k) once and stream it every step (the
                                   reference's BB array) instead of evaluating the Green's
                                   function on the fly inside the x pass */
#define PFDD_GL_FLAG_AUTOTUNE 2 /* run pfdd_gl_autotune(h, 4, NULL) at the end of create */

enum pfdd_gl_status {
    PFDD_OK = 0,
    PFDD_ERR_ARG = 1,         /* invalid argument */
    PFDD_ERR_UNSUPPORTED = 2, /* grid size / system count not built in */
    PFDD_ERR_HIP = 3,         /* HIP runtime error */
    PFDD_ERR_NONFINITE = 4,   /* a reduction produced NaN/Inf (field diverged) */
    PFDD_ERR_STATE = 5        /* call out of order */
};

enum pfdd_gl_precision { PFDD_FP64 = 0, PFDD_FP32 = 1 };

/* Core-energy models (App::core_energy, app.cpp:663-672).  All coefficients
 * are already divided by mu*dslip*b, as the reference does inside each
 * core_energy_* routine. */
enum pfdd_gl_core {
    PFDD_CORE_NONE = 0,
    PFDD_CORE_PERFECT = 1, /* flag 1, fft_fftw_slab.cpp:1130-1191: coef[0] = An'            */
    PFDD_CORE_EXTENDED = 2,/* flag 2, :1243-1354: gamma surface coupling the 3 systems of
                              each plane (ns % 3 == 0): coef[0..6] = c0..c4, a1, a3 over
                              (mu dslip b)                                                  */
    PFDD_CORE_SINE = 3,    /* flag 3, :1360-1418: coef[0] = isf', coef[1] = Cn'             */
    PFDD_CORE_BCC = 5,     /* flag 5, :1071-1125 with gradient() :869-1065: coef[0] = usf',
                              scaled per element by a quadratic in the dislocation
                              character angle from the gradient of Re xi (no amp[])       */
    PFDD_CORE_PYRII = 4,   /* flag 4, :1429-1493 as executed: Fourier density
                              aa0 + sum_m aa_m cos(2 pi m x) + bb_m sin(2 pi m x), m <= 4,
                              coef[0..8] = aa0..aa4, bb1..bb4 over (mu dslip b); its force
                              reaches only the last cell of each system (the reference's
                              dE_core assignment follows the cell loop, :1491)             */
    PFDD_CORE_MPEA = 7,    /* flag 7, :1196-1235: amp[] = A_mpea/(mu dslip b) per cell      */
    PFDD_CORE_SERIES = 8,  /* sine series sum_m c_{a,m} sin^2(m pi x), optional amp[]:
                              coef[a*ncoef + m-1], generalises flags 1/3/7 (and 6, whose
                              An' the host recomputes per stress state)                   */
    PFDD_CORE_FOURIER = 9  /* the Fourier density of flag 4 with its force on every cell
                              (the model core_energy_pyrII describes); optional amp[]      */
};

typedef struct pfdd_gl pfdd_gl;

/* Everything FFTW_Slab::init/setup (fft_fftw_slab.cpp:52-119) and the app's
 * set_slip/constructor supply to the GL loop. */
typedef struct pfdd_gl_desc {
    int nx, ny, nz;      /* grid: fft->nx, ny, nz (create_sites.cpp:228-232); each one of
                            8 16 32 40 48 64 80 96 128 160 192 256 512                  */
    int ns;              /* slip systems: app->slip_systems                             */
    int precision;       /* PFDD_FP64 (plain) or PFDD_FP32                              */
    double four[9];      /* reciprocal matrix fft->four, row-major (app.cpp:396-431)    */
    double C44, C12, C11;/* material stiffness (material.cpp:156-160)                  */
    double S12, S44;     /* material compliance (material.cpp:162-164), diagnostics     */
    double dslip;        /* material->dslip                                             */
    const double *xn;    /* ns*3 slip-plane normals, already rotated by set_slip        */
    const double *xb;    /* ns*3 Burgers vectors, already rotated by set_slip           */
    double cddt;         /* app->CD * app->timestep                                     */
    int core_model;      /* enum pfdd_gl_core                                           */
    int ncoef;           /* coefficients per system for PFDD_CORE_SERIES (<= 8)         */
    const double *coef;  /* see enum pfdd_gl_core                                       */
    const double *amp;   /* optional ns*N per-cell amplitude (flags 7, 8), may be NULL  */
    const double *xo;    /* optional ns*N obstacle mask fft->xo (nonzero = frozen)      */
    int device;          /* HIP device ordinal                                          */
    void *stream;        /* hipStream_t to order all work on; NULL = handle's own       */
    int flags;           /* PFDD_GL_FLAG_* (0: on-the-fly B(k), FP64 arithmetic)         */
} pfdd_gl_desc;

/* Scalars of the last executed GL iteration (update_order_parameter,
 * fft_fftw_slab.cpp:1746-1790, plus the energies of the pre-update state). */
typedef struct pfdd_gl_reduce {
    double xinorm;      /* fft->xinorm: sqrt(sum |dxi|^2), quirk Q3 under masks        */
    double xiave;       /* fft->xiave: sum over systems and cells of Re xi (a sum)     */
    double e_elastic;   /* (1/2N) sum_k xihat^* B xihat                                 */
    double e_core;      /* sum_r sum_a core energy density                              */
    double e_stress;    /* sum_a tau_a sum_r Re xi_a                                    */
    double e_total;     /* e_elastic + e_core - e_stress                                */
    double e_core_ref;  /* fft->E_core as the reference accumulates it (sum / nx)       */
    int steps_done;     /* GL iterations executed by the last pfdd_gl_step call        */
    int converged;      /* 1 if it stopped because xinorm < tol                         */
} pfdd_gl_reduce;

/* Library version, (major << 16) | minor. */
int pfdd_gl_version(void);

/* Practical HBM streaming rates of GPU `device` (no counterpart in the
 * reference; context for roofline figures): GB/s of a copy (1 read + 1
 * write) and of an add (2 reads + 1 write) over arrays of `bytes` each, with
 * the nontemporal 16-byte accesses of the field passes, averaged over `reps`
 * launches after 20 untimed ones.  Allocates and frees three arrays. */
int pfdd_gl_stream_rates(int device, long long bytes, int reps, double *copy_GBps, double *add_GBps);

/* Message of the last failing call on this thread ("" if none). */
const char *pfdd_gl_last_error(void);

/* Replaces FFTW_Slab::init + allocate + setup (fft_fftw_slab.cpp:52-237,
 * 111-119): allocates device fields, builds B(k) (Bmatrix, :466-638) on the
 * device in FP64.  The initial xi is zero. */
int pfdd_gl_create(const pfdd_gl_desc *desc, pfdd_gl **out);

/* A batch of `nbatch` independent runs of the same grid, slip systems and
 * material in one handle (no counterpart in the reference, which runs one
 * grid per process): the fields of all members are concatenated, member m's
 * systems being m*ns .. m*ns+ns-1, and every kernel of a GL iteration covers
 * all members in one launch.  Per member: initial field, tau (set_tau takes
 * nbatch*ns values), amp and xo (nbatch*ns*N values); shared: core
 * coefficients.  Reductions of pfdd_gl_step are sums over all members;
 * per-member sums of Re xi come from pfdd_gl_xi_sums (nbatch*ns values).
 * Output-step diagnostics and reverse mode are single-run only. */
int pfdd_gl_create_batch(const pfdd_gl_desc *desc, int nbatch, pfdd_gl **out);

/* Frees everything the handle owns.  (The reference never frees its FFT
 * arrays: ~FFTW_Slab is empty, fft_fftw_slab.cpp:43-46.) */
int pfdd_gl_destroy(pfdd_gl *h);

/* Upload / download the order parameter in the host layout of fft->xi[0]
 * (the reference's initial_sxtal_* writers and compute_* readers). */
int pfdd_gl_set_xi(pfdd_gl *h, const double *xi_host);
int pfdd_gl_get_xi(pfdd_gl *h, double *xi_host);

/* Device-resident access: pointer to the current field on the device (layout
 * as above, element type double2 for FP64 or float2 for FP32).  The pointer
 * changes after every GL iteration (ping-pong buffers); fetch it again. */
int pfdd_gl_xi_device(pfdd_gl *h, void **dptr);
/* Copy a device buffer (same element type/layout) into the current field,
 * and the current field into a device buffer. */
int pfdd_gl_set_xi_device(pfdd_gl *h, const void *dptr);
int pfdd_gl_get_xi_device(pfdd_gl *h, void *dptr);

/* Resolved shear stresses tau_a in units of mu: the result of
 * App::resolSS -> FFTW_Slab::resolSS_Schmid (fft_fftw_slab.cpp:737-769). */
int pfdd_gl_set_tau(pfdd_gl *h, const double *tau);

/* Replace the core-energy coefficients (same meaning as in the descriptor),
 * e.g. per stress state for flag 6 (core_energy_USFE_angle_tau, :1500-1545). */
int pfdd_gl_set_core_coef(pfdd_gl *h, int ncoef, const double *coef);

/* Runs up to `nsteps` GL iterations at the current tau: the inner loop of
 * SolveGL::iterate (solve_GL.cpp:86-148) -- core_energy, prep_forward,
 * internal_energy, prep_backward, update_order_parameter, prepare_next_itr --
 * stopping (device-side, no host round trip) after the first iteration with
 * xinorm < tol (tol <= 0: never).  If `out` is non-NULL the call waits for
 * the GPU and reports the last executed iteration; otherwise it only
 * enqueues the work.  Runs of 20 or more iterations replay a graph of 10
 * iterations (captured once per launch configuration and tol, for both
 * ping-pong parities together, at the end of autotuning or in the first
 * long run); PFDD_GL_GRAPH=0 in the environment at pfdd_gl_create launches
 * every iteration eagerly. */
int pfdd_gl_step(pfdd_gl *h, int nsteps, double tol, pfdd_gl_reduce *out);

/* Per-system sums of Re xi over all cells of the current field (used by the
 * stress-free strain average, stressfree_strain :3252-3326). */
int pfdd_gl_xi_sums(pfdd_gl *h, double *sum_re);

/* pfdd_gl_xi_sums without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under stream capture,
 * tight loops, harnesses that forbid hipMalloc).  Same result as
 * pfdd_gl_xi_sums, bit for bit.  Its block partials live in a fixed device
 * array of (256 + 1) * PFDD_GL_NOALLOC_MAX_SYSTEMS doubles (2.06 MB, held on
 * each GPU the library runs on until the process ends); handles with more
 * than PFDD_GL_NOALLOC_MAX_SYSTEMS systems (nbatch * ns) get
 * PFDD_ERR_UNSUPPORTED.  All calls on one GPU share that array: calls that
 * overlap in time on the same GPU (other threads or streams) race. */
#define PFDD_GL_NOALLOC_MAX_SYSTEMS 1024
int pfdd_gl_xi_sums_noalloc(pfdd_gl *h, double *sum_re);

/* Output-step averages: DiagStrain::compute + DiagStress::compute
 * (diag_strain.cpp:44-61, diag_stress.cpp:43-73) i.e. strain(),
 * total_average_strain(), stress() (fft_fftw_slab.cpp:3252-3553), for the
 * current (post-update) field.  sigma_rot in units of mu; results in Pa-based
 * units exactly as fft->ave_eps / fft->ave_sigma (row-major 3x3). */
int pfdd_gl_diag(pfdd_gl *h, const double *sigma_rot, double ave_eps[9],
                 double ave_sigma[9]);

/* Same as pfdd_gl_diag but also returns the fields the dump command reads:
 * data_eps and data_sigma (2*9*N doubles each, layout of fft->data_eps /
 * fft->data_sigma: [2*(r + i*N + j*3N) + {re,im}]).  The strain field is built
 * from the k-space field of the last step's pre-update xi, as the reference's
 * output step does (strain() reads data_fftw); before any step it uses the
 * current field.  Either output pointer may be NULL. */
int pfdd_gl_diag_fields(pfdd_gl *h, const double *sigma_rot, double ave_eps[9],
                        double ave_sigma[9], double *data_eps, double *data_sigma);

/* Reverse mode (no counterpart in the reference: it has no derivatives).
 * pfdd_gl_tape_record runs `nsteps` GL iterations from the current field at
 * the current tau and core coefficients, exactly as pfdd_gl_step(h, nsteps,
 * 0, out) does, and keeps a checkpoint of xi every `every` iterations
 * (ceil(nsteps/every) fields on the device).  pfdd_gl_tape_backward then takes
 * dL/dxi_T for the final field (layout of fft->xi[0]: Re and Im parts are
 * independent real variables, i.e. the torch convention for a real loss)
 * and returns, by recomputing each segment from its checkpoint (`every`-1
 * more fields) and running the adjoint iteration backwards:
 *   dxi_0[2*S*N]        dL/dxi_0 (may be NULL)
 *   dtau[S]             dL/dtau_a (may be NULL)
 *   dcoef[S*ncoef]      (ncoef of the handle's core model, <= 8)
 *                       dL/dc_{a,m} of the core-energy coefficients in their
 *                       sine-series form dE = A sum_m c_m m pi sin(2 m pi x)
 *                       (PERFECT: m = 1 -> An'; SINE: m = 1, 2 -> isf', Cn';
 *                       FOURIER / PYRII: aa1..aa4, bb1..bb4, ncoef 8 -- aa0
 *                       does not enter the force; EXTENDED: c0..c4, a1, a3,
 *                       ncoef 7 -- c0 does not enter the force;
 *                       per system -- sum over a for a shared coefficient)
 * Core model 5 (bcc_perfect) has no reverse mode.
 * Requires nx*ny to be a multiple of the inverse-z tile (every power-of-two
 * and 48/96 grid).  The _device variant takes device pointers (double2
 * layout, any precision handle). */
int pfdd_gl_tape_record(pfdd_gl *h, int nsteps, int every, pfdd_gl_reduce *out);
int pfdd_gl_tape_backward(pfdd_gl *h, const double *dxi_T, double *dxi_0, double *dtau, double *dcoef);
int pfdd_gl_tape_backward_device(pfdd_gl *h, const void *dxi_T, void *dxi_0, double *dtau, double *dcoef);

/* gradient() (fft_fftw_slab.cpp:869-1065): the dislocation character angle
 * theta of every system and cell (S*N doubles, [a][i][j][k]; -1 where the
 * line direction |grad Re xi x n| <= 0.1), computed as the reference does at
 * the start of the last GL iteration, i.e. from the field before its update
 * (the current field before the first step); the in-place transform to
 * global axes of :1020-1029 included.  What the apps' compute_theta dumps. */
int pfdd_gl_theta(pfdd_gl *h, double *theta);

/* Autotuner (no counterpart in the reference): times each candidate LDS tile
 * shape (0, 1, 2: about 1K/2K/4K FP64 or 2K/4K/8K FP32 elements per tile) of
 * the four line passes -- forward z, forward y, inverse y, inverse z + update
 * -- plus, for the first three, the FFT radix plan: choice 3, 4, 5 are shapes
 * 0, 1, 2 with radix-16 stages wherever they save a stage (lengths 16, 48,
 * 80, 128, 256: 128 = 16 x 8 instead of 8 x 8 x 2); 0-2 use radix 8 first.
 * choice[4] is the iteration variant:
 *   0  five kernels, x pass one tile per workgroup;
 *   1  five kernels, x pass on persistent workgroups that load the next tile
 *      while transforming the current one;
 *   2  three kernels (the "plane" iteration: 2-D FFT of each y-z plane in LDS,
 *      x pass, 2-D inverse FFT fused with the update), for square y-z planes
 *      that fit in LDS (FP32 up to 128 x 128, FP64 up to 80 x 80), nx of
 *      16..128, on-the-fly B.
 * Timing runs on this handle's own buffers: after 20 warm-up iterations, two
 * interleaved rounds of `nsteps` GL iterations per candidate, keeping each
 * kernel's best time; the fastest shape and plan per line pass and the
 * faster x-pass variant form a mix, which is timed as a whole iteration
 * against the two uniform candidates with the smallest iteration sums (the
 * fastest of the three wins), and variant 2 replaces the winner if its whole
 * iteration is faster (choice[5], may be NULL).  The field and step counters
 * are restored; the choice is cached for the process per grid, system count,
 * batch, precision and options.  Tile shapes and iteration variants give the
 * same xi bit for bit; the two radix plans round differently (relative
 * differences of a few ulp per transform).  The update of core models 2, 4
 * and 9 always uses tile shape 0. */
int pfdd_gl_autotune(pfdd_gl *h, int nsteps, int choice[5]);
/* Sets the choices explicitly (same meaning as pfdd_gl_autotune's output). */
int pfdd_gl_set_tuning(pfdd_gl *h, const int choice[5]);

/* Average device time (ms) of each kernel of one GL iteration over `nsteps`
 * iterations, measured with HIP events on the handle's stream that the kernel
 * dispatches record themselves (start and end of each kernel).  ms[0..4] =
 * forward z pass, forward y pass, x pass with contraction, inverse y pass,
 * inverse z pass with update; ms[5] = reduction; ms[6] = whole iteration.
 * nkernels receives the number of kernels of one iteration. */
int pfdd_gl_profile(pfdd_gl *h, int nsteps, double *ms, int *nkernels);

/* Bytes the GL iteration moves, by the kernel split above (algorithmic
 * figure: each field element read/written once per kernel, B once). */
int pfdd_gl_traffic(pfdd_gl *h, double *bytes_per_kernel, double *bytes_per_step);

/* Several handles (e.g. batches of different grid sizes) stepped together by
 * one hipGraph of `iters` (even) GL iterations, replayed nsteps/iters times
 * (the remainder runs eagerly); no host synchronisation inside the run.  No
 * counterpart in the reference.  The graph is captured on first use and again
 * after set_core_coef, a single-run handle's set_tau, a change of tol or an
 * odd number of steps taken outside the graph.  pfdd_gl_graph_run reports
 * each handle's last iteration in out[i] (may be NULL); tol as in
 * pfdd_gl_step, per handle. */
typedef struct pfdd_gl_graph pfdd_gl_graph;
int pfdd_gl_graph_create(pfdd_gl *const *handles, int n, int iters, pfdd_gl_graph **out);
int pfdd_gl_graph_run(pfdd_gl_graph *g, int nsteps, double tol, pfdd_gl_reduce *out);
int pfdd_gl_graph_destroy(pfdd_gl_graph *g);

/* Conductivity sub-problem of solve_style GLCond (solve_GL+cond.cpp:70-150):
 * a scalar complex field c (fft->condxi, host layout [i][j][k][re,im]) on
 * the grid, advanced by one call per iteration of
 *   forwardcond                 fft_fftw_slab.cpp:1999-2006  (FFT of c)
 *   divcurrent                  :1704-1741   a = i c f_y/f^2, b = c f_y^2/f^2
 *   backwardcond                :2057-2191   unnormalised inverse FFTs of a, b;
 *                                            j = sig (El - a/eps0), rhs = (sig/eps0) b
 *                                            + (a/eps0 - El) dsig, conductivity
 *   update_conduct_order_param  :1794-1835   c -= rate rhs / N
 *   prepare_next_conditr        :1866-1886   (c is read directly: no copy)
 * It does not read the dislocation field (the reference couples the two only
 * through the material constant rho_disloc).  The per-row layer
 * conductivities of backwardcond are computed on the host
 * (pfdd_host_cond_layers in include/pfdd_host.h). */
typedef struct pfdd_gl_cond pfdd_gl_cond;
typedef struct pfdd_gl_cond_desc {
    int nx, ny, nz;
    double four[9];        /* reciprocal matrix fft->four, row-major                       */
    const double *sigma;   /* ny: condsigma of row j (backwardcond :2118-2159)             */
    const double *dsigma;  /* ny: the condsigma_array difference of row j (:2171-2182)     */
    double El;             /* material->El[1], applied field along y (nonzero)            */
    double epsilonzero;    /* material->epsilonzero                                        */
    double rate;           /* 1e-22 * app->timestep (:1821)                                */
    int device;
    void *stream;          /* hipStream_t; NULL = handle's own                             */
} pfdd_gl_cond_desc;
typedef struct pfdd_gl_cond_reduce {
    double conductivity;        /* fft->conductivity (one rank): javerage_y.re / El      */
    double javerage_re, javerage_im;
    double condxinorm, condxiave;
    double conductivity_repeat; /* what the trailing backwardcond of SolveGLCond::iterate
                                   (solve_GL+cond.cpp:136-139) prints: backwardcond run again
                                   on its own real-space output                          */
} pfdd_gl_cond_reduce;
int pfdd_gl_cond_create(const pfdd_gl_cond_desc *desc, pfdd_gl_cond **out); /* c = 0 (FFTW_Slab::allocate) */
int pfdd_gl_cond_destroy(pfdd_gl_cond *h);
int pfdd_gl_cond_set(pfdd_gl_cond *h, const double *c_host);
int pfdd_gl_cond_get(pfdd_gl_cond *h, double *c_host);
/* nsteps iterations; out receives the scalars of the last one. */
int pfdd_gl_cond_step(pfdd_gl_cond *h, int nsteps, pfdd_gl_cond_reduce *out);

#ifdef __cplusplus
}
#endif

#endif /* PFDD_GL_H */
