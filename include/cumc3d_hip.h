/*
 * cumc3d_hip.h - C-ABI of the MI355X implementation of CUMC3D's solver/mhd
 * update step (libcumc3d_hip.so).
 *
 * The reference runs the step through argument-less Fortran subroutines that
 * share module globals (solver/mhd/definition.f90:8-160) and keeps device
 * copies alive in an OpenACC data region (solver/mhd/openacc.f90:5-47).  This
 * library owns the device state instead; a context handle replaces the module
 * globals.  Plain pointers and sizes only; every entry point returns 0 on
 * success or a CUMC3D_E* code (cumc3d_error_string() gives the message).
 *
 * Host arrays use the reference's layout: prim is (imin:imax, -2:nx+3,
 * -2:ny+3, -2:nz+3) in Fortran order, i.e. the 8 variables
 * (rho, vx, vy, vz, p, Bx, By, Bz) of one cell are contiguous and x runs
 * fastest over cells; scalar fields (epsilon, cs, floors) are
 * (-2:nx+3, -2:ny+3, -2:nz+3); bcell is (ibx:ibz, ...) like prim.
 * Face fields follow the reference's staggering: prim(ibx,j,k,l) is Bx on
 * face j+1/2 (solver/mhd/fromrvetou.f90:54-56).
 */
#ifndef CUMC3D_HIP_H
#define CUMC3D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CUMC3D_OK = 0,
    CUMC3D_EINVAL = 1,      /* invalid descriptor or argument */
    CUMC3D_EHIP = 2,        /* HIP runtime error */
    CUMC3D_ENOMEM = 3,      /* device allocation failed */
    CUMC3D_ESTATE = 4,      /* call out of order (e.g. step before upload) */
    CUMC3D_ENONFINITE = 5,  /* dt not finite or not positive: the state holds NaN/negative energy */
    CUMC3D_ESTOP = 6        /* the model stopped the run where the reference STOPs (collapsar:
                               horizon beyond the inner face, model/collapsar/src/custom.f90:482-484;
                               Poisson solve converged on its last sweep, :644-647) */
};

/* Riemann solvers: parameter.h:91-101 (LF_flag, HLL_flag, HLLC_flag, HLLD_flag) */
enum { CUMC3D_LF = 0, CUMC3D_HLL = 1, CUMC3D_HLLC = 2, CUMC3D_HLLD = 3 };
/* Reconstructions: parameter.h:103-115 (tvdmm, tvdvl, tvdmc, ppmc, weno flags) */
enum { CUMC3D_TVDMM = 0, CUMC3D_TVDVL = 1, CUMC3D_TVDMC = 2, CUMC3D_PPM = 3, CUMC3D_WENO = 4 };
/* In-step model hooks: model/riemann, model/torus-cyn, model/torus-sph, model/collapsar */
enum { CUMC3D_MODEL_RIEMANN = 0, CUMC3D_MODEL_TORUS = 1, CUMC3D_MODEL_TORUS_SPH = 2, CUMC3D_MODEL_COLLAPSAR = 3 };
/* Arithmetic of the spatial operator */
enum { CUMC3D_FP64 = 0, CUMC3D_FP32 = 1 };
/* Value of HLLD's unassigned usstar(ibn) (rsolver_module.f90:704-741):
 * CUMC3D_BN_CARRY reproduces the serial reference, CUMC3D_BN_FIELD uses the
 * normal field (Miyoshi & Kusano 2005), CUMC3D_BN_ZERO uses 0. */
enum { CUMC3D_BN_ZERO = 0, CUMC3D_BN_FIELD = 1, CUMC3D_BN_CARRY = 2 };

/* GETGRID outputs the step reads (getgrid.f90:43-98, after CUSTOM_GRID and
 * the FIXPOLE override), in the reference's index ranges.  Optional: with
 * cumc3d_desc.metrics == NULL the library derives them from the face
 * coordinates by the same formulas, taking the faces as given before the
 * FIXPOLE override and applying it itself.  A caller holding the reference's
 * module arrays passes them here instead: GETGRID takes dy, y and the sines
 * before FIXPOLE moves yF(0) and yF(ny) to 0 and pi, so the overridden faces
 * alone do not reproduce them.  The spherical arrays may be NULL on other
 * geometries. */
typedef struct cumc3d_grid_metrics {
    const double *dx, *dy, *dz;          /* dx(-2:nx+3), dy(-2:ny+3), dz(-2:nz+3) */
    const double *x, *y, *z;             /* cell centres, same ranges */
    const double *dx_cb, *xbar;          /* dx_cb(-2:nx+3), xbar(-2:nx+3); entry -2 unused */
    const double *dcose, *dsine, *sine;  /* (-2:ny+3); entry -2 of dcose/dsine unused */
    const double *sinf;                  /* sinf(-3:ny+3) */
} cumc3d_grid_metrics;

/* The collapsar's self-gravity and black hole (module CUSTOM_DEF,
 * model/collapsar/src/def.f90:6-46, and param.h:37-44), as GET_MODEL leaves
 * them (model/collapsar/src/model.f90:44-52,239-242).  Read during
 * cumc3d_init only. */
typedef struct cumc3d_gravity_params {
    double prim_a[4];           /* atmosphere prim_a(irho:ivz): the floor CUSTOM_CHECKRHO resets to */
    double eps_a;               /* atmosphere epsilon eps_a */
    double m_bh;                /* black-hole mass m_bh at the start */
    double c_code;              /* clight*vel2code: r_sh = 2 m_bh/c_code^2 (custom.f90:481) */
    double omega;               /* SOR weight omega_weight (def.f90:31, 1.90) */
    double tolerance;           /* SOR relative tolerance (param.h:44, 1e-10) */
    int relax_max;              /* SOR sweep cap (param.h:41, 100000) */
    int n_pot;                  /* potential solved every n_pot steps (param.h:38, 10) */
} cumc3d_gravity_params;

typedef struct cumc3d_desc {
    int nx, ny, nz;             /* parameter.h:70-72 */
    int n_dim;                  /* parameter.h:37 */
    int coordinate_flag;        /* parameter.h:34: 0 Cartesian, 1 cylindrical (r,phi,z), 2 spherical (r,theta,phi) */
    int boundary_flag[6];       /* parameter.h:52: x-in,x-out,y-in,y-out,z-in,z-out; 0..4 */
    int riemann;                /* CUMC3D_LF .. CUMC3D_HLLD */
    int recon;                  /* CUMC3D_TVDMM .. CUMC3D_WENO */
    int model;                  /* CUMC3D_MODEL_* */
    int rt_gravity;             /* model/riemann/src/custom.f90:143: 0.1 g downward in y */
    int hlld_bn_mode;           /* CUMC3D_BN_* */
    int precision;              /* CUMC3D_FP64 or CUMC3D_FP32 */
    double cfl;                 /* parameter.h:81 */
    double ggas;                /* adiabatic index (DEFINITION ggas) */
    double r_sh;                /* model/torus-cyn/src/param.h:19, model/torus-sph/src/param.h:19 */
    const double *xF, *yF, *zF; /* face coordinates xF(-3:nx+3) etc. (n+7 values; GETGRID output) */
    int fixpole;                /* -DFIXPOLE (make.inc:8-21), spherical only: polar faces at theta
                                   0 and pi with sinf = 0 (getgrid.f90:91-98) and the radial edge
                                   EMFs on them averaged over phi (mhd_module.f90:437-461) */
    const cumc3d_grid_metrics *metrics;   /* optional, see cumc3d_grid_metrics */
    const cumc3d_gravity_params *gravity; /* CUMC3D_MODEL_COLLAPSAR only (FP64, spherical) */
} cumc3d_desc;

typedef struct cumc3d_ctx cumc3d_ctx;

/* SETUP_EQNS + BUILD_HYDRO + BUILD_RIEMANN + BUILD_MHD + GETGRID metrics and
 * the reconstruction weights (solver/mhd/initial.f90:7-61); allocates the
 * device state.  The xF/yF/zF arrays are copied. */
int cumc3d_init(const cumc3d_desc *desc, cumc3d_ctx **out);

/* POPULATE_DEVICE (openacc.f90:5-24): copy the initial model (GET_MODEL
 * output: prim and epsilon, plus rho/p/eps floor arrays for the torus
 * models, NULL otherwise) to the device. */
int cumc3d_upload(cumc3d_ctx *ctx, const double *prim, const double *eps, const double *floors3);

/* INITIAL_UPDATE (initial.f90:86-116): floors, cons, BOUNDARY, UPDATE(0).
 * For the collapsar UPDATE(0) includes CUSTOM_UPDATE(0): r_sh from m_bh and
 * the potential solved from a zero guess (model/collapsar/src/custom.f90:481-649). */
int cumc3d_initial_update(cumc3d_ctx *ctx);

/* FINDDT (rungekutta.f90:221-332): CFL time step, returned to the host.
 * CUMC3D_ENONFINITE when a cell's candidate is NaN or not positive (NaN or
 * negative energy in the state), where the reference would step on. */
int cumc3d_finddt(cumc3d_ctx *ctx, double *dt);

/* RUNGEKUTTA (rungekutta.f90:12-208): one SSP-RK3 step with the given dt. */
int cumc3d_rk3_step(cumc3d_ctx *ctx, double dt);

/* One stage of RUNGEKUTTA, for checking a step stage by stage: stage 1
 * (rungekutta.f90:37-103: u_old backup, SPATIAL, U1 = u_old + dt L,
 * FROMUTORVE, CUSTOM_CHECKRHO, FROMRVETOU, BOUNDARY, UPDATE(1)), stage 2
 * (:105-147) or stage 3 (:149-199, which ends the step).  The stages of a
 * step run 1, 2, 3 in order with the step's dt (CUMC3D_ESTATE otherwise);
 * cumc3d_rk3_step and cumc3d_run refuse to start while a step is part done.
 * Stages 1-3 in a row give exactly cumc3d_rk3_step. */
int cumc3d_stage(cumc3d_ctx *ctx, int stage, double dt);

/* SPATIAL (spatial.f90:11-162) on the current state, for checking the
 * operator on its own: l_rk(1:8, -2:nx+3, -2:ny+3, -2:nz+3) in the AoS layout
 * above.  Written: l_rk(1:5) of cells 1..n, and l_rk(6:8) of the faces the
 * update advances (index 0..n along the field component, 1..n across);
 * every other element keeps the value it has in l_rk on entry.  The state
 * is not changed.  Like the reference's SPATIAL it rewrites the fluxes and
 * EMFs and advances the HLLD carry (CUMC3D_BN_CARRY), which the next stage
 * reads. */
int cumc3d_spatial(cumc3d_ctx *ctx, double *l_rk);

/* The time loop of main.f90:97-139 on the device: up to max_steps steps of
 * FINDDT, dt clip to total_time, RUNGEKUTTA, global_time += dt, stopping once
 * global_time >= total_time.  With sync_each_step = 0 the steps are queued
 * without host round trips (steps past total_time become no-ops).  *t is the
 * global time on entry and on return; dts (may be NULL) receives each dt. */
int cumc3d_run(cumc3d_ctx *ctx, int max_steps, double total_time, int sync_each_step, double *t,
               double *dts, int *steps_done);

/* The largest max_steps cumc3d_run_noalloc logs dts for. */
#define CUMC3D_NOALLOC_MAX_STEPS 65536

/* cumc3d_run without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, steps and results as cumc3d_run; the per-step dt log that
 * cumc3d_run allocates lives in a static device array instead (one per GPU,
 * 512 KiB, held until the process ends), so dts may receive at most
 * CUMC3D_NOALLOC_MAX_STEPS entries (CUMC3D_EINVAL beyond that; with dts =
 * NULL any max_steps runs).  It calls no hipMalloc/hipFree and refuses kernel
 * timing (cumc3d_set_timing on: CUMC3D_ESTATE), whose events allocate.  All
 * calls on one GPU share that array: calls that overlap on one GPU (two host
 * threads) race on it. */
int cumc3d_run_noalloc(cumc3d_ctx *ctx, int max_steps, double total_time, int sync_each_step, double *t,
                       double *dts, int *steps_done);

/* Copy the device state back (replaces !$ACC UPDATE HOST, dataio.f90:26).
 * Any pointer may be NULL.  prim and bcell use the AoS layout above. */
int cumc3d_download(cumc3d_ctx *ctx, double *prim, double *eps, double *bcell, double *cs);

/* The three arrays print_hydroprofile writes (dataio.f90:178-233), packed on
 * the device in its order (Fortran column-major) and copied to the host:
 *   primitive  prim(1:5, 0:nx+1, 0:ny+1, 0:nz+1)   (rho, vx, vy, vz, p)
 *   epsilon    epsilon(0:nx+1, 0:ny+1, 0:nz+1)
 *   bfield     prim(ibx:ibz, -1:nx+1, -1:ny+1, -1:nz+1)   (face field)
 * i.e. 5(nx+2)(ny+2)(nz+2), (nx+2)(ny+2)(nz+2) and 3(nx+3)(ny+3)(nz+3)
 * doubles.  Any pointer may be NULL.  Replaces the host update and array
 * sections of dataio.f90:26,220,196; libcumc3d_io writes the file
 * (include/cumc3d_io.h). */
int cumc3d_profile(cumc3d_ctx *ctx, double *primitive, double *epsilon, double *bfield);

/* The collapsar's gravity state (module CUSTOM_DEF, model/collapsar/src/def.f90:16-36):
 * phi(-2:nx+3, -2:ny+3, -2:nz+3) in Fortran order (cells 0..n+1 are the
 * solved potential and its boundary values, the outer ghost layers are 0),
 * m_bh and r_sh after the last UPDATE, and the DO variable n of the last
 * Poisson solve (relax_max + 1 when it ran out without converging).  Any
 * pointer may be NULL; CUMC3D_EINVAL on other models. */
int cumc3d_gravity(cumc3d_ctx *ctx, double *phi, double *m_bh, double *r_sh, int *iterations);

/* FIND_DIVB (mhd_module.f90:110-201): max |div B| over the interior. */
int cumc3d_divb(cumc3d_ctx *ctx, double *maxdivb);

/* CLEAR_DEVICE (openacc.f90:30-47): release the device state. */
int cumc3d_finalize(cumc3d_ctx *ctx);

const char *cumc3d_error_string(int code);
/* Message of the last error on this thread (empty if none). */
const char *cumc3d_last_error(void);

/* Measurement helpers (no reference counterpart). */
/* Mean device time in ms per launch of each kernel class over the last
 * cumc3d_run call with timing enabled: [0] x-flux, [1] y-flux, [2] z-flux,
 * [3] edge EMF, [4] update, [5] ghost fill, [6] whole step, [7] selector
 * inference (per refresh); for the collapsar [8] one Poisson solve and [9]
 * its mean number of SOR sweeps (a count, not ms).  n up to 10. */
int cumc3d_set_timing(cumc3d_ctx *ctx, int enable);
int cumc3d_get_timing(cumc3d_ctx *ctx, double *ms_per_launch, int n);

/* Learned reconstruction selector (no reference counterpart: the reference's
 * reconstruction is one compile-time choice, spatial.f90:33-43).  A 15->32->32->3
 * MLP (BF16 weights on the matrix cores, FP32 accumulation) picks TVD-MC, PPM
 * or WENO-Z per cell and sweep direction; the flux kernels run the highest
 * class among a wave's cells.  Parameters: CUMC3D_SEL_NPARAM floats, layout
 * W1[16][32], b1[32], W2[32][32], b2[32], W3[32][3], b3[3] (input 16 is a zero pad). */
enum { CUMC3D_SEL_NPARAM = 1699 };
/* Allocate the selector (sample buffer of `capacity` samples) and set fixed-seed
 * He-uniform initial weights.  Until the first inference every cell is WENO-Z. */
int cumc3d_selector_init(cumc3d_ctx *ctx, int capacity, unsigned seed);
int cumc3d_selector_set_params(cumc3d_ctx *ctx, const float *w, int n);
int cumc3d_selector_get_params(cumc3d_ctx *ctx, float *w, int n);
/* Offer training samples from the current state: every stride-th cell of each
 * sweep region, labelled 2 (WENO-Z) where the relative gap between TVD-MC and
 * WENO-Z faces of rho or p exceeds thr_weno, 1 (PPM) above thr_ppm, else 0.
 * The buffer is a reservoir over every candidate offered since
 * cumc3d_selector_init: once full, each new candidate replaces a random held
 * sample with probability capacity / offered, so later snapshots are held as
 * well as earlier ones.  counts (may be NULL) receives {samples held, and the
 * class-0, class-1, class-2 counts of the samples held}. */
int cumc3d_selector_collect(cumc3d_ctx *ctx, int stride, double thr_ppm, double thr_weno, unsigned seed,
                            int *counts);
/* iters Adam steps on minibatches of blocks*128 samples, class-weighted
 * cross-entropy; loss (may be NULL) receives the mean loss of each step. */
int cumc3d_selector_train(cumc3d_ctx *ctx, int iters, int blocks, double lr, const double *class_weight,
                          unsigned seed, float *loss);
/* Use the selector in the flux sweeps (refreshing the class masks every `every`
 * steps, at stage 1), or the descriptor's reconstruction (enable = 0). */
int cumc3d_selector_enable(cumc3d_ctx *ctx, int enable, int every);
/* Logits of n feature rows (16 floats each) through the MFMA inference path. */
int cumc3d_selector_mlp(cumc3d_ctx *ctx, const float *features, int n, float *logits);
/* Class mask (one byte per ghosted-box cell) of sweep direction dir, after a
 * fresh inference on the current state if refresh != 0. */
int cumc3d_selector_mask(cumc3d_ctx *ctx, int dir, int refresh, unsigned char *out);

#ifdef __cplusplus
}
#endif
#endif
