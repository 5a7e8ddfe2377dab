/*
 * CPU restatement of the reference's FP64 MHD update step (TEST INFRASTRUCTURE).
 *
 * This oracle is the checker for the MI355X library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, never
 * the product path.  It follows the reference loop by loop (ranges, ghost-fill
 * order and floating-point operation order) so it can be pinned bit-for-bit
 * against the reference compiled from its own sources (oracle/_ref).
 *
 * Variable indices follow SETUP_EQNS (solver/mhd/setup.f90:39-96):
 * irho=1 ivx=2 ivy=3 ivz=4 itau=5 ibx=6 iby=7 ibz=8, iex=9 iey=10 iez=11.
 * Arrays are SoA over the ghosted box -2..n+3 per axis, x fastest:
 * cell (j,k,l) -> (j+2) + (nx+6)*((k+2) + (ny+6)*(l+2)).
 */
#ifndef MHD_ORACLE_H
#define MHD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_desc {
    int nx, ny, nz, n_dim, coordinate_flag;
    int boundary_flag[6];
    int riemann;   /* 0 LF, 1 HLL, 2 HLLC, 3 HLLD            (parameter.h:91-101) */
    int recon;     /* 0 minmod, 1 van Leer, 2 MC, 3 PPM, 4 WENO (parameter.h:103-115) */
    int model;     /* 0 riemann hooks, 1 torus-cyn hooks, 2 torus-sph hooks, 3 collapsar hooks */
    int rt_gravity;
    int hlld_bn_mode; /* value of the unassigned usstar(ibn) in HLLDNM: 0 -> 0.0, 1 -> Bn,
                      2 -> carried from earlier double-star interfaces (serial reference) */
    double cfl, ggas, r_sh;
    const double *xF, *yF, *zF;  /* face coordinates -3..n+3 (n+7 values), before the FIXPOLE override */
    int fixpole;                 /* -DFIXPOLE: polar sinf = 0 and phi-averaged polar EMFs (spherical only);
                                    2 = the same with a double accumulator, a planted mistake for tests */
    /* model 3 (model/collapsar): self-gravity and black hole */
    double prim_a[4];            /* atmosphere rho, vx, vy, vz (prim_a(irho:ivz)) */
    double eps_a;                /* atmosphere epsilon */
    double m_bh;                 /* initial black-hole mass */
    double c_code;               /* clight*vel2code, the speed of light in code units */
    double omega;                /* SOR weight omega_weight */
    double tolerance;            /* SOR tolerance */
    int relax_max, n_pot;        /* SOR iteration cap; potential solved every n_pot steps */
    int sor_fold;                /* SOR error fold: 0 the one-thread OpenMP reduction of the
                                    reference build; 1 a plain serial fold, a planted mistake for tests */
} oracle_desc;

typedef struct oracle oracle;

oracle *oracle_create(const oracle_desc *d);
void oracle_destroy(oracle *o);
/* prim: Fortran-order (8, nx+6, ny+6, nz+6) as written by GET_MODEL; eps and
 * floors (rho, p, eps; may be NULL) over the ghosted box. */
int oracle_load(oracle *o, const double *prim_aos, const double *eps, const double *floors);
void oracle_initial_update(oracle *o);      /* initial.f90:86-116 */
double oracle_finddt(oracle *o);            /* rungekutta.f90:221-332 */
void oracle_rungekutta(oracle *o, double dt);  /* rungekutta.f90:12-208 */
void oracle_stage(oracle *o, int stage, double dt);   /* one stage of it, 1..3 in order */
void oracle_spatial(oracle *o);             /* spatial.f90:11-162: l_rk (oracle_get "lrk1".."lrk8") */
/* main.f90:97-139 loop: returns steps taken; *t is advanced */
int oracle_run(oracle *o, int max_steps, double total_time, double *t, double *dts);
double oracle_find_divb(oracle *o);         /* mhd_module.f90:110-201 */
/* GETGRID metric arrays, index j+3 for j = -3..n+3: "dx","dy","dz","x","y","z",
 * "dx_cb","xbar","dcose","dsine","sine","sinf"; returns the length or -1. */
int oracle_grid(const oracle *o, const char *name, double *out);
/* Copy one named array (SoA, ghosted box) out: "rho","vx","vy","vz","p","bx","by",
 * "bz","eps","cs","bcx","bcy","bcz","u1".."u8","lrk1".."lrk8","flux1".."flux8". */
int oracle_get(oracle *o, const char *name, double *out);
long oracle_cells(const oracle *o);
/* model 3: the potential phi over the ghosted box, black-hole mass and
 * Schwarzschild radius, SOR iterations of the last solve, and the status
 * (0 ok, 1 horizon beyond the inner face, 2 converged at the relax cap). */
int oracle_gravity(const oracle *o, double *phi, double *m_bh, double *r_sh, int *iterations);
/* model 3: poisson_coef's coefficients (model/collapsar/src/poisson.f90:35-64):
 * "ajp1","ajm1" at [j] (nx+1 values), "bkp1","bkm1" at [k*(nx+1)+j]
 * ((nx+1)*(ny+1)), "clp1","clm1","epsc" over the ghosted box (cell index as
 * for oracle_get; interior cells only set).  Returns the length or -1. */
long oracle_poisson(const oracle *o, const char *name, double *out);

/* Point kernels for unit tests. */
void oracle_tvd(int limiter, const double w[2], double vm, double vc, double vp,
                double *vm_out, double *vp_out);
void oracle_ppm(const double w[14], const double v[5], double *vm_out, double *vp_out);
void oracle_weno(const double w[10], const double v[5], double *vm_out, double *vp_out);
void oracle_ppm_weight(double dm2, double dm1, double dc, double dp1, double dp2, double w[14]);
void oracle_weno_weight(double dm2, double dm1, double dc, double dp1, double dp2, double w[10]);
/* One interface: qL/qR are primitive face states indexed 1..8 (entry 0
 * unused, normal B must agree), dir 1..3; flux out indexed 1..8. */
void oracle_riemann_point(int solver, int dir, double ggas, int hlld_bn_mode,
                          const double qL[9], const double qR[9], double flux[9]);

#ifdef __cplusplus
}
#endif
#endif
