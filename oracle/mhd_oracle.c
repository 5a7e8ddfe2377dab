/*
 * CPU restatement of the reference's FP64 MHD update step.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker.  It may be loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
 * the product library (cumc3d_amd/csrc).
 *
 * Every routine follows one reference routine (cited file:line), keeping its
 * loop ranges, its ghost-fill order and its floating-point evaluation order
 * (Fortran evaluates a*b*c as (a*b)*c and a+b+c as (a+b)+c; x**2 is x*x).
 * Build with -ffp-contract=off so no multiply-add is fused.
 */
#include "mhd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* parameter.h:24 small_num = TINY(1.0D0) */
#define SMALL_NUM DBL_MIN

enum { IRHO = 1, IVX, IVY, IVZ, ITAU, IBX, IBY, IBZ, IEX, IEY, IEZ };  /* setup.f90:198-255 */
enum { IYX = 1, IZX, IXY, IZY, IXZ, IYZ };                           /* mhd_module.f90:19-24 */
enum { XDIR = 1, YDIR = 2, ZDIR = 3 };

struct oracle {
    oracle_desc d;
    int nx, ny, nz, NX, NY, NZ;
    long G;
    /* 1D grid arrays, index j+3 for j = -3..n+3 */
    double *xF, *yF, *zF, *dx, *dy, *dz, *x, *y, *z, *xbar, *dx_cb, *dsine, *dcose, *sine, *sinf;
    double *wx, *wy, *wz;
    int nw;
    int bxin[12], bxout[12], byin[12], byout[12], bzin[12], bzout[12];
    double *prim[9], *cons[9], *l_rk[9], *u_old[9], *sc[9], *flux[9];
    double *primL[9], *primR[9], *uL[9], *uR[9], *fluxL[9], *fluxR[9];
    double *epsL, *epsR, *csL, *csR, *eps, *cs, *dpdrho, *dpdeps;
    double *bcell[9];
    double *ecell[12];
    double *eface[7];
    double *efx, *efy, *efz;
    double *rho_floor, *p_floor, *eps_floor;
    double maxdivb;
    double uss_keep[2][9];   /* usstarL/usstarR carried across interfaces (hlld_bn_mode 2) */
    /* model 3 (model/collapsar/src/def.f90:6-46) */
    double *phi, *phi_old, *ajp1, *ajm1, *bkp1, *bkm1, *clp1, *clm1, *epsc, *mflux;
    double m_bh, mbh_old, r_sh;
    int nstep, sor_iter, status;
    double sor_err;
};

#define CI(o, j, k, l) ((long)((j) + 2) + (long)(o)->NX * ((long)((k) + 2) + (long)(o)->NY * ((l) + 2)))
#define A1(a, j) ((a)[(j) + 3])
#define W(w, j, i) ((w)[((j) + 3) * 14 + (i) - 1])

/* Fortran intrinsics */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }
/* flang evaluates MIN(a,b) as a < b ? a : b (a NaN or a tie gives b) */
static inline double fmin2(double a, double b) { return (a < b) ? a : b; }
static inline double fmax2(double a, double b) { return (a > b) ? a : b; }

/* ------------------------------------------------------------------------ */
/* allocation                                                               */
/* ------------------------------------------------------------------------ */

static double *alloc_cells(oracle *o) { return (double *)calloc((size_t)o->G, sizeof(double)); }

oracle *oracle_create(const oracle_desc *d)
{
    oracle *o = (oracle *)calloc(1, sizeof(oracle));
    o->d = *d;
    o->nx = d->nx; o->ny = d->ny; o->nz = d->nz;
    o->NX = d->nx + 6; o->NY = d->ny + 6; o->NZ = d->nz + 6;
    o->G = (long)o->NX * o->NY * o->NZ;

    /* SETUP_EQNS parity factors, setup.f90:102-326 */
    int *fin[3] = {o->bxin, o->byin, o->bzin}, *fout[3] = {o->bxout, o->byout, o->bzout};
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 12; ++i) fin[a][i] = fout[a][i] = 1;
    const int *bf = d->boundary_flag;
    if (bf[0] == 2) { o->bxin[IVX] = -1; o->bxin[IBX] = -1; }
    else if (bf[0] == 3) { o->bxin[IVX] = -1; o->bxin[IBX] = -1; o->bxin[IVZ] = -1; o->bxin[IBZ] = -1; }
    else if (bf[0] == 4) { free(o); return NULL; }
    if (bf[1] == 2) { o->bxout[IVX] = -1; o->bxout[IBX] = -1; }
    else if (bf[1] == 3) { o->bxout[IVX] = -1; o->bxout[IBX] = -1; o->bxout[IVZ] = -1; o->bxout[IBZ] = -1; }
    else if (bf[1] == 4) { free(o); return NULL; }
    if (bf[2] == 2) { o->byin[IVY] = -1; o->byin[IBY] = -1; }
    else if (bf[2] == 3) { o->byin[IVY] = -1; o->byin[IBY] = -1; o->byin[IVZ] = -1; o->byin[IBZ] = -1; }
    else if (bf[2] == 4) { o->byin[IVY] = -1; o->byin[IBX] = -1; o->byin[IBZ] = -1; }
    if (bf[3] == 2) { o->byout[IVY] = -1; o->byout[IBY] = -1; }
    else if (bf[3] == 3) { o->byout[IVY] = -1; o->byout[IBY] = -1; o->byout[IVZ] = -1; o->byout[IBZ] = -1; }
    else if (bf[3] == 4) { o->byout[IVY] = -1; o->byout[IBX] = -1; o->byout[IBZ] = -1; }
    if (bf[4] == 2) { o->bzin[IVZ] = -1; o->bzin[IBZ] = -1; }
    else if (bf[4] == 3) { free(o); return NULL; }
    else if (bf[4] == 4) { o->bzin[IVZ] = -1; o->bzin[IBX] = -1; o->bzin[IBY] = -1; }
    if (bf[5] == 2) { o->bzout[IVZ] = -1; o->bzout[IBZ] = -1; }
    else if (bf[5] == 3) { free(o); return NULL; }
    else if (bf[5] == 4) { o->bzout[IVZ] = -1; o->bzout[IBX] = -1; o->bzout[IBY] = -1; }
    int *all[6] = {o->bxin, o->bxout, o->byin, o->byout, o->bzin, o->bzout};
    for (int s = 0; s < 6; ++s) {
        int *b = all[s];
        if (b[IVZ] * b[IBY] * b[IVY] * b[IBZ] < 0) { free(o); return NULL; }
        b[IEX] = b[IVZ] * b[IBY];
        if (b[IVX] * b[IBZ] * b[IVZ] * b[IBX] < 0) { free(o); return NULL; }
        b[IEY] = b[IVX] * b[IBZ];
        if (b[IVY] * b[IBX] * b[IVX] * b[IBY] < 0) { free(o); return NULL; }
        b[IEZ] = b[IVY] * b[IBX];
    }

    /* GetGrid, getgrid.f90:26-86 (faces come from the caller) */
    int n3[3] = {o->nx, o->ny, o->nz};
    const double *F[3] = {d->xF, d->yF, d->zF};
    double **faces[3] = {&o->xF, &o->yF, &o->zF};
    double **wid[3] = {&o->dx, &o->dy, &o->dz};
    double **cen[3] = {&o->x, &o->y, &o->z};
    for (int a = 0; a < 3; ++a) {
        int n = n3[a];
        *faces[a] = (double *)calloc(n + 7, sizeof(double));
        *wid[a] = (double *)calloc(n + 7, sizeof(double));
        *cen[a] = (double *)calloc(n + 7, sizeof(double));
        memcpy(*faces[a], F[a], (n + 7) * sizeof(double));
        for (int j = -2; j <= n + 3; ++j) {
            A1(*wid[a], j) = A1(*faces[a], j) - A1(*faces[a], j - 1);
            A1(*cen[a], j) = (A1(*faces[a], j) + A1(*faces[a], j - 1)) / 2;
        }
    }
    o->xbar = (double *)calloc(o->nx + 7, sizeof(double));
    o->dx_cb = (double *)calloc(o->nx + 7, sizeof(double));
    for (int j = -1; j <= o->nx + 3; ++j) {
        double a = A1(o->xF, j), b = A1(o->xF, j - 1);
        A1(o->dx_cb, j) = a * a * a - b * b * b;
        A1(o->xbar, j) = (A1(o->dx_cb, j) / 3.0) / (A1(o->x, j) * A1(o->dx, j));
    }
    o->dsine = (double *)calloc(o->ny + 7, sizeof(double));
    o->dcose = (double *)calloc(o->ny + 7, sizeof(double));
    o->sine = (double *)calloc(o->ny + 7, sizeof(double));
    o->sinf = (double *)calloc(o->ny + 7, sizeof(double));
    for (int k = -1; k <= o->ny + 3; ++k) {
        A1(o->dcose, k) = -(cos(A1(o->yF, k)) - cos(A1(o->yF, k - 1)));
        A1(o->dsine, k) = sin(A1(o->yF, k)) - sin(A1(o->yF, k - 1));
    }
    for (int k = -2; k <= o->ny + 3; ++k) A1(o->sine, k) = sin(A1(o->y, k));
    for (int k = -3; k <= o->ny + 3; ++k) A1(o->sinf, k) = sin(A1(o->yF, k));
    /* FIXPOLE, getgrid.f90:91-98: after dy, y and the sines are taken */
    if (d->coordinate_flag == 2 && d->fixpole) {
        A1(o->yF, 0) = 0.0;
        A1(o->yF, o->ny) = 4.0 * atan(1.0);
        A1(o->sinf, 0) = 0.0;
        A1(o->sinf, o->ny) = 0.0;
    }

    /* reconstruction weights */
    o->wx = (double *)calloc((o->nx + 7) * 14, sizeof(double));
    o->wy = (double *)calloc((o->ny + 7) * 14, sizeof(double));
    o->wz = (double *)calloc((o->nz + 7) * 14, sizeof(double));
    double *ws[3] = {o->wx, o->wy, o->wz};
    for (int a = 0; a < 3; ++a) {
        int n = n3[a];
        double *dd = *wid[a], *w = ws[a];
        if (d->recon == 4) {            /* WENO_WEIGHT, weno_module.f90:62-74 */
            for (int j = 0; j <= n + 1; ++j)
                oracle_weno_weight(A1(dd, j - 2), A1(dd, j - 1), A1(dd, j), A1(dd, j + 1), A1(dd, j + 2),
                                   &W(w, j, 1));
        } else if (d->recon == 3) {     /* PPM_WEIGHT, ppmc_module.f90:36-48 */
            for (int j = 0; j <= n + 1; ++j)
                oracle_ppm_weight(A1(dd, j - 2), A1(dd, j - 1), A1(dd, j), A1(dd, j + 1), A1(dd, j + 2),
                                  &W(w, j, 1));
        } else {                        /* TVD_WEIGHT, tvd_module.f90:42-53 */
            for (int j = -1; j <= n + 2; ++j) {
                W(w, j, 1) = (A1(dd, j) + A1(dd, j + 1)) / A1(dd, j);
                W(w, j, 2) = (A1(dd, j) + A1(dd, j - 1)) / A1(dd, j);
            }
        }
    }

    for (int i = 1; i <= 8; ++i) {
        o->prim[i] = alloc_cells(o); o->cons[i] = alloc_cells(o); o->l_rk[i] = alloc_cells(o);
        o->u_old[i] = alloc_cells(o); o->sc[i] = alloc_cells(o); o->flux[i] = alloc_cells(o);
        o->primL[i] = alloc_cells(o); o->primR[i] = alloc_cells(o); o->uL[i] = alloc_cells(o);
        o->uR[i] = alloc_cells(o); o->fluxL[i] = alloc_cells(o); o->fluxR[i] = alloc_cells(o);
    }
    o->epsL = alloc_cells(o); o->epsR = alloc_cells(o); o->csL = alloc_cells(o); o->csR = alloc_cells(o);
    o->eps = alloc_cells(o); o->cs = alloc_cells(o); o->dpdrho = alloc_cells(o); o->dpdeps = alloc_cells(o);
    for (int i = IBX; i <= IBZ; ++i) o->bcell[i] = alloc_cells(o);
    for (int i = IEX; i <= IEZ; ++i) o->ecell[i] = alloc_cells(o);
    for (int i = 1; i <= 6; ++i) o->eface[i] = alloc_cells(o);
    o->efx = alloc_cells(o); o->efy = alloc_cells(o); o->efz = alloc_cells(o);
    o->rho_floor = alloc_cells(o); o->p_floor = alloc_cells(o); o->eps_floor = alloc_cells(o);
    o->phi = alloc_cells(o); o->phi_old = alloc_cells(o);
    o->mflux = (double *)calloc((size_t)o->NY * o->NZ, sizeof(double));
    if (d->model == 3) {
        /* get_poisson / poisson_coef, model/collapsar/src/poisson.f90:6-64 */
        int nx = o->nx, ny = o->ny, nz = o->nz;
        o->ajp1 = (double *)calloc(nx + 1, sizeof(double));
        o->ajm1 = (double *)calloc(nx + 1, sizeof(double));
        o->bkp1 = (double *)calloc((size_t)(nx + 1) * (ny + 1), sizeof(double));
        o->bkm1 = (double *)calloc((size_t)(nx + 1) * (ny + 1), sizeof(double));
        o->clp1 = alloc_cells(o); o->clm1 = alloc_cells(o); o->epsc = alloc_cells(o);
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny; ++j)
                for (int k = 1; k <= nz; ++k) {
                    double xm1 = A1(o->x, i - 1), xc = A1(o->x, i), xp1 = A1(o->x, i + 1);
                    double ym1 = A1(o->y, j - 1), yc = A1(o->y, j), yp1 = A1(o->y, j + 1);
                    double zm1 = A1(o->z, k - 1), zc = A1(o->z, k), zp1 = A1(o->z, k + 1);
                    double t = tan(yc), sn = sin(yc);
                    long c = CI(o, i, j, k);
                    o->epsc[c] = 2.0 * (xp1 + xm1 - 3.0 * xc) / xc / (xp1 - xc) / (xc - xm1)
                               + (yp1 + ym1 - 2.0 * yc - 2.0 * t) / (xc * xc) / t / (yp1 - yc) / (yc - ym1)
                               - 2.0 / (xc * xc) / (sn * sn) / (zp1 - zc) / (zc - zm1);
                    o->ajp1[i] = 2.0 * (2.0 * xc - xm1) / (xp1 - xm1) / (xp1 - xc) / xc;
                    o->ajm1[i] = 2.0 * (2.0 * xc - xp1) / (xp1 - xm1) / (xc - xm1) / xc;
                    o->bkp1[(size_t)j * (nx + 1) + i] = (2.0 * t + yc - ym1) / (xc * xc) / t / (yp1 - yc) / (yp1 - ym1);
                    o->bkm1[(size_t)j * (nx + 1) + i] = (2.0 * t + yc - yp1) / (xc * xc) / t / (yp1 - ym1) / (yc - ym1);
                    o->clp1[c] = 2.0 / (xc * xc) / (sn * sn) / (zp1 - zm1) / (zp1 - zc);
                    o->clm1[c] = 2.0 / (xc * xc) / (sn * sn) / (zp1 - zm1) / (zc - zm1);
                }
        o->m_bh = d->m_bh;
    }
    return o;
}

void oracle_destroy(oracle *o)
{
    if (!o) return;
    double *one[] = {o->xF, o->yF, o->zF, o->dx, o->dy, o->dz, o->x, o->y, o->z, o->xbar, o->dx_cb,
                     o->dsine, o->dcose, o->sine, o->sinf, o->wx, o->wy, o->wz, o->epsL, o->epsR,
                     o->csL, o->csR, o->eps, o->cs, o->dpdrho, o->dpdeps, o->efx, o->efy, o->efz,
                     o->rho_floor, o->p_floor, o->eps_floor};
    for (size_t i = 0; i < sizeof(one) / sizeof(one[0]); ++i) free(one[i]);
    for (int i = 1; i <= 8; ++i) {
        free(o->prim[i]); free(o->cons[i]); free(o->l_rk[i]); free(o->u_old[i]); free(o->sc[i]);
        free(o->flux[i]); free(o->primL[i]); free(o->primR[i]); free(o->uL[i]); free(o->uR[i]);
        free(o->fluxL[i]); free(o->fluxR[i]);
    }
    for (int i = IBX; i <= IBZ; ++i) free(o->bcell[i]);
    for (int i = IEX; i <= IEZ; ++i) free(o->ecell[i]);
    for (int i = 1; i <= 6; ++i) free(o->eface[i]);
    double *g[] = {o->phi, o->phi_old, o->ajp1, o->ajm1, o->bkp1, o->bkm1, o->clp1, o->clm1, o->epsc, o->mflux};
    for (size_t i = 0; i < sizeof(g) / sizeof(g[0]); ++i) free(g[i]);
    free(o);
}

long oracle_cells(const oracle *o) { return o->G; }

long oracle_poisson(const oracle *o, const char *name, double *out)
{
    if (o->d.model != 3) return -1;
    long n1 = o->nx + 1, n2 = (long)(o->nx + 1) * (o->ny + 1);
    const char *nm[7] = {"ajp1", "ajm1", "bkp1", "bkm1", "clp1", "clm1", "epsc"};
    const double *a[7] = {o->ajp1, o->ajm1, o->bkp1, o->bkm1, o->clp1, o->clm1, o->epsc};
    long len[7] = {n1, n1, n2, n2, o->G, o->G, o->G};
    for (int i = 0; i < 7; ++i)
        if (!strcmp(name, nm[i])) {
            memcpy(out, a[i], (size_t)len[i] * sizeof(double));
            return len[i];
        }
    return -1;
}

int oracle_gravity(const oracle *o, double *phi, double *m_bh, double *r_sh, int *iterations)
{
    if (o->d.model != 3) return -1;
    if (phi) memcpy(phi, o->phi, (size_t)o->G * sizeof(double));
    if (m_bh) *m_bh = o->m_bh;
    if (r_sh) *r_sh = o->r_sh;
    if (iterations) *iterations = o->sor_iter;
    return o->status;
}

int oracle_load(oracle *o, const double *prim_aos, const double *eps, const double *floors)
{
    for (long c = 0; c < o->G; ++c) {
        for (int i = 1; i <= 8; ++i) o->prim[i][c] = prim_aos[c * 8 + (i - 1)];
        o->eps[c] = eps[c];
    }
    if (floors) {
        memcpy(o->rho_floor, floors, o->G * sizeof(double));
        memcpy(o->p_floor, floors + o->G, o->G * sizeof(double));
        memcpy(o->eps_floor, floors + 2 * o->G, o->G * sizeof(double));
    }
    /* INITIAL_MODEL sets cons = 0 before GET_MODEL (initial.f90:63-64) */
    for (int i = 1; i <= 8; ++i) memset(o->cons[i], 0, o->G * sizeof(double));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* EOS hooks (model/riemann/src/getepsilon.f90:4-17, soundspeed.f90:4-21;    */
/* identical in model/torus-cyn/src)                                         */
/* ------------------------------------------------------------------------ */

static inline double eos_eps(double g, double rho, double p) { return p / rho / (g - 1.0); }
static inline double eos_cs(double g, double p, double rho, double eps)
{
    double dpdden = eps * (g - 1.0);
    double dpdint = rho * (g - 1.0);
    return sqrt(dpdden + dpdint * p / (rho * rho));
}

/* ------------------------------------------------------------------------ */
/* reconstruction point kernels                                             */
/* ------------------------------------------------------------------------ */

/* TVD, tvd_module.f90:279-344 */
void oracle_tvd(int lim, const double w[2], double vm, double vc, double vp, double *vm_out,
                double *vp_out)
{
    double cf = w[0], cb = w[1], slope = 0.0;
    double dqf = 2.0 * (vp - vc) / cf;
    double dqb = 2.0 * (vc - vm) / cb;
    if (lim == 0) {
        slope = 0.5 * (fsign(1.0, dqf) + fsign(1.0, dqb)) * fmin2(fabs(dqf), fabs(dqb));
    } else if (lim == 1) {
        double dq2 = dqf * dqb;
        if (dq2 <= 0.0) slope = 0.0;
        else slope = (dq2 * (cf * dqb + cb * dqf) / (dqb * dqb + dqf * dqf + dq2 * (cf + cb - 2.0)));
    } else if (lim == 2) {
        if (dqf == 0.0) {
            slope = 0.0;
        } else {
            double v = dqb / dqf;
            slope = fmax2(0.0, fmin2(fmin2(0.5 * (1.0 + v), cf), cb * v));
            slope = dqf * slope;
        }
    }
    *vp_out = vc + 0.5 * slope;
    *vm_out = vc - 0.5 * slope;
}

/* PPM_INTERPOLANT, ppmc_module.f90:91-145 */
void oracle_ppm_weight(double dm2, double dm1, double dc, double dp1, double dp2, double w[14])
{
    double a1R = dc / (dc + dp1), a1L = dm1 / (dm1 + dc);
    double aR = 2.0 * dp1 * dc / (dp1 + dc), aL = 2.0 * dc * dm1 / (dc + dm1);
    double deltaXR = dm1 + dc + dp1 + dp2, deltaXL = dm2 + dm1 + dc + dp1;
    double z1R = (dm1 + dc) / (2.0 * dc + dp1), z1L = (dm2 + dm1) / (2.0 * dm1 + dc);
    double zR = (dp2 + dp1) / (2.0 * dp1 + dc), zL = (dp1 + dc) / (2.0 * dc + dm1);
    double cc1 = dc / (dm1 + dc + dp1), cp1 = dp1 / (dc + dp1 + dp2), cm1 = dm1 / (dm2 + dm1 + dc);
    double cc2 = (2.0 * dm1 + dc) / (dp1 + dc), cp2 = (2.0 * dc + dp1) / (dp2 + dp1);
    double cm2 = (2.0 * dm2 + dm1) / (dc + dm1);
    double cc3 = (dc + 2.0 * dp1) / (dm1 + dc), cp3 = (dp1 + 2.0 * dp2) / (dc + dp1);
    double cm3 = (dm1 + 2.0 * dc) / (dm2 + dm1);
    w[0] = a1L;
    w[1] = aL * (z1L - zL) / deltaXL;
    w[2] = (-dm1 * z1L) / deltaXL;
    w[3] = dc * zL / deltaXL;
    w[4] = a1R;
    w[5] = aR * (z1R - zR) / deltaXR;
    w[6] = (-dc * z1R) / deltaXR;
    w[7] = dp1 * zR / deltaXR;
    w[8] = cm1 * cm2;
    w[9] = cm1 * cm3;
    w[10] = cc1 * cc2;
    w[11] = cc1 * cc3;
    w[12] = cp1 * cp2;
    w[13] = cp1 * cp3;
}

/* PPMC, ppmc_module.f90:359-464 */
void oracle_ppm(const double w[14], const double v[5], double *vm_out, double *vp_out)
{
    double vm2 = v[0], vm1 = v[1], vc = v[2], vp1 = v[3], vp2 = v[4];
    double a0m = w[0], a1m = w[1], a2m = w[2], a3m = w[3], a0p = w[4], a1p = w[5], a2p = w[6],
           a3p = w[7], b0m = w[8], b1m = w[9], b0c = w[10], b1c = w[11], b0p = w[12], b1p = w[13];
    double deltap1 = b0p * (vp2 - vp1) + b1p * (vp1 - vc);
    double deltac = b0c * (vp1 - vc) + b1c * (vc - vm1);
    double deltam1 = b0m * (vc - vm1) + b1m * (vm1 - vm2);
    double dmc, dmp1, dmm1, cond;
    cond = (vp1 - vc) * (vc - vm1);
    if (cond > 0.0)
        dmc = fmin2(fmin2(fabs(deltac), 2.0 * fabs(vc - vm1)), 2.0 * fabs(vc - vp1)) * fsign(1.0, deltac);
    else dmc = 0.0;
    cond = (vp2 - vp1) * (vp1 - vc);
    if (cond > 0.0)
        dmp1 = fmin2(fmin2(fabs(deltap1), 2.0 * fabs(vp1 - vc)), 2.0 * fabs(vp1 - vp2)) * fsign(1.0, deltap1);
    else dmp1 = 0.0;
    cond = (vc - vm1) * (vm1 - vm2);
    if (cond > 0.0)
        dmm1 = fmin2(fmin2(fabs(deltam1), 2.0 * fabs(vm1 - vm2)), 2.0 * fabs(vm1 - vc)) * fsign(1.0, deltam1);
    else dmm1 = 0.0;
    double vp = vc + a0p * (vp1 - vc) + a1p * (vp1 - vc) + a2p * dmp1 + a3p * dmc;
    double vm = vm1 + a0m * (vc - vm1) + a1m * (vc - vm1) + a2m * dmc + a3m * dmm1;
    double vl = vm, vr = vp;
    if ((vr - vc) * (vc - vl) <= 0.0) {
        vp = vc;
        vm = vc;
    } else {
        cond = (vr - vl) * (vl - 3.0 * vc + 2.0 * vr);
        if (cond < 0.0) vm = 3.0 * vc - 2.0 * vr;
        cond = (vp - vm) * (3.0 * vc - 2.0 * vl - vr);
        if (cond < 0.0) vp = 3.0 * vc - 2.0 * vl;
    }
    *vm_out = vm;
    *vp_out = vp;
}

/* WENO_INTERPOLANT, weno_module.f90:117-149 */
void oracle_weno_weight(double dm2, double dm1, double dc, double dp1, double dp2, double w[10])
{
    double a12 = dp2 / (dp1 + dp2), a01 = dp1 / (dc + dp1), a10 = dc / (dm1 + dc), a21 = dm1 / (dm2 + dm1);
    w[0] = 1.0 - a01; w[1] = a01; w[2] = 1.0 - a12; w[3] = a12; w[4] = 1.0 - a10;
    w[5] = a10; w[6] = 1.0 - a21; w[7] = a21; w[8] = 1.0 - a21; w[9] = a12;
}

/* WENO (WENO-Z), weno_module.f90:315-408 */
void oracle_weno(const double w[10], const double v[5], double *vm_out, double *vp_out)
{
    const double smallpara = 1.0e-40, g1 = 1.0 / 10.0, g2 = 6.0 / 10.0, g3 = 3.0 / 10.0;
    const double r1 = 13.0 / 12.0, r2 = 1.0 / 4.0;
    double vm2 = v[0], vm1 = v[1], vc = v[2], vp1 = v[3], vp2 = v[4];
    double U01 = w[0] * vc + w[1] * vp1;
    double U12 = w[2] * vp1 + w[3] * vp2;
    double U10 = w[4] * vm1 + w[5] * vc;
    double U21 = w[6] * vm2 + w[7] * vm1;
    double U00 = vm1 + w[8] * (vm1 - vm2);
    double U11 = vp1 + w[9] * (vp1 - vp2);
    double P1, P2, P3, B1, B2, B3, t, s, tau5, q, w1, w2, w3;

    P1 = (6.0 * vc - 1.0 * U10 - 2.0 * U00) / 3.0;
    P2 = (-1.0 * U10 + 2.0 * vc + 2.0 * U01) / 3.0;
    P3 = (2.0 * U01 + 2.0 * vp1 - 1.0 * U12) / 3.0;
    t = 2.0 * U10 - 2.0 * U00; s = 4.0 * vc - 2.0 * U10 - 2.0 * U00;
    B1 = r1 * (t * t) + r2 * (s * s);
    t = 2.0 * U10 - 4.0 * vc + 2.0 * U01; s = -2.0 * U10 + 2.0 * U01;
    B2 = r1 * (t * t) + r2 * (s * s);
    t = 2.0 * U01 - 4.0 * vp1 + 2.0 * U12; s = -6.0 * U01 + 8.0 * vp1 - 2.0 * U12;
    B3 = r1 * (t * t) + r2 * (s * s);
    tau5 = fabs(B1 - B3);
    q = tau5 / (smallpara + B1); w1 = g1 * (1.0 + q * q);
    q = tau5 / (smallpara + B2); w2 = g2 * (1.0 + q * q);
    q = tau5 / (smallpara + B3); w3 = g3 * (1.0 + q * q);
    *vp_out = (w1 * P1 + w2 * P2 + w3 * P3) / (w1 + w2 + w3);

    P1 = (6.0 * vc - 1.0 * U01 - 2.0 * U11) / 3.0;
    P2 = (-1.0 * U01 + 2.0 * vc + 2.0 * U10) / 3.0;
    P3 = (2.0 * U10 + 2.0 * vm1 - 1.0 * U21) / 3.0;
    t = 2.0 * U01 - 2.0 * U11; s = 4.0 * vc - 2.0 * U01 - 2.0 * U11;
    B1 = r1 * (t * t) + r2 * (s * s);
    t = 2.0 * U01 - 4.0 * vc + 2.0 * U10; s = 2.0 * U01 - 2.0 * U10;
    B2 = r1 * (t * t) + r2 * (s * s);
    t = 2.0 * U10 - 4.0 * vm1 + 2.0 * U21; s = 8.0 * vm1 - 6.0 * U10 - 2.0 * U21;
    B3 = r1 * (t * t) + r2 * (s * s);
    tau5 = fabs(B1 - B3);
    q = tau5 / (smallpara + B1); w1 = g1 * (1.0 + q * q);
    q = tau5 / (smallpara + B2); w2 = g2 * (1.0 + q * q);
    q = tau5 / (smallpara + B3); w3 = g3 * (1.0 + q * q);
    *vm_out = (w1 * P1 + w2 * P2 + w3 * P3) / (w1 + w2 + w3);
}

/* one cell: faces of a 5-point stencil with the scheme selected at build time
 * (spatial.f90:33-43). */
static void recon_cell(const oracle *o, const double *w, const double v[5], double *vm, double *vp)
{
    int r = o->d.recon;
    if (r <= 2) oracle_tvd(r, w, v[1], v[2], v[3], vm, vp);
    else if (r == 3) oracle_ppm(w, v, vm, vp);
    else oracle_weno(w, v, vm, vp);
}

/* TVD_recon{x,y,z} / PPMC_recon{x,y,z} / WENO_recon{x,y,z}
 * (tvd_module.f90:62-274, ppmc_module.f90:151-357, weno_module.f90:154-308):
 * loops over cells c = 0..n_d+1 along the sweep (1..n across), writing
 * primR(c-1) (left face of c) and primL(c) (right face of c). */
static void reconstruct(oracle *o, int dir)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    int j0 = 1, j1 = nx, k0 = 1, k1 = ny, l0 = 1, l1 = nz, ibn, ibt1, ibt2;
    const double *w;
    long st;
    if (dir == XDIR) { j0 = 0; j1 = nx + 1; w = o->wx; st = 1; ibn = IBX; ibt1 = IBY; ibt2 = IBZ; }
    else if (dir == YDIR) { k0 = 0; k1 = ny + 1; w = o->wy; st = o->NX; ibn = IBY; ibt1 = IBX; ibt2 = IBZ; }
    else { l0 = 0; l1 = nz + 1; w = o->wz; st = (long)o->NX * o->NY; ibn = IBZ; ibt1 = IBX; ibt2 = IBY; }
    const double g = o->d.ggas;
    for (int l = l0; l <= l1; ++l)
        for (int k = k0; k <= k1; ++k)
            for (int j = j0; j <= j1; ++j) {
                long c = CI(o, j, k, l), cm = c - st;
                int s = (dir == XDIR) ? j : (dir == YDIR) ? k : l;
                const double *ww = &W(w, s, 1);
                double v[5];
                for (int i = IRHO; i <= ITAU; ++i) {
                    const double *a = o->prim[i];
                    if (o->d.recon <= 2) { v[1] = a[c - st]; v[2] = a[c]; v[3] = a[c + st]; }
                    else { v[0] = a[c - 2 * st]; v[1] = a[c - st]; v[2] = a[c]; v[3] = a[c + st]; v[4] = a[c + 2 * st]; }
                    recon_cell(o, ww, v, &o->primR[i][cm], &o->primL[i][c]);
                }
                o->epsR[cm] = eos_eps(g, o->primR[IRHO][cm], o->primR[ITAU][cm]);
                o->epsL[c] = eos_eps(g, o->primL[IRHO][c], o->primL[ITAU][c]);
                o->csR[cm] = eos_cs(g, o->primR[ITAU][cm], o->primR[IRHO][cm], o->epsR[cm]);
                o->csL[c] = eos_cs(g, o->primL[ITAU][c], o->primL[IRHO][c], o->epsL[c]);
                int tb[2] = {ibt1, ibt2};
                for (int m = 0; m < 2; ++m) {
                    const double *a = o->bcell[tb[m]];
                    if (o->d.recon <= 2) { v[1] = a[c - st]; v[2] = a[c]; v[3] = a[c + st]; }
                    else { v[0] = a[c - 2 * st]; v[1] = a[c - st]; v[2] = a[c]; v[3] = a[c + st]; v[4] = a[c + 2 * st]; }
                    recon_cell(o, ww, v, &o->primR[tb[m]][cm], &o->primL[tb[m]][c]);
                }
                o->primR[ibn][cm] = o->prim[ibn][cm];
                o->primL[ibn][c] = o->prim[ibn][c];
            }
}

/* ------------------------------------------------------------------------ */
/* states and Riemann solvers                                               */
/* ------------------------------------------------------------------------ */

typedef struct { int ivn, ivt1, ivt2, ibn, ibt1, ibt2; } perm_t;

static perm_t perm_of(int dir)
{
    perm_t p;
    if (dir == XDIR) { p.ivn = IVX; p.ivt1 = IVY; p.ivt2 = IVZ; p.ibn = IBX; p.ibt1 = IBY; p.ibt2 = IBZ; }
    else if (dir == YDIR) { p.ivn = IVY; p.ivt1 = IVZ; p.ivt2 = IVX; p.ibn = IBY; p.ibt1 = IBZ; p.ibt2 = IBX; }
    else { p.ivn = IVZ; p.ivt1 = IVX; p.ivt2 = IVY; p.ibn = IBZ; p.ibt1 = IBX; p.ibt2 = IBY; }
    return p;
}

static inline double dot3(const double *a, int ia, const double *b, int ib)
{
    return a[ia] * b[ib] + a[ia + 1] * b[ib + 1] + a[ia + 2] * b[ib + 2];
}

/* BUILDSTATES, spatial.f90:342-388, for one face state q (index 1..8) */
static void build_state(const perm_t *p, const double q[9], double eps, double u[9], double f[9])
{
    double b2 = dot3(q, IBX, q, IBX), v2 = dot3(q, IVX, q, IVX), vb = dot3(q, IBX, q, IVX);
    for (int i = IRHO; i <= ITAU; ++i) u[i] = q[i] * q[IRHO];
    u[IRHO] = q[IRHO];
    for (int i = IBX; i <= IBZ; ++i) u[i] = q[i];
    u[ITAU] = q[IRHO] * (0.5 * v2 + eps) + 0.5 * b2;
    for (int i = IRHO; i <= ITAU; ++i) f[i] = u[i] * q[p->ivn];
    f[p->ivn] = f[p->ivn] + q[ITAU] + 0.5 * (b2) - q[p->ibn] * q[p->ibn];
    f[p->ivt1] = f[p->ivt1] - q[p->ibn] * q[p->ibt1];
    f[p->ivt2] = f[p->ivt2] - q[p->ibn] * q[p->ibt2];
    f[ITAU] = f[ITAU] + (q[ITAU] + 0.5 * (b2)) * q[p->ivn] - q[p->ibn] * vb;
    f[p->ibn] = 0.0;
    f[p->ibt1] = (q[p->ivn] * q[p->ibt1] - q[p->ivt1] * q[p->ibn]);
    f[p->ibt2] = (q[p->ivn] * q[p->ibt2] - q[p->ivt2] * q[p->ibn]);
}

/* rsolver_module.f90:32-93 */
static inline double signalspeed(double cs, double bn, double bt1, double bt2, double rho)
{
    double a2 = cs * cs, a4 = a2 * a2;
    double b2n = (bn * bn / rho), b2t1 = (bt1 * bt1 / rho), b2t2 = (bt2 * bt2 / rho);
    double b2 = b2n + b2t1 + b2t2, b4 = b2 * b2;
    return sqrt(0.5 * (a2 + b2 + sqrt(fmax2((a4 + b4 + 2.0 * b2 * a2) - 4.0 * a2 * b2n, 0.0))));
}
static inline double hll_state(double ul, double ur, double fl, double fr, double sl, double sr)
{
    return (sr * ur - sl * ul - (fr - fl)) / (sr - sl);
}
static inline double hll_flux(double yl, double yr, double xl, double xr, double sl, double sr)
{
    return (sr * yl - sl * yr + sl * sr * (xr - xl)) / (sr - sl);
}
static inline double roe(double xl, double xr, double rl, double rr)
{
    return (sqrt(rl) * xl + sqrt(rr) * xr) / (sqrt(rl) + sqrt(rr));
}
static inline double sstar_hllc(double pl, double pr, double rl, double rr, double vl, double vr,
                                double sl, double sr, double bl, double br)
{
    return (rr * vr * (sr - vr) - rl * vl * (sl - vl) + pl - pr + br * br - bl * bl) /
           (rr * (sr - vr) - rl * (sl - vl));
}

/* One interface with the selected solver: LFNM (rsolver_module.f90:199-207),
 * HLLNM (:294-311), HLLCNM (:410-499), HLLDNM (:600-752). */
static void riemann_core(int solver, const perm_t *p, int bn_mode, double keep[2][9], const double qL[9],
                         const double qR[9], double csL, double csR, const double uL[9],
                         const double uR[9], const double fL[9], const double fR[9], double flux[9])
{
    const int ivn = p->ivn, ivt1 = p->ivt1, ivt2 = p->ivt2, ibn = p->ibn, ibt1 = p->ibt1,
              ibt2 = p->ibt2;
    double cfsL = signalspeed(csL, qL[ibn], qL[ibt1], qL[ibt2], qL[IRHO]);
    double cfsR = signalspeed(csR, qR[ibn], qR[ibt1], qR[ibt2], qR[IRHO]);
    if (solver == 0) {
        double sL = fabs(qL[ivn]) + cfsL, sR = fabs(qR[ivn]) + cfsR, splus = fmax2(sL, sR);
        for (int i = 1; i <= 8; ++i) flux[i] = 0.5 * (fL[i] + fR[i] - splus * (uR[i] - uL[i]));
        return;
    }
    double ubar = roe(qL[ivn], qR[ivn], qL[IRHO], qR[IRHO]);
    double cbar = roe(cfsL, cfsR, qL[IRHO], qR[IRHO]);
    double sL = fmin2(qL[ivn] - cfsL, ubar - cbar);
    double sR = fmax2(qR[ivn] + cfsR, ubar + cbar);
    if (solver == 1) {
        if (sL >= 0.0) { for (int i = 1; i <= 8; ++i) flux[i] = fL[i]; }
        else if (sL <= 0.0 && sR >= 0.0) { for (int i = 1; i <= 8; ++i) flux[i] = hll_flux(fL[i], fR[i], uL[i], uR[i], sL, sR); }
        else if (sR <= 0.0) { for (int i = 1; i <= 8; ++i) flux[i] = fR[i]; }
        return;
    }
    if (sL >= 0.0) { for (int i = 1; i <= 8; ++i) flux[i] = fL[i]; return; }
    if (sR <= 0.0) { for (int i = 1; i <= 8; ++i) flux[i] = fR[i]; return; }
    double b2L = dot3(qL, IBX, qL, IBX), b2R = dot3(qR, IBX, qR, IBX);
    double pLt = qL[ITAU] + 0.5 * b2L, pRt = qR[ITAU] + 0.5 * b2R;
    if (solver == 2) {
        double sstar = sstar_hllc(pLt, pRt, qL[IRHO], qR[IRHO], qL[ivn], qR[ivn], sL, sR, qL[ibn], qR[ibn]);
        double u_hll[9] = {0}, p_hll[9] = {0}, us[9];
        u_hll[ibn] = qL[ibn];
        u_hll[ibt1] = hll_state(qL[ibt1], qR[ibt1], fL[ibt1], fR[ibt1], sL, sR);
        u_hll[ibt2] = hll_state(qL[ibt2], qR[ibt2], fL[ibt2], fR[ibt2], sL, sR);
        u_hll[IRHO] = hll_state(qL[IRHO], qR[IRHO], fL[IRHO], fR[IRHO], sL, sR);
        u_hll[ivt1] = hll_state(uL[ivt1], uR[ivt1], fL[ivt1], fR[ivt1], sL, sR);
        u_hll[ivt2] = hll_state(uL[ivt2], uR[ivt2], fL[ivt2], fR[ivt2], sL, sR);
        p_hll[ivn] = sstar;
        p_hll[ivt1] = u_hll[ivt1] / u_hll[IRHO];
        p_hll[ivt2] = u_hll[ivt2] / u_hll[IRHO];
        const double *q = (sstar >= 0.0) ? qL : qR, *u = (sstar >= 0.0) ? uL : uR, *f = (sstar >= 0.0) ? fL : fR;
        double s = (sstar >= 0.0) ? sL : sR, pt = (sstar >= 0.0) ? pLt : pRt;
        us[IRHO] = u[IRHO] * (s - q[ivn]) / (s - sstar);
        double vb = dot3(q, IBX, q, IVX);
        double vbhll = dot3(u_hll, IBX, p_hll, IVX);
        double ps = q[IRHO] * (s - q[ivn]) * (sstar - q[ivn]) + pt;
        us[ibn] = u_hll[ibn];
        us[ibt1] = u_hll[ibt1];
        us[ibt2] = u_hll[ibt2];
        us[ivn] = us[IRHO] * sstar;
        us[ivt1] = (u[ivt1] * (s - q[ivn]) - (us[ibn] * us[ibt1] - q[ibn] * q[ibt1])) / (s - sstar);
        us[ivt2] = (u[ivt2] * (s - q[ivn]) - (us[ibn] * us[ibt2] - q[ibn] * q[ibt2])) / (s - sstar);
        us[ITAU] = (u[ITAU] * (s - q[ivn]) + (ps * sstar - pt * q[ivn]) - ((vbhll * us[ibn] - vb * q[ibn]))) / (s - sstar);
        for (int i = 1; i <= 8; ++i) flux[i] = f[i] + s * (us[i] - u[i]);
        return;
    }
    /* HLLD */
    double sM = sstar_hllc(pLt, pRt, qL[IRHO], qR[IRHO], qL[ivn], qR[ivn], sL, sR, qL[ibn], qR[ibn]);
    double usL[9], usR[9], psL[9] = {0}, psR[9] = {0}, ussL[9], ussR[9], pssL[9] = {0}, pssR[9] = {0};
    for (int i = IRHO; i <= ITAU; ++i) usL[i] = uL[i] * (sL - qL[ivn]) / (sL - sM);
    for (int i = IRHO; i <= ITAU; ++i) usR[i] = uR[i] * (sR - qR[ivn]) / (sR - sM);
    double sstarL = sM - fabs(qL[ibn]) / sqrt(usL[IRHO]);
    double sstarR = sM + fabs(qR[ibn]) / sqrt(usR[IRHO]);
    if (((sstarL - sL) < 1.0e-4 * (sM - sL)) || ((sstarR - sR) > -1.0e-4 * (sR - sM))) {
        usL[ibn] = qL[ibn];
        usL[ibt1] = hll_state(qL[ibt1], qR[ibt1], fL[ibt1], fR[ibt1], sL, sR);
        usL[ibt2] = hll_state(qL[ibt2], qR[ibt2], fL[ibt2], fR[ibt2], sL, sR);
        usR[ibn] = qR[ibn];
        usR[ibt1] = usL[ibt1];
        usR[ibt2] = usL[ibt2];
        sstarL = sM;
        sstarR = sM;
    } else {
        double fac = (qL[IRHO] * (sL - qL[ivn]) * (sL - sM) - qL[ibn] * qL[ibn]);
        usL[ibn] = qL[ibn];
        usL[ibt1] = (qL[ibt1] * (qL[IRHO] * (sL - qL[ivn]) * (sL - qL[ivn]) - qL[ibn] * qL[ibn])) / fac;
        usL[ibt2] = (qL[ibt2] * (qL[IRHO] * (sL - qL[ivn]) * (sL - qL[ivn]) - qL[ibn] * qL[ibn])) / fac;
        fac = (qR[IRHO] * (sR - qR[ivn]) * (sR - sM) - qR[ibn] * qR[ibn]);
        usR[ibn] = qR[ibn];
        usR[ibt1] = (qR[ibt1] * (qR[IRHO] * (sR - qR[ivn]) * (sR - qR[ivn]) - qR[ibn] * qR[ibn])) / fac;
        usR[ibt2] = (qR[ibt2] * (qR[IRHO] * (sR - qR[ivn]) * (sR - qR[ivn]) - qR[ibn] * qR[ibn])) / fac;
    }
    usL[ivn] = usL[IRHO] * sM;
    usL[ivt1] = usL[ivt1] - (usL[ibn] * usL[ibt1] - qL[ibn] * qL[ibt1]) / (sL - sM);
    usL[ivt2] = usL[ivt2] - (usL[ibn] * usL[ibt2] - qL[ibn] * qL[ibt2]) / (sL - sM);
    usR[ivn] = usR[IRHO] * sM;
    usR[ivt1] = usR[ivt1] - (usR[ibn] * usR[ibt1] - qR[ibn] * qR[ibt1]) / (sR - sM);
    usR[ivt2] = usR[ivt2] - (usR[ibn] * usR[ibt2] - qR[ibn] * qR[ibt2]) / (sR - sM);
    psL[ivn] = sM;
    psL[ivt1] = usL[ivt1] / usL[IRHO];
    psL[ivt2] = usL[ivt2] / usL[IRHO];
    psR[ivn] = sM;
    psR[ivt1] = usR[ivt1] / usR[IRHO];
    psR[ivt2] = usR[ivt2] / usR[IRHO];
    double pLs = qL[IRHO] * (sL - qL[ivn]) * (sM - qL[ivn]) + pLt;
    double pRs = qR[IRHO] * (sR - qR[ivn]) * (sM - qR[ivn]) + pRt;
    double bdotu = dot3(qL, IBX, qL, IVX);
    double bdotus = dot3(usL, IBX, psL, IVX);
    usL[ITAU] = usL[ITAU] + ((pLs * sM - pLt * qL[ivn]) - (bdotus - bdotu) * qL[ibn]) / (sL - sM);
    bdotu = dot3(qR, IBX, qR, IVX);
    bdotus = dot3(usR, IBX, psR, IVX);
    usR[ITAU] = usR[ITAU] + ((pRs * sM - pRt * qR[ivn]) - (bdotus - bdotu) * qR[ibn]) / (sR - sM);
    if (sstarL >= 0.0) {
        for (int i = 1; i <= 8; ++i) flux[i] = fL[i] + sL * (usL[i] - uL[i]);
    } else if (sstarR <= 0.0) {
        for (int i = 1; i <= 8; ++i) flux[i] = fR[i] + sR * (usR[i] - uR[i]);
    } else {
        /* usstar(ibn) is never assigned in the reference (rsolver_module.f90:704-741).
         * bn_mode selects the value it holds: 0 -> 0.0, 1 -> Bn (Miyoshi-Kusano),
         * 2 -> the serial reference's behaviour: usstarL/usstarR keep the values
         * last written at an earlier double-star interface (loop order, earlier
         * sweeps), starting from 0. */
        if (bn_mode == 2) {
            for (int i = 1; i <= 8; ++i) { ussL[i] = keep[0][i]; ussR[i] = keep[1][i]; }
        } else {
            ussL[ibn] = (bn_mode == 1) ? qL[ibn] : 0.0;
            ussR[ibn] = (bn_mode == 1) ? qR[ibn] : 0.0;
        }
        ussL[IRHO] = usL[IRHO];
        ussR[IRHO] = usR[IRHO];
        double f1 = 1.0 / (sqrt(usL[IRHO]) + sqrt(usR[IRHO]));
        double f2 = fsign(1.0, qL[ibn]);
        pssL[ivn] = sM;
        pssL[ivt1] = f1 * (sqrt(usL[IRHO]) * psL[ivt1] + sqrt(usR[IRHO]) * psR[ivt1] + (usR[ibt1] - usL[ibt1]) * f2);
        pssL[ivt2] = f1 * (sqrt(usL[IRHO]) * psL[ivt2] + sqrt(usR[IRHO]) * psR[ivt2] + (usR[ibt2] - usL[ibt2]) * f2);
        ussL[ibt1] = f1 * (sqrt(usL[IRHO]) * usR[ibt1] + sqrt(usR[IRHO]) * usL[ibt1] + sqrt(usL[IRHO] * usR[IRHO]) * (psR[ivt1] - psL[ivt1]) * f2);
        ussL[ibt2] = f1 * (sqrt(usL[IRHO]) * usR[ibt2] + sqrt(usR[IRHO]) * usL[ibt2] + sqrt(usL[IRHO] * usR[IRHO]) * (psR[ivt2] - psL[ivt2]) * f2);
        ussL[ivn] = ussL[IRHO] * pssL[ivn];
        ussL[ivt1] = ussL[IRHO] * pssL[ivt1];
        ussL[ivt2] = ussL[IRHO] * pssL[ivt2];
        bdotus = dot3(usL, IBX, psL, IVX);
        double bdotuss = dot3(ussL, IBX, pssL, IVX);
        ussL[ITAU] = usL[ITAU] - sqrt(usL[IRHO]) * (bdotus - bdotuss) * f2;
        pssR[ivn] = sM;
        pssR[ivt1] = pssL[ivt1];
        pssR[ivt2] = pssL[ivt2];
        ussR[ibt1] = ussL[ibt1];
        ussR[ibt2] = ussL[ibt2];
        ussR[ivn] = ussR[IRHO] * pssR[ivn];
        ussR[ivt1] = ussR[IRHO] * pssR[ivt1];
        ussR[ivt2] = ussR[IRHO] * pssR[ivt2];
        bdotus = dot3(usR, IBX, psR, IVX);
        bdotuss = dot3(ussR, IBX, pssR, IVX);
        ussR[ITAU] = usR[ITAU] + sqrt(usR[IRHO]) * (bdotus - bdotuss) * f2;
        if (bn_mode == 2) {
            for (int i = 1; i <= 8; ++i) { keep[0][i] = ussL[i]; keep[1][i] = ussR[i]; }
        }
        if (sM >= 0.0) {
            for (int i = 1; i <= 8; ++i) flux[i] = fL[i] + sL * (usL[i] - uL[i]) + sstarL * (ussL[i] - usL[i]);
        } else {
            for (int i = 1; i <= 8; ++i) flux[i] = fR[i] + sR * (usR[i] - uR[i]) + sstarR * (ussR[i] - usR[i]);
        }
    }
}

void oracle_riemann_point(int solver, int dir, double ggas, int bn_mode, const double qL[9],
                          const double qR[9], double flux[9])
{
    perm_t p = perm_of(dir);
    double uL[9], uR[9], fL[9], fR[9];
    double epsL = eos_eps(ggas, qL[IRHO], qL[ITAU]), epsR = eos_eps(ggas, qR[IRHO], qR[ITAU]);
    double csL = eos_cs(ggas, qL[ITAU], qL[IRHO], epsL), csR = eos_cs(ggas, qR[ITAU], qR[IRHO], epsR);
    build_state(&p, qL, epsL, uL, fL);
    build_state(&p, qR, epsR, uR, fR);
    double keep[2][9] = {{0}};
    riemann_core(solver, &p, bn_mode, keep, qL, qR, csL, csR, uL, uR, fL, fR, flux);
}

/* BUILDSTATES over its loop range (spatial.f90:338-392) */
static void buildstates(oracle *o, int dir)
{
    perm_t p = perm_of(dir);
    int kx = 1, ky = 1, kz = 1;
    if (dir == XDIR) kx = 0; else if (dir == YDIR) ky = 0; else kz = 0;
    double q[9], u[9], f[9];
    for (int l = 0; l <= o->nz + kz; ++l)
        for (int k = 0; k <= o->ny + ky; ++k)
            for (int j = 0; j <= o->nx + kx; ++j) {
                long c = CI(o, j, k, l);
                for (int i = 1; i <= 8; ++i) q[i] = o->primL[i][c];
                build_state(&p, q, o->epsL[c], u, f);
                for (int i = 1; i <= 8; ++i) { o->uL[i][c] = u[i]; o->fluxL[i][c] = f[i]; }
                for (int i = 1; i <= 8; ++i) q[i] = o->primR[i][c];
                build_state(&p, q, o->epsR[c], u, f);
                for (int i = 1; i <= 8; ++i) { o->uR[i][c] = u[i]; o->fluxR[i][c] = f[i]; }
            }
}

/* Riemann solver loop over I_d (rsolver_module.f90:195-197 etc.) */
static void riemann(oracle *o, int dir)
{
    perm_t p = perm_of(dir);
    int kx = 1, ky = 1, kz = 1;
    if (dir == XDIR) kx = 0; else if (dir == YDIR) ky = 0; else kz = 0;
    double qL[9], qR[9], uL[9], uR[9], fL[9], fR[9], fl[9];
    for (int l = kz; l <= o->nz; ++l)
        for (int k = ky; k <= o->ny; ++k)
            for (int j = kx; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                for (int i = 1; i <= 8; ++i) {
                    qL[i] = o->primL[i][c]; qR[i] = o->primR[i][c];
                    uL[i] = o->uL[i][c]; uR[i] = o->uR[i][c];
                    fL[i] = o->fluxL[i][c]; fR[i] = o->fluxR[i][c];
                    /* HLLNM's IF/ELSEIF chain (rsolver_module.f90:303-311) has no
                     * ELSE: with a NaN signal speed no branch is taken and
                     * flux(:,j,k,l) keeps what the shared array held, i.e. the
                     * last sweep (of any direction, stage or step) that wrote
                     * this index */
                    fl[i] = o->flux[i][c];
                }
                riemann_core(o->d.riemann, &p, o->d.hlld_bn_mode, o->uss_keep, qL, qR, o->csL[c], o->csR[c],
                             uL, uR, fL, fR, fl);
                for (int i = 1; i <= 8; ++i) o->flux[i][c] = fl[i];
            }
}

/* ------------------------------------------------------------------------ */
/* sources, divergence, EMF                                                 */
/* ------------------------------------------------------------------------ */

/* GET_SOURCE, spatial.f90:195-256 */
static void get_source(oracle *o)
{
    for (int i = 1; i <= 8; ++i)
        for (int l = 1; l <= o->nz; ++l)
            for (int k = 1; k <= o->ny; ++k)
                for (int j = 1; j <= o->nx; ++j) o->sc[i][CI(o, j, k, l)] = 0.0;
    if (o->d.coordinate_flag == 1) {
        for (int l = 1; l <= o->nz; ++l)
            for (int k = 1; k <= o->ny; ++k)
                for (int j = 1; j <= o->nx; ++j) {
                    long c = CI(o, j, k, l);
                    double bx = o->bcell[IBX][c], by = o->bcell[IBY][c], bz = o->bcell[IBZ][c];
                    double bsq = bx * bx + by * by + bz * bz;
                    double rho = o->prim[IRHO][c], vx = o->prim[IVX][c], vy = o->prim[IVY][c];
                    o->sc[IVX][c] = o->sc[IVX][c] + (o->prim[ITAU][c] + rho * vy * vy + 0.5 * bsq - by * by) / (A1(o->x, j));
                    o->sc[IVY][c] = o->sc[IVY][c] - (rho * vx * vy - bx * by) / (A1(o->x, j));
                }
    } else if (o->d.coordinate_flag == 2) {
        for (int l = 1; l <= o->nz; ++l)
            for (int k = 1; k <= o->ny; ++k)
                for (int j = 1; j <= o->nx; ++j) {
                    long c = CI(o, j, k, l);
                    double bx = o->bcell[IBX][c], by = o->bcell[IBY][c], bz = o->bcell[IBZ][c];
                    double bsq = bx * bx + by * by + bz * bz;
                    double rho = o->prim[IRHO][c], vx = o->prim[IVX][c], vy = o->prim[IVY][c], vz = o->prim[IVZ][c];
                    double p = o->prim[ITAU][c];
                    double g = 3.0 * A1(o->x, j) * A1(o->dx, j) / A1(o->dx_cb, j);
                    double cot = A1(o->dsine, k) / A1(o->dcose, k);
                    o->sc[IVX][c] = o->sc[IVX][c] + (2.0 * p + bx * bx + rho * (vy * vy + vz * vz)) * g;
                    o->sc[IVY][c] = o->sc[IVY][c] + (p + rho * vz * vz + 0.5 * bsq - bz * bz) * g * cot
                                  - (rho * vx * vy - bx * by) * g;
                    o->sc[IVZ][c] = o->sc[IVZ][c] - (rho * vx * vz - bx * bz) * g
                                  - (rho * vy * vz - by * bz) * g * cot;
                }
    }
}

/* first_derivative, model/collapsar/src/custom.f90:402-409 */
static inline double first_derivative(double xm1, double xc, double xp1, double fm1, double fc, double fp1)
{
    double h2 = xp1 - xc, h1 = xc - xm1;
    return ((fp1 - fc) * h1 * h1 + (fc - fm1) * h2 * h2) / (h1 * h2 * (h1 + h2));
}

/* CUSTOM_SOURCE: model/riemann/src/custom.f90:143-155 (RT gravity),
 * model/torus-cyn/src/custom.f90:297-313 and model/torus-sph/src/custom.f90:283-296
 * (Paczynski-Wiita gravity with floor mask) */
static void custom_source(oracle *o)
{
    if (o->d.model == 0) {
        if (!o->d.rt_gravity) return;
        for (int l = 1; l <= o->nz; ++l)
            for (int k = 1; k <= o->ny; ++k)
                for (int j = 1; j <= o->nx; ++j) {
                    long c = CI(o, j, k, l);
                    o->sc[IVY][c] = o->sc[IVY][c] - 0.1 * o->prim[IRHO][c];
                }
        return;
    }
    double rsh = o->d.r_sh;
    if (o->d.model == 3) {
        /* model/collapsar/src/custom.f90:363-396: the matter potential's
         * gradient (three-point, non-uniform) plus the black hole */
        const double *phi = o->phi;
        long sy = o->NX, sz = (long)o->NX * o->NY;
        for (int l = 1; l <= o->nz; ++l)
            for (int k = 1; k <= o->ny; ++k)
                for (int j = 1; j <= o->nx; ++j) {
                    long c = CI(o, j, k, l);
                    double rho = o->prim[IRHO][c];
                    double diff = rho - o->d.prim_a[0];
                    double factor = fmax2(fsign(1.0, diff), 0.0);
                    double xj = A1(o->x, j), sk = A1(o->sine, k);
                    double dphidx = first_derivative(A1(o->x, j - 1), xj, A1(o->x, j + 1), phi[c - 1], phi[c], phi[c + 1]);
                    double dphidy = first_derivative(A1(o->y, k - 1), A1(o->y, k), A1(o->y, k + 1), phi[c - sy], phi[c],
                                                     phi[c + sy]);
                    double dphidz = first_derivative(A1(o->z, l - 1), A1(o->z, l), A1(o->z, l + 1), phi[c - sz], phi[c],
                                                     phi[c + sz]);
                    dphidx = dphidx + o->m_bh / ((xj - o->r_sh) * (xj - o->r_sh));
                    double vx = o->prim[IVX][c], vy = o->prim[IVY][c], vz = o->prim[IVZ][c];
                    o->sc[IVX][c] = o->sc[IVX][c] - factor * rho * dphidx;
                    o->sc[IVY][c] = o->sc[IVY][c] - factor * rho * dphidy / xj;
                    o->sc[IVZ][c] = o->sc[IVZ][c] - factor * rho * dphidz / xj / sk;
                    o->sc[ITAU][c] = o->sc[ITAU][c] - factor * rho * (vx * dphidx + vy * dphidy / xj + vz * dphidz / xj / sk);
                }
        return;
    }
    if (o->d.model == 2) {
        for (int l = 1; l <= o->nz; ++l)
            for (int k = 1; k <= o->ny; ++k)
                for (int j = 1; j <= o->nx; ++j) {
                    long c = CI(o, j, k, l);
                    double diff = o->prim[IRHO][c] - o->rho_floor[c];
                    double factor = fmax2(fsign(1.0, diff), 0.0);
                    double xj = A1(o->x, j);
                    double dphidr = 1.0 / ((xj - rsh) * (xj - rsh));
                    double rho = o->prim[IRHO][c];
                    o->sc[IVX][c] = o->sc[IVX][c] + (-factor * rho * dphidr);
                    o->sc[ITAU][c] = o->sc[ITAU][c] + (-factor * rho * o->prim[IVX][c] * dphidr);
                }
        return;
    }
    for (int l = 1; l <= o->nz; ++l)
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                double diff = o->prim[IRHO][c] - o->rho_floor[c];
                double factor = fmax2(fsign(1.0, diff), 0.0);
                double xj = A1(o->x, j), zl = A1(o->z, l);
                double radius = sqrt(xj * xj + zl * zl);
                double dphidr = xj / radius / ((radius - rsh) * (radius - rsh));
                double dphidz = zl / radius / ((radius - rsh) * (radius - rsh));
                double rho = o->prim[IRHO][c];
                o->sc[IVX][c] = o->sc[IVX][c] + (-factor * rho * dphidr);
                o->sc[IVZ][c] = o->sc[IVZ][c] + (-factor * rho * dphidz);
                o->sc[ITAU][c] = o->sc[ITAU][c] + (-factor * rho * (o->prim[IVX][c] * dphidr + o->prim[IVZ][c] * dphidz));
            }
}

/* FLUX_DIFF, spatial.f90:436-583 */
static void flux_diff(oracle *o, int dir)
{
    int cyl = o->d.coordinate_flag == 1, sph = o->d.coordinate_flag == 2;
    long sy = o->NX, sz = (long)o->NX * o->NY;
    for (int l = 1; l <= o->nz; ++l)
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j)
                for (int i = IRHO; i <= ITAU; ++i) {
                    long c = CI(o, j, k, l);
                    const double *f = o->flux[i];
                    double dflux;
                    if (dir == XDIR) {
                        if (cyl) dflux = (A1(o->xF, j) * f[c] - A1(o->xF, j - 1) * f[c - 1]) / (A1(o->x, j) * A1(o->dx, j));
                        else if (sph) {
                            double a = A1(o->xF, j), b = A1(o->xF, j - 1);
                            dflux = (a * a * f[c] - b * b * f[c - 1]) / (A1(o->dx_cb, j) / 3.0);
                        } else dflux = (f[c] - f[c - 1]) / (A1(o->dx, j));
                        o->l_rk[i][c] = -dflux;
                    } else if (dir == YDIR) {
                        if (cyl) dflux = (f[c] - f[c - sy]) / (A1(o->x, j) * A1(o->dy, k));
                        else if (sph) dflux = (A1(o->sinf, k) * f[c] - A1(o->sinf, k - 1) * f[c - sy]) /
                                              (A1(o->xbar, j) * A1(o->dcose, k));
                        else dflux = (f[c] - f[c - sy]) / (A1(o->dy, k));
                        o->l_rk[i][c] = o->l_rk[i][c] - dflux;
                    } else {
                        if (sph) dflux = (f[c] - f[c - sz]) * A1(o->dy, k) /
                                         (A1(o->xbar, j) * A1(o->dz, l) * A1(o->dcose, k));
                        else dflux = (f[c] - f[c - sz]) / A1(o->dz, l);
                        o->l_rk[i][c] = o->l_rk[i][c] - dflux;
                    }
                }
}

/* MHD_FLUX, mhd_module.f90:229-271 */
static void mhd_flux(oracle *o, int dir)
{
    if (dir == XDIR) {
        for (int l = 1; l <= o->nz; ++l)
            for (int k = 1; k <= o->ny; ++k)
                for (int j = 0; j <= o->nx; ++j) {
                    long c = CI(o, j, k, l);
                    o->eface[IYX][c] = o->flux[IBZ][c];
                    o->eface[IZX][c] = -o->flux[IBY][c];
                }
    } else if (dir == YDIR) {
        for (int l = 1; l <= o->nz; ++l)
            for (int k = 0; k <= o->ny; ++k)
                for (int j = 1; j <= o->nx; ++j) {
                    long c = CI(o, j, k, l);
                    o->eface[IXY][c] = -o->flux[IBZ][c];
                    o->eface[IZY][c] = o->flux[IBX][c];
                }
    } else {
        for (int l = 0; l <= o->nz; ++l)
            for (int k = 1; k <= o->ny; ++k)
                for (int j = 1; j <= o->nx; ++j) {
                    long c = CI(o, j, k, l);
                    o->eface[IXZ][c] = o->flux[IBY][c];
                    o->eface[IYZ][c] = -o->flux[IBX][c];
                }
    }
}

/* BOUNDARY_EFIELD, mhd_module.f90:478-840 */
static void boundary_efield(oracle *o)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    const int *bf = o->d.boundary_flag;
    double **E = o->ecell, **F = o->eface;
    /* x: ecell(iex:iez), eface ixz, ixy, iyz, izy */
    static const int xcomp[4] = {IXZ, IXY, IYZ, IZY}, xsgn[4] = {IEX, IEX, IEY, IEZ};
    static const int ycomp[4] = {IXZ, IYX, IYZ, IZX}, ysgn[4] = {IEX, IEY, IEY, IEZ};
    static const int zcomp[4] = {IXY, IYX, IZX, IZY}, zsgn[4] = {IEX, IEY, IEZ, IEZ};
    for (int side = 0; side < 2; ++side) {
        int flag = bf[side];
        const int *bfac = side ? o->bxout : o->bxin;
        for (int l = -2; l <= nz + 3; ++l)
            for (int k = -2; k <= ny + 3; ++k)
                for (int m = 1; m <= 3; ++m) {
                    int dst = side ? nx + m : 1 - m, src;
                    double s[12];
                    for (int i = IEX; i <= IEZ; ++i) s[i] = 1.0;
                    if (flag == 0) src = side ? m : nx + 1 - m;
                    else if (flag == 1) src = side ? nx : 1;
                    else { src = side ? nx + 1 - m : m; for (int i = IEX; i <= IEZ; ++i) s[i] = bfac[i]; }
                    long cd = CI(o, dst, k, l), cs = CI(o, src, k, l);
                    for (int i = IEX; i <= IEZ; ++i) E[i][cd] = (flag >= 2) ? s[i] * E[i][cs] : E[i][cs];
                    for (int q = 0; q < 4; ++q)
                        F[xcomp[q]][cd] = (flag >= 2) ? s[xsgn[q]] * F[xcomp[q]][cs] : F[xcomp[q]][cs];
                }
    }
    for (int side = 0; side < 2; ++side) {
        int flag = bf[2 + side];
        const int *bfac = side ? o->byout : o->byin;
        for (int l = -2; l <= nz + 3; ++l)
            for (int m = 1; m <= 3; ++m)
                for (int j = -2; j <= nx + 3; ++j) {
                    int dst = side ? ny + m : 1 - m, src;
                    if (flag == 0) src = side ? m : ny + 1 - m;
                    else if (flag == 1) src = side ? ny : 1;
                    else src = side ? ny + 1 - m : m;
                    long cd = CI(o, j, dst, l), cs = CI(o, j, src, l);
                    for (int i = IEX; i <= IEZ; ++i) E[i][cd] = (flag >= 2) ? bfac[i] * E[i][cs] : E[i][cs];
                    for (int q = 0; q < 4; ++q)
                        F[ycomp[q]][cd] = (flag >= 2) ? bfac[ysgn[q]] * F[ycomp[q]][cs] : F[ycomp[q]][cs];
                }
    }
    for (int side = 0; side < 2; ++side) {
        int flag = bf[4 + side];
        const int *bfac = side ? o->bzout : o->bzin;
        for (int m = 1; m <= 3; ++m)
            for (int k = -2; k <= ny + 3; ++k)
                for (int j = -2; j <= nx + 3; ++j) {
                    int dst = side ? nz + m : 1 - m, src;
                    if (flag == 0) src = side ? m : nz + 1 - m;
                    else if (flag == 1) src = side ? nz : 1;
                    else src = side ? nz + 1 - m : m;
                    long cd = CI(o, j, k, dst), cs = CI(o, j, k, src);
                    for (int i = IEX; i <= IEZ; ++i) E[i][cd] = (flag >= 2) ? bfac[i] * E[i][cs] : E[i][cs];
                    for (int q = 0; q < 4; ++q)
                        F[zcomp[q]][cd] = (flag >= 2) ? bfac[zsgn[q]] * F[zcomp[q]][cs] : F[zcomp[q]][cs];
                }
    }
}

/* flux_ct, mhd_module.f90:285-473 */
static void flux_ct(oracle *o)
{
    double **E = o->ecell, **F = o->eface;
    for (int l = 1; l <= o->nz; ++l)
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                double vx = o->prim[IVX][c], vy = o->prim[IVY][c], vz = o->prim[IVZ][c];
                double bx = o->bcell[IBX][c], by = o->bcell[IBY][c], bz = o->bcell[IBZ][c];
                E[IEX][c] = vz * by - vy * bz;
                E[IEY][c] = vx * bz - vz * bx;
                E[IEZ][c] = vy * bx - vx * by;
            }
    if (o->d.n_dim == 1) {
        for (int j = 1; j <= o->nx; ++j) {
            long c11 = CI(o, j, 1, 1);
            F[IXY][c11] = E[IEX][c11];
            F[IZY][c11] = E[IEZ][c11];
            F[IXY][CI(o, j, 0, 1)] = E[IEX][c11];
            F[IZY][CI(o, j, 0, 1)] = E[IEZ][c11];
            F[IXZ][c11] = E[IEX][c11];
            F[IYZ][c11] = E[IEY][c11];
            F[IXZ][CI(o, j, 1, 0)] = E[IEX][c11];
            F[IYZ][CI(o, j, 1, 0)] = E[IEY][c11];
        }
    } else if (o->d.n_dim == 2) {
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j) {
                long c1 = CI(o, j, k, 1), c0 = CI(o, j, k, 0);
                F[IXZ][c1] = E[IEX][c1];
                F[IYZ][c1] = E[IEY][c1];
                F[IXZ][c0] = E[IEX][c1];
                F[IYZ][c0] = E[IEY][c1];
            }
    }
    boundary_efield(o);
    long sy = o->NX, sz = (long)o->NX * o->NY;
    for (int l = 0; l <= o->nz; ++l)
        for (int k = 0; k <= o->ny; ++k)
            for (int j = 0; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                o->efx[c] = 0.50 * (F[IXY][c] + F[IXY][c + sz] + F[IXZ][c] + F[IXZ][c + sy])
                          - 0.25 * (E[IEX][c] + E[IEX][c + sz] + E[IEX][c + sy] + E[IEX][c + sy + sz]);
                o->efy[c] = 0.50 * (F[IYX][c] + F[IYX][c + sz] + F[IYZ][c] + F[IYZ][c + 1])
                          - 0.25 * (E[IEY][c] + E[IEY][c + sz] + E[IEY][c + 1] + E[IEY][c + 1 + sz]);
                o->efz[c] = 0.50 * (F[IZX][c] + F[IZX][c + sy] + F[IZY][c] + F[IZY][c + 1])
                          - 0.25 * (E[IEZ][c] + E[IEZ][c + sy] + E[IEZ][c + 1] + E[IEZ][c + 1 + sy]);
            }
    /* FIXPOLE, mhd_module.f90:437-461: the radial EMF on the polar faces is
     * replaced by its phi average.  The accumulators sum1/sum2 are declared
     * REAL (:293), so each addition is done in double and rounded to single;
     * the order is the serial l loop of a one-thread run. */
    if (o->d.coordinate_flag == 2 && o->d.fixpole) {
        for (int j = 1; j <= o->nx; ++j) {
            float sum1 = 0.0f, sum2 = 0.0f;
            double dsum1 = 0.0, dsum2 = 0.0;   /* fixpole == 2: a planted mistake for tests */
            for (int l = 1; l <= o->nz; ++l) {
                sum1 = (float)((double)sum1 + o->efx[CI(o, j, 0, l)]);
                sum2 = (float)((double)sum2 + o->efx[CI(o, j, o->ny, l)]);
                dsum1 += o->efx[CI(o, j, 0, l)];
                dsum2 += o->efx[CI(o, j, o->ny, l)];
            }
            double a1 = o->d.fixpole == 2 ? dsum1 : (double)sum1, a2 = o->d.fixpole == 2 ? dsum2 : (double)sum2;
            for (int l = -2; l <= o->nz + 3; ++l) {
                o->efx[CI(o, j, 0, l)] = a1 / (double)o->nz;
                o->efx[CI(o, j, o->ny, l)] = a2 / (double)o->nz;
            }
        }
    }
}

/* l_operator, spatial.f90:620-702 */
static void l_operator(oracle *o)
{
    for (int l = 1; l <= o->nz; ++l)
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j)
                for (int i = IRHO; i <= ITAU; ++i) {
                    long c = CI(o, j, k, l);
                    o->l_rk[i][c] = o->l_rk[i][c] + o->sc[i][c];
                }
    long sy = o->NX, sz = (long)o->NX * o->NY;
    const double *ex = o->efx, *ey = o->efy, *ez = o->efz;
    for (int l = 0; l <= o->nz; ++l)
        for (int k = 0; k <= o->ny; ++k)
            for (int j = 0; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                double dxj = A1(o->dx, j), dyk = A1(o->dy, k), dzl = A1(o->dz, l);
                if (o->d.coordinate_flag == 0) {
                    o->l_rk[IBX][c] = -((ez[c] - ez[c - sy]) / (dyk) - (ey[c] - ey[c - sz]) / (dzl));
                    o->l_rk[IBY][c] = -((ex[c] - ex[c - sz]) / (dzl) - (ez[c] - ez[c - 1]) / (dxj));
                    o->l_rk[IBZ][c] = -((ey[c] - ey[c - 1]) / (dxj) - (ex[c] - ex[c - sy]) / (dyk));
                } else if (o->d.coordinate_flag == 2) {
                    double xFj = A1(o->xF, j), xFm = A1(o->xF, j - 1), xj = A1(o->x, j), xb = A1(o->xbar, j);
                    double sk = A1(o->sinf, k), skm = A1(o->sinf, k - 1), dc = A1(o->dcose, k);
                    o->l_rk[IBX][c] = (-(sk * ez[c] - skm * ez[c - sy]) + (ey[c] - ey[c - sz]) * (dyk / dzl)) /
                                      (xFj * dc + SMALL_NUM);
                    o->l_rk[IBY][c] = (-(ex[c] - ex[c - sz]) / (xb * sk * dzl + SMALL_NUM) +
                                       (xFj * ez[c] - xFm * ez[c - 1]) / (xj * dxj));
                    o->l_rk[IBZ][c] = (-(xFj * ey[c] - xFm * ey[c - 1]) / (xj * dxj) + (ex[c] - ex[c - sy]) / (xb * dyk));
                } else {
                    double xFj = A1(o->xF, j), xFm = A1(o->xF, j - 1), xj = A1(o->x, j);
                    o->l_rk[IBX][c] = -(ez[c] - ez[c - sy]) / (xFj * dyk) + (ey[c] - ey[c - sz]) / (dzl);
                    o->l_rk[IBY][c] = -(ex[c] - ex[c - sz]) / (dzl) + (ez[c] - ez[c - 1]) / (dxj);
                    o->l_rk[IBZ][c] = -(xFj * ey[c] - xFm * ey[c - 1]) / (xj * dxj) + (ex[c] - ex[c - sy]) / (xj * dyk);
                }
            }
}

/* SPATIAL, spatial.f90:11-162 */
static void spatial(oracle *o)
{
    get_source(o);
    custom_source(o);
    for (int dir = XDIR; dir <= o->d.n_dim; ++dir) {
        reconstruct(o, dir);
        buildstates(o, dir);
        riemann(o, dir);
        flux_diff(o, dir);
        mhd_flux(o, dir);
        if (dir == XDIR)    /* GETFLUX_X of the collapsar, custom.f90:220-240: mflux_x = flux(irho,0,k,l) */
            for (int l = 1; l <= o->nz; ++l)
                for (int k = 1; k <= o->ny; ++k) o->mflux[(size_t)(l + 2) * o->NY + (k + 2)] = o->flux[IRHO][CI(o, 0, k, l)];
    }
    flux_ct(o);
    l_operator(o);
}

/* ------------------------------------------------------------------------ */
/* conversions, floors, boundaries, EOS update                              */
/* ------------------------------------------------------------------------ */

static void bcell_from(oracle *o, double **B)
{
    long sy = o->NX, sz = (long)o->NX * o->NY;
    for (int l = 1; l <= o->nz; ++l)
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                if (o->d.coordinate_flag == 2) {
                    /* fromrvetou.f90:76-88 and :181-193 */
                    double a = A1(o->xF, j), b = A1(o->xF, j - 1);
                    o->bcell[IBX][c] = 1.5 * A1(o->dx, j) / A1(o->dx_cb, j) * (a * a * B[IBX][c] + b * b * B[IBX][c - 1]);
                    o->bcell[IBY][c] = 0.5 * A1(o->dy, k) / A1(o->dcose, k) *
                                       (A1(o->sinf, k) * B[IBY][c] + A1(o->sinf, k - 1) * B[IBY][c - sy]);
                    o->bcell[IBZ][c] = 0.5 * (B[IBZ][c] + B[IBZ][c - sz]);
                    continue;
                }
                if (o->d.coordinate_flag == 0)
                    o->bcell[IBX][c] = 0.5 * (B[IBX][c] + B[IBX][c - 1]);
                else
                    o->bcell[IBX][c] = (A1(o->xF, j) * B[IBX][c] + A1(o->xF, j - 1) * B[IBX][c - 1]) /
                                       (A1(o->xF, j) + A1(o->xF, j - 1));
                o->bcell[IBY][c] = 0.5 * (B[IBY][c] + B[IBY][c - sy]);
                o->bcell[IBZ][c] = 0.5 * (B[IBZ][c] + B[IBZ][c - sz]);
            }
}

/* FROMRVETOU, fromrvetou.f90:21-121 */
static void fromrvetou(oracle *o)
{
    bcell_from(o, o->prim);
    for (int l = 0; l <= o->nz; ++l)
        for (int k = 0; k <= o->ny; ++k)
            for (int j = 0; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                double vx = o->prim[IVX][c], vy = o->prim[IVY][c], vz = o->prim[IVZ][c];
                double bx = o->bcell[IBX][c], by = o->bcell[IBY][c], bz = o->bcell[IBZ][c];
                double vsq = vx * vx + vy * vy + vz * vz, bsq = bx * bx + by * by + bz * bz;
                double rho = o->prim[IRHO][c];
                o->cons[IRHO][c] = rho;
                for (int i = IVX; i <= IVZ; ++i) o->cons[i][c] = o->prim[i][c] * rho;
                o->cons[ITAU][c] = rho * (o->eps[c] + 0.5 * vsq) + 0.5 * bsq;
                for (int i = IBX; i <= IBZ; ++i) o->cons[i][c] = o->prim[i][c];
            }
}

/* FROMUTORVE, fromrvetou.f90:126-227 */
static void fromutorve(oracle *o)
{
    bcell_from(o, o->cons);
    for (int l = 1; l <= o->nz; ++l)
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                double rho = o->cons[IRHO][c];
                o->prim[IRHO][c] = rho;
                for (int i = IVX; i <= IVZ; ++i) o->prim[i][c] = o->cons[i][c] / rho;
                for (int i = IBX; i <= IBZ; ++i) o->prim[i][c] = o->cons[i][c];
                double bx = o->bcell[IBX][c], by = o->bcell[IBY][c], bz = o->bcell[IBZ][c];
                double vx = o->prim[IVX][c], vy = o->prim[IVY][c], vz = o->prim[IVZ][c];
                double bsq = bx * bx + by * by + bz * bz, vsq = vx * vx + vy * vy + vz * vz;
                o->eps[c] = (o->cons[ITAU][c] - 0.5 * bsq) / rho - 0.5 * vsq;
            }
}

/* CUSTOM_CHECKRHO: model/riemann/src/custom.f90:115-121, model/torus-cyn/src/custom.f90:254-271
 * (model/torus-sph/src/custom.f90:240-259 is the same) */
static void checkrho(oracle *o)
{
    for (int l = 1; l <= o->nz; ++l)
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                if (o->d.model == 0) {
                    o->eps[c] = fmax2(o->eps[c], 0.0);
                } else if (o->d.model == 3) {
                    /* model/collapsar/src/custom.f90:266-282 */
                    double diff = o->prim[IRHO][c] - o->d.prim_a[0];
                    double f = fmax2(fsign(1.0, diff), 0.0);
                    for (int i = IRHO; i <= IVZ; ++i)
                        o->prim[i][c] = f * o->prim[i][c] + (1.0 - f) * o->d.prim_a[i - IRHO];
                    o->eps[c] = f * o->eps[c] + (1.0 - f) * o->d.eps_a;
                    if (o->eps[c] < 0.0) o->eps[c] = o->d.eps_a;
                } else {
                    double diff = o->prim[IRHO][c] - o->rho_floor[c];
                    double f = fmax2(fsign(1.0, diff), 0.0);
                    o->prim[IRHO][c] = f * o->prim[IRHO][c] + (1.0 - f) * o->rho_floor[c];
                    for (int i = IVX; i <= IVZ; ++i) o->prim[i][c] = f * o->prim[i][c];
                    o->eps[c] = f * o->eps[c] + (1.0 - f) * o->eps_floor[c];
                    if (o->eps[c] < 0.0) o->eps[c] = o->eps_floor[c];
                }
            }
}

/* one ghost-fill pass of BOUNDARYP_NM (boundary.f90:423-745) for axis a */
static int ghost_src(int flag, int side, int n, int m)
{
    if (flag == 0) return side ? m : n + 1 - m;
    if (flag == 1) return side ? n : 1;
    return side ? n + 1 - m : m;
}

static void boundaryp(oracle *o)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    const int *bf = o->d.boundary_flag;
    for (int side = 0; side < 2; ++side) {          /* x, boundary.f90:423-523 */
        int flag = bf[side];
        const int *b = side ? o->bxout : o->bxin;
        for (int l = -2; l <= nz + 3; ++l)
            for (int k = -2; k <= ny + 3; ++k)
                for (int m = 1; m <= 3; ++m) {
                    int dst = side ? nx + m : 1 - m, src = ghost_src(flag, side, nx, m);
                    long cd = CI(o, dst, k, l), cs = CI(o, src, k, l);
                    for (int i = IRHO; i <= ITAU; ++i) o->prim[i][cd] = (flag >= 2) ? b[i] * o->prim[i][cs] : o->prim[i][cs];
                    for (int i = IBX; i <= IBZ; ++i) o->bcell[i][cd] = (flag >= 2) ? b[i] * o->bcell[i][cs] : o->bcell[i][cs];
                    o->prim[IBY][cd] = (flag >= 2) ? b[IBY] * o->prim[IBY][cs] : o->prim[IBY][cs];
                    o->prim[IBZ][cd] = (flag >= 2) ? b[IBZ] * o->prim[IBZ][cs] : o->prim[IBZ][cs];
                }
    }
    for (int side = 0; side < 2; ++side) {          /* y, boundary.f90:534-634 */
        int flag = bf[2 + side];
        const int *b = side ? o->byout : o->byin;
        for (int l = -2; l <= nz + 3; ++l)
            for (int m = 1; m <= 3; ++m)
                for (int j = -2; j <= nx + 3; ++j) {
                    int dst = side ? ny + m : 1 - m, src = ghost_src(flag, side, ny, m);
                    long cd = CI(o, j, dst, l), cs = CI(o, j, src, l);
                    for (int i = IRHO; i <= ITAU; ++i) o->prim[i][cd] = (flag >= 2) ? b[i] * o->prim[i][cs] : o->prim[i][cs];
                    for (int i = IBX; i <= IBZ; ++i) o->bcell[i][cd] = (flag >= 2) ? b[i] * o->bcell[i][cs] : o->bcell[i][cs];
                    o->prim[IBX][cd] = (flag >= 2) ? b[IBX] * o->prim[IBX][cs] : o->prim[IBX][cs];
                    o->prim[IBZ][cd] = (flag >= 2) ? b[IBZ] * o->prim[IBZ][cs] : o->prim[IBZ][cs];
                }
    }
    for (int side = 0; side < 2; ++side) {          /* z, boundary.f90:645-745 */
        int flag = bf[4 + side];
        const int *b = side ? o->bzout : o->bzin;
        for (int m = 1; m <= 3; ++m)
            for (int k = -2; k <= ny + 3; ++k)
                for (int j = -2; j <= nx + 3; ++j) {
                    int dst = side ? nz + m : 1 - m, src = ghost_src(flag, side, nz, m);
                    long cd = CI(o, j, k, dst), cs = CI(o, j, k, src);
                    for (int i = IRHO; i <= ITAU; ++i) o->prim[i][cd] = (flag >= 2) ? b[i] * o->prim[i][cs] : o->prim[i][cs];
                    for (int i = IBX; i <= IBZ; ++i) o->bcell[i][cd] = (flag >= 2) ? b[i] * o->bcell[i][cs] : o->bcell[i][cs];
                    o->prim[IBX][cd] = (flag >= 2) ? b[IBX] * o->prim[IBX][cs] : o->prim[IBX][cs];
                    o->prim[IBY][cd] = (flag >= 2) ? b[IBY] * o->prim[IBY][cs] : o->prim[IBY][cs];
                }
    }
    if (o->d.model >= 1) {   /* CUSTOM_BOUNDARY_X, model/torus-cyn/src/custom.f90:146-163,
                                model/torus-sph/src/custom.f90:150-174 */
        for (int l = -2; l <= nz + 3; ++l)
            for (int k = -2; k <= ny + 3; ++k)
                for (int m = 1; m <= 3; ++m) {
                    long ci = CI(o, 1 - m, k, l), co = CI(o, nx + m, k, l);
                    o->prim[IVX][ci] = fmin2(o->prim[IVX][ci], 0.0);
                    o->prim[IVX][co] = fmax2(o->prim[IVX][co], 0.0);
                }
    }
    if (o->d.model == 1) {   /* CUSTOM_BOUNDARY_Z, model/torus-cyn/src/custom.f90:186-201 */
        for (int m = 1; m <= 3; ++m)
            for (int k = -2; k <= ny + 3; ++k)
                for (int j = -2; j <= nx + 3; ++j) {
                    long ci = CI(o, j, k, 1 - m), co = CI(o, j, k, nz + m);
                    o->prim[IVZ][ci] = fmin2(o->prim[IVZ][ci], 0.0);
                    o->prim[IVZ][co] = fmax2(o->prim[IVZ][co], 0.0);
                }
    }
}

/* BOUNDARY1D_NM(epsilon, even...), boundary.f90:29-389, including the
 * y-outer branch keyed on boundary_flag(3) (:243) */
static void boundary_eps(oracle *o)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    const int *bf = o->d.boundary_flag;
    double *a = o->eps;
    for (int side = 0; side < 2; ++side) {
        int flag = bf[side];
        for (int l = -2; l <= nz + 3; ++l)
            for (int k = -2; k <= ny + 3; ++k)
                for (int m = 1; m <= 3; ++m) {
                    int dst = side ? nx + m : 1 - m, src = ghost_src(flag, side, nx, m);
                    a[CI(o, dst, k, l)] = a[CI(o, src, k, l)];
                }
    }
    for (int side = 0; side < 2; ++side) {
        int flag = bf[2 + side];
        if (side == 1) {
            if (bf[2] == 0) flag = 0;
            else if (bf[3] == 1) flag = 1;
            else if (bf[3] >= 2) flag = 2;
            else continue;
        }
        for (int l = -2; l <= nz + 3; ++l)
            for (int m = 1; m <= 3; ++m)
                for (int j = -2; j <= nx + 3; ++j) {
                    int dst = side ? ny + m : 1 - m, src = ghost_src(flag, side, ny, m);
                    a[CI(o, j, dst, l)] = a[CI(o, j, src, l)];
                }
    }
    for (int side = 0; side < 2; ++side) {
        int flag = bf[4 + side];
        for (int m = 1; m <= 3; ++m)
            for (int k = -2; k <= ny + 3; ++k)
                for (int j = -2; j <= nx + 3; ++j) {
                    int dst = side ? nz + m : 1 - m, src = ghost_src(flag, side, nz, m);
                    a[CI(o, j, k, dst)] = a[CI(o, j, k, src)];
                }
    }
}

static void boundary(oracle *o) { boundaryp(o); boundary_eps(o); }   /* boundary.f90:5-13 */

/* UPDATE -> FINDPRESSURE: model/riemann/src/findpressure.f90:17-28,
 * model/torus-cyn/src/findpressure.f90:18-32 */
static void update(oracle *o)
{
    double g = o->d.ggas;
    for (long c = 0; c < o->G; ++c) {
        double rho = o->prim[IRHO][c];
        if (o->d.model == 0 || o->d.model == 3 || rho > o->rho_floor[c]) {
            o->prim[ITAU][c] = rho * o->eps[c] * (g - 1.0);
            o->dpdrho[c] = o->eps[c] * (g - 1.0);
            o->dpdeps[c] = rho * (g - 1.0);
        } else {
            o->prim[ITAU][c] = o->p_floor[c];
            o->dpdrho[c] = o->eps_floor[c] * (g - 1.0);
            o->dpdeps[c] = o->rho_floor[c] * (g - 1.0);
        }
        o->cs[c] = sqrt(o->dpdrho[c] + o->dpdeps[c] * o->prim[ITAU][c] / (rho * rho));
    }
}

/* The red-black SOR relaxation of the potential, CUSTOM_UPDATE's loop at
 * model/collapsar/src/custom.f90:511-647.  Red cells ((-1)**(j+k+l) > 0) are
 * swept first, then black; each reads only cells of the other colour, so the
 * order inside a sweep does not matter.
 * The error is the OpenMP reduction of :588-598 as the flang -fopenmp build
 * computes it on one thread: the private copy folds
 * MAX(abserror, |dphi/phi_old|) in l, k, j order with flang's MAX
 * (a > b ? a : b), so a NaN term replaces the running value and the next
 * term replaces the NaN; the combine with the shared 1e-50 drops a NaN
 * result.  A NaN last term (0/0 where phi stays 0, as on the first sweep
 * from the zero guess) therefore gives abserror = 1e-50 and ends the loop;
 * a serial build would keep the NaN and go on.  Returns the iteration at
 * which the loop ended (relax_max + 1 when it ran out without converging). */
static int sor_solve(oracle *o)
{
    int nx = o->nx, ny = o->ny, nz = o->nz;
    long sy = o->NX, sz = (long)o->NX * o->NY;
    double *phi = o->phi, *old = o->phi_old;
    double pa = o->d.prim_a[0], w = o->d.omega;
    const double pi = 4.0 * atan(1.0);   /* parameter.h:21 */
    int n;
    for (n = 1; n <= o->d.relax_max; ++n) {
        for (int l = 1; l <= nz; ++l)
            for (int k = 1; k <= ny; ++k)
                for (int j = 1; j <= nx; ++j) old[CI(o, j, k, l)] = phi[CI(o, j, k, l)];
        for (int colour = 0; colour < 2; ++colour)
            for (int l = 1; l <= nz; ++l)
                for (int k = 1; k <= ny; ++k)
                    for (int j = 1; j <= nx; ++j) {
                        if (((j + k + l) & 1) != colour) continue;
                        long c = CI(o, j, k, l);
                        double rho = o->prim[IRHO][c];
                        double diff = rho - pa;
                        double factor = diff > 0.0 ? 1.0 : 0.0;
                        double rho_in = factor * rho;
                        size_t jk = (size_t)k * (nx + 1) + j;
                        double s = o->ajp1[j] * phi[c + 1] + o->ajm1[j] * phi[c - 1] + o->bkp1[jk] * phi[c + sy] +
                                   o->bkm1[jk] * phi[c - sy] + o->clp1[c] * phi[c + sz] + o->clm1[c] * phi[c - sz];
                        double rhs = (4.0 * pi * rho_in - s) / o->epsc[c];
                        phi[c] = (1.0 - w) * phi[c] + w * rhs;
                    }
        double priv = -DBL_MAX;
        for (int l = 1; l <= nz; ++l)
            for (int k = 1; k <= ny; ++k)
                for (int j = 1; j <= nx; ++j) {
                    long c = CI(o, j, k, l);
                    priv = fmax2(priv, fabs((phi[c] - old[c]) / old[c]));
                }
        double err = priv != priv ? 1.0e-50 : fmax2(1.0e-50, priv);
        if (o->d.sor_fold == 1) {   /* planted: the serial fold from 1e-50 keeps a NaN last term */
            err = 1.0e-50;
            for (int l = 1; l <= nz; ++l)
                for (int k = 1; k <= ny; ++k)
                    for (int j = 1; j <= nx; ++j) {
                        long c = CI(o, j, k, l);
                        err = fmax2(err, fabs((phi[c] - old[c]) / old[c]));
                    }
        }
        o->sor_err = err;
        for (int l = 1; l <= nz; ++l)
            for (int k = 1; k <= ny; ++k) {
                phi[CI(o, 0, k, l)] = phi[CI(o, 1, k, l)];
                phi[CI(o, nx + 1, k, l)] = -1.0;
            }
        for (int l = 1; l <= nz; ++l)
            for (int j = 1; j <= nx; ++j) {
                phi[CI(o, j, 0, l)] = phi[CI(o, j, 1, l)];
                phi[CI(o, j, ny + 1, l)] = phi[CI(o, j, ny, l)];
            }
        for (int k = 1; k <= ny; ++k)
            for (int j = 1; j <= nx; ++j) {
                phi[CI(o, j, k, 0)] = phi[CI(o, j, k, nz)];
                phi[CI(o, j, k, nz + 1)] = phi[CI(o, j, k, 1)];
            }
        if (err <= o->d.tolerance) break;
    }
    return n;
}

/* CUSTOM_UPDATE of the collapsar, model/collapsar/src/custom.f90:416-658:
 * black-hole accretion through the inner face (stages 1-3, from the x mass
 * flux GETFLUX_X saved), the horizon check, and the potential solve at the
 * initial update and on stage 3 of every n_pot-th step.  A STOP of the
 * reference sets o->status (1 horizon, 2 convergence) and ends oracle_run. */
static void custom_update(oracle *o, int p, double dt)
{
    if (o->d.model != 3 || o->status) return;
    if (p > 0) {
        const double rk20 = 3.0 / 4.0, rk21 = 1.0 / 4.0, rk22 = 1.0 / 4.0;
        const double rk30 = 1.0 / 3.0, rk31 = 2.0 / 3.0, rk32 = 2.0 / 3.0;
        double xf0 = A1(o->xF, 0), mdot = 0.0;
        for (int l = 1; l <= o->nz; ++l)
            for (int k = 1; k <= o->ny; ++k)
                mdot = mdot + (-o->mflux[(size_t)(l + 2) * o->NY + (k + 2)]) * xf0 * xf0 * A1(o->dcose, k) * A1(o->dz, l);
        mdot = fmax2(mdot, 0.0);
        if (p == 1) {
            o->mbh_old = o->m_bh;
            o->m_bh = o->mbh_old + dt * mdot;
        } else if (p == 2) {
            o->m_bh = rk20 * o->mbh_old + rk21 * o->m_bh + rk22 * dt * mdot;
        } else {
            o->m_bh = rk30 * o->mbh_old + rk31 * o->m_bh + rk32 * dt * mdot;
        }
    }
    o->r_sh = 2.0 * o->m_bh / (o->d.c_code * o->d.c_code);
    if (o->r_sh >= A1(o->xF, 0)) {
        o->status = 1;
        return;
    }
    if (p == 0 || (p == 3 && o->nstep % o->d.n_pot == 0)) {
        if (p == 0)
            for (int l = 0; l <= o->nz + 1; ++l)
                for (int k = 0; k <= o->ny + 1; ++k)
                    for (int j = 0; j <= o->nx + 1; ++j) o->phi[CI(o, j, k, l)] = 0.0;
        int n = sor_solve(o);
        o->sor_iter = n;
        if (n == o->d.relax_max) o->status = 2;
    }
}

void oracle_initial_update(oracle *o)
{
    checkrho(o);
    fromrvetou(o);
    boundary(o);
    update(o);
    custom_update(o, 0, 0.0);
}

/* FINDDT, rungekutta.f90:256-323 */
double oracle_finddt(oracle *o)
{
    double dt_out = 1.0e5;
    for (int l = 1; l <= o->nz; ++l)
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                double cs = o->cs[c], rho = o->prim[IRHO][c];
                double a2 = cs * cs, a4 = a2 * a2;
                double bx = o->bcell[IBX][c], by = o->bcell[IBY][c], bz = o->bcell[IBZ][c];
                double b2x = (bx * bx / rho), b2y = (by * by / rho), b2z = (bz * bz / rho);
                double b2 = b2x + b2y + b2z, b4 = b2 * b2;
                double cfx = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2x)));
                double cfy = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2y)));
                double cfz = sqrt(0.5 * (a2 + b2 + sqrt((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2z)));
                double l1 = fabs(o->prim[IVX][c]) + cfx, l2 = fabs(o->prim[IVY][c]) + cfy,
                       l3 = fabs(o->prim[IVZ][c]) + cfz;
                double lambda = fmax2(fmax2(l1, l2), l3);
                double h = A1(o->dx, j);
                if (o->d.coordinate_flag == 0) {
                    if (o->d.n_dim > 1) h = fmin2(h, A1(o->dy, k));
                    if (o->d.n_dim > 2) h = fmin2(h, A1(o->dz, l));
                } else if (o->d.coordinate_flag == 1) {
                    if (o->d.n_dim > 1 && o->ny > 1) h = fmin2(h, A1(o->x, j) * A1(o->dy, k));
                    if (o->d.n_dim > 2) h = fmin2(h, A1(o->dz, l));
                } else {
                    if (o->d.n_dim > 1) h = fmin2(h, A1(o->x, j) * A1(o->dy, k));
                    if (o->d.n_dim > 2) h = fmin2(h, A1(o->x, j) * A1(o->sine, k) * A1(o->dz, l));
                }
                h = h * o->d.cfl / lambda;
                dt_out = fmin2(dt_out, h);
            }
    return fmin2(1.0e5, dt_out);
}

/* SPATIAL on its own (spatial.f90:11-162): l_rk of the current state, with
 * SPATIAL's side effects (the fluxes, the HLLD carry, the collapsar's mflux) */
void oracle_spatial(oracle *o) { spatial(o); }

/* One stage of RUNGEKUTTA (rungekutta.f90:37-199): stage 1 backs up u_old
 * (:37-50), every stage runs SPATIAL, its combine, FROMUTORVE, CUSTOM_CHECKRHO
 * (twice at stage 3, around the empty OPERATOR_SPLIT), FROMRVETOU, BOUNDARY
 * and UPDATE(stage).  Stage 3 ends the step (n_step, main.f90:112). */
void oracle_stage(oracle *o, int stage, double dt)
{
    const double rk20 = 3.0 / 4.0, rk21 = 1.0 / 4.0, rk22 = 1.0 / 4.0;
    const double rk30 = 1.0 / 3.0, rk31 = 2.0 / 3.0, rk32 = 2.0 / 3.0;
    if (stage == 1) {
        for (int l = 0; l <= o->nz; ++l)
            for (int k = 0; k <= o->ny; ++k)
                for (int j = 0; j <= o->nx; ++j)
                    for (int i = 1; i <= 8; ++i) {
                        long c = CI(o, j, k, l);
                        o->u_old[i][c] = o->cons[i][c];
                    }
    }
    spatial(o);
    for (int l = 0; l <= o->nz; ++l)
        for (int k = 0; k <= o->ny; ++k)
            for (int j = 0; j <= o->nx; ++j)
                for (int i = 1; i <= 8; ++i) {
                    long c = CI(o, j, k, l);
                    double uo = o->u_old[i][c], u = o->cons[i][c], lr = o->l_rk[i][c];
                    if (stage == 1) o->cons[i][c] = uo + dt * lr;
                    else if (stage == 2) o->cons[i][c] = rk20 * uo + rk21 * u + rk22 * dt * lr;
                    else o->cons[i][c] = rk30 * uo + rk31 * u + rk32 * dt * lr;
                }
    fromutorve(o);
    checkrho(o);
    if (stage == 3) checkrho(o);   /* OPERATOR_SPLIT is empty in both models */
    fromrvetou(o);
    boundary(o);
    update(o);
    custom_update(o, stage, dt);
    if (stage == 3) o->nstep = o->nstep + 1;   /* main.f90:112, after the step */
}

/* RUNGEKUTTA, rungekutta.f90:12-208 */
void oracle_rungekutta(oracle *o, double dt)
{
    for (int stage = 1; stage <= 3; ++stage) oracle_stage(o, stage, dt);
}

int oracle_run(oracle *o, int max_steps, double total_time, double *t, double *dts)
{
    int n = 0;
    while (*t < total_time && n < max_steps && !o->status) {
        double dt = oracle_finddt(o);
        if (*t + dt > total_time) dt = total_time - *t;
        oracle_rungekutta(o, dt);
        *t = *t + dt;
        if (dts) dts[n] = dt;
        ++n;
    }
    return n;
}

/* FIND_DIVB, mhd_module.f90:133-189 */
double oracle_find_divb(oracle *o)
{
    double m = 0.0;
    long sy = o->NX, sz = (long)o->NX * o->NY;
    for (int l = 1; l <= o->nz; ++l)
        for (int k = 1; k <= o->ny; ++k)
            for (int j = 1; j <= o->nx; ++j) {
                long c = CI(o, j, k, l);
                double d;
                const double *bx = o->prim[IBX], *by = o->prim[IBY], *bz = o->prim[IBZ];
                if (o->d.coordinate_flag == 0)
                    d = (bx[c] - bx[c - 1]) / (A1(o->dx, j)) + (by[c] - by[c - sy]) / (A1(o->dy, k)) +
                        (bz[c] - bz[c - sz]) / (A1(o->dz, l));
                else if (o->d.coordinate_flag == 1)
                    d = (A1(o->xF, j) * bx[c] - A1(o->xF, j - 1) * bx[c - 1]) / (A1(o->x, j) * A1(o->dx, j)) +
                        (by[c] - by[c - sy]) / (A1(o->x, j) * A1(o->dy, k)) + (bz[c] - bz[c - sz]) / (A1(o->dz, l));
                else {
                    double a = A1(o->xF, j), b = A1(o->xF, j - 1), dcb = A1(o->dx_cb, j), dc = A1(o->dcose, k);
                    double xdx = A1(o->x, j) * A1(o->dx, j);
                    d = (a * a * bx[c] - b * b * bx[c - 1]) / (dcb / 3.0) +
                        (A1(o->sinf, k) * by[c] - A1(o->sinf, k - 1) * by[c - sy]) * xdx / (dcb * dc / 3.0) +
                        (bz[c] - bz[c - sz]) * (xdx * A1(o->dy, k)) / (dcb * dc * A1(o->dz, l) / 3.0);
                }
                m = fmax2(m, fabs(d));
            }
    o->maxdivb = m;
    return m;
}

int oracle_grid(const oracle *o, const char *name, double *out)
{
    const char *names[12] = {"dx", "dy", "dz", "x", "y", "z", "dx_cb", "xbar", "dcose", "dsine", "sine", "sinf"};
    const double *arr[12] = {o->dx, o->dy, o->dz, o->x, o->y, o->z, o->dx_cb, o->xbar, o->dcose, o->dsine,
                             o->sine, o->sinf};
    const int len[12] = {o->nx + 7, o->ny + 7, o->nz + 7, o->nx + 7, o->ny + 7, o->nz + 7, o->nx + 7, o->nx + 7,
                         o->ny + 7, o->ny + 7, o->ny + 7, o->ny + 7};
    for (int i = 0; i < 12; ++i)
        if (!strcmp(name, names[i])) {
            memcpy(out, arr[i], (size_t)len[i] * sizeof(double));
            return len[i];
        }
    return -1;
}

int oracle_get(oracle *o, const char *name, double *out)
{
    const double *src = NULL;
    static const char *pn[8] = {"rho", "vx", "vy", "vz", "p", "bx", "by", "bz"};
    for (int i = 0; i < 8; ++i) if (!strcmp(name, pn[i])) src = o->prim[i + 1];
    if (!strcmp(name, "eps")) src = o->eps;
    else if (!strcmp(name, "cs")) src = o->cs;
    else if (!strcmp(name, "bcx")) src = o->bcell[IBX];
    else if (!strcmp(name, "bcy")) src = o->bcell[IBY];
    else if (!strcmp(name, "bcz")) src = o->bcell[IBZ];
    else if (!strncmp(name, "u", 1) && name[1] >= '1' && name[1] <= '8' && !name[2]) src = o->cons[name[1] - '0'];
    else if (!strncmp(name, "lrk", 3) && name[3] >= '1' && name[3] <= '8' && !name[4]) src = o->l_rk[name[3] - '0'];
    else if (!strncmp(name, "flux", 4) && name[4] >= '1' && name[4] <= '8' && !name[5]) src = o->flux[name[4] - '0'];
    if (!src) return -1;
    memcpy(out, src, o->G * sizeof(double));
    return 0;
}
