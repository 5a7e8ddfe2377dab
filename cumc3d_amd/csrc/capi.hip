// Host side of libcumc3d_hip.so: the C-ABI declared in include/cumc3d_hip.h,
// step sequencing and the FP64 mode's kernels (the FP32 split's are in
// fp32.hip and fp32_state.hip).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cumc3d_hip.h"
#include "context.hpp"
#include "gravity.hpp"

using namespace cumc3d;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// A device temporary of one entry point, freed on every return path
// (hipFree waits for the work queued on it).
struct DevTmp {
    void *p = nullptr;
    DevTmp() = default;
    DevTmp(const DevTmp &) = delete;
    DevTmp &operator=(const DevTmp &) = delete;
    ~DevTmp() {
        if (p) hipFree(p);
    }
};

// Two timing events of one call, destroyed on every return path.
struct EvPair {
    hipEvent_t a = nullptr, b = nullptr;
    EvPair() = default;
    EvPair(const EvPair &) = delete;
    EvPair &operator=(const EvPair &) = delete;
    ~EvPair() {
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
    }
};

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            return fail(CUMC3D_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

// ---------------------------------------------------------------------------
// Allocation-free time loop (cumc3d_run_noalloc): its per-step dt log lives in
// this file-scope device array instead of a per-call hipMalloc.  Each GPU the
// library runs on holds its own copy until the process ends (512 KiB).  Host
// code never passes `scratch` itself to a kernel or a copy (on the host the
// name is only a stand-in for the symbol): cumc3d_init resolves the array's
// device address once per device, with that device current, and every use
// goes through that address.  All calls on one GPU share the array and it
// keeps the last call's contents; every slot a call reads back was written
// by that call, and calls that overlap on one GPU race on it.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[CUMC3D_NOALLOC_MAX_STEPS];
constexpr int MAX_DEVICES = 64;
static double *g_scratch[MAX_DEVICES];
static std::mutex g_scratch_mu;

// The current device's address of `scratch`, resolved on first use per device.
static int scratch_address(double **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(CUMC3D_EHIP, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CUMC3D_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return 0;
}

// dt = MIN(1e5, cell minimum), clipped to total_time (main.f90:101-106); a
// step that starts at or past total_time gets dt = 0 and is a no-op in time.
__global__ void k_dt_finalize(DevScalars *S) {
    S->active = S->t < S->total_time && !S->grav_stop;
    if (!S->active) {
        // keep the CFL minimum of the current state for a later continuation
        S->dt = 0.0;
        return;
    }
    double m = __longlong_as_double((long long)S->dtmin_bits);
    double dt = fmin2(1.0e5, m);
    if (S->t + dt > S->total_time) dt = S->total_time - S->t;
    S->dt = dt;
    S->dtmin_bits = (unsigned long long)__double_as_longlong(1.0e300);
}

__global__ void k_time_advance(DevScalars *S, double *dtlog, int slot) {
    if (S->active) {   // the step ran: t < total_time and no model STOP during it
        S->t = S->t + S->dt;
        S->nstep += 1;
    }
    if (dtlog) dtlog[slot] = S->dt;
}

__global__ void k_set_dt(DevScalars *S, double dt) {
    S->dt = dt;
    S->active = !S->grav_stop;
}

// a new initial model starts a run afresh: no model STOP pending
__global__ void k_clear_stop(DevScalars *S) { S->grav_stop = 0; }

// SPATIAL on its own runs the stage kernels whatever the run's end time
__global__ void k_activate(DevScalars *S) { S->active = !S->grav_stop; }

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
// SETUP_EQNS parity factors (setup.f90:102-326); returns false on the
// combinations the reference rejects with STOP.
static bool setup_bfac(const int *bf, int bfac[6][12]) {
    for (int s = 0; s < 6; ++s)
        for (int i = 0; i < 12; ++i) bfac[s][i] = 1;
    enum { IVX = 2, IVY = 3, IVZ = 4, IBX = 6, IBY = 7, IBZ = 8, IEX = 9, IEY = 10, IEZ = 11 };
    for (int side = 0; side < 2; ++side) {
        int *b = bfac[side], f = bf[side];
        if (f == 2) { b[IVX] = -1; b[IBX] = -1; }
        else if (f == 3) { b[IVX] = -1; b[IBX] = -1; b[IVZ] = -1; b[IBZ] = -1; }
        else if (f == 4) return false;
    }
    for (int side = 0; side < 2; ++side) {
        int *b = bfac[2 + side], f = bf[2 + side];
        if (f == 2) { b[IVY] = -1; b[IBY] = -1; }
        else if (f == 3) { b[IVY] = -1; b[IBY] = -1; b[IVZ] = -1; b[IBZ] = -1; }
        else if (f == 4) { b[IVY] = -1; b[IBX] = -1; b[IBZ] = -1; }
    }
    for (int side = 0; side < 2; ++side) {
        int *b = bfac[4 + side], f = bf[4 + side];
        if (f == 2) { b[IVZ] = -1; b[IBZ] = -1; }
        else if (f == 3) return false;
        else if (f == 4) { b[IVZ] = -1; b[IBX] = -1; b[IBY] = -1; }
    }
    for (int s = 0; s < 6; ++s) {
        int *b = bfac[s];
        if (b[IVZ] * b[IBY] * b[IVY] * b[IBZ] < 0) return false;
        b[IEX] = b[IVZ] * b[IBY];
        if (b[IVX] * b[IBZ] * b[IVZ] * b[IBX] < 0) return false;
        b[IEY] = b[IVX] * b[IBZ];
        if (b[IVY] * b[IBX] * b[IVX] * b[IBY] < 0) return false;
        b[IEZ] = b[IVY] * b[IBX];
    }
    return true;
}

// Weights of TVD_WEIGHT / PPM_INTERPOLANT / WENO_INTERPOLANT
// (tvd_module.f90:42-53, ppmc_module.f90:91-145, weno_module.f90:117-149)
static void weights(int recon, const std::vector<double> &dd, int n, std::vector<double> &w) {
    w.assign((size_t)(n + 7) * NW, 0.0);
    auto D = [&](int j) { return dd[j + 3]; };
    if (recon <= 2) {
        for (int j = -1; j <= n + 2; ++j) {
            w[(j + 3) * NW + 0] = (D(j) + D(j + 1)) / D(j);
            w[(j + 3) * NW + 1] = (D(j) + D(j - 1)) / D(j);
        }
    } else if (recon == 3) {
        for (int j = 0; j <= n + 1; ++j) {
            double dm2 = D(j - 2), dm1 = D(j - 1), dc = D(j), dp1 = D(j + 1), dp2 = D(j + 2);
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
            double *o = &w[(j + 3) * NW];
            o[0] = a1L; o[1] = aL * (z1L - zL) / deltaXL; o[2] = (-dm1 * z1L) / deltaXL; o[3] = dc * zL / deltaXL;
            o[4] = a1R; o[5] = aR * (z1R - zR) / deltaXR; o[6] = (-dc * z1R) / deltaXR; o[7] = dp1 * zR / deltaXR;
            o[8] = cm1 * cm2; o[9] = cm1 * cm3; o[10] = cc1 * cc2; o[11] = cc1 * cc3; o[12] = cp1 * cp2;
            o[13] = cp1 * cp3;
        }
    } else {
        for (int j = 0; j <= n + 1; ++j) {
            double dm2 = D(j - 2), dm1 = D(j - 1), dc = D(j), dp1 = D(j + 1), dp2 = D(j + 2);
            double a12 = dp2 / (dp1 + dp2), a01 = dp1 / (dc + dp1), a10 = dc / (dm1 + dc), a21 = dm1 / (dm2 + dm1);
            double *o = &w[(j + 3) * NW];
            o[0] = 1.0 - a01; o[1] = a01; o[2] = 1.0 - a12; o[3] = a12; o[4] = 1.0 - a10;
            o[5] = a10; o[6] = 1.0 - a21; o[7] = a21; o[8] = 1.0 - a21; o[9] = a12;
        }
    }
}

// class masks from the stage-1 state of a step (kept for stages 2 and 3)
static void sel_refresh(cumc3d_ctx *x) {
    if (!x->sel_on) return;
    if (x->sel_tick++ % x->sel_every != 0) return;
    tmark(x, 7);
    if (x->fp32) f32::sel_infer(x);
    else sel_infer_t<double>(x);
    tmark(x, 8);
    if (x->timing && x->tstage < TSTAGES) x->sel_timed[x->tstage] = 1;
}

// ---------------------------------------------------------------------------
// collapsar self-gravity (model/collapsar/src/custom.f90:416-658)
// ---------------------------------------------------------------------------
static SorArgs sor_args(cumc3d_ctx *x) {
    SorArgs a;
    memset(&a, 0, sizeof(a));
    a.nx = x->nx; a.ny = x->ny; a.nz = x->nz;
    a.NX = x->NX; a.NY = x->NY;
    a.rin = x->sor_phi[0][0];
    a.bin = x->sor_phi[0][1];
    a.rout = x->sor_phi[1][0];
    a.bout = x->sor_phi[1][1];
    a.rho_in = x->sor_rho;
    a.a2 = x->sor_a2; a.e2 = x->sor_e2; a.zpm = x->sor_z[0]; a.zpc = x->sor_z[1]; a.zcm = x->sor_z[2];
    a.ajp1 = x->ajp1; a.ajm1 = x->ajm1; a.bkp1 = x->bkp1; a.bkm1 = x->bkm1;
    a.part_nan = x->part_nan;
    a.part_max = x->part_max;
    a.nblk = x->sor_nblk;
    a.relax_max = x->grav_p.relax_max;
    a.pa = x->grav_p.prim_a[0];
    a.omega = x->grav_p.omega;
    a.tol = x->grav_p.tolerance;
    a.pi4 = 4.0 * (4.0 * std::atan(1.0));   // 4 pi, pi = 4 atan(1) (parameter.h:21)
    a.S = x->S;
    return a;
}

// The RBSOR loop (custom.f90:510-647): phi moved into the colour-split
// arrays, sweep pairs queued in growing batches with the end flag read
// between batches, and the arrays the last pair wrote moved back into x->phi
// with its boundary values set.  A solve that converged on sweep relax_max sets
// grav_error (the reference STOPs there); one that ran out continues.
static int sor_solve(cumc3d_ctx *x, bool zero_guess) {
    hipStream_t st = x->stream;
    EvPair ev;
    if (x->timing) {
        HIPCHK(hipEventCreate(&ev.a));
        HIPCHK(hipEventCreate(&ev.b));
        HIPCHK(hipEventRecord(ev.a, st));
    }
    if (zero_guess) hipLaunchKernelGGL(k_fill0, dim3(nblk(x->G)), dim3(256), 0, st, x->phi, x->G);
    hipLaunchKernelGGL(k_sor_begin, dim3(1), dim3(1), 0, st, x->S);
    SorArgs a = sor_args(x);
    hipLaunchKernelGGL(k_sor_rho, dim3(nblk((long)x->nx * x->ny), x->nz), dim3(256), 0, st, a,
                       (const double *)x->P[0], x->sor_rho);
    const long hx = (x->nx + 1) / 2;
    // red: one more block column, whose block (0, 0) decides the previous pair
    const dim3 gr(nblk(hx * x->ny) + 1, x->nz), gb(nblk(hx * x->ny, SOR_SLOTS), x->nz);
    const dim3 gc(nblk((long)x->nx * x->ny), x->nz);
    hipLaunchKernelGGL(k_sor_copy, gc, dim3(256), 0, st, a, x->phi, x->sor_phi[0][0], x->sor_phi[0][1], 1);
    long queued = 0;
    int batch = 4, flags[3] = {0, 0, 0};
    for (;;) {
        for (int i = 0; i < batch; ++i) {
            a.rin = x->sor_phi[queued & 1][0];
            a.bin = x->sor_phi[queued & 1][1];
            a.rout = x->sor_phi[(queued + 1) & 1][0];
            a.bout = x->sor_phi[(queued + 1) & 1][1];
            a.zero_ghosts = zero_guess && queued == 0;
            a.combine = i > 0;
            hipLaunchKernelGGL(k_sor_red, gr, dim3(256), 0, st, a);
            hipLaunchKernelGGL(k_sor_black, gb, dim3(256), 0, st, a);
            ++queued;
        }
        hipLaunchKernelGGL(k_sor_decide, dim3(1), dim3(256), 0, st, a);
        if (ev.b) HIPCHK(hipEventRecord(ev.b, st));
        HIPCHK(hipMemcpyAsync(flags, &x->S->sor_state, sizeof(flags), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (flags[0]) break;
        if (queued > (long)x->grav_p.relax_max + 1)   // the device caps the sweeps at relax_max
            return fail(CUMC3D_EHIP, "Poisson solve: sweep count not reported by the device");
        if (batch < 256) batch *= 2;
    }
    if (flags[1] > 0) {   // the arrays the last pair wrote
        const int f = flags[1] & 1;
        hipLaunchKernelGGL(k_sor_copy, gc, dim3(256), 0, st, a, x->phi, x->sor_phi[f][0], x->sor_phi[f][1], 0);
        hipLaunchKernelGGL(k_sor_bc, dim3(nblk((long)x->ny * x->nz + (long)x->nx * x->nz + (long)x->nx * x->ny)),
                           dim3(256), 0, st, x->phi, x->nx, x->ny, x->nz, x->NX, x->NY);
    }
    if (flags[2] == x->grav_p.relax_max) x->grav_error = 1;
    if (ev.a) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev.a, ev.b) == hipSuccess && flags[1] > 0) {
            x->tsum[8] += ms;
            x->tcnt[8] += 1;
            x->tsum[9] += flags[1];
            x->tcnt[9] += 1;
        }
    }
    return 0;
}

// UPDATE(p)'s CUSTOM_UPDATE after the stage's update and ghost fill
static int grav_stage(cumc3d_ctx *x, int stage) {
    if (x->d.model != CUMC3D_MODEL_COLLAPSAR) return 0;
    const ArgsT<double> a = make_args<double>(x);
    const double c = x->grav_p.c_code;
    // mdot on the side stream from the end of this stage's x sweep, beside
    // the y/z sweeps, EMFs, update and ghost fill; m_bh after both
    HIPCHK(hipStreamWaitEvent(x->side, x->ev_x, 0));
    hipLaunchKernelGGL(k_mdot, dim3(1), dim3(256), 0, x->side, a);
    HIPCHK(hipEventRecord(x->ev_m, x->side));
    HIPCHK(hipStreamWaitEvent(x->stream, x->ev_m, 0));
    if (stage == 1) hipLaunchKernelGGL(k_mbh<1>, dim3(1), dim3(1), 0, x->stream, x->S, c, x->xF0);
    else if (stage == 2) hipLaunchKernelGGL(k_mbh<2>, dim3(1), dim3(1), 0, x->stream, x->S, c, x->xF0);
    else hipLaunchKernelGGL(k_mbh<3>, dim3(1), dim3(1), 0, x->stream, x->S, c, x->xF0);
    if (stage == 3 && x->cur_step % x->grav_p.n_pot == 0) return sor_solve(x, false);
    return 0;
}

// after a sync: the reference's STOPs as errors
static int grav_check(cumc3d_ctx *x) {
    if (x->d.model != CUMC3D_MODEL_COLLAPSAR) return 0;
    int stop = 0;
    HIPCHK(hipMemcpy(&stop, &x->S->grav_stop, sizeof(stop), hipMemcpyDeviceToHost));
    if (stop) return fail(CUMC3D_ESTOP, "Black hole event horizon exceeds computational grid");
    if (x->grav_error) return fail(CUMC3D_ESTOP, "Convergence error in poisson solver");
    return 0;
}

static int run_stage(cumc3d_ctx *x, int stage) {
    if (stage == 1) sel_refresh(x);
    if (x->fp32) f32::run_stage(x, stage);
    else run_stage_t<double>(x, stage);
    return grav_stage(x, stage);
}

// per-kernel-class device times from the recorded events (after a sync)
static void accumulate_timing(cumc3d_ctx *x) {
    int dims = x->d.n_dim;
    for (int st = 0; st < x->tstage; ++st) {
        hipEvent_t *e = &x->ev[st * TMARKS];
        for (int s = 0; s < 6; ++s) {
            if ((s == 1 && dims < 2) || (s == 2 && dims < 3)) continue;
            float ms = 0;
            if (hipEventElapsedTime(&ms, e[s], e[s + 1]) != hipSuccess) continue;
            x->tsum[s] += ms;
            x->tcnt[s] += 1;
        }
        float ms = 0;
        if (x->sel_timed[st] && hipEventElapsedTime(&ms, e[7], e[8]) == hipSuccess) {
            x->tsum[7] += ms;
            x->tcnt[7] += 1;
        }
    }
}

static int launch_finddt(cumc3d_ctx *x) {
    const long nn = (long)x->nx * x->ny * x->nz;
    if (x->fp32) f32::finddt(x);
    else hipLaunchKernelGGL(k_finddt<double>, dim3(nblk(nn)), dim3(256), 0, x->stream, make_args<double>(x));
    return 0;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *cumc3d_error_string(int code) {
    switch (code) {
        case CUMC3D_OK: return "ok";
        case CUMC3D_EINVAL: return "invalid argument";
        case CUMC3D_EHIP: return "HIP runtime error";
        case CUMC3D_ENOMEM: return "device allocation failed";
        case CUMC3D_ESTATE: return "call out of order";
        case CUMC3D_ENONFINITE: return "non-finite or non-positive time step";
        case CUMC3D_ESTOP: return "the model stopped the run";
        default: return "unknown error";
    }
}

const char *cumc3d_last_error(void) { return g_err.c_str(); }

int cumc3d_init(const cumc3d_desc *desc, cumc3d_ctx **out) {
    g_err.clear();
    if (!desc || !out) return fail(CUMC3D_EINVAL, "null descriptor or output pointer");
    *out = nullptr;
    const cumc3d_desc &d = *desc;
    if (d.nx < 1 || d.ny < 1 || d.nz < 1) return fail(CUMC3D_EINVAL, "grid sizes must be >= 1");
    if (d.n_dim < 1 || d.n_dim > 3) return fail(CUMC3D_EINVAL, "n_dim must be 1, 2 or 3");
    if (d.n_dim == 1 && (d.ny != 1 || d.nz != 1)) return fail(CUMC3D_EINVAL, "n_dim=1 needs ny=nz=1");
    if (d.n_dim == 2 && d.nz != 1) return fail(CUMC3D_EINVAL, "n_dim=2 needs nz=1");
    if (d.coordinate_flag < 0 || d.coordinate_flag > 2)
        return fail(CUMC3D_EINVAL, "coordinate_flag must be 0 (Cartesian), 1 (cylindrical) or 2 (spherical)");
    if (d.riemann < 0 || d.riemann > 3) return fail(CUMC3D_EINVAL, "unknown Riemann solver");
    if (d.recon < 0 || d.recon > 4) return fail(CUMC3D_EINVAL, "unknown reconstruction");
    if (d.model < 0 || d.model > 3) return fail(CUMC3D_EINVAL, "unknown model");
    if (d.model == CUMC3D_MODEL_COLLAPSAR) {
        if (d.coordinate_flag != 2) return fail(CUMC3D_EINVAL, "the collapsar model needs spherical coordinates");
        if (!d.gravity) return fail(CUMC3D_EINVAL, "the collapsar model needs cumc3d_desc.gravity");
        if (d.precision != CUMC3D_FP64) return fail(CUMC3D_EINVAL, "the collapsar model runs in FP64 only");
        if (d.gravity->relax_max < 1 || d.gravity->n_pot < 1)
            return fail(CUMC3D_EINVAL, "gravity relax_max and n_pot must be positive");
    }
    if (d.model == CUMC3D_MODEL_TORUS && d.coordinate_flag != 1)
        return fail(CUMC3D_EINVAL, "the torus-cyn model needs cylindrical coordinates");
    if (d.model == CUMC3D_MODEL_TORUS_SPH && d.coordinate_flag != 2)
        return fail(CUMC3D_EINVAL, "the torus-sph model needs spherical coordinates");
    if (d.metrics && (!d.metrics->dx || !d.metrics->dy || !d.metrics->dz || !d.metrics->x || !d.metrics->y ||
                      !d.metrics->z))
        return fail(CUMC3D_EINVAL, "grid metrics need dx, dy, dz, x, y and z");
    if (d.metrics && d.coordinate_flag == 2 &&
        (!d.metrics->dx_cb || !d.metrics->xbar || !d.metrics->dcose || !d.metrics->dsine || !d.metrics->sine ||
         !d.metrics->sinf))
        return fail(CUMC3D_EINVAL, "spherical grid metrics need dx_cb, xbar, dcose, dsine, sine and sinf");
    if (d.hlld_bn_mode < 0 || d.hlld_bn_mode > 2) return fail(CUMC3D_EINVAL, "unknown hlld_bn_mode");
    if (d.precision != CUMC3D_FP64 && d.precision != CUMC3D_FP32) return fail(CUMC3D_EINVAL, "unknown precision");
    if (!d.xF || !d.yF || !d.zF) return fail(CUMC3D_EINVAL, "face coordinate arrays are required");
    for (int i = 0; i < 6; ++i)
        if (d.boundary_flag[i] < 0 || d.boundary_flag[i] > 4) return fail(CUMC3D_EINVAL, "boundary flag out of range");
    int nmin = d.recon >= 3 ? 2 : 1;
    int nd3[3] = {d.nx, d.ny, d.nz};
    for (int a = 0; a < d.n_dim; ++a)
        for (int s = 0; s < 2; ++s)
            if (d.boundary_flag[2 * a + s] != 1 && nd3[a] < 3)
                return fail(CUMC3D_EINVAL, "periodic or mirror boundaries need at least 3 cells on that axis");
    (void)nmin;

    if ((long)(d.nx + 6) * (d.ny + 6) * (d.nz + 6) * 8 >= (1L << 32))
        return fail(CUMC3D_EINVAL, "grid too large: each field must stay below 4 GiB (about 800^3 cells)");
    cumc3d_ctx *x = new cumc3d_ctx();
    x->d = d;
    x->d.metrics = nullptr;   // caller-owned: read during init only
    x->d.gravity = nullptr;
    if (d.model == CUMC3D_MODEL_COLLAPSAR) x->grav_p = *d.gravity;
    x->xF0 = d.xF[3];
    x->nx = d.nx; x->ny = d.ny; x->nz = d.nz;
    x->NX = d.nx + 6; x->NY = d.ny + 6; x->NZ = d.nz + 6;
    x->G = x->NX * x->NY * x->NZ;
    if (!setup_bfac(d.boundary_flag, x->bfac)) {
        delete x;
        return fail(CUMC3D_EINVAL, "boundary flag combination rejected by SETUP_EQNS (setup.f90:120-322)");
    }
    // grid metrics: GetGrid (getgrid.f90:43-98) from the face coordinates,
    // or the caller's GETGRID arrays
    const double *Fc[3] = {d.xF, d.yF, d.zF};
    const cumc3d_grid_metrics *gm = d.metrics;
    const double *Dm[3] = {gm ? gm->dx : nullptr, gm ? gm->dy : nullptr, gm ? gm->dz : nullptr};
    const double *Cm[3] = {gm ? gm->x : nullptr, gm ? gm->y : nullptr, gm ? gm->z : nullptr};
    std::vector<double> hm;
    long off[3][7];
    std::vector<double> widths[3], centres[3];
    for (int a = 0; a < 3; ++a) {
        int n = nd3[a];
        std::vector<double> f(Fc[a], Fc[a] + n + 7), dd(n + 7, 0.0), cc(n + 7, 0.0), w;
        for (int j = -2; j <= n + 3; ++j) {
            dd[j + 3] = Dm[a] ? Dm[a][j + 2] : f[j + 3] - f[j + 2];
            cc[j + 3] = Cm[a] ? Cm[a][j + 2] : (f[j + 3] + f[j + 2]) / 2;
        }
        widths[a] = dd;
        centres[a] = cc;
        weights(d.recon, dd, n, w);
        off[a][0] = (long)hm.size(); hm.insert(hm.end(), f.begin(), f.end());
        off[a][1] = (long)hm.size(); hm.insert(hm.end(), dd.begin(), dd.end());
        off[a][2] = (long)hm.size(); hm.insert(hm.end(), cc.begin(), cc.end());
        off[a][3] = (long)hm.size(); hm.insert(hm.end(), w.begin(), w.end());
        // the selector's three schemes: TVD-MC, PPM, WENO-Z
        for (int sch = 0; sch < 3; ++sch) {
            weights(sch == 0 ? 2 : (sch == 1 ? 3 : 4), dd, n, w);
            off[a][4 + sch] = (long)hm.size(); hm.insert(hm.end(), w.begin(), w.end());
        }
    }
    // spherical metrics (getgrid.f90:68-98) and the per-index factors the
    // kernels multiply by, each formed exactly as the reference's expression
    // forms it (spatial.f90:227-256,467-581; fromrvetou.f90:181-188)
    long soff = -1;
    if (d.coordinate_flag == 2) {
        const int nx = d.nx, ny = d.ny;
        const std::vector<double> xF(d.xF, d.xF + nx + 7), yF(d.yF, d.yF + ny + 7);
        const std::vector<double> &dx = widths[0], &xc = centres[0], &dy = widths[1], &yc = centres[1];
        std::vector<double> t[10];
        for (int i = 0; i < 5; ++i) t[i].assign(nx + 7, 0.0);
        for (int i = 5; i < 10; ++i) t[i].assign(ny + 7, 0.0);
        for (int j = -1; j <= nx + 3; ++j) {
            double a = xF[j + 3], b = xF[j + 2];
            double dcb = gm ? gm->dx_cb[j + 2] : a * a * a - b * b * b;
            double xb = gm ? gm->xbar[j + 2] : (dcb / 3.0) / (xc[j + 3] * dx[j + 3]);
            t[0][j + 3] = dcb;
            t[1][j + 3] = dcb / 3.0;
            t[2][j + 3] = xb;
            t[3][j + 3] = 3.0 * xc[j + 3] * dx[j + 3] / dcb;
            t[4][j + 3] = 1.5 * dx[j + 3] / dcb;
        }
        for (int k = -3; k <= ny + 3; ++k) t[5][k + 3] = gm ? gm->sinf[k + 3] : std::sin(yF[k + 3]);
        if (!gm && d.fixpole) {
            t[5][3] = 0.0;
            t[5][ny + 3] = 0.0;
        }
        for (int k = -1; k <= ny + 3; ++k) {
            double dc = gm ? gm->dcose[k + 2] : -(std::cos(yF[k + 3]) - std::cos(yF[k + 2]));
            double ds = gm ? gm->dsine[k + 2] : std::sin(yF[k + 3]) - std::sin(yF[k + 2]);
            t[6][k + 3] = dc;
            t[7][k + 3] = ds / dc;
            t[8][k + 3] = 0.5 * dy[k + 3] / dc;
        }
        for (int k = -2; k <= ny + 3; ++k) t[9][k + 3] = gm ? gm->sine[k + 2] : std::sin(yc[k + 3]);
        soff = (long)hm.size();
        for (int i = 0; i < 10; ++i) hm.insert(hm.end(), t[i].begin(), t[i].end());
    }
    // get_poisson / poisson_coef (model/collapsar/src/poisson.f90:6-64) from
    // the cell centres, in its operation order; libm tan and sin
    // ajp1, ajm1, bkp1, bkm1 as arrays; clp1, clm1 and epsc as the (j, k)
    // factor 2/xc^2/sin^2, epsc's radial + polar terms and the phi spacings,
    // from which the kernels form them with poisson_coef's divisions
    // (gravity.hpp); then room for rho_in, split by colour
    std::vector<double> pc;
    long pco[10] = {};
    if (d.model == CUMC3D_MODEL_COLLAPSAR) {
        const int nx = d.nx, ny = d.ny, nz = d.nz;
        const long nb = (long)(nx + 1) * (ny + 1), hx = (nx + 1) / 2;
        const long len[9] = {nx + 1, nx + 1, nb, nb, nb, nb, nz + 1, nz + 1, nz + 1};
        for (int i = 0; i < 9; ++i) {
            pco[i] = (long)pc.size();
            pc.resize(pc.size() + len[i], 0.0);
        }
        pco[9] = (long)pc.size();
        pc.resize(pc.size() + 2 * hx * ny * nz, 0.0);
        double *ajp1 = &pc[pco[0]], *ajm1 = &pc[pco[1]], *bkp1 = &pc[pco[2]], *bkm1 = &pc[pco[3]];
        double *a2 = &pc[pco[4]], *e2 = &pc[pco[5]], *zpm = &pc[pco[6]], *zpc = &pc[pco[7]], *zcm = &pc[pco[8]];
        const std::vector<double> &X = centres[0], &Y = centres[1], &Z = centres[2];
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny; ++j) {
                const double xm1 = X[i + 2], xc = X[i + 3], xp1 = X[i + 4];
                const double ym1 = Y[j + 2], yc = Y[j + 3], yp1 = Y[j + 4];
                const double t = std::tan(yc), sn = std::sin(yc);
                const long jk = (long)j * (nx + 1) + i;
                e2[jk] = 2.0 * (xp1 + xm1 - 3.0 * xc) / xc / (xp1 - xc) / (xc - xm1) +
                         (yp1 + ym1 - 2.0 * yc - 2.0 * t) / (xc * xc) / t / (yp1 - yc) / (yc - ym1);
                a2[jk] = 2.0 / (xc * xc) / (sn * sn);
                ajp1[i] = 2.0 * (2.0 * xc - xm1) / (xp1 - xm1) / (xp1 - xc) / xc;
                ajm1[i] = 2.0 * (2.0 * xc - xp1) / (xp1 - xm1) / (xc - xm1) / xc;
                bkp1[jk] = (2.0 * t + yc - ym1) / (xc * xc) / t / (yp1 - yc) / (yp1 - ym1);
                bkm1[jk] = (2.0 * t + yc - yp1) / (xc * xc) / t / (yp1 - ym1) / (yc - ym1);
            }
        for (int k = 1; k <= nz; ++k) {
            const double zm1 = Z[k + 2], zc = Z[k + 3], zp1 = Z[k + 4];
            zpm[k] = zp1 - zm1;
            zpc[k] = zp1 - zc;
            zcm[k] = zc - zm1;
        }
    }
    int rc = 0;
    auto cleanup = [&]() { cumc3d_finalize(x); };
    if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess) {
        delete x;
        return fail(CUMC3D_EHIP, "hipStreamCreate failed");
    }
    if (int e = scratch_address(&x->scratch)) { cleanup(); return e; }
    if (hipMalloc(&x->met, hm.size() * sizeof(double)) != hipSuccess) { cleanup(); return fail(CUMC3D_ENOMEM, "metrics"); }
    if (hipMemcpy(x->met, hm.data(), hm.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return fail(CUMC3D_EHIP, "metric upload");
    }
    for (int a = 0; a < 3; ++a) {
        x->ax[a].xF = x->met + off[a][0];
        x->ax[a].d = x->met + off[a][1];
        x->ax[a].c = x->met + off[a][2];
        x->ax[a].w = x->met + off[a][3];
        for (int sch = 0; sch < 3; ++sch) x->ax[a].wsel[sch] = x->met + off[a][4 + sch];
    }
    if (soff >= 0) {
        const double *b = x->met + soff;
        const long nx7 = d.nx + 7, ny7 = d.ny + 7;
        for (int i = 0; i < 5; ++i) x->sph[i] = b + i * nx7;
        for (int i = 5; i < 10; ++i) x->sph[i] = b + 5 * nx7 + (i - 5) * ny7;
    }
    // per-cell fields: the working-type block, then the double block, each
    // field G entries, fields 256-byte aligned
    x->fp32 = d.precision == CUMC3D_FP32;
    const size_t wsz = x->fp32 ? sizeof(float) : sizeof(double);
    const size_t wfield = (x->G * wsz + 255) / 256 * 256, dfield = (x->G * sizeof(double) + 255) / 256 * 256;
    const int nw = 5 + 2 + 3 + 21 + 3 + 3, nd = 6 + 8 + (x->fp32 ? 5 : 0) + 3;
    size_t bytes = nw * wfield + nd * dfield;
    if (hipMalloc(&x->pool, bytes) != hipSuccess) {
        cleanup();
        return fail(CUMC3D_ENOMEM, "state allocation of " + std::to_string(bytes) + " bytes failed");
    }
    if (hipMemset(x->pool, 0, bytes) != hipSuccess) { cleanup(); return fail(CUMC3D_EHIP, "memset"); }
    char *p = (char *)x->pool;
    auto takew = [&]() { void *r = p; p += wfield; return r; };
    auto taked = [&]() { double *r = (double *)p; p += dfield; return r; };
    for (int i = 0; i < 5; ++i) x->P[i] = takew();
    x->eps = takew();
    x->cs = takew();
    for (int i = 0; i < 3; ++i) x->Bc[i] = takew();
    for (int dd = 0; dd < 3; ++dd)
        for (int i = 0; i < 7; ++i) x->F[dd][i] = takew();
    for (int i = 0; i < 3; ++i) x->E[i] = takew();
    for (int i = 0; i < 3; ++i) x->Ec[i] = takew();
    for (int b = 0; b < 2; ++b)
        for (int i = 0; i < 3; ++i) x->B[b][i] = taked();
    for (int i = 0; i < 8; ++i) x->Uo[i] = taked();
    for (int i = 0; i < 5; ++i) x->U[i] = x->fp32 ? taked() : nullptr;
    for (int i = 0; i < 3; ++i) x->floors[i] = taked();
    if (hipMalloc(&x->S, sizeof(DevScalars)) != hipSuccess) { cleanup(); return fail(CUMC3D_ENOMEM, "scalars"); }
    DevScalars s0;
    memset(&s0, 0, sizeof(s0));
    s0.total_time = 1.0e300;
    s0.active = 1;
    s0.dtmin_bits = (unsigned long long)0x7E37E43C8800759CULL;   // bits of 1e300
    s0.m_bh = s0.mbh_old = x->grav_p.m_bh;
    if (hipMemcpy(x->S, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess) { cleanup(); return fail(CUMC3D_EHIP, "scalars"); }
    if (const char *mt = getenv("CUMC3D_MARCH_THREADS")) {
        long v = atol(mt);
        if (v >= 1) x->march_threads = v;
    }
    if (d.riemann == CUMC3D_HLL) {
        if (hipMalloc(&x->stale, (size_t)3 * x->G) != hipSuccess) { cleanup(); return fail(CUMC3D_ENOMEM, "stale flags"); }
        if (hipMemset(x->stale, 0, (size_t)3 * x->G) != hipSuccess) { cleanup(); return fail(CUMC3D_EHIP, "memset"); }
    }
    if (d.model == CUMC3D_MODEL_COLLAPSAR) {
        if (hipStreamCreateWithFlags(&x->side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&x->ev_x, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&x->ev_m, hipEventDisableTiming) != hipSuccess) {
            cleanup();
            return fail(CUMC3D_EHIP, "accretion stream");
        }
        // phi and its second buffer (G each), the coefficients, the error blocks' partials
        x->sor_nblk = (int)nblk((long)((d.nx + 1) / 2) * d.ny, SOR_SLOTS) * d.nz;
        const size_t Gp = ((size_t)x->G + 3) / 4 * 4;
        const size_t Hp = ((size_t)((d.nx + 1) / 2) * d.ny * d.nz + 3) / 4 * 4;   // one colour of phi
        const size_t bytes = (Gp + 4 * Hp + pc.size() + 2 * (size_t)x->sor_nblk) * sizeof(double);
        if (hipMalloc(&x->grav, bytes) != hipSuccess) { cleanup(); return fail(CUMC3D_ENOMEM, "gravity arrays"); }
        if (hipMemset(x->grav, 0, bytes) != hipSuccess) { cleanup(); return fail(CUMC3D_EHIP, "memset"); }
        double *g = (double *)x->grav;
        x->phi = g;
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) x->sor_phi[b][c] = g + Gp + (2 * b + c) * Hp;
        double *c0 = g + Gp + 4 * Hp;
        if (hipMemcpy(c0, pc.data(), pc.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            cleanup();
            return fail(CUMC3D_EHIP, "poisson coefficient upload");
        }
        x->ajp1 = c0 + pco[0]; x->ajm1 = c0 + pco[1]; x->bkp1 = c0 + pco[2]; x->bkm1 = c0 + pco[3];
        x->sor_a2 = c0 + pco[4]; x->sor_e2 = c0 + pco[5];
        for (int i = 0; i < 3; ++i) x->sor_z[i] = c0 + pco[6 + i];
        x->sor_rho = c0 + pco[9];
        x->part_max = c0 + pc.size();
        x->part_nan = (long long *)(x->part_max + x->sor_nblk);
    }
    *out = x;
    return rc;
}

int cumc3d_upload(cumc3d_ctx *x, const double *prim, const double *eps, const double *floors3) {
    if (!x || !prim || !eps) return fail(CUMC3D_EINVAL, "null argument");
    if ((x->d.model == CUMC3D_MODEL_TORUS || x->d.model == CUMC3D_MODEL_TORUS_SPH) && !floors3)
        return fail(CUMC3D_EINVAL, "torus models need floor arrays");
    DevTmp buf;
    HIPCHK(hipMalloc(&buf.p, (size_t)9 * x->G * sizeof(double)));
    double *tmp = (double *)buf.p;
    HIPCHK(hipMemcpyAsync(tmp, prim, (size_t)8 * x->G * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(tmp + (size_t)8 * x->G, eps, x->G * sizeof(double), hipMemcpyHostToDevice, x->stream));
    x->cur = 0;
    if (x->fp32)
        f32::split_prim(x, tmp, tmp + (size_t)8 * x->G);
    else
        hipLaunchKernelGGL(k_split_prim<double>, dim3(nblk(x->G)), dim3(256), 0, x->stream, tmp, tmp + (size_t)8 * x->G,
                           make_args<double>(x));
    if (floors3)
        for (int i = 0; i < 3; ++i)
            HIPCHK(hipMemcpyAsync(x->floors[i], floors3 + (size_t)i * x->G, x->G * sizeof(double),
                                  hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    x->uploaded = true;
    x->initialized = false;
    return 0;
}

int cumc3d_initial_update(cumc3d_ctx *x) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    if (!x->uploaded) return fail(CUMC3D_ESTATE, "upload an initial model first");
    hipLaunchKernelGGL(k_clear_stop, dim3(1), dim3(1), 0, x->stream, x->S);
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, x->stream, x->S, 0.0);   // also re-activates the stage kernels
    const unsigned gi = nblk((long)x->nx * x->ny * x->nz), gg = nblk(x->G);
    if (x->fp32) {
        f32::initial_update(x);
    } else {
        hipLaunchKernelGGL(k_init_interior<double>, dim3(gi), dim3(256), 0, x->stream, make_args<double>(x));
        hipLaunchKernelGGL(k_ghost<double>, dim3(nblk(ghost_count(*x))), dim3(256), 0, x->stream, make_args<double>(x));
    }
    x->nstep_total = 0;
    x->grav_error = 0;
    if (x->d.model == CUMC3D_MODEL_COLLAPSAR) {   // UPDATE(0): CUSTOM_UPDATE(0)
        hipLaunchKernelGGL(k_grav_start, dim3(1), dim3(1), 0, x->stream, x->S, x->grav_p.m_bh, x->grav_p.c_code,
                           x->xF0);
        int rc = sor_solve(x, true);
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    if (int rc = grav_check(x)) return rc;
    x->initialized = true;
    x->dt_pending = false;
    x->next_stage = 1;
    return 0;
}

int cumc3d_finddt(cumc3d_ctx *x, double *dt) {
    if (!x || !dt) return fail(CUMC3D_EINVAL, "null argument");
    if (!x->initialized) return fail(CUMC3D_ESTATE, "call cumc3d_initial_update first");
    if (!x->dt_pending) launch_finddt(x);
    x->dt_pending = false;
    unsigned long long bits = 0;
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bits, &x->S->dtmin_bits, sizeof(bits), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(&bad, &x->S->nonfinite, sizeof(bad), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    double m;
    memcpy(&m, &bits, sizeof(m));
    *dt = m < 1.0e5 ? m : 1.0e5;
    unsigned long long reset = 0x7E37E43C8800759CULL;
    const int clear = 0;
    HIPCHK(hipMemcpyAsync(&x->S->dtmin_bits, &reset, sizeof(reset), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(&x->S->nonfinite, &clear, sizeof(clear), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (bad)
        return fail(CUMC3D_ENONFINITE,
                    "a cell's CFL time step was NaN or non-positive (NaN or negative energy in the state)");
    if (!(*dt > 0.0) || !std::isfinite(*dt)) return fail(CUMC3D_ENONFINITE, "FINDDT produced a non-positive dt");
    return 0;
}

int cumc3d_rk3_step(cumc3d_ctx *x, double dt) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    if (!x->initialized) return fail(CUMC3D_ESTATE, "call cumc3d_initial_update first");
    if (x->next_stage != 1) return fail(CUMC3D_ESTATE, "a step run by cumc3d_stage is not finished");
    if (int rc = grav_check(x)) return rc;
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, x->stream, x->S, dt);
    x->cur_step = x->nstep_total;
    for (int s = 1; s <= 3; ++s)
        if (int rc = run_stage(x, s)) return rc;
    x->nstep_total += 1;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    if (int rc = grav_check(x)) return rc;
    // stage 3 reduced the CFL minimum of the new state (FINDDT of the next step)
    x->dt_pending = true;
    return 0;
}

int cumc3d_stage(cumc3d_ctx *x, int stage, double dt) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    if (!x->initialized) return fail(CUMC3D_ESTATE, "call cumc3d_initial_update first");
    if (stage < 1 || stage > 3) return fail(CUMC3D_EINVAL, "stage must be 1, 2 or 3");
    if (stage != x->next_stage) return fail(CUMC3D_ESTATE, "the stages of a step run in the order 1, 2, 3");
    if (int rc = grav_check(x)) return rc;
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, x->stream, x->S, dt);
    if (stage == 1) {
        x->cur_step = x->nstep_total;
        x->dt_pending = false;
    }
    if (int rc = run_stage(x, stage)) return rc;
    x->next_stage = stage % 3 + 1;
    if (stage == 3) {
        x->nstep_total += 1;
        x->dt_pending = true;   // stage 3 reduced the CFL minimum of the new state
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    return grav_check(x);
}

int cumc3d_spatial(cumc3d_ctx *x, double *l_rk) {
    if (!x || !l_rk) return fail(CUMC3D_EINVAL, "null argument");
    if (!x->initialized) return fail(CUMC3D_ESTATE, "call cumc3d_initial_update first");
    if (int rc = grav_check(x)) return rc;
    const size_t bytes = (size_t)8 * x->G * sizeof(double);
    DevTmp buf;
    HIPCHK(hipMalloc(&buf.p, bytes));
    double *d = (double *)buf.p;
    // the entries SPATIAL does not form keep the caller's values
    HIPCHK(hipMemcpyAsync(d, l_rk, bytes, hipMemcpyHostToDevice, x->stream));
    hipLaunchKernelGGL(k_activate, dim3(1), dim3(1), 0, x->stream, x->S);
    if (x->fp32) f32::spatial(x, d);
    else spatial_t<double>(x, d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(l_rk, d, bytes, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

// The time loop of cumc3d_run and cumc3d_run_noalloc: dlog (device, at
// least max_steps doubles, or null when the caller wants no dt log) receives
// each step's dt.
static int run_loop(cumc3d_ctx *x, int max_steps, double total_time, int sync_each_step, double *t, double *dts,
                    int *steps_done, double *dlog) {
    // set t and total_time
    DevScalars h;
    HIPCHK(hipMemcpy(&h, x->S, sizeof(h), hipMemcpyDeviceToHost));
    h.t = *t;
    h.total_time = total_time;
    h.nstep = 0;
    HIPCHK(hipMemcpy(x->S, &h, sizeof(h), hipMemcpyHostToDevice));
    if (x->timing) {
        if (x->ev.empty()) {
            x->ev.resize((size_t)TSTAGES * TMARKS);
            x->sel_timed.assign(TSTAGES, 0);
            for (auto &e : x->ev) HIPCHK(hipEventCreate(&e));
        }
        for (int i = 0; i < 10; ++i) { x->tsum[i] = 0; x->tcnt[i] = 0; }
        x->sel_timed.assign(TSTAGES, 0);
        x->tstage = 0;
    }
    EvPair ev;
    if (x->timing) {
        HIPCHK(hipEventCreate(&ev.a));
        HIPCHK(hipEventCreate(&ev.b));
        HIPCHK(hipEventRecord(ev.a, x->stream));
    }
    int n = 0, grc = 0;
    double tcur = *t;
    for (; n < max_steps && !grc; ++n) {
        if (sync_each_step && !(tcur < total_time)) break;
        // FINDDT: the previous stage 3 already reduced the minimum
        if (!x->dt_pending) launch_finddt(x);
        hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(1), 0, x->stream, x->S);
        x->cur_step = x->nstep_total + n;
        for (int s = 1; s <= 3 && !grc; ++s) grc = run_stage(x, s);
        x->dt_pending = true;
        hipLaunchKernelGGL(k_time_advance, dim3(1), dim3(1), 0, x->stream, x->S, dlog, n);
        if (sync_each_step) {
            HIPCHK(hipMemcpyAsync(&tcur, &x->S->t, sizeof(double), hipMemcpyDeviceToHost, x->stream));
            HIPCHK(hipStreamSynchronize(x->stream));
        }
    }
    if (x->timing) {
        HIPCHK(hipEventRecord(ev.b, x->stream));
        HIPCHK(hipEventSynchronize(ev.b));
        float ms = 0;
        hipEventElapsedTime(&ms, ev.a, ev.b);
        accumulate_timing(x);
        x->tsum[6] = ms;
        x->tcnt[6] = n > 0 ? n : 1;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(&h, x->S, sizeof(h), hipMemcpyDeviceToHost));
    *t = h.t;
    int done = h.nstep;
    x->nstep_total += done;
    if (steps_done) *steps_done = done;
    if (dts && dlog) {
        HIPCHK(hipMemcpy(dts, dlog, (size_t)(done < max_steps ? done : max_steps) * sizeof(double),
                         hipMemcpyDeviceToHost));
    }
    if (grc) return grc;
    if (h.nonfinite) {
        h.nonfinite = 0;
        HIPCHK(hipMemcpy(x->S, &h, sizeof(h), hipMemcpyHostToDevice));
        return fail(CUMC3D_ENONFINITE, "a cell's CFL time step was NaN or non-positive (NaN or negative energy in the state)");
    }
    if (!std::isfinite(h.dt) || h.dt < 0.0) return fail(CUMC3D_ENONFINITE, "time step became non-finite");
    return grav_check(x);
}

int cumc3d_run(cumc3d_ctx *x, int max_steps, double total_time, int sync_each_step, double *t, double *dts,
               int *steps_done) {
    if (!x || !t) return fail(CUMC3D_EINVAL, "null argument");
    if (!x->initialized) return fail(CUMC3D_ESTATE, "call cumc3d_initial_update first");
    if (x->next_stage != 1) return fail(CUMC3D_ESTATE, "a step run by cumc3d_stage is not finished");
    DevTmp dbuf;
    if (dts && max_steps > 0) HIPCHK(hipMalloc(&dbuf.p, (size_t)max_steps * sizeof(double)));
    return run_loop(x, max_steps, total_time, sync_each_step, t, dts, steps_done, (double *)dbuf.p);
}

int cumc3d_run_noalloc(cumc3d_ctx *x, int max_steps, double total_time, int sync_each_step, double *t,
                       double *dts, int *steps_done) {
    if (!x || !t) return fail(CUMC3D_EINVAL, "null argument");
    if (!x->initialized) return fail(CUMC3D_ESTATE, "call cumc3d_initial_update first");
    if (x->next_stage != 1) return fail(CUMC3D_ESTATE, "a step run by cumc3d_stage is not finished");
    if (dts && max_steps > CUMC3D_NOALLOC_MAX_STEPS)
        return fail(CUMC3D_EINVAL, "cumc3d_run_noalloc logs at most CUMC3D_NOALLOC_MAX_STEPS (65536) dts per call");
    if (x->timing) return fail(CUMC3D_ESTATE, "cumc3d_run_noalloc does not time kernels (the events allocate): "
                                              "turn timing off with cumc3d_set_timing");
    if (!x->scratch) return fail(CUMC3D_ESTATE, "the scratch array's device address was not resolved");
    return run_loop(x, max_steps, total_time, sync_each_step, t, dts, steps_done,
                    dts && max_steps > 0 ? x->scratch : nullptr);
}

int cumc3d_download(cumc3d_ctx *x, double *prim, double *eps, double *bcell, double *cs) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    DevTmp buf;
    if (prim || bcell) HIPCHK(hipMalloc(&buf.p, (size_t)8 * x->G * sizeof(double)));
    if (!buf.p && x->fp32 && (eps || cs)) HIPCHK(hipMalloc(&buf.p, (size_t)x->G * sizeof(double)));
    double *tmp = (double *)buf.p;
    const unsigned gg = nblk(x->G);
    if (prim) {
        if (x->fp32) f32::join_prim(x, tmp);
        else hipLaunchKernelGGL(k_join_prim<double>, dim3(gg), dim3(256), 0, x->stream, tmp, make_args<double>(x));
        HIPCHK(hipMemcpyAsync(prim, tmp, (size_t)8 * x->G * sizeof(double), hipMemcpyDeviceToHost, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
    }
    if (bcell) {
        if (x->fp32)
            f32::join_bcell(x, tmp);
        else
            hipLaunchKernelGGL(k_join3<double>, dim3(gg), dim3(256), 0, x->stream, tmp, (const double *)x->Bc[0],
                               (const double *)x->Bc[1], (const double *)x->Bc[2], x->G);
        HIPCHK(hipMemcpyAsync(bcell, tmp, (size_t)3 * x->G * sizeof(double), hipMemcpyDeviceToHost, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
    }
    void *scal[2] = {x->eps, x->cs};
    double *outs[2] = {eps, cs};
    for (int i = 0; i < 2; ++i) {
        if (!outs[i]) continue;
        if (x->fp32) {
            f32::widen(x, tmp, scal[i]);
            HIPCHK(hipMemcpyAsync(outs[i], tmp, x->G * sizeof(double), hipMemcpyDeviceToHost, x->stream));
        } else {
            HIPCHK(hipMemcpyAsync(outs[i], scal[i], x->G * sizeof(double), hipMemcpyDeviceToHost, x->stream));
        }
        HIPCHK(hipStreamSynchronize(x->stream));
    }
    return 0;
}

int cumc3d_profile(cumc3d_ctx *x, double *primitive, double *epsilon, double *bfield) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    if (!x->uploaded) return fail(CUMC3D_ESTATE, "upload an initial model first");
    const size_t np = (size_t)(x->nx + 2) * (x->ny + 2) * (x->nz + 2);
    const size_t nb = (size_t)(x->nx + 3) * (x->ny + 3) * (x->nz + 3);
    DevTmp buf;
    HIPCHK(hipMalloc(&buf.p, (5 * np + np + 3 * nb) * sizeof(double)));
    double *tmp = (double *)buf.p;
    double *dp = primitive ? tmp : nullptr, *de = epsilon ? tmp + 5 * np : nullptr,
           *db = bfield ? tmp + 6 * np : nullptr;
    if (x->fp32)
        hipLaunchKernelGGL(k_pack_profile<float>, dim3(nblk((long)nb)), dim3(256), 0, x->stream, make_args<float>(x),
                           dp, de, db);
    else
        hipLaunchKernelGGL(k_pack_profile<double>, dim3(nblk((long)nb)), dim3(256), 0, x->stream,
                           make_args<double>(x), dp, de, db);
    if (primitive) HIPCHK(hipMemcpyAsync(primitive, dp, 5 * np * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    if (epsilon) HIPCHK(hipMemcpyAsync(epsilon, de, np * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    if (bfield) HIPCHK(hipMemcpyAsync(bfield, db, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int cumc3d_gravity(cumc3d_ctx *x, double *phi, double *m_bh, double *r_sh, int *iterations) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    if (x->d.model != CUMC3D_MODEL_COLLAPSAR) return fail(CUMC3D_EINVAL, "not a collapsar run");
    HIPCHK(hipStreamSynchronize(x->stream));
    if (phi) HIPCHK(hipMemcpy(phi, x->phi, (size_t)x->G * sizeof(double), hipMemcpyDeviceToHost));
    DevScalars h;
    HIPCHK(hipMemcpy(&h, x->S, sizeof(h), hipMemcpyDeviceToHost));
    if (m_bh) *m_bh = h.m_bh;
    if (r_sh) *r_sh = h.r_sh;
    if (iterations) *iterations = h.sor_iter;
    return 0;
}

int cumc3d_divb(cumc3d_ctx *x, double *maxdivb) {
    if (!x || !maxdivb) return fail(CUMC3D_EINVAL, "null argument");
    DevTmp buf;
    unsigned long long h = 0;
    HIPCHK(hipMalloc(&buf.p, sizeof(h)));
    unsigned long long *d = (unsigned long long *)buf.p;
    HIPCHK(hipMemsetAsync(d, 0, sizeof(*d), x->stream));
    hipLaunchKernelGGL(k_divb<double>, dim3(nblk((long)x->nx * x->ny * x->nz)), dim3(256), 0, x->stream,
                       make_args<double>(x), d);
    HIPCHK(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    memcpy(maxdivb, &h, sizeof(double));
    return 0;
}

int cumc3d_finalize(cumc3d_ctx *x) {
    if (!x) return 0;
    if (x->stream) hipStreamSynchronize(x->stream);
    if (x->side) hipStreamSynchronize(x->side);
    for (auto &e : x->ev) hipEventDestroy(e);
    if (x->pool) hipFree(x->pool);
    if (x->met) hipFree(x->met);
    if (x->S) hipFree(x->S);
    if (x->stale) hipFree(x->stale);
    if (x->grav) hipFree(x->grav);
    if (x->ev_x) hipEventDestroy(x->ev_x);
    if (x->ev_m) hipEventDestroy(x->ev_m);
    if (x->side) hipStreamDestroy(x->side);
    void *sel[] = {x->sel_w, x->sel_m, x->sel_v, x->sel_frag, x->mask, x->samp_feat, x->samp_label,
                   x->samp_owner, x->samp_counts, x->grad_part, x->loss_part, x->loss_dev};
    for (void *q : sel)
        if (q) hipFree(q);
    if (x->stream) hipStreamDestroy(x->stream);
    delete x;
    return 0;
}

// ---------------------------------------------------------------------------
// learned reconstruction selector
// ---------------------------------------------------------------------------
static int sel_pack(cumc3d_ctx *x) {
    hipLaunchKernelGGL(k_sel_pack, dim3(1), dim3(64 * SEL_NFRAG), 0, x->stream, x->sel_w, x->sel_frag);
    HIPCHK(hipGetLastError());
    return 0;
}

int cumc3d_selector_init(cumc3d_ctx *x, int capacity, unsigned seed) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    if (capacity < SEL_TB) return fail(CUMC3D_EINVAL, "sample capacity must be at least 128");
    if (!x->sel_w) {
        HIPCHK(hipMalloc(&x->sel_w, SEL_NPARAM * sizeof(float)));
        HIPCHK(hipMalloc(&x->sel_m, SEL_NPARAM * sizeof(float)));
        HIPCHK(hipMalloc(&x->sel_v, SEL_NPARAM * sizeof(float)));
        HIPCHK(hipMalloc(&x->sel_frag, (size_t)64 * SEL_NFRAG * 4 * sizeof(unsigned short)));
        HIPCHK(hipMalloc(&x->mask, (size_t)3 * x->G));
        HIPCHK(hipMalloc(&x->samp_counts, 4 * sizeof(int)));
        HIPCHK(hipMalloc(&x->loss_dev, sizeof(float)));
    }
    if (x->samp_cap != capacity) {
        if (x->samp_feat) HIPCHK(hipFree(x->samp_feat));
        if (x->samp_label) HIPCHK(hipFree(x->samp_label));
        if (x->samp_owner) HIPCHK(hipFree(x->samp_owner));
        x->samp_feat = nullptr;
        x->samp_label = nullptr;
        x->samp_owner = nullptr;
        x->samp_cap = 0;
        HIPCHK(hipMalloc(&x->samp_feat, (size_t)capacity * SEL_IN * sizeof(float)));
        HIPCHK(hipMalloc(&x->samp_label, (size_t)capacity));
        HIPCHK(hipMalloc(&x->samp_owner, (size_t)capacity * sizeof(int)));
        x->samp_cap = capacity;
    }
    // He-uniform initial weights from a fixed-seed generator, zero biases
    std::vector<float> w(SEL_NPARAM, 0.0f);
    unsigned long long st = 0x9E3779B97F4A7C15ULL ^ seed;
    auto uni = [&]() {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return (float)((st >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
    };
    for (int i = 0; i < SEL_IN * SEL_H; ++i) w[SEL_W1 + i] = uni() * sqrtf(6.0f / 15.0f);
    for (int i = 0; i < SEL_H * SEL_H; ++i) w[SEL_W2 + i] = uni() * sqrtf(6.0f / SEL_H);
    for (int i = 0; i < SEL_H * SEL_OUT; ++i) w[SEL_W3 + i] = uni() * sqrtf(6.0f / SEL_H);
    std::vector<float> zero(SEL_NPARAM, 0.0f);
    HIPCHK(hipMemcpy(x->sel_w, w.data(), SEL_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x->sel_m, zero.data(), SEL_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x->sel_v, zero.data(), SEL_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(x->samp_counts, 0, 4 * sizeof(int)));
    x->samp_seen = 0;
    HIPCHK(hipMemset(x->mask, 2, (size_t)3 * x->G));   // WENO-Z until the first inference
    x->adam_step = 0;
    if (int rc = sel_pack(x)) return rc;
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int cumc3d_selector_set_params(cumc3d_ctx *x, const float *w, int n) {
    if (!x || !w) return fail(CUMC3D_EINVAL, "null argument");
    if (!x->sel_w) return fail(CUMC3D_ESTATE, "call cumc3d_selector_init first");
    if (n != SEL_NPARAM) return fail(CUMC3D_EINVAL, "the selector has " + std::to_string(SEL_NPARAM) + " parameters");
    HIPCHK(hipMemcpy(x->sel_w, w, SEL_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    if (int rc = sel_pack(x)) return rc;
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int cumc3d_selector_get_params(cumc3d_ctx *x, float *w, int n) {
    if (!x || !w) return fail(CUMC3D_EINVAL, "null argument");
    if (!x->sel_w) return fail(CUMC3D_ESTATE, "call cumc3d_selector_init first");
    if (n != SEL_NPARAM) return fail(CUMC3D_EINVAL, "the selector has " + std::to_string(SEL_NPARAM) + " parameters");
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(w, x->sel_w, SEL_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"

extern "C" {

int cumc3d_selector_collect(cumc3d_ctx *x, int stride, double thr_ppm, double thr_weno, unsigned seed, int *counts) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    if (!x->sel_w) return fail(CUMC3D_ESTATE, "call cumc3d_selector_init first");
    if (!x->initialized) return fail(CUMC3D_ESTATE, "call cumc3d_initial_update first");
    if (stride < 1) return fail(CUMC3D_EINVAL, "stride must be >= 1");
    if (x->fp32) f32::sel_collect(x, stride, seed, (float)thr_ppm, (float)thr_weno);
    else sel_collect_t<double>(x, stride, seed, (float)thr_ppm, (float)thr_weno);
    HIPCHK(hipGetLastError());
    // counts: samples held (at most the capacity; a reservoir over every
    // candidate offered so far once full) and the class histogram of those held
    int h[4];
    HIPCHK(hipMemcpyAsync(h, x->samp_counts, sizeof(h), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (counts) memcpy(counts, h, sizeof(h));
    return 0;
}

int cumc3d_selector_train(cumc3d_ctx *x, int iters, int blocks, double lr, const double *class_weight, unsigned seed,
                          float *loss) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    if (!x->sel_w) return fail(CUMC3D_ESTATE, "call cumc3d_selector_init first");
    if (iters < 0 || blocks < 1 || blocks > 65536) return fail(CUMC3D_EINVAL, "bad iteration or block count");
    int h[4];
    HIPCHK(hipMemcpy(h, x->samp_counts, sizeof(h), hipMemcpyDeviceToHost));
    int n = h[0] < x->samp_cap ? h[0] : x->samp_cap;
    if (n < 1) return fail(CUMC3D_ESTATE, "no training samples: call cumc3d_selector_collect first");
    if (x->grad_blocks < blocks) {
        if (x->grad_part) HIPCHK(hipFree(x->grad_part));
        if (x->loss_part) HIPCHK(hipFree(x->loss_part));
        HIPCHK(hipMalloc(&x->grad_part, (size_t)blocks * SEL_NPARAM * sizeof(float)));
        HIPCHK(hipMalloc(&x->loss_part, (size_t)blocks * sizeof(float)));
        x->grad_blocks = blocks;
    }
    float cw[3] = {1.0f, 1.0f, 1.0f};
    if (class_weight)
        for (int i = 0; i < 3; ++i) cw[i] = (float)class_weight[i];
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_sel_grad, dim3(blocks), dim3(SEL_TB), 0, x->stream, x->sel_w, x->samp_feat,
                           x->samp_label, n, seed + 7919u * (unsigned)(x->adam_step + 1), cw[0], cw[1], cw[2],
                           x->grad_part, x->loss_part);
        x->adam_step += 1;
        hipLaunchKernelGGL(k_sel_adam, dim3((SEL_NPARAM + 255) / 256), dim3(256), 0, x->stream, x->sel_w, x->sel_m,
                           x->sel_v, x->grad_part, x->loss_part, blocks, (float)lr, x->adam_step, x->loss_dev);
        if (loss) HIPCHK(hipMemcpyAsync(&loss[it], x->loss_dev, sizeof(float), hipMemcpyDeviceToHost, x->stream));
    }
    if (int rc = sel_pack(x)) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int cumc3d_selector_enable(cumc3d_ctx *x, int enable, int every) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    if (enable && !x->sel_w) return fail(CUMC3D_ESTATE, "call cumc3d_selector_init first");
    if (every < 1) return fail(CUMC3D_EINVAL, "every must be >= 1");
    x->sel_on = enable != 0;
    x->sel_every = every;
    x->sel_tick = 0;
    return 0;
}

int cumc3d_selector_mlp(cumc3d_ctx *x, const float *feat, int n, float *logits) {
    if (!x || !feat || !logits) return fail(CUMC3D_EINVAL, "null argument");
    if (!x->sel_w) return fail(CUMC3D_ESTATE, "call cumc3d_selector_init first");
    if (n < 1) return fail(CUMC3D_EINVAL, "n must be >= 1");
    float *df = nullptr, *dl = nullptr;
    HIPCHK(hipMalloc(&df, (size_t)n * SEL_IN * sizeof(float)));
    HIPCHK(hipMalloc(&dl, (size_t)n * 3 * sizeof(float)));
    HIPCHK(hipMemcpyAsync(df, feat, (size_t)n * SEL_IN * sizeof(float), hipMemcpyHostToDevice, x->stream));
    unsigned g = (unsigned)((n + 15) / 16);
    hipLaunchKernelGGL(k_sel_mlp_rows, dim3(g < 4096 ? g : 4096), dim3(64), 0, x->stream, x->sel_frag, x->sel_w, df,
                       (long)n, dl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(logits, dl, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipFree(df));
    HIPCHK(hipFree(dl));
    return 0;
}

int cumc3d_selector_mask(cumc3d_ctx *x, int dir, int refresh, unsigned char *out) {
    if (!x || !out) return fail(CUMC3D_EINVAL, "null argument");
    if (!x->sel_w) return fail(CUMC3D_ESTATE, "call cumc3d_selector_init first");
    if (dir < 0 || dir >= x->d.n_dim) return fail(CUMC3D_EINVAL, "direction out of range");
    if (refresh) {
        if (x->fp32) f32::sel_infer(x);
        else sel_infer_t<double>(x);
    }
    HIPCHK(hipMemcpyAsync(out, x->mask + (size_t)dir * x->G, x->G, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

int cumc3d_set_timing(cumc3d_ctx *x, int enable) {
    if (!x) return fail(CUMC3D_EINVAL, "null context");
    x->timing = enable != 0;
    return 0;
}

int cumc3d_get_timing(cumc3d_ctx *x, double *ms, int n) {
    if (!x || !ms) return fail(CUMC3D_EINVAL, "null argument");
    for (int i = 0; i < n && i < 10; ++i) ms[i] = x->tcnt[i] ? x->tsum[i] / x->tcnt[i] : 0.0;
    return 0;
}

}  // extern "C"
