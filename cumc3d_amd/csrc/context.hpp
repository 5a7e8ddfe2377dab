// Device context and the launch sequence of one RK stage, shared by the
// translation units of libcumc3d_hip.so: capi.hip (C-ABI, the FP64 mode),
// fp32.hip (the FP32 split's sweeps and selector, compiled with contraction,
// the short single-precision divide and square root and reciprocal divides)
// and fp32_state.hip (the FP32 split's update, ghost fill and FINDDT, with
// IEEE division in double).
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/cumc3d_hip.h"
#include "kernels.hpp"
#include "selector.hpp"

struct cumc3d_ctx {
    cumc3d_desc d;
    int nx, ny, nz;
    long NX, NY, NZ, G;
    hipStream_t stream = nullptr;
    void *pool = nullptr;         // all per-cell fields
    bool fp32 = false;            // CUMC3D_FP32: working fields are float
    double *met = nullptr;        // grid metrics and weights
    cumc3d::DevScalars *S = nullptr;
    // working-type fields (double, or float in the FP32 split)
    void *P[5], *eps, *cs, *Bc[3], *F[3][7], *E[3], *Ec[3];
    // always double: face field (ping-pong), u_old, conserved state (FP32 split), floors
    double *B[2][3], *Uo[8], *U[5], *floors[3];
    int cur = 0;                  // B[cur] holds the current face field
    long march_threads = 1L << 19;  // target thread count of the y/z sweeps
    // learned reconstruction selector
    bool sel_on = false;
    int sel_every = 1, sel_tick = 0;
    float *sel_w = nullptr, *sel_m = nullptr, *sel_v = nullptr;   // FP32 parameters + Adam moments
    unsigned short *sel_frag = nullptr;                            // BF16 MFMA fragments
    unsigned char *mask = nullptr;                                 // 3 x G class mask
    unsigned char *stale = nullptr;                                // HLL: 3 x G unassigned-face flags
    const double *sph[10] = {};   // spherical tables in met (see ArgsT), null on other geometries
    float *samp_feat = nullptr;
    unsigned char *samp_label = nullptr;
    int samp_cap = 0, *samp_counts = nullptr, adam_step = 0;
    int *samp_owner = nullptr;              // reservoir slot claims of one collect
    double *scratch = nullptr;              // this device's address of capi.hip's scratch array
    unsigned long long samp_seen = 0;       // candidates offered since cumc3d_selector_init
    float *grad_part = nullptr, *loss_part = nullptr, *loss_dev = nullptr;
    int grad_blocks = 0;
    int bfac[6][12];
    cumc3d::Axis ax[3];
    bool uploaded = false, initialized = false, dt_pending = false;
    int next_stage = 1;           // cumc3d_stage: the stage that runs next (1 between steps)
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;   // TMARKS marks per stage, TSTAGES stages
    std::vector<char> sel_timed;  // stage refreshed the selector masks
    int tstage = 0;               // stages recorded in the current run
    double tsum[10] = {0};        // [8] Poisson solve ms, [9] SOR sweeps per solve
    long tcnt[10] = {0};
    // collapsar self-gravity (model/collapsar/src/def.f90:16-36): one
    // allocation holding the potential, its colour-split copies for the SOR,
    // the poisson_coef coefficients and the error blocks' partials
    cumc3d_gravity_params grav_p{};
    void *grav = nullptr;
    double *phi = nullptr;                  // the potential (ghosted box)
    double *sor_phi[2][2] = {};             // during a solve: [buffer][red, black] (colour-split)
    double *ajp1 = nullptr, *ajm1 = nullptr, *bkp1 = nullptr, *bkm1 = nullptr;
    double *sor_rho = nullptr, *part_max = nullptr;   // rho_in split by colour; error blocks' maxima
    double *sor_a2 = nullptr, *sor_e2 = nullptr, *sor_z[3] = {};   // poisson_coef factors (gravity.hpp)
    long long *part_nan = nullptr;
    int sor_nblk = 0;
    double xF0 = 0.0;             // xF(0), the inner radial face (horizon check)
    long nstep_total = 0;         // steps since the initial update (main.f90's n_step)
    long cur_step = 0;            // n_step of the step being queued
    int grav_error = 0;           // 1: a solve converged on its last allowed sweep (the reference STOPs)
    hipStream_t side = nullptr;   // the accretion sum runs here beside the stage's later kernels
    hipEvent_t ev_x = nullptr, ev_m = nullptr;   // x sweep done (main stream); mdot done (side)
};

namespace cumc3d {


// ---------------------------------------------------------------------------
// layout conversion kernels (upload/download only)
// ---------------------------------------------------------------------------
template <class W> __global__ void k_split_prim(const double *aos, const double *eps64, ArgsT<W> a) {
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.G) return;
    const double *q = aos + 8 * c;
    for (int i = 0; i < 5; ++i) a.P[i][c] = (W)q[i];
    for (int d = 0; d < 3; ++d) {
        a.Bin[d][c] = q[5 + d];
        a.Bout[d][c] = q[5 + d];
    }
    a.eps[c] = (W)eps64[c];
    if constexpr (sizeof(W) == 4) {
        // double copies of rho, v, eps for INITIAL_UPDATE's conserved state
        for (int i = 0; i < 4; ++i) a.U[i][c] = q[i];
        a.U[4][c] = eps64[c];
    }
}

// prim AoS (rho, vx, vy, vz, p, Bx, By, Bz per cell) from the device fields
template <class W> __global__ void k_join_prim(double *aos, ArgsT<W> a) {
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.G) return;
    double *q = aos + 8 * c;
    for (int i = 0; i < 5; ++i) q[i] = a.P[i][c];
    for (int d = 0; d < 3; ++d) q[5 + d] = a.Bin[d][c];
}

// print_hydroprofile's three arrays (dataio.f90:178-233) in its order, Fortran
// column-major: primitive(1:5, 0:nx+1, 0:ny+1, 0:nz+1), epsilon(0:nx+1, ...)
// and bfield(1:3, -1:nx+1, -1:ny+1, -1:nz+1) (the face field).  One thread
// per cell of the bfield box; any output pointer may be null.
template <class W> __global__ void k_pack_profile(ArgsT<W> a, double *prim5, double *eps, double *bf3) {
    const long bx = a.nx + 3, by = a.ny + 3;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= bx * by * (a.nz + 3)) return;
    const int j = -1 + (int)(t % bx), k = -1 + (int)((t / bx) % by), l = -1 + (int)(t / (bx * by));
    const long c = cidx(a, j, k, l);
    if (bf3)
        for (int d = 0; d < 3; ++d) bf3[3 * t + d] = a.Bin[d][c];
    if (j >= 0 && k >= 0 && l >= 0) {
        const long q = j + (long)(a.nx + 2) * (k + (long)(a.ny + 2) * l);
        if (prim5)
            for (int v = 0; v < 5; ++v) prim5[5 * q + v] = (double)a.P[v][c];
        if (eps) eps[q] = (double)a.eps[c];
    }
}

template <class W> __global__ void k_join3(double *aos, const W *f0, const W *f1, const W *f2, long G) {
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= G) return;
    aos[3 * c] = f0[c];
    aos[3 * c + 1] = f1[c];
    aos[3 * c + 2] = f2[c];
}

template <class W> __global__ void k_widen(double *out, const W *in, long G) {
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < G) out[c] = in[c];
}

// INITIAL_UPDATE interior part: CUSTOM_CHECKRHO, bcell of FROMRVETOU and
// FINDPRESSURE on the interior (initial.f90:91-108).  The FP32 split starts
// from the double copies k_split_prim left in U and writes the conserved
// state there (FROMRVETOU, fromrvetou.f90:94-108).
template <class W> __global__ void k_init_interior(ArgsT<W> a) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)a.nx * a.ny * a.nz) return;
    const int j = 1 + (int)(t % a.nx), k = 1 + (int)((t / a.nx) % a.ny), l = 1 + (int)(t / ((long)a.nx * a.ny));
    const long c = cidx(a, j, k, l);
    constexpr bool split = sizeof(W) == 4;
    double rho, eps, v[3];
    if constexpr (split) {
        rho = a.U[0][c]; v[0] = a.U[1][c]; v[1] = a.U[2][c]; v[2] = a.U[3][c]; eps = a.U[4][c];
    } else {
        rho = a.P[0][c]; v[0] = a.P[1][c]; v[1] = a.P[2][c]; v[2] = a.P[3][c]; eps = a.eps[c];
    }
    checkrho(a, c, rho, v, eps);
    a.P[0][c] = (W)rho; a.P[1][c] = (W)v[0]; a.P[2][c] = (W)v[1]; a.P[3][c] = (W)v[2];
    a.eps[c] = (W)eps;
    double b[3] = {bavg_rt<0>(a, j, k, a.Bin[0][c], a.Bin[0][c - 1]),
                   bavg_rt<1>(a, j, k, a.Bin[1][c], a.Bin[1][c - a.NX]),
                   bavg_rt<2>(a, j, k, a.Bin[2][c], a.Bin[2][c - a.NX * a.NY])};
    for (int d = 0; d < 3; ++d) a.Bc[d][c] = (W)b[d];
    cell_emf(a, c, (W)v[0], (W)v[1], (W)v[2], (W)b[0], (W)b[1], (W)b[2]);
    double p, cs;
    pressure(a, c, rho, eps, p, cs);
    a.P[4][c] = (W)p;
    a.cs[c] = (W)cs;
    if constexpr (split) {
        double vsq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2], bsq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        a.U[0][c] = rho;
        a.U[1][c] = v[0] * rho;
        a.U[2][c] = v[1] * rho;
        a.U[3][c] = v[2] * rho;
        a.U[4][c] = rho * (eps + 0.5 * vsq) + 0.5 * bsq;
    }
}

// standalone FINDDT (rungekutta.f90:256-323)
template <class W> __global__ __launch_bounds__(256) void k_finddt(ArgsT<W> a) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double dtloc = 1.0e5;
    if (t < (long)a.nx * a.ny * a.nz) {
        const int j = 1 + (int)(t % a.nx), k = 1 + (int)((t / a.nx) % a.ny), l = 1 + (int)(t / ((long)a.nx * a.ny));
        const long c = cidx(a, j, k, l);
        double v[3] = {a.P[1][c], a.P[2][c], a.P[3][c]};
        double b[3] = {a.Bc[0][c], a.Bc[1][c], a.Bc[2][c]};
        dtloc = cfl_dt_rt(a, j, k, l, a.cs[c], a.P[0][c], v, b);
    }
    dtloc = block_min_f64(dtloc, a.S);
    if (threadIdx.x == 0) atomicMin(&a.S->dtmin_bits, (unsigned long long)__double_as_longlong(dtloc));
}

// FIND_DIVB (mhd_module.f90:133-189) as a max reduction on bit patterns
template <class W> __global__ __launch_bounds__(256) void k_divb(ArgsT<W> a, unsigned long long *out) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (t < (long)a.nx * a.ny * a.nz) {
        const int j = 1 + (int)(t % a.nx), k = 1 + (int)((t / a.nx) % a.ny), l = 1 + (int)(t / ((long)a.nx * a.ny));
        const long c = cidx(a, j, k, l), sy = a.NX, sz = a.NX * a.NY;
        const double *bx = a.Bin[0], *by = a.Bin[1], *bz = a.Bin[2];
        double dx = a.ax[0].d[j + 3], dy = a.ax[1].d[k + 3], dz = a.ax[2].d[l + 3];
        if (a.coord == 0)
            d = (bx[c] - bx[c - 1]) / (dx) + (by[c] - by[c - sy]) / (dy) + (bz[c] - bz[c - sz]) / (dz);
        else if (a.coord == 1)
            d = (a.ax[0].xF[j + 3] * bx[c] - a.ax[0].xF[j + 2] * bx[c - 1]) / (a.ax[0].c[j + 3] * dx) +
                (by[c] - by[c - sy]) / (a.ax[0].c[j + 3] * dy) + (bz[c] - bz[c - sz]) / (dz);
        else {   // mhd_module.f90:171-186
            double xF = a.ax[0].xF[j + 3], xFm = a.ax[0].xF[j + 2], dcb = a.dcb[j + 3], dc = a.dcose[k + 3];
            double xdx = a.ax[0].c[j + 3] * dx;
            d = (xF * xF * bx[c] - xFm * xFm * bx[c - 1]) / (dcb / 3.0) +
                (a.sinf[k + 3] * by[c] - a.sinf[k + 2] * by[c - sy]) * xdx / (dcb * dc / 3.0) +
                (bz[c] - bz[c - sz]) * (xdx * dy) / (dcb * dc * dz / 3.0);
        }
        d = fabsT(d);
    }
    // max of non-negative doubles == max of their bit patterns
    unsigned long long bits = (unsigned long long)__double_as_longlong(d);
    bits = block_max_u64(bits);
    if (threadIdx.x == 0) atomicMax(out, bits);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
// k_update for the grid's geometry and the stage
template <class W, int G>
static void launch_update_g(int stage, dim3 grid, hipStream_t st, const ArgsT<W> &a) {
    if (stage == 0) hipLaunchKernelGGL((k_update<0, W, G>), grid, dim3(256), 0, st, a);
    else if (stage == 1) hipLaunchKernelGGL((k_update<1, W, G>), grid, dim3(256), 0, st, a);
    else if (stage == 2) hipLaunchKernelGGL((k_update<2, W, G>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_update<3, W, G>), grid, dim3(256), 0, st, a);
}
template <class W>
static void launch_update(int geo, int stage, dim3 grid, hipStream_t st, const ArgsT<W> &a) {
    if (geo == 1) launch_update_g<W, 1>(stage, grid, st, a);
    else if (geo == 2) launch_update_g<W, 2>(stage, grid, st, a);
    else launch_update_g<W, 0>(stage, grid, st, a);
}

template <class W> static ArgsT<W> make_args(cumc3d_ctx *x) {
    ArgsT<W> a;
    memset(&a, 0, sizeof(a));
    a.nx = x->nx; a.ny = x->ny; a.nz = x->nz;
    a.ndim = x->d.n_dim;
    a.coord = x->d.coordinate_flag;
    a.model = x->d.model;
    a.rt_gravity = x->d.rt_gravity;
    a.bn_mode = x->d.hlld_bn_mode;
    for (int i = 0; i < 6; ++i) a.bc[i] = x->d.boundary_flag[i];
    memcpy(a.bfac, x->bfac, sizeof(a.bfac));
    a.NX = x->NX; a.NY = x->NY; a.G = x->G;
    a.g1 = x->d.ggas - 1.0;
    a.cfl = x->d.cfl;
    a.r_sh = x->d.r_sh;
    for (int d = 0; d < 3; ++d) a.ax[d] = x->ax[d];
    for (int i = 0; i < 5; ++i) a.P[i] = (W *)x->P[i];
    a.eps = (W *)x->eps;
    a.cs = (W *)x->cs;
    for (int d = 0; d < 3; ++d) {
        a.Bc[d] = (W *)x->Bc[d];
        a.Bin[d] = x->B[x->cur][d];
        a.Bout[d] = x->B[1 - x->cur][d];
        a.E[d] = (W *)x->E[d];
        a.Ec[d] = (W *)x->Ec[d];
        for (int i = 0; i < 7; ++i) a.F[d][i] = (W *)x->F[d][i];
    }
    for (int i = 0; i < 8; ++i) a.Uo[i] = x->Uo[i];
    for (int i = 0; i < 5; ++i) a.U[i] = x->U[i];
    a.rfl = x->floors[0];
    a.pfl = x->floors[1];
    a.efl = x->floors[2];
    a.S = x->S;
    for (int d = 0; d < 3; ++d) a.mask[d] = x->mask ? x->mask + (size_t)d * x->G : nullptr;
    for (int d = 0; d < 3; ++d) a.stale[d] = x->stale ? x->stale + (size_t)d * x->G : nullptr;
    a.dcb = x->sph[0]; a.dcb3 = x->sph[1]; a.xbar = x->sph[2]; a.gsrc = x->sph[3]; a.hbx = x->sph[4];
    a.sinf = x->sph[5]; a.dcose = x->sph[6]; a.cot = x->sph[7]; a.hby = x->sph[8]; a.sine = x->sph[9];
    a.fixpole = x->d.coordinate_flag == 2 && x->d.fixpole;
    a.phi = x->phi;
    for (int i = 0; i < 4; ++i) a.pa[i] = x->grav_p.prim_a[i];
    a.eps_a = x->grav_p.eps_a;
    return a;
}

static inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Segments per column of the y/z sweeps: each thread marches one segment of
// interfaces along the sweep axis, reusing the reconstruction of the cell it
// shares with the previous interface.  About march_threads threads in flight
// (2^19 = 8k waves by default); CUMC3D_MARCH_THREADS overrides it at init.
static int nseg_for(const cumc3d_ctx *x, int nx, int na, int nd) {
    long cols = (long)nx * na;
    long want = x->march_threads / (cols > 0 ? cols : 1);
    if (want < 1) want = 1;
    if (want > nd + 1) want = nd + 1;
    return (int)want;
}

// ---------------------------------------------------------------------------
// flux dispatch over (reconstruction, Riemann solver)
// ---------------------------------------------------------------------------
template <class W, int R, int S>
static void launch_sweep_rs(cumc3d_ctx *x, const ArgsT<W> &a, int D, hipStream_t st) {
    if (D == 0) {
        long waves = ((long)(x->nx + 2) * x->ny * x->nz + 62) / 63 + 1;
        unsigned grid = (unsigned)((waves * 64 + 255) / 256);
        hipLaunchKernelGGL((k_flux_x<W, R, S>), dim3(grid), dim3(256), 0, st, a);
        if (S == 3 && a.bn_mode == 2)
            hipLaunchKernelGGL((k_carry<W, R, 0>), dim3(1), dim3(64), 0, st, a);
    } else if (D == 1) {
        int ns = nseg_for(x, x->nx, x->nz, x->ny);
        unsigned grid = nblk((long)x->nx * x->nz * ns);
        hipLaunchKernelGGL((k_flux_yz<W, R, S, 1>), dim3(grid), dim3(256), 0, st, a, ns);
        if (S == 3 && a.bn_mode == 2)
            hipLaunchKernelGGL((k_carry<W, R, 1>), dim3(1), dim3(64), 0, st, a);
    } else {
        int ns = nseg_for(x, x->nx, x->ny, x->nz);
        unsigned grid = nblk((long)x->nx * x->ny * ns);
        hipLaunchKernelGGL((k_flux_yz<W, R, S, 2>), dim3(grid), dim3(256), 0, st, a, ns);
        if (S == 3 && a.bn_mode == 2)
            hipLaunchKernelGGL((k_carry<W, R, 2>), dim3(1), dim3(64), 0, st, a);
    }
}

template <class W, int R>
static void launch_sweep_r(cumc3d_ctx *x, const ArgsT<W> &a, int D, hipStream_t st) {
    switch (x->d.riemann) {
        case 0: launch_sweep_rs<W, R, 0>(x, a, D, st); break;
        case 1: launch_sweep_rs<W, R, 1>(x, a, D, st); break;
        case 2: launch_sweep_rs<W, R, 2>(x, a, D, st); break;
        default: launch_sweep_rs<W, R, 3>(x, a, D, st); break;
    }
}

template <class W> static void launch_sweep(cumc3d_ctx *x, const ArgsT<W> &a, int D, hipStream_t st) {
    if (x->sel_on) {
        launch_sweep_r<W, RSEL>(x, a, D, st);
        return;
    }
    switch (x->d.recon) {
        case 0: launch_sweep_r<W, 0>(x, a, D, st); break;
        case 1: launch_sweep_r<W, 1>(x, a, D, st); break;
        case 2: launch_sweep_r<W, 2>(x, a, D, st); break;
        case 3: launch_sweep_r<W, 3>(x, a, D, st); break;
        default: launch_sweep_r<W, 4>(x, a, D, st); break;
    }
}

// ---------------------------------------------------------------------------
// stage / step
// ---------------------------------------------------------------------------
static constexpr int TSTAGES = 192;
static constexpr int TMARKS = 9;   // 0..6 stage kernels, 7..8 selector inference
static inline void tmark(cumc3d_ctx *x, int slot) {
    if (x->timing && x->tstage < TSTAGES) hipEventRecord(x->ev[x->tstage * TMARKS + slot], x->stream);
}

// SPATIAL up to l_operator: the three sweeps (fluxes, face and cell EMFs)
// and the edge EMFs of the current state; `marks` records the stage's
// timing events
// The FP32 split's state kernels (update, ghost fill) live in their own
// translation unit, whose double arithmetic keeps IEEE division.
namespace f32 {
void update(cumc3d_ctx *x, const ArgsT<float> &a, int stage, dim3 grid);
void ghost(cumc3d_ctx *x, const ArgsT<float> &g);
}  // namespace f32

template <class W> static void stage_update(cumc3d_ctx *x, const ArgsT<W> &a, int stage, dim3 grid) {
    if constexpr (sizeof(W) == 4) f32::update(x, a, stage, grid);
    else launch_update<W>(x->d.coordinate_flag, stage, grid, x->stream, a);
}
template <class W> static void stage_ghost(cumc3d_ctx *x, const ArgsT<W> &g) {
    if constexpr (sizeof(W) == 4) f32::ghost(x, g);
    else hipLaunchKernelGGL(k_ghost<W>, dim3(nblk(ghost_count(*x))), dim3(256), 0, x->stream, g);
}

template <class W> static void sweeps_t(cumc3d_ctx *x, const ArgsT<W> &a, hipStream_t st, bool marks) {
    auto mark = [&](int slot) { if (marks) tmark(x, slot); };
    mark(0);
    launch_sweep(x, a, 0, st);
    if (x->ev_x) hipEventRecord(x->ev_x, st);   // the face-0 density flux the accretion sum reads
    mark(1);
    if (x->d.n_dim > 1) launch_sweep(x, a, 1, st);
    mark(2);
    if (x->d.n_dim > 2) launch_sweep(x, a, 2, st);
    mark(3);
    hipLaunchKernelGGL(k_emf_prep<W>, dim3(nblk(emf_prep_count(a))), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_edge<W>, dim3(nblk((long)(x->nx + 1) * (x->ny + 1) * (x->nz + 1))), dim3(256), 0, st, a);
    if (a.fixpole) hipLaunchKernelGGL(k_fixpole<W>, dim3(2 * x->nx), dim3(64), 0, st, a);
    mark(4);
}

// SPATIAL alone: l_rk of the current state into lrk (device, (8, G), var
// fastest); the state is not changed
template <class W> static void spatial_t(cumc3d_ctx *x, double *lrk) {
    ArgsT<W> a = make_args<W>(x);
    a.lrk = lrk;
    sweeps_t<W>(x, a, x->stream, false);
    stage_update<W>(x, a, 0, dim3(nblk((long)x->nx * x->ny * x->nz)));
}

template <class W> static void run_stage_t(cumc3d_ctx *x, int stage) {
    hipStream_t st = x->stream;
    ArgsT<W> a = make_args<W>(x);
    const long nn = (long)x->nx * x->ny * x->nz;
    sweeps_t<W>(x, a, st, true);
    stage_update<W>(x, a, stage, dim3(nblk(nn)));
    tmark(x, 5);
    x->cur = 1 - x->cur;
    ArgsT<W> g = make_args<W>(x);
    stage_ghost<W>(x, g);
    tmark(x, 6);
    if (x->timing && x->tstage < TSTAGES) x->tstage += 1;
}

template <class W> static void sel_infer_t(cumc3d_ctx *x) {
    ArgsT<W> a = make_args<W>(x);
    for (int D = 0; D < x->d.n_dim; ++D) {
        unsigned char *m = x->mask + (size_t)D * x->G;
        if (D == 0) hipLaunchKernelGGL((k_sel_infer<W, 0>), dim3(2048), dim3(256), 0, x->stream, a, x->sel_frag, x->sel_w, m);
        else if (D == 1) hipLaunchKernelGGL((k_sel_infer<W, 1>), dim3(2048), dim3(256), 0, x->stream, a, x->sel_frag, x->sel_w, m);
        else hipLaunchKernelGGL((k_sel_infer<W, 2>), dim3(2048), dim3(256), 0, x->stream, a, x->sel_frag, x->sel_w, m);
    }
}

template <class W> static void sel_collect_t(cumc3d_ctx *x, int stride, unsigned seed, float t1, float t2) {
    ArgsT<W> a = make_args<W>(x);
    const long cap = x->samp_cap;
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(cap)), dim3(256), 0, x->stream, x->samp_owner, cap, 0);
    long nt[3] = {0, 0, 0};
    unsigned sd[3];
    for (int D = 0; D < x->d.n_dim; ++D) {
        long nq = (long)(x->nx + (D == 0 ? 2 : 0)) * (x->ny + (D == 1 ? 2 : 0)) * (x->nz + (D == 2 ? 2 : 0));
        nt[D] = (nq + stride - 1) / stride;
        sd[D] = seed * 3u + (unsigned)D;
    }
    // pass 1 (store or claim), then pass 2 (replace), over the three regions
    for (int pass = 0; pass < 2; ++pass) {
        unsigned long long base = x->samp_seen;
        long id0 = 0;
        for (int D = 0; D < x->d.n_dim; ++D) {
            const dim3 g(nblk(nt[D]));
#define SEL_COLLECT_ARGS a, stride, sd[D], t1, t2, x->samp_feat, x->samp_label, cap, base, id0
            if (pass == 0) {
                if (D == 0) hipLaunchKernelGGL((k_sel_collect<W, 0>), g, dim3(256), 0, x->stream, SEL_COLLECT_ARGS, x->samp_owner);
                else if (D == 1) hipLaunchKernelGGL((k_sel_collect<W, 1>), g, dim3(256), 0, x->stream, SEL_COLLECT_ARGS, x->samp_owner);
                else hipLaunchKernelGGL((k_sel_collect<W, 2>), g, dim3(256), 0, x->stream, SEL_COLLECT_ARGS, x->samp_owner);
            } else {
                if (D == 0) hipLaunchKernelGGL((k_sel_place<W, 0>), g, dim3(256), 0, x->stream, SEL_COLLECT_ARGS, x->samp_owner);
                else if (D == 1) hipLaunchKernelGGL((k_sel_place<W, 1>), g, dim3(256), 0, x->stream, SEL_COLLECT_ARGS, x->samp_owner);
                else hipLaunchKernelGGL((k_sel_place<W, 2>), g, dim3(256), 0, x->stream, SEL_COLLECT_ARGS, x->samp_owner);
            }
#undef SEL_COLLECT_ARGS
            base += (unsigned long long)nt[D];
            id0 += nt[D];
        }
    }
    x->samp_seen += (unsigned long long)(nt[0] + nt[1] + nt[2]);
    const long held = (long)(x->samp_seen < (unsigned long long)cap ? x->samp_seen : (unsigned long long)cap);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, x->stream, x->samp_counts, 4L, 0);
    unsigned hb = nblk(held > 0 ? held : 1);
    if (hb > 1024) hb = 1024;
    hipLaunchKernelGGL(k_sel_hist, dim3(hb), dim3(256), 0, x->stream, x->samp_label, held, x->samp_counts);
}


// Entry points of the FP32 translation units: fp32.hip (stage sequence,
// sweeps, selector) and fp32_state.hip (FINDDT, the layout kernels and the
// initial update).
namespace f32 {
void run_stage(cumc3d_ctx *x, int stage);
void spatial(cumc3d_ctx *x, double *lrk);
void sel_infer(cumc3d_ctx *x);
void sel_collect(cumc3d_ctx *x, int stride, unsigned seed, float t1, float t2);
void finddt(cumc3d_ctx *x);
void split_prim(cumc3d_ctx *x, const double *aos, const double *eps64);
void initial_update(cumc3d_ctx *x);
void join_prim(cumc3d_ctx *x, double *tmp);
void join_bcell(cumc3d_ctx *x, double *tmp);
void widen(cumc3d_ctx *x, double *tmp, const void *src);
}  // namespace f32

}  // namespace cumc3d
