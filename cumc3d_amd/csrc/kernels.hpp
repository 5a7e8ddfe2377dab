// Kernels of one SSP-RK3 stage of the MHD step on MI355X (gfx950).
//
// Data layout: every field is a separate array (SoA) over the ghosted box
// -2..n+3 per axis with x fastest, cell (j,k,l) at (j+2) + NX*((k+2) + NY*(l+2)),
// NX=nx+6, NY=ny+6 -- the reference's cell ordering (setup.f90:348-351) with
// the variable index moved out to the slowest position.
//
// Per stage: x/y/z interface fluxes (reconstruction + face states + Riemann
// solver), the HLLD carry fix-up per sweep, edge EMFs, the fused update
// (flux divergence, sources, RK combine, CT, cons->prim->cons, floors, EOS,
// CFL minimum in stage 3) and the ghost fill.
#pragma once
#include <hip/hip_runtime.h>

#include "physics.hpp"

namespace cumc3d {

constexpr int NW = 14;   // weight slots per cell (PPM needs 14, WENO 10, TVD 2)

struct DevScalars {
    double dt;                        // dt of the current step
    double t;                         // global_time
    double total_time;
    unsigned long long dtmin_bits;    // CFL min (bit pattern of a positive double)
    double carry[3];                  // HLLD usstar carried per field component (x,y,z)
    unsigned long long carry_key;     // max serial key of the sweep's double-star interfaces (0: none)
    int nstep;
    int active;                       // 0 once global_time >= total_time: stage kernels exit
    int nonfinite;                    // set when a cell's CFL time step is NaN or not positive
    int pad;
    // collapsar (model/collapsar/src/def.f90:34-39, custom.f90:416-649)
    double m_bh, mbh_old, r_sh;       // black-hole mass, its value at step start, Schwarzschild radius
    double mdot;                      // the stage's accretion rate
    double sor_err;                   // abserror of the last SOR sweep pair
    int sor_state;                    // 0 while the Poisson solve iterates, 1 once it ended
    int sor_n;                        // sweeps done in the running solve
    int sor_iter;                     // the DO variable n when the solve ended
    int grav_stop;                    // 1: r_sh reached xF(0) (the reference STOPs)
    int pad2;
};

struct Axis {
    const double *xF;   // face coordinate, index j+3 (j=-3..n+3)
    const double *d;    // width, index j+3
    const double *c;    // centre, index j+3
    const double *w;    // reconstruction weights of the configured scheme, (j+3)*NW
    const double *wsel[3];   // TVD-MC, PPM and WENO-Z weights for the learned selector
};

// Reconstruction ids: 0 minmod, 1 van Leer, 2 MC, 3 PPM, 4 WENO-Z, and
// RSEL: per cell from the learned selector's class mask (0 MC, 1 PPM, 2 WENO-Z).
constexpr int RSEL = 5;

// Field pointers of one context.  W is the working type of the spatial
// operator: double in the FP64 mode; float in the FP32 split, where
// primitives, interface fluxes and EMFs are single precision and the face
// field, u_old and the conserved state U (the RK accumulators) stay double.
template <class W> struct ArgsT {
    int nx, ny, nz, ndim, coord, model, rt_gravity, bn_mode, fixpole;
    int bc[6];
    int bfac[6][12];        // [2*axis+side][var 1..11], SETUP_EQNS parity factors
    long NX, NY, G;
    double g1, cfl, r_sh;
    Axis ax[3];
    W *P[5];                // rho, vx, vy, vz, p
    W *eps, *cs;
    W *Bc[3];               // cell-centred field (bcell)
    double *Bin[3];         // face field at stage start (prim(ibx..ibz))
    double *Bout[3];        // face field written by the update
    double *Uo[8];          // u_old
    double *U[5];           // conserved hydro state (FP32 split only)
    W *F[3][7];             // interface fluxes: 5 hydro + 2 face EMFs
    W *E[3];                // edge EMFs efield_x/y/z
    W *Ec[3];               // cell EMFs ecell after the ghost fill, cells 0..n+1
    const double *rfl, *pfl, *efl;   // torus floors
    DevScalars *S;
    const unsigned char *mask[3];    // selector class per cell and sweep direction
    unsigned char *stale[3];         // HLL only: 1 where a sweep assigned no flux (see store_flux_hll)
    // spherical grids (getgrid.f90:68-98), index j+3 / k+3: dx_cb, dx_cb/3,
    // xbar, 3 x dx/dx_cb, 1.5 dx/dx_cb along r; sinf, dcose, dsine/dcose,
    // 0.5 dy/dcose, sine along theta.  Null on other geometries.
    const double *dcb, *dcb3, *xbar, *gsrc, *hbx;
    const double *sinf, *dcose, *cot, *hby, *sine;
    // collapsar: potential over the ghosted box, atmosphere prim_a(irho:ivz) and eps_a
    const double *phi;
    double pa[4], eps_a;
    // SPATIAL alone (k_update<0>): l_rk(1:8) in the host layout, var fastest
    double *lrk;
};

// Load field[idx] as scalar base + 32-bit byte offset (the global_load saddr
// form): one 32-bit register per address instead of a 64-bit pointer.  Every
// field is below 4 GiB (checked in cumc3d_init).
__device__ __forceinline__ double ldf(const double *__restrict__ base, long idx) {
    unsigned off = (unsigned)idx << 3;
    return *(const double *)((const char *)base + off);
}
__device__ __forceinline__ float ldf(const float *__restrict__ base, long idx) {
    unsigned off = (unsigned)idx << 2;
    return *(const float *)((const char *)base + off);
}

template <class A> __device__ __forceinline__ long cidx(const A &a, int j, int k, int l) {
    return (long)(j + 2) + a.NX * ((long)(k + 2) + a.NY * (long)(l + 2));
}

// Ghost-cell source along one axis (BOUNDARYP_NM, boundary.f90:423-745):
// flag 0 periodic, 1 outflow, >=2 parity mirror.  side: -1 interior, 0 inner, 1 outer.
// The source is clamped into 1..n: on an axis with fewer cells than ghost
// layers (the inactive axes of 1D/2D runs, n = 1) the reference's sequential
// fill copies ghost from ghost, and for periodic and outflow flags the chain
// ends at the same interior cell; reading that cell directly keeps a
// one-pass fill free of reads of ghosts it is writing.  (A mirror chain on
// such an axis reads the previous fill's outer ghost in the reference; those
// outer layers of an inactive axis feed no computation.)
__device__ __forceinline__ int gmap(int i, int n, int fin, int fout, int &side) {
    int r;
    if (i < 1) {
        side = 0;
        int m = 1 - i;
        r = fin == 0 ? n + 1 - m : (fin == 1 ? 1 : m);
    } else if (i > n) {
        side = 1;
        int m = i - n;
        r = fout == 0 ? m : (fout == 1 ? n : n + 1 - m);
    } else {
        side = -1;
        return i;
    }
    return r < 1 ? 1 : (r > n ? n : r);
}

// parity factor of variable var for a ghost on (axis, side); constant array
// indices only, so the kernel argument block is never copied to scratch
template <class A> __device__ __forceinline__ double gsign(const A &a, int axis, int side, int var) {
    if (side < 0) return 1.0;
    int f, b;
    if (axis == 0) { f = side ? a.bc[1] : a.bc[0]; b = side ? a.bfac[1][var] : a.bfac[0][var]; }
    else if (axis == 1) { f = side ? a.bc[3] : a.bc[2]; b = side ? a.bfac[3][var] : a.bfac[2][var]; }
    else { f = side ? a.bc[5] : a.bc[4]; b = side ? a.bfac[5][var] : a.bfac[4][var]; }
    return f >= 2 ? (double)b : 1.0;
}

// ---------------------------------------------------------------------------
// interface fluxes
// ---------------------------------------------------------------------------

// Serial position of interface (j,k,l) in the reference's Riemann loop
// (rsolver_module.f90:596-598: l outer, j inner; the sweep axis starts at 0,
// the others at 1), plus one.
template <int D, class A>
__device__ __forceinline__ unsigned long long serial_key(const A &a, int j, int k, int l) {
    if constexpr (D == 0) return ((unsigned long long)(l - 1) * a.ny + (k - 1)) * (a.nx + 1) + j + 1;
    else if constexpr (D == 1) return ((unsigned long long)(l - 1) * (a.ny + 1) + k) * a.nx + (j - 1) + 1;
    else return ((unsigned long long)l * a.ny + (k - 1)) * a.nx + (j - 1) + 1;
}

// 7-variable stencil of cell c along D: rho, vx, vy, vz, p and the two
// transverse cell-centred field components.
template <class T, int R, int D>
__device__ __forceinline__ void load_stencil(const ArgsT<T> &a, long c, long st, T s[7][5]) {
    constexpr int t1 = (D == 0) ? 1 : 0;
    constexpr int t2 = (D == 2) ? 1 : 2;
    const T *src[7] = {a.P[0], a.P[1], a.P[2], a.P[3], a.P[4], a.Bc[t1], a.Bc[t2]};
#pragma unroll
    for (int v = 0; v < 7; ++v) {
        if constexpr (R >= 3) {
            s[v][0] = (T)ldf(src[v], c - 2 * st);
            s[v][4] = (T)ldf(src[v], c + 2 * st);
        }
        s[v][1] = (T)ldf(src[v], c - st);
        s[v][2] = (T)ldf(src[v], c);
        s[v][3] = (T)ldf(src[v], c + st);
    }
}

template <class T> __device__ __forceinline__ void load_w(const double *w, T out[NW], int n) {
#pragma unroll
    for (int i = 0; i < NW; ++i)
        if (i < n) out[i] = (T)w[i];
}

template <int R> constexpr int nweights() { return R <= 2 ? 2 : (R == 3 ? 14 : 10); }

// Stencil width a reconstruction loads: 3 cells for TVD, 5 otherwise.
template <int R> constexpr int stencil_r() { return R == RSEL ? 3 : R; }

// Slide the stencil of the previous cell along the sweep to cell c: the
// values move down one place and only the cell entering the stencil is loaded.
template <class T, int R, int D>
__device__ __forceinline__ void slide_stencil(const ArgsT<T> &a, long c, long st, T s[7][5]) {
    constexpr int t1 = (D == 0) ? 1 : 0;
    constexpr int t2 = (D == 2) ? 1 : 2;
    const T *src[7] = {a.P[0], a.P[1], a.P[2], a.P[3], a.P[4], a.Bc[t1], a.Bc[t2]};
#pragma unroll
    for (int v = 0; v < 7; ++v) {
        if constexpr (stencil_r<R>() >= 3) {
            s[v][0] = s[v][1];
            s[v][1] = s[v][2];
            s[v][2] = s[v][3];
            s[v][3] = s[v][4];
            s[v][4] = (T)ldf(src[v], c + 2 * st);
        } else {
            s[v][1] = s[v][2];
            s[v][2] = s[v][3];
            s[v][3] = (T)ldf(src[v], c + st);
        }
    }
}

// Both faces of cell c (index i along D) from its loaded stencil.
template <class T, int R, int D>
__device__ __forceinline__ void recon_stencil(const ArgsT<T> &a, long c, int i, const T s[7][5], Face<T> &fm,
                                              Face<T> &fp) {
    if constexpr (R == RSEL) {
        // the highest class among the wave's active lanes, so the wave runs
        // one reconstruction (ballots see active lanes only)
        const int own = a.mask[D][c];
        const unsigned long long b2 = __ballot(own == 2), b1 = __ballot(own >= 1);
        const int cls = b2 ? 2 : (b1 ? 1 : 0);
        T w[NW];
        if (cls == 0) {
            load_w<T>(a.ax[D].wsel[0] + (long)(i + 3) * NW, w, 2);
            recon_cell<T, 2>(w, s, (T)a.g1, fm, fp);
        } else if (cls == 1) {
            load_w<T>(a.ax[D].wsel[1] + (long)(i + 3) * NW, w, 14);
            recon_cell<T, 3>(w, s, (T)a.g1, fm, fp);
        } else {
            load_w<T>(a.ax[D].wsel[2] + (long)(i + 3) * NW, w, 10);
            recon_cell<T, 4>(w, s, (T)a.g1, fm, fp);
        }
        // positivity fallback (the selector's own, no reference counterpart):
        // a PPM or WENO-Z face with non-positive density or pressure gives
        // the Riemann solver a NaN sound speed; such a cell takes its TVD-MC
        // faces instead, which stay within the neighbouring cell values
        if (cls > 0) {
            const bool ok = fm.rho > T(0) && fp.rho > T(0) && fm.p > T(0) && fp.p > T(0);
            if (__ballot(!ok)) {
                Face<T> gm, gp;
                load_w<T>(a.ax[D].wsel[0] + (long)(i + 3) * NW, w, 2);
                recon_cell<T, 2>(w, s, (T)a.g1, gm, gp);
                if (!ok) {
                    fm = gm;
                    fp = gp;
                }
            }
        }
    } else {
        T w[NW];
        load_w<T>(a.ax[D].w + (long)(i + 3) * NW, w, nweights<R>());
        recon_cell<T, R>(w, s, (T)a.g1, fm, fp);
    }
}

template <class T, int R, int D>
__device__ __forceinline__ void recon_at(const ArgsT<T> &a, int j, int k, int l, long st, Face<T> &fm, Face<T> &fp) {
    const long c = cidx(a, j, k, l);
    T s[7][5];
    load_stencil<T, stencil_r<R>(), D>(a, c, st, s);
    recon_stencil<T, R, D>(a, c, (D == 0) ? j : (D == 1 ? k : l), s, fm, fp);
}

template <int D, class A> __device__ __forceinline__ long stride(const A &a) {
    return D == 0 ? 1L : (D == 1 ? a.NX : a.NX * a.NY);
}

// Store the flux of interface (j,k,l) along D: 5 hydro fluxes and the two face
// EMFs of MHD_FLUX (mhd_module.f90:229-271).  Slot 5 holds the face EMF with
// the lower field index, slot 6 the other.
template <class T, int D>
__device__ __forceinline__ void store_flux(const ArgsT<T> &a, long c, const RiemannOut<T> &o) {
#pragma unroll
    for (int i = 0; i < 5; ++i) a.F[D][i][c] = o.f[i];
    if constexpr (D == 0) { a.F[0][5][c] = o.f[7]; a.F[0][6][c] = -o.f[6]; }
    else if constexpr (D == 1) { a.F[1][5][c] = -o.f[7]; a.F[1][6][c] = o.f[5]; }
    else { a.F[2][5][c] = o.f[6]; a.F[2][6][c] = -o.f[5]; }
}

// HLLNM's IF/ELSEIF chain (rsolver_module.f90:303-311) has no ELSE: with a
// NaN signal speed (a face state with negative pressure) no branch is taken
// and the reference's single flux array, shared by the three sweeps, keeps
// the last value any sweep wrote at that index.  The sweeps here have arrays
// of their own, so such a face takes what the shared array would hold: the
// flux of the previous sweep whose loop range contains the index -- x from z
// of the previous stage (from y in 2D), y from x, z from y; the lowest face of
// the sweep's own axis is in no other range and keeps its own value --
// converted between the component layouts of store_flux.  A sweep stores
// only the two transverse field fluxes; its normal-field flux is +0 where it
// assigned the face (fluxL(ibn) = 0, spatial.f90:381-382) and inherited from
// its own predecessor otherwise, which the stale flags record.  The oracle's
// single flux array does the same by construction.
template <class T, int D>
__device__ __forceinline__ void store_flux_hll(const ArgsT<T> &a, long c, const RiemannOut<T> &o, bool lowest) {
    a.stale[D][c] = o.none ? 1 : 0;
    if (!o.none) {
        store_flux<T, D>(a, c, o);
        return;
    }
    if (lowest) return;
    if constexpr (D == 0) {
        if (a.ndim == 1) return;
        T fbz, fby;
        const int p = a.ndim == 3 ? 2 : 1;
        if (a.ndim == 3) {
            fbz = a.stale[2][c] ? -a.F[1][5][c] : T(0);
            fby = a.F[2][5][c];
        } else {
            fbz = -a.F[1][5][c];
            fby = a.stale[1][c] ? -a.F[0][6][c] : T(0);
        }
#pragma unroll
        for (int i = 0; i < 5; ++i) a.F[0][i][c] = a.F[p][i][c];
        a.F[0][5][c] = fbz;
        a.F[0][6][c] = -fby;
    } else if constexpr (D == 1) {
        T fbz = a.F[0][5][c];
        T fbx = a.stale[0][c] ? (a.ndim == 3 ? -a.F[2][6][c] : a.F[1][6][c]) : T(0);
#pragma unroll
        for (int i = 0; i < 5; ++i) a.F[1][i][c] = a.F[0][i][c];
        a.F[1][5][c] = -fbz;
        a.F[1][6][c] = fbx;
    } else {
        T fbx = a.F[1][6][c];
        T fby = a.stale[1][c] ? -a.F[0][6][c] : T(0);
#pragma unroll
        for (int i = 0; i < 5; ++i) a.F[2][i][c] = a.F[1][i][c];
        a.F[2][5][c] = fby;
        a.F[2][6][c] = -fbx;
    }
}

template <class T, int D, int S>
__device__ __forceinline__ void store_face(const ArgsT<T> &a, long c, const RiemannOut<T> &o, bool lowest) {
    if constexpr (S == 1) store_flux_hll<T, D>(a, c, o, lowest);
    else store_flux<T, D>(a, c, o);
}

__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long v) {
    __shared__ unsigned long long red[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_down(v, off, 64);
        v = o > v ? o : v;
    }
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        for (int i = 1; i < nw; ++i) v = red[i] > v ? red[i] : v;
    }
    return v;
}

// x sweep.  The (nx+2) cells j=0..nx+1 of every row are laid end to end; each
// 64-lane wave reconstructs 64 consecutive cells and solves the 63 interfaces
// between them, taking the right neighbour's left face with a lane shuffle.
// Waves overlap by one cell, so every cell is reconstructed about 64/63 times.
// Blocks take the rows from the last to the first: the first blocks hold the
// largest serial keys of the HLLD carry, and a later block whose key does not
// exceed the value already folded skips its atomic.  (In order, nearly every
// one of the ~68k blocks of a 256^3 sweep with many double-star interfaces
// raised the maximum, and the atomics on the one address serialised: FP32
// TVD-MC x sweep 0.79 ms against 0.39 ms for the z sweep on turbulence.)
// The single-precision selector sweep is held at 5 waves per SIMD (86
// VGPRs, 0.41 ms at 256^3).  Left to itself the compiler gives it 112 VGPRs
// and 4 waves (0.47 ms); at 6 waves it spills (0.44 ms).
template <class T, int R, int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && R == RSEL ? 5 : 1)))
void k_flux_x(ArgsT<T> a) {
    if (!a.S->active) return;
    const int lane = threadIdx.x & 63;
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    const long wave = nwaves - 1 - (((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const long g = wave * 63 + lane;
    const long RL = a.nx + 2;
    const long gtot = RL * a.ny * a.nz;
    const bool live = g < gtot;
    int j = 0, k = 1, l = 1;
    if (live) {
        long r = g / RL;
        j = (int)(g - r * RL);
        k = 1 + (int)(r % a.ny);
        l = 1 + (int)(r / a.ny);
    }
    Face<T> fm, fp;
    if (live) recon_at<T, R, 0>(a, j, k, l, 1, fm, fp);
    Face<T> nm;
    nm.rho = __shfl_down(fm.rho, 1, 64);
    nm.v[0] = __shfl_down(fm.v[0], 1, 64);
    nm.v[1] = __shfl_down(fm.v[1], 1, 64);
    nm.v[2] = __shfl_down(fm.v[2], 1, 64);
    nm.p = __shfl_down(fm.p, 1, 64);
    nm.bt1 = __shfl_down(fm.bt1, 1, 64);
    nm.bt2 = __shfl_down(fm.bt2, 1, 64);
    nm.eps = __shfl_down(fm.eps, 1, 64);
    nm.cs = __shfl_down(fm.cs, 1, 64);
    unsigned long long key = 0;
    if (live && lane < 63 && j <= a.nx && g + 1 < gtot) {
        long c = cidx(a, j, k, l);
        T bn = (T)a.Bin[0][c];
        Q<T> qL = to_q<T, 0>(fp, bn), qR = to_q<T, 0>(nm, bn);
        RiemannOut<T> o;
        riemann<T, 0, S>(qL, qR, fp.eps, nm.eps, fp.cs, nm.cs, a.bn_mode, (T)a.S->carry[0], o);
        store_face<T, 0, S>(a, c, o, j == 0);
        if (o.ds) key = serial_key<0>(a, j, k, l);
    }
    if constexpr (S == 3) {
        if (a.bn_mode == 2) {
            key = block_max_u64(key);
            if (threadIdx.x == 0 && key &&
                key > __hip_atomic_load(&a.S->carry_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(&a.S->carry_key, key);
        }
    }
}

// y and z sweeps.  Threads run along x (coalesced); each thread marches along
// the sweep axis over a segment of interfaces [i0,i1), reconstructing cells
// i0..i1 once and keeping the previous cell's right face in registers.
template <class T, int R, int S, int D>
__global__ __launch_bounds__(256) void k_flux_yz(ArgsT<T> a, int nseg) {
    if (!a.S->active) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int na = (D == 1) ? a.nz : a.ny;    // the other transverse axis
    const int nd = (D == 1) ? a.ny : a.nz;    // sweep axis
    const long total = (long)a.nx * na * nseg;
    unsigned long long key = 0;
    if (t < total) {
        int j = 1 + (int)(t % a.nx);
        long rest = t / a.nx;
        int o2 = 1 + (int)(rest % na);
        int s = (int)(rest / na);
        int i0 = (int)(((long)s * (nd + 1)) / nseg);
        int i1 = (int)(((long)(s + 1) * (nd + 1)) / nseg);
        const long st = stride<D>(a);
        Face<T> prev, fm;
        // with the selector (single precision) the stencil stays in registers
        // and slides one cell per step: one new load per field instead of
        // five.  The fixed-scheme kernels reload it from cache; the window's
        // registers cost them more occupancy than the loads they save
        // (measured on MI355X at 256^3)
        constexpr bool slide = sizeof(T) == 4 && R == RSEL;
        T win[slide ? 7 : 1][5];
        long cc = cidx(a, j, (D == 1) ? i0 : o2, (D == 1) ? o2 : i0);
        if constexpr (slide) {
            load_stencil<T, stencil_r<R>(), D>(a, cc, st, win);
            recon_stencil<T, R, D>(a, cc, i0, win, fm, prev);
        } else {
            recon_at<T, R, D>(a, j, (D == 1) ? i0 : o2, (D == 1) ? o2 : i0, st, fm, prev);
        }
        const T carry = (T)a.S->carry[D];
#pragma unroll 1
        for (int i = i0 + 1; i <= i1; ++i) {
            int k = (D == 1) ? i : o2;
            int l = (D == 1) ? o2 : i;
            Face<T> fp;
            cc += st;
            if constexpr (slide) {
                slide_stencil<T, R, D>(a, cc, st, win);
                recon_stencil<T, R, D>(a, cc, i, win, fm, fp);
            } else {
                recon_at<T, R, D>(a, j, k, l, st, fm, fp);
            }
            int kk = (D == 1) ? i - 1 : k;
            int ll = (D == 1) ? l : i - 1;
            long c = cidx(a, j, kk, ll);
            T bn = (T)a.Bin[D][c];
            Q<T> qL = to_q<T, D>(prev, bn), qR = to_q<T, D>(fm, bn);
            RiemannOut<T> o;
            riemann<T, D, S>(qL, qR, prev.eps, fm.eps, prev.cs, fm.cs, a.bn_mode, carry, o);
            store_face<T, D, S>(a, c, o, i == 1);
            if (o.ds) {
                unsigned long long kk2 = serial_key<D>(a, j, kk, ll);
                key = kk2 > key ? kk2 : key;
            }
            prev = fp;
        }
    }
    if constexpr (S == 3) {
        if (a.bn_mode == 2) {
            key = block_max_u64(key);
            if (threadIdx.x == 0 && key) atomicMax(&a.S->carry_key, key);
        }
    }
}

// After a sweep: find the last double-star interface in the reference's serial
// loop order and recompute its double-star transverse field, which the serial
// reference leaves behind in usstarL/usstarR and reads back as usstar(ibn) in
// later sweeps (rsolver_module.f90:704-741).
template <class T, int R, int D>
__global__ __launch_bounds__(64) void k_carry(ArgsT<T> a) {
    if (!a.S->active || threadIdx.x != 0) return;
    const unsigned long long key = a.S->carry_key;
    a.S->carry_key = 0;     // the next sweep starts from no double-star interface
    if (key == 0) return;
    unsigned long long q = key - 1;
    int j, k, l;
    if constexpr (D == 0) {
        j = (int)(q % (a.nx + 1)); q /= (a.nx + 1);
        k = 1 + (int)(q % a.ny); l = 1 + (int)(q / a.ny);
    } else if constexpr (D == 1) {
        j = 1 + (int)(q % a.nx); q /= a.nx;
        k = (int)(q % (a.ny + 1)); l = 1 + (int)(q / (a.ny + 1));
    } else {
        j = 1 + (int)(q % a.nx); q /= a.nx;
        k = 1 + (int)(q % a.ny); l = (int)(q / a.ny);
    }
    const long st = stride<D>(a);
    Face<T> lm, lp, rm, rp;
    recon_at<T, R, D>(a, j, k, l, st, lm, lp);
    recon_at<T, R, D>(a, j + (D == 0), k + (D == 1), l + (D == 2), st, rm, rp);
    T bn = (T)a.Bin[D][cidx(a, j, k, l)];
    Q<T> qL = to_q<T, D>(lp, bn), qR = to_q<T, D>(rm, bn);
    RiemannOut<T> o;
    riemann<T, D, 3>(qL, qR, lp.eps, rm.eps, lp.cs, rm.cs, a.bn_mode, (T)a.S->carry[D], o);
    a.S->carry[(D + 1) % 3] = (double)o.uss_t1;
    a.S->carry[(D + 2) % 3] = (double)o.uss_t2;
}

// ---------------------------------------------------------------------------
// constrained transport: edge EMFs (flux_ct, mhd_module.f90:310-426)
// ---------------------------------------------------------------------------

// cell-centred EMF -v x B of an interior cell (mhd_module.f90:316-318)
template <int E, class W> __device__ __forceinline__ W ecell_raw(const ArgsT<W> &a, long c) {
    W vx = a.P[1][c], vy = a.P[2][c], vz = a.P[3][c];
    W bx = a.Bc[0][c], by = a.Bc[1][c], bz = a.Bc[2][c];
    if constexpr (E == 0) return vz * by - vy * bz;
    else if constexpr (E == 1) return vx * bz - vz * bx;
    else return vy * bx - vx * by;
}

// Store the cell EMFs of cell c from its velocity and cell-centred field
// (times a parity factor per component for mapped ghost cells).
template <class W>
__device__ __forceinline__ void cell_emf(const ArgsT<W> &a, long c, W vx, W vy, W vz, W bx, W by, W bz, W s0 = W(1),
                                         W s1 = W(1), W s2 = W(1)) {
    a.Ec[0][c] = s0 * (vz * by - vy * bz);
    a.Ec[1][c] = s1 * (vx * bz - vz * bx);
    a.Ec[2][c] = s2 * (vy * bx - vx * by);
}

// ecell after BOUNDARY_EFIELD (mhd_module.f90:478-840): ghost cells take the
// mapped interior value times the parity factors of component E.
template <int E, class W> __device__ __forceinline__ W ecell_at(const ArgsT<W> &a, int j, int k, int l) {
    int sx, sy, sz;
    int jj = gmap(j, a.nx, a.bc[0], a.bc[1], sx);
    int kk = gmap(k, a.ny, a.bc[2], a.bc[3], sy);
    int ll = gmap(l, a.nz, a.bc[4], a.bc[5], sz);
    W s = (W)(gsign(a, 0, sx, 9 + E) * gsign(a, 1, sy, 9 + E) * gsign(a, 2, sz, 9 + E));
    W v = ecell_raw<E>(a, cidx(a, jj, kk, ll));
    return (sx < 0 && sy < 0 && sz < 0) ? v : s * v;
}

// Face EMF eface of component E at faces normal to D, after the 1D/2D fill
// (mhd_module.f90:328-356) and BOUNDARY_EFIELD: face index along D is real,
// the two cell-centred axes are ghost-mapped.
template <int E, int D, class W> __device__ __forceinline__ W eface_at(const ArgsT<W> &a, int j, int k, int l) {
    int sx = -1, sy = -1, sz = -1;
    if constexpr (D != 0) j = gmap(j, a.nx, a.bc[0], a.bc[1], sx);
    if constexpr (D != 1) k = gmap(k, a.ny, a.bc[2], a.bc[3], sy);
    if constexpr (D != 2) l = gmap(l, a.nz, a.bc[4], a.bc[5], sz);
    W v;
    if (D < a.ndim) {
        constexpr int slot = (D == 0) ? (E == 1 ? 5 : 6) : (D == 1 ? (E == 0 ? 5 : 6) : (E == 0 ? 5 : 6));
        v = a.F[D][slot][cidx(a, j, k, l)];
    } else {
        if constexpr (D == 0) j = 1;
        else if constexpr (D == 1) k = 1;
        else l = 1;
        v = ecell_raw<E>(a, cidx(a, j, k, l));
    }
    if (sx < 0 && sy < 0 && sz < 0) return v;
    return (W)(gsign(a, 0, sx, 9 + E) * gsign(a, 1, sy, 9 + E) * gsign(a, 2, sz, 9 + E)) * v;
}

// Face EMFs on the cells 0..n+1 the edge stencil reads (flux_ct,
// mhd_module.f90:310-360, ghosts mapped as in BOUNDARY_EFIELD): the ghost rows
// along their two cell-centred axes plus the faces of inactive directions
// (1D/2D fill).  The cell EMFs (ecell) are stored by the update for interior
// cells and by the ghost fill (mapped) for ghost cells.
// Cells of the (n+2)^3 box 0..n+1 the face-EMF fill visits: in 3D only its
// outer shell (a ghost index on some axis), in 1D/2D every cell (the
// inactive directions' faces are filled everywhere).
template <class A> __host__ __device__ __forceinline__ long emf_prep_count(const A &a) {
    const long m = (long)(a.nx + 2) * (a.ny + 2) * (a.nz + 2);
    return a.ndim == 3 ? m - (long)a.nx * a.ny * a.nz : m;
}

template <class W> __global__ __launch_bounds__(256) void k_emf_prep(ArgsT<W> a) {
    if (!a.S->active) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= emf_prep_count(a)) return;
    int j, k, l;
    if (a.ndim == 3) {
        // shell of the box: the two z planes, the two y rows of each inner
        // plane, the two x cells of each inner row
        const long n0 = a.nx + 2, n1 = a.ny + 2, pz = 2 * n0 * n1, py = 2 * n0 * (long)a.nz;
        long q = t;
        if (q < pz) {
            l = (q / (n0 * n1)) == 0 ? 0 : a.nz + 1;
            k = (int)((q % (n0 * n1)) / n0);
            j = (int)(q % n0);
        } else if (q < pz + py) {
            q -= pz;
            l = 1 + (int)(q / (2 * n0));
            k = ((q % (2 * n0)) / n0) == 0 ? 0 : a.ny + 1;
            j = (int)(q % n0);
        } else {
            q -= pz + py;
            l = 1 + (int)(q / (2L * a.ny));
            k = 1 + (int)((q % (2L * a.ny)) / 2);
            j = (q % 2) == 0 ? 0 : a.nx + 1;
        }
    } else {
        const long n0 = a.nx + 2, n1 = a.ny + 2;
        j = (int)(t % n0);
        k = (int)((t / n0) % n1);
        l = (int)(t / (n0 * n1));
    }
    const long c = cidx(a, j, k, l);
    const bool gx = j < 1 || j > a.nx, gy = k < 1 || k > a.ny, gz = l < 1 || l > a.nz;
    // x faces (face index j <= nx), cell axes y, z
    if (j <= a.nx && (a.ndim < 1 || gy || gz)) {
        a.F[0][5][c] = eface_at<1, 0>(a, j, k, l);
        a.F[0][6][c] = eface_at<2, 0>(a, j, k, l);
    }
    if (k <= a.ny && (a.ndim < 2 || gx || gz)) {
        a.F[1][5][c] = eface_at<0, 1>(a, j, k, l);
        a.F[1][6][c] = eface_at<2, 1>(a, j, k, l);
    }
    if (l <= a.nz && (a.ndim < 3 || gx || gy)) {
        a.F[2][5][c] = eface_at<0, 2>(a, j, k, l);
        a.F[2][6][c] = eface_at<1, 2>(a, j, k, l);
    }
}

// Edge EMFs (mhd_module.f90:414-423) on the edges the CT update reads:
// efield_x for j=1..nx, efield_y for k=1..ny, efield_z for l=1..nz
// (0..n along the other axes), from the arrays k_emf_prep completed.
template <class W> __global__ __launch_bounds__(256) void k_edge(ArgsT<W> a) {
    if (!a.S->active) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n1 = a.nx + 1, n2 = a.ny + 1;
    if (t >= n1 * n2 * (a.nz + 1)) return;
    const int j = (int)(t % n1), k = (int)((t / n1) % n2), l = (int)(t / (n1 * n2));
    const long c = cidx(a, j, k, l), sy = a.NX, sz = a.NX * a.NY;
    const W *ex = a.Ec[0], *ey = a.Ec[1], *ez = a.Ec[2];
    const W h = W(0.50), q = W(0.25);
    // the three EMFs are formed before any is stored, so that all their loads
    // are in flight together
    W e0 = W(0), e1 = W(0), e2 = W(0);
    if (j >= 1) {
        const W *fxy = a.F[1][5], *fxz = a.F[2][5];
        e0 = h * (fxy[c] + fxy[c + sz] + fxz[c] + fxz[c + sy]) - q * (ex[c] + ex[c + sz] + ex[c + sy] + ex[c + sy + sz]);
    }
    if (k >= 1) {
        const W *fyx = a.F[0][5], *fyz = a.F[2][6];
        e1 = h * (fyx[c] + fyx[c + sz] + fyz[c] + fyz[c + 1]) - q * (ey[c] + ey[c + sz] + ey[c + 1] + ey[c + 1 + sz]);
    }
    if (l >= 1) {
        const W *fzx = a.F[0][6], *fzy = a.F[1][6];
        e2 = h * (fzx[c] + fzx[c + sy] + fzy[c] + fzy[c + 1]) - q * (ez[c] + ez[c + sy] + ez[c + 1] + ez[c + 1 + sy]);
    }
    if (j >= 1) a.E[0][c] = e0;
    if (k >= 1) a.E[1][c] = e1;
    if (l >= 1) a.E[2][c] = e2;
}

// FIXPOLE (mhd_module.f90:437-461): the radial edge EMF on each polar face
// (theta faces 0 and ny) becomes its average over phi, written to every phi
// edge l=-2..nz+3.  The reference's accumulators sum1/sum2 are REAL (:293):
// each addition is done in double and rounded to single, in the order of the
// l loop of a one-thread run.  One wave per (radius j, pole): the lanes load
// the ring into LDS, every lane runs the serial sum from it.
template <class W> __global__ __launch_bounds__(64) void k_fixpole(ArgsT<W> a) {
    if (!a.S->active) return;
    constexpr int CH = 1024;
    __shared__ W buf[CH];
    const int j = 1 + (blockIdx.x >> 1), k = (blockIdx.x & 1) ? a.ny : 0;
    W *ex = a.E[0];
    float sum = 0.0f;
    for (int l0 = 1; l0 <= a.nz; l0 += CH) {
        const int m = a.nz - l0 + 1 < CH ? a.nz - l0 + 1 : CH;
        for (int i = threadIdx.x; i < m; i += 64) buf[i] = ex[cidx(a, j, k, l0 + i)];   // all loads in flight
        __syncthreads();
        int i = 0;
        for (; i + 8 <= m; i += 8) {
            W r[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) r[u] = buf[i + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) sum = (float)((double)sum + (double)r[u]);
        }
        for (; i < m; ++i) sum = (float)((double)sum + (double)buf[i]);
        __syncthreads();
    }
    const W avg = (W)((double)sum / (double)a.nz);
    for (int l = -2 + (int)threadIdx.x; l <= a.nz + 3; l += 64) ex[cidx(a, j, k, l)] = avg;
}

// ---------------------------------------------------------------------------
// update: FLUX_DIFF + sources + l_operator + RK combine + FROMUTORVE +
// CUSTOM_CHECKRHO + FROMRVETOU + FINDPRESSURE (+ FINDDT) for one interior cell
// ---------------------------------------------------------------------------

// l_rk of the face fields (l_operator, spatial.f90:639-702) from the four
// edge EMFs around the face (read in the working type, differenced in double):
// x face jf: ez at (jf,k,l),(jf,k-1,l), ey at (jf,k,l),(jf,k,l-1);
// y face kf: ex at (j,kf,l),(j,kf,l-1), ez at (j,kf,l),(j-1,kf,l);
// z face lf: ey at (j,k,lf),(j-1,k,lf), ex at (j,k,lf),(j,k-1,lf).
// parameter.h:24 small_num = TINY(1.0D0), the smallest normal double
constexpr double SMALL_NUM = 2.2250738585072014e-308;

// G: geometry (coordinate_flag), a template parameter so that each grid type
// compiles only its own terms (the update kernel's registers depend on it).
template <int G, class A>
__device__ __forceinline__ double lB_x(const A &a, int jf, int k, int l, double ez0, double ez1, double ey0, double ey1) {
    double dy = a.ax[1].d[k + 3], dz = a.ax[2].d[l + 3];
    if constexpr (G == 0) return -((ez0 - ez1) / (dy) - (ey0 - ey1) / (dz));
    double xF = a.ax[0].xF[jf + 3];
    if constexpr (G == 2)   // spatial.f90:689
        return (-(a.sinf[k + 3] * ez0 - a.sinf[k + 2] * ez1) + (ey0 - ey1) * (dy / dz)) / (xF * a.dcose[k + 3] + SMALL_NUM);
    return -(ez0 - ez1) / (xF * dy) + (ey0 - ey1) / (dz);
}
template <int G, class A>
__device__ __forceinline__ double lB_y(const A &a, int j, int kf, int l, double ex0, double ex1, double ez0, double ez1) {
    double dx = a.ax[0].d[j + 3], dz = a.ax[2].d[l + 3];
    if constexpr (G == 0) return -((ex0 - ex1) / (dz) - (ez0 - ez1) / (dx));
    if constexpr (G == 2) {   // spatial.f90:692
        double xF = a.ax[0].xF[j + 3], xFm = a.ax[0].xF[j + 2], x = a.ax[0].c[j + 3];
        return (-(ex0 - ex1) / (a.xbar[j + 3] * a.sinf[kf + 3] * dz + SMALL_NUM) + (xF * ez0 - xFm * ez1) / (x * dx));
    }
    return -(ex0 - ex1) / (dz) + (ez0 - ez1) / (dx);
}
template <int G, class A>
__device__ __forceinline__ double lB_z(const A &a, int j, int k, double ey0, double ey1, double ex0, double ex1) {
    double dx = a.ax[0].d[j + 3], dy = a.ax[1].d[k + 3];
    if constexpr (G == 0) return -((ey0 - ey1) / (dx) - (ex0 - ex1) / (dy));
    double xF = a.ax[0].xF[j + 3], xFm = a.ax[0].xF[j + 2], x = a.ax[0].c[j + 3];
    if constexpr (G == 2)   // spatial.f90:695
        return (-(xF * ey0 - xFm * ey1) / (x * dx) + (ex0 - ex1) / (a.xbar[j + 3] * dy));
    return -(xF * ey0 - xFm * ey1) / (x * dx) + (ex0 - ex1) / (x * dy);
}

// first_derivative of model/collapsar/src/custom.f90:402-409: the
// three-point derivative at xc on a non-uniform grid
__device__ __forceinline__ double first_derivative(double xm1, double xc, double xp1, double fm1, double fc, double fp1) {
    double h2 = xp1 - xc, h1 = xc - xm1;
    return ((fp1 - fc) * h1 * h1 + (fc - fm1) * h2 * h2) / (h1 * h2 * (h1 + h2));
}

template <int STAGE>
__device__ __forceinline__ double rk(double uo, double u, double dt, double L) {
    if constexpr (STAGE == 1) return uo + dt * L;
    else if constexpr (STAGE == 2) return (3.0 / 4.0) * uo + (1.0 / 4.0) * u + (1.0 / 4.0) * dt * L;
    else return (1.0 / 3.0) * uo + (2.0 / 3.0) * u + (2.0 / 3.0) * dt * L;
}

// CUSTOM_CHECKRHO of the model (model/riemann/src/custom.f90:115-121,
// model/torus-cyn/src/custom.f90:254-271, model/collapsar/src/custom.f90:295-320)
template <class A> __device__ __forceinline__ void checkrho(const A &a, long c, double &rho, double v[3], double &eps) {
    if (a.model == 0) {
        eps = fmax2(eps, 0.0);
    } else if (a.model == 3) {   // reset to the atmosphere prim_a, eps_a
        double diff = rho - a.pa[0];
        double f = fmax2(fsign1(diff), 0.0);
        rho = f * rho + (1.0 - f) * a.pa[0];
        v[0] = f * v[0] + (1.0 - f) * a.pa[1];
        v[1] = f * v[1] + (1.0 - f) * a.pa[2];
        v[2] = f * v[2] + (1.0 - f) * a.pa[3];
        eps = f * eps + (1.0 - f) * a.eps_a;
        if (eps < 0.0) eps = a.eps_a;
    } else {
        double diff = rho - a.rfl[c];
        double f = fmax2(fsign1(diff), 0.0);
        rho = f * rho + (1.0 - f) * a.rfl[c];
        v[0] = f * v[0]; v[1] = f * v[1]; v[2] = f * v[2];
        eps = f * eps + (1.0 - f) * a.efl[c];
        if (eps < 0.0) eps = a.efl[c];
    }
}

// FINDPRESSURE of the model (model/riemann/src/findpressure.f90:17-28,
// model/torus-cyn/src/findpressure.f90:18-32; the collapsar's,
// model/collapsar/src/findpressure.f90:27-38, is the ideal gas of riemann's)
template <class A> __device__ __forceinline__ void pressure(const A &a, long c, double rho, double eps, double &p, double &cs) {
    double dpdrho, dpdeps;
    if (a.model == 0 || a.model == 3 || rho > a.rfl[c]) {
        p = rho * eps * a.g1;
        dpdrho = eps * a.g1;
        dpdeps = rho * a.g1;
    } else {
        p = a.pfl[c];
        dpdrho = a.efl[c] * a.g1;
        dpdeps = a.rfl[c] * a.g1;
    }
    cs = sqrtT(dpdrho + dpdeps * p / (rho * rho));
}

// FINDDT per cell (rungekutta.f90:272-311)
template <int G, class A>
__device__ __forceinline__ double cfl_dt(const A &a, int j, int k, int l, double cs, double rho, const double v[3],
                                         const double b[3]) {
    double a2 = cs * cs, a4 = a2 * a2;
    double b2x = (b[0] * b[0] / rho), b2y = (b[1] * b[1] / rho), b2z = (b[2] * b[2] / rho);
    double b2 = b2x + b2y + b2z, b4 = b2 * b2;
    double cfx = sqrtT(0.5 * (a2 + b2 + sqrtT((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2x)));
    double cfy = sqrtT(0.5 * (a2 + b2 + sqrtT((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2y)));
    double cfz = sqrtT(0.5 * (a2 + b2 + sqrtT((a4 + 2.0 * a2 * b2 + b4) - 4.0 * a2 * b2z)));
    double lam = fmax2(fmax2(fabsT(v[0]) + cfx, fabsT(v[1]) + cfy), fabsT(v[2]) + cfz);
    double h = a.ax[0].d[j + 3];
    if constexpr (G == 0) {
        if (a.ndim > 1) h = fmin2(h, a.ax[1].d[k + 3]);
        if (a.ndim > 2) h = fmin2(h, a.ax[2].d[l + 3]);
    } else if constexpr (G == 1) {
        if (a.ndim > 1 && a.ny > 1) h = fmin2(h, a.ax[0].c[j + 3] * a.ax[1].d[k + 3]);
        if (a.ndim > 2) h = fmin2(h, a.ax[2].d[l + 3]);
    } else {
        if (a.ndim > 1) h = fmin2(h, a.ax[0].c[j + 3] * a.ax[1].d[k + 3]);
        if (a.ndim > 2) h = fmin2(h, a.ax[0].c[j + 3] * a.sine[k + 3] * a.ax[2].d[l + 3]);
    }
    return h * a.cfl / lam;
}

// Block minimum of per-cell time steps.  A NaN or non-positive candidate
// (a cell with NaN or negative energy) raises S->nonfinite and is left out.
__device__ __forceinline__ double block_min_f64(double v, DevScalars *S) {
    if (!(v > 0.0)) {
        atomicOr(&S->nonfinite, 1);
        v = 1.0e300;
    }
    __shared__ double red[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_down(v, off, 64);
        v = fmin2(v, o);
    }
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        for (int i = 1; i < nw; ++i) v = fmin2(v, red[i]);
    }
    return v;
}

// geometry-weighted cell average of face field component D (fromrvetou.f90:153-195)
template <int D, int G, class A>
__device__ __forceinline__ double bavg(const A &a, int j, int k, double up, double lo) {
    if constexpr (D == 0 && G == 1) {
        double xF = a.ax[0].xF[j + 3], xFm = a.ax[0].xF[j + 2];
        return (xF * up + xFm * lo) / (xF + xFm);
    }
    if constexpr (D == 0 && G == 2) {
        double xF = a.ax[0].xF[j + 3], xFm = a.ax[0].xF[j + 2];
        return a.hbx[j + 3] * (xF * xF * up + xFm * xFm * lo);
    }
    if constexpr (D == 1 && G == 2) return a.hby[k + 3] * (a.sinf[k + 3] * up + a.sinf[k + 2] * lo);
    return 0.5 * (up + lo);
}

// the same with the geometry chosen at run time (kernels that run once per step or less)
template <int D, class A>
__device__ __forceinline__ double bavg_rt(const A &a, int j, int k, double up, double lo) {
    if (a.coord == 1) return bavg<D, 1>(a, j, k, up, lo);
    if (a.coord == 2) return bavg<D, 2>(a, j, k, up, lo);
    return bavg<D, 0>(a, j, k, up, lo);
}
template <class A>
__device__ __forceinline__ double cfl_dt_rt(const A &a, int j, int k, int l, double cs, double rho, const double v[3],
                                            const double b[3]) {
    if (a.coord == 1) return cfl_dt<1>(a, j, k, l, cs, rho, v, b);
    if (a.coord == 2) return cfl_dt<2>(a, j, k, l, cs, rho, v, b);
    return cfl_dt<0>(a, j, k, l, cs, rho, v, b);
}

// Stores of the update's full-row double-precision outputs (the new face
// field, the hydro u_old, the FP32 split's conserved state), which no thread
// of the same launch reads: non-temporal, so they do not displace the
// neighbouring rows and planes the update reads from L2.  (Measured on
// MI355X at 256^3: update 1.15 -> 1.12 ms and the ghost fill after it
// 0.109 -> 0.100 ms in the FP32 split.  The single-precision outputs and the
// u_old face stores stay ordinary: non-temporal, they were slower.)
__device__ __forceinline__ void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
// The primitives, cell field and cell EMFs the update writes: non-temporal in
// the FP64 mode (update 1.57 -> 1.53 ms at 256^3), ordinary in single
// precision, where non-temporal was slower.
template <class W> __device__ __forceinline__ void st_w(W *p, W v) {
    if constexpr (sizeof(W) == 8) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// One interior cell: FLUX_DIFF + sources + l_operator + RK combine +
// FROMUTORVE + CUSTOM_CHECKRHO + FROMRVETOU + FINDPRESSURE; returns the
// cell's CFL time step in stage 3.  STAGE 0 is SPATIAL alone: it stores the
// cell's l_rk (and that of its lowest faces) in a.lrk and changes no state.
template <int STAGE, class W, int G>
__device__ __forceinline__ double update_cell(const ArgsT<W> &a, int j, int k, int l, double dt) {
    constexpr bool split = sizeof(W) == 4;   // FP32 split: U holds the conserved state
    const long c = cidx(a, j, k, l), sy = a.NX, sz = a.NX * a.NY;
    // stage-start primitives: FROMRVETOU input (fromrvetou.f90:94-108) in the
    // FP64 mode, source terms only in the FP32 split
    double rho = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
    if (!split || G != 0 || a.model != 0 || a.rt_gravity) {
        rho = a.P[0][c]; vx = a.P[1][c]; vy = a.P[2][c]; vz = a.P[3][c];
        bx = a.Bc[0][c]; by = a.Bc[1][c]; bz = a.Bc[2][c];
    }
    // FP32 split: U holds the conserved state at step start and Uo[0..4] the
    // stage results in between -- stage 1 reads U and writes Uo, stage 2
    // reads both and writes Uo, stage 3 reads both and writes U -- so u_old
    // is never copied.  FP64: the state is carried as primitives, and stage 1
    // saves u_old in Uo.
    double u[5], uo[5];
    if constexpr (split) {
        if (STAGE == 1) {
            for (int i = 0; i < 5; ++i) uo[i] = u[i] = a.U[i][c];
        } else {
            for (int i = 0; i < 5; ++i) { uo[i] = a.U[i][c]; u[i] = a.Uo[i][c]; }
        }
    } else {
        double eps = a.eps[c];
        double vsq = vx * vx + vy * vy + vz * vz, bsq = bx * bx + by * by + bz * bz;
        u[0] = rho;
        u[1] = vx * rho;
        u[2] = vy * rho;
        u[3] = vz * rho;
        u[4] = rho * (eps + 0.5 * vsq) + 0.5 * bsq;
        if (STAGE == 1) {
            for (int i = 0; i < 5; ++i) { uo[i] = u[i]; st_nt(&a.Uo[i][c], u[i]); }
        } else {
            for (int i = 0; i < 5; ++i) uo[i] = a.Uo[i][c];
        }
    }
    // sources: GET_SOURCE (spatial.f90:195-226) + CUSTOM_SOURCE
    double sc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if constexpr (G == 1) {
        double bsq = bx * bx + by * by + bz * bz, xj = a.ax[0].c[j + 3];
        sc[1] = sc[1] + (a.P[4][c] + rho * vy * vy + 0.5 * bsq - by * by) / (xj);
        sc[2] = sc[2] - (rho * vx * vy - bx * by) / (xj);
    } else if constexpr (G == 2) {   // spatial.f90:227-256
        double bsq = bx * bx + by * by + bz * bz, p = a.P[4][c];
        double g = a.gsrc[j + 3], cot = a.cot[k + 3];
        sc[1] = sc[1] + (2.0 * p + bx * bx + rho * (vy * vy + vz * vz)) * g;
        sc[2] = sc[2] + (p + rho * vz * vz + 0.5 * bsq - bz * bz) * g * cot - (rho * vx * vy - bx * by) * g;
        sc[3] = sc[3] - (rho * vx * vz - bx * bz) * g - (rho * vy * vz - by * bz) * g * cot;
    }
    if (G == 2 && a.model == 3) {   // model/collapsar/src/custom.f90:361-387
        // the matter potential's three-point gradient on the non-uniform
        // grid (first_derivative, :402-409) plus the black hole's pull
        double diff = rho - a.pa[0];
        double f = fmax2(fsign1(diff), 0.0);
        const double *phi = a.phi, *xc = a.ax[0].c, *yc = a.ax[1].c, *zc = a.ax[2].c;
        const double pc = phi[c];
        double dphidx = first_derivative(xc[j + 2], xc[j + 3], xc[j + 4], phi[c - 1], pc, phi[c + 1]);
        double dphidy = first_derivative(yc[k + 2], yc[k + 3], yc[k + 4], phi[c - sy], pc, phi[c + sy]);
        double dphidz = first_derivative(zc[l + 2], zc[l + 3], zc[l + 4], phi[c - sz], pc, phi[c + sz]);
        const double xj = xc[j + 3], sk = a.sine[k + 3], rs = a.S->r_sh;
        dphidx = dphidx + a.S->m_bh / ((xj - rs) * (xj - rs));
        sc[1] = sc[1] - f * rho * dphidx;
        sc[2] = sc[2] - f * rho * dphidy / xj;
        sc[3] = sc[3] - f * rho * dphidz / xj / sk;
        sc[4] = sc[4] - f * rho * (vx * dphidx + vy * dphidy / xj + vz * dphidz / xj / sk);
    } else if (a.model == 2) {   // model/torus-sph/src/custom.f90:283-296
        double diff = rho - a.rfl[c];
        double f = fmax2(fsign1(diff), 0.0);
        double xj = a.ax[0].c[j + 3];
        double dphidr = 1.0 / ((xj - a.r_sh) * (xj - a.r_sh));
        sc[1] = sc[1] + (-f * rho * dphidr);
        sc[4] = sc[4] + (-f * rho * vx * dphidr);
    } else if (a.model == 1) {
        double diff = rho - a.rfl[c];
        double f = fmax2(fsign1(diff), 0.0);
        double xj = a.ax[0].c[j + 3], zl = a.ax[2].c[l + 3];
        double radius = sqrtT(xj * xj + zl * zl);
        double dphidr = xj / radius / ((radius - a.r_sh) * (radius - a.r_sh));
        double dphidz = zl / radius / ((radius - a.r_sh) * (radius - a.r_sh));
        sc[1] = sc[1] + (-f * rho * dphidr);
        sc[3] = sc[3] + (-f * rho * dphidz);
        sc[4] = sc[4] + (-f * rho * (vx * dphidr + vz * dphidz));
    } else if (a.rt_gravity) {
        sc[2] = sc[2] - 0.1 * rho;
    }
    // FLUX_DIFF (spatial.f90:436-583) + l_operator hydro part; fluxes are read
    // in the working type and differenced in double
    double L[5];
    {
        double dx = a.ax[0].d[j + 3];
        for (int i = 0; i < 5; ++i) {
            const double f1 = a.F[0][i][c], f0 = a.F[0][i][c - 1];
            double df;
            if constexpr (G == 1)
                df = (a.ax[0].xF[j + 3] * f1 - a.ax[0].xF[j + 2] * f0) / (a.ax[0].c[j + 3] * dx);
            else if constexpr (G == 2) {
                double xF = a.ax[0].xF[j + 3], xFm = a.ax[0].xF[j + 2];
                df = (xF * xF * f1 - xFm * xFm * f0) / a.dcb3[j + 3];
            } else df = (f1 - f0) / (dx);
            L[i] = -df;
        }
        if (a.ndim > 1) {
            double dy = a.ax[1].d[k + 3];
            for (int i = 0; i < 5; ++i) {
                const double f1 = a.F[1][i][c], f0 = a.F[1][i][c - sy];
                double df;
                if constexpr (G == 1) df = (f1 - f0) / (a.ax[0].c[j + 3] * dy);
                else if constexpr (G == 2)
                    df = (a.sinf[k + 3] * f1 - a.sinf[k + 2] * f0) / (a.xbar[j + 3] * a.dcose[k + 3]);
                else df = (f1 - f0) / (dy);
                L[i] = L[i] - df;
            }
        }
        if (a.ndim > 2) {
            double dz = a.ax[2].d[l + 3];
            for (int i = 0; i < 5; ++i) {
                const double f1 = a.F[2][i][c], f0 = a.F[2][i][c - sz];
                if constexpr (G == 2)
                    L[i] = L[i] - (f1 - f0) * a.ax[1].d[k + 3] / (a.xbar[j + 3] * dz * a.dcose[k + 3]);
                else L[i] = L[i] - (f1 - f0) / dz;
            }
        }
        for (int i = 0; i < 5; ++i) L[i] = L[i] + sc[i];
    }
    double un[5];
    for (int i = 0; i < 5; ++i) un[i] = rk<STAGE>(uo[i], u[i], dt, L[i]);
    // the edge EMFs around this cell's six faces, each loaded once
    const double ex_c = a.E[0][c], ex_my = a.E[0][c - sy], ex_mz = a.E[0][c - sz], ex_mymz = a.E[0][c - sy - sz];
    const double ey_c = a.E[1][c], ey_mx = a.E[1][c - 1], ey_mz = a.E[1][c - sz], ey_mxmz = a.E[1][c - 1 - sz];
    const double ez_c = a.E[2][c], ez_my = a.E[2][c - sy], ez_mx = a.E[2][c - 1], ez_mxmy = a.E[2][c - 1 - sy];
    // CT update of the six faces of this cell (faces j-1/j, k-1/k, l-1/l)
    double bn_up[3], bn_lo[3], bprim_lo[3];
    {
        const long cl[3] = {c - 1, c - sy, c - sz};
        const int fidx[3] = {j, k, l};
        const double Lup[3] = {lB_x<G>(a, j, k, l, ez_c, ez_my, ey_c, ey_mz), lB_y<G>(a, j, k, l, ex_c, ex_mz, ez_c, ez_mx),
                               lB_z<G>(a, j, k, ey_c, ey_mx, ex_c, ex_my)};
        const double Llo[2] = {lB_x<G>(a, j - 1, k, l, ez_mx, ez_mxmy, ey_mx, ey_mxmz),
                               lB_y<G>(a, j, k - 1, l, ex_my, ex_mymz, ez_my, ez_mxmy)};
        if constexpr (STAGE == 0) {
            // l_rk(1:5) of the cell, l_rk(6:8) of its upper faces and of the
            // faces at index 0 (spatial.f90:620-702); a face shared with a
            // lower cell is that cell's upper face
            double *o = a.lrk;
            for (int i = 0; i < 5; ++i) o[8 * c + i] = L[i];
            for (int d = 0; d < 3; ++d) o[8 * c + 5 + d] = Lup[d];
            if (j == 1) o[8 * (c - 1) + 5] = Llo[0];
            if (k == 1) o[8 * (c - sy) + 6] = Llo[1];
            if (l == 1) o[8 * (c - sz) + 7] = lB_z<G>(a, j, k, ey_mz, ey_mxmz, ex_mz, ex_mymz);
            return 1.0e5;
        }
        // all six face loads first: the stage-1 stores to Uo below would
        // otherwise hold each next load back (the compiler cannot tell Uo
        // from Bin), one memory round trip per face
        double bup6[3], blo6[3];
        for (int d = 0; d < 3; ++d) {
            bup6[d] = a.Bin[d][c];
            blo6[d] = a.Bin[d][cl[d]];
        }
        for (int d = 0; d < 3; ++d) {
            double bup = bup6[d];
            double uoup;
            if (STAGE == 1) {
                uoup = bup;
                a.Uo[5 + d][c] = bup;
            } else {
                uoup = a.Uo[5 + d][c];
            }
            bn_up[d] = rk<STAGE>(uoup, bup, dt, Lup[d]);
            double blo = blo6[d], uolo;
            if (STAGE == 1) {
                uolo = blo;
                if (fidx[d] == 1) a.Uo[5 + d][cl[d]] = blo;
            } else {
                uolo = a.Uo[5 + d][cl[d]];
            }
            const double llo = d < 2 ? Llo[d] : lB_z<G>(a, j, k, ey_mz, ey_mxmz, ex_mz, ex_mymz);
            bn_lo[d] = rk<STAGE>(uolo, blo, dt, llo);
            // the lowest face (index 0) keeps its stored value in prim
            // (FROMUTORVE copies faces 1..n only, fromrvetou.f90:201-207)
            bprim_lo[d] = (fidx[d] == 1) ? blo : bn_lo[d];
        }
    }
    // FROMUTORVE (fromrvetou.f90:153-211)
    double bce[3] = {bavg<0, G>(a, j, k, bn_up[0], bn_lo[0]), bavg<1, G>(a, j, k, bn_up[1], bn_lo[1]),
                     bavg<2, G>(a, j, k, bn_up[2], bn_lo[2])};
    double rn = un[0];
    double v[3] = {un[1] / rn, un[2] / rn, un[3] / rn};
    double bsq = bce[0] * bce[0] + bce[1] * bce[1] + bce[2] * bce[2];
    double vsq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double en = (un[4] - 0.5 * bsq) / rn - 0.5 * vsq;
    checkrho(a, c, rn, v, en);
    if (STAGE == 3) checkrho(a, c, rn, v, en);
    // FROMRVETOU bcell (fromrvetou.f90:49-89)
    double bcf[3] = {bavg<0, G>(a, j, k, bn_up[0], bprim_lo[0]), bavg<1, G>(a, j, k, bn_up[1], bprim_lo[1]),
                     bavg<2, G>(a, j, k, bn_up[2], bprim_lo[2])};
    double p, cs;
    pressure(a, c, rn, en, p, cs);
    // the CFL step (stage 3) before the stores, so that its grid-spacing
    // loads do not wait behind them
    const double dtc = STAGE == 3 ? cfl_dt<G>(a, j, k, l, cs, rn, v, bcf) : 1.0e5;
    st_w(&a.P[0][c], (W)rn);
    st_w(&a.P[1][c], (W)v[0]);
    st_w(&a.P[2][c], (W)v[1]);
    st_w(&a.P[3][c], (W)v[2]);
    st_w(&a.P[4][c], (W)p);
    st_w(&a.eps[c], (W)en);
    st_w(&a.cs[c], (W)cs);
    for (int d = 0; d < 3; ++d) {
        st_w(&a.Bc[d][c], (W)bcf[d]);
        st_nt(&a.Bout[d][c], bn_up[d]);
    }
    // cell EMFs -v x B of the new state (flux_ct, mhd_module.f90:316-318),
    // from the values just stored, for the next stage's edge EMFs
    if constexpr (sizeof(W) == 8) {
        const W vx = v[0], vy = v[1], vz = v[2], bx = bcf[0], by = bcf[1], bz = bcf[2];
        st_w(&a.Ec[0][c], vz * by - vy * bz);
        st_w(&a.Ec[1][c], vx * bz - vz * bx);
        st_w(&a.Ec[2][c], vy * bx - vx * by);
    } else {
        cell_emf(a, c, (W)v[0], (W)v[1], (W)v[2], (W)bcf[0], (W)bcf[1], (W)bcf[2]);
    }
    if constexpr (split) {
        // FROMRVETOU of the double-precision primitives: the next stage's u
        // (Uo after stages 1 and 2, U after stage 3)
        double vsq2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        double bsq2 = bcf[0] * bcf[0] + bcf[1] * bcf[1] + bcf[2] * bcf[2];
        double *const *dst = STAGE == 3 ? a.U : a.Uo;
        const double un2[5] = {rn, v[0] * rn, v[1] * rn, v[2] * rn, rn * (en + 0.5 * vsq2) + 0.5 * bsq2};
        for (int i = 0; i < 5; ++i) {
            if (STAGE == 1) dst[i][c] = un2[i];   // ordinary: non-temporal was slower at stage 1
            else st_nt(&dst[i][c], un2[i]);
        }
    }
    return dtc;
}

// One thread per interior cell, x fastest.  (Measured on MI355X at 256^3:
// this beats a z march carrying the plane below in registers, whose extra
// live values cut occupancy, 64 x 4 cell tiles, and an XCD-grouped block
// order; rows in plain round-robin block order stream best.  y/z sweeps that
// store each cell's flux difference instead of its interface fluxes, so that
// the update reads no neighbouring row or plane, ran 25-50% slower for their
// extra registers and divisions, and the update no faster.)
template <int STAGE, class W, int G>
__global__ __launch_bounds__(256) void k_update(ArgsT<W> a) {
    const long nn = (long)a.nx * a.ny * a.nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < nn;
    const int j = 1 + (int)(t % a.nx), k = 1 + (int)((t / a.nx) % a.ny), l = 1 + (int)(t / ((long)a.nx * a.ny));
    if (!a.S->active) {
        // a step past total_time: keep the state, but the host swaps the face
        // buffers after every stage, so carry the interior faces across
        if (STAGE != 0 && live) {
            const long c = cidx(a, j, k, l);
            for (int d = 0; d < 3; ++d) a.Bout[d][c] = a.Bin[d][c];
        }
        return;
    }
    double dtloc = 1.0e5;
    if (live) dtloc = update_cell<STAGE, W, G>(a, j, k, l, a.S->dt);
    if constexpr (STAGE == 3) {
        dtloc = block_min_f64(dtloc, a.S);
        if (threadIdx.x == 0) atomicMin(&a.S->dtmin_bits, (unsigned long long)__double_as_longlong(dtloc));
    }
}

// ---------------------------------------------------------------------------
// ghost fill: BOUNDARYP_NM + BOUNDARY1D_NM(epsilon) + custom BCs, then
// FINDPRESSURE on the ghost cells
// ---------------------------------------------------------------------------
// Number of ghost cells: the box minus the interior.
template <class A> __host__ __device__ __forceinline__ long ghost_count(const A &a) {
    return a.G - (long)a.nx * a.ny * a.nz;
}

// q-th ghost cell: the six z planes first, then the six y rows of each
// interior plane, then the six x cells of each interior row.
template <class A> __device__ __forceinline__ void ghost_cell(const A &a, long q, int &j, int &k, int &l) {
    const long NX = a.NX, NY = a.NY, pz = 6 * NX * NY, py = 6 * NX * (long)a.nz;
    if (q < pz) {
        const int p = (int)(q / (NX * NY));
        const long r = q % (NX * NY);
        l = p < 3 ? p - 2 : a.nz + 1 + (p - 3);
        k = (int)(r / NX) - 2;
        j = (int)(r % NX) - 2;
    } else if (q < pz + py) {
        q -= pz;
        l = 1 + (int)(q / (6 * NX));
        const long r = q % (6 * NX);
        const int p = (int)(r / NX);
        k = p < 3 ? p - 2 : a.ny + 1 + (p - 3);
        j = (int)(r % NX) - 2;
    } else {
        q -= pz + py;
        l = 1 + (int)(q / (6L * a.ny));
        const long r = q % (6L * a.ny);
        k = 1 + (int)(r / 6);
        const int p = (int)(r % 6);
        j = p < 3 ? p - 2 : a.nx + 1 + (p - 3);
    }
}

template <class W> __global__ __launch_bounds__(256) void k_ghost(ArgsT<W> a) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ghost_count(a)) return;
    int j, k, l;
    ghost_cell(a, q, j, k, l);
    const long t = cidx(a, j, k, l);
    int sx, sy, sz;
    const int jj = gmap(j, a.nx, a.bc[0], a.bc[1], sx);
    const int kk = gmap(k, a.ny, a.bc[2], a.bc[3], sy);
    const int ll = gmap(l, a.nz, a.bc[4], a.bc[5], sz);
    const bool gx = sx >= 0, gy = sy >= 0, gz = sz >= 0;
    if (!gx && !gy && !gz) return;
    if (!a.S->active) {
        // inactive step: Bin is the buffer the host just switched to and Bout
        // the one holding the state; copy the ghost faces back
        for (int d = 0; d < 3; ++d) a.Bin[d][t] = a.Bout[d][t];
        return;
    }
    const long c = t;
    // Every load comes first, then every store.  A load placed after a store
    // would wait for the loads feeding that store (the compiler cannot tell
    // the arrays apart), one memory round trip each.  The sources are
    // interior cells or ghosts of another axis that this launch does not
    // write, so the order changes no value.
    // cell-centred primitives and bcell from the fully mapped interior cell
    const long cs_ = cidx(a, jj, kk, ll);
    const W P0 = a.P[0][cs_], P1 = a.P[1][cs_], P2 = a.P[2][cs_], P3 = a.P[3][cs_];
    const W C0 = a.Bc[0][cs_], C1 = a.Bc[1][cs_], C2 = a.Bc[2][cs_];
    // epsilon: BOUNDARY1D_NM takes the y-outer branch from boundary_flag(3) (boundary.f90:243)
    int ey = kk;
    bool eps_y = gy;
    if (sy == 1) {
        int f3 = a.bc[2], f4 = a.bc[3];
        int m = k - a.ny;
        if (f3 == 0) ey = m;
        else if (f4 == 1) ey = a.ny;
        else if (f4 >= 2) ey = a.ny + 1 - m;
        else { ey = k; eps_y = false; }
        if (eps_y) ey = ey < 1 ? 1 : (ey > a.ny ? a.ny : ey);   // as gmap's clamp
    }
    const bool eps_w = gx || eps_y || gz;
    const bool eps_same_cell = eps_w && ey == kk;
    const W epsv = eps_w ? a.eps[cidx(a, jj, ey, ll)] : a.eps[c];
    // transverse face fields: Bx is filled in y/z ghosts, By in x/z, Bz in x/y
    double F0 = 0.0, F1 = 0.0, F2 = 0.0;
    if (gy || gz) F0 = a.Bin[0][cidx(a, j, kk, ll)];
    if (gx || gz) F1 = a.Bin[1][cidx(a, jj, k, ll)];
    if (gx || gy) F2 = a.Bin[2][cidx(a, jj, kk, l)];
    // FP32 split: the source cell's p and c_s were computed from its
    // double-precision rho and eps; the same FINDPRESSURE of the
    // single-precision copies could differ in the last bit
    const bool copy_pc = sizeof(W) == 4 && a.model == 0 && eps_same_cell;
    W Pp = W(0), Pcs = W(0);
    if (copy_pc) {
        Pp = a.P[4][cs_];
        Pcs = a.cs[cs_];
    }
    double rho = P0 * (gsign(a, 0, sx, 1) * gsign(a, 1, sy, 1) * gsign(a, 2, sz, 1));
    double v[3] = {P1 * (gsign(a, 0, sx, 2) * gsign(a, 1, sy, 2) * gsign(a, 2, sz, 2)),
                   P2 * (gsign(a, 0, sx, 3) * gsign(a, 1, sy, 3) * gsign(a, 2, sz, 3)),
                   P3 * (gsign(a, 0, sx, 4) * gsign(a, 1, sy, 4) * gsign(a, 2, sz, 4))};
    // cell EMFs of the ghost: the mapped interior value times the parity of
    // each component (BOUNDARY_EFIELD, mhd_module.f90:478-840), computed from
    // the source cell before the no-inflow clamps below touch v
    cell_emf(a, c, P1, P2, P3, C0, C1, C2,
             (W)(gsign(a, 0, sx, 9) * gsign(a, 1, sy, 9) * gsign(a, 2, sz, 9)),
             (W)(gsign(a, 0, sx, 10) * gsign(a, 1, sy, 10) * gsign(a, 2, sz, 10)),
             (W)(gsign(a, 0, sx, 11) * gsign(a, 1, sy, 11) * gsign(a, 2, sz, 11)));
    a.Bc[0][c] = (W)(C0 * (gsign(a, 0, sx, 6) * gsign(a, 1, sy, 6) * gsign(a, 2, sz, 6)));
    a.Bc[1][c] = (W)(C1 * (gsign(a, 0, sx, 7) * gsign(a, 1, sy, 7) * gsign(a, 2, sz, 7)));
    a.Bc[2][c] = (W)(C2 * (gsign(a, 0, sx, 8) * gsign(a, 1, sy, 8) * gsign(a, 2, sz, 8)));
    if (eps_w) a.eps[c] = epsv;
    if (gy || gz) a.Bin[0][c] = F0 * (gsign(a, 1, sy, 6) * gsign(a, 2, sz, 6));
    if (gx || gz) a.Bin[1][c] = F1 * (gsign(a, 0, sx, 7) * gsign(a, 2, sz, 7));
    if (gx || gy) a.Bin[2][c] = F2 * (gsign(a, 0, sx, 8) * gsign(a, 1, sy, 8));
    // torus no-inflow clamps: radial in both torus models
    // (model/torus-cyn/src/custom.f90:156-161, model/torus-sph/src/custom.f90:160-167),
    // vertical in the cylindrical one (model/torus-cyn/src/custom.f90:194-199)
    if (a.model != 0) {
        if (sx == 0) v[0] = fmin_zero(v[0]);
        if (sx == 1) v[0] = fmax2(v[0], 0.0);
    }
    if (a.model == 1) {
        if (sz == 0) v[2] = fmin_zero(v[2]);
        if (sz == 1) v[2] = fmax2(v[2], 0.0);
    }
    a.P[0][c] = (W)rho;
    a.P[1][c] = (W)v[0];
    a.P[2][c] = (W)v[1];
    a.P[3][c] = (W)v[2];
    if (copy_pc) {
        a.P[4][c] = Pp;
        a.cs[c] = Pcs;
        return;
    }
    double p, cs;
    pressure(a, c, rho, (double)epsv, p, cs);
    a.P[4][c] = (W)p;
    a.cs[c] = (W)cs;
}

}  // namespace cumc3d
