// Self-gravity and black-hole accretion of the collapsar model: the parts of
// CUSTOM_UPDATE (model/collapsar/src/custom.f90:416-658) that run after
// FINDPRESSURE in UPDATE(p).
//
//   k_accrete<STAGE>  mass accretion rate through the inner face from the x
//                     sweep's density flux at face 0 (GETFLUX_X, :222-245),
//                     the RK update of m_bh, r_sh and the horizon check
//   k_sor_sweep<C>    one colour of a red-black SOR sweep of the potential
//   k_sor_post        the error fold over all cells and the boundary values
//
// FP64 only.  The SOR sweeps are bit-identical to the serial loop: a red cell
// ((-1)**(j+k+l) > 0) reads only black neighbours and vice versa, so the order
// within a sweep does not matter, and each sweep computes the relative change
// of its own cells (the value read before the update is phi_old).
// The error is the OpenMP MAX reduction of :588-598 as the reference's
// one-thread CPU build evaluates it: a private fold of
// MAX(abserror, |dphi/phi_old|) in l, k, j order with flang's MAX
// (a > b ? a : b) -- a NaN term replaces the running value and the next term
// replaces the NaN -- then a combine with the shared 1e-50 that drops a NaN.
// So the result is 1e-50 when the last term is NaN (0/0 where phi stays 0,
// as on the first sweep from the zero guess), else the larger of 1e-50 and
// the maximum of the terms after the last NaN.  k_sor_post forms it exactly
// from per-block partials (the last NaN index of the block and the maximum of
// the terms after it), combined by the last block to finish.  The host runs
// sweeps in batches and reads the end flag between batches; sweeps queued
// after the solve ended return at once.
#pragma once
#include "kernels.hpp"

namespace cumc3d {

struct SorArgs {
    int nx, ny, nz;
    long NX, NY;
    double *phi;                        // ghosted box, cidx layout
    double *err;                        // |dphi/phi_old| per interior cell, q = (j-1) + nx((k-1) + ny(l-1))
    const double *rho;                  // prim(irho), ghosted box
    const double *ajp1, *ajm1;          // [j], j = 1..nx
    const double *bkp1, *bkm1;          // [k(nx+1) + j]
    const double *clp1, *clm1, *epsc;   // [q]
    long long *part_nan;                // per error block: index of its last NaN term, -1 if none
    double *part_max;                   // per error block: max of the terms after it, -1 if none
    int nerr;                           // error blocks (256 terms each)
    int relax_max;
    double pa, omega, tol, pi4;         // prim_a(irho), omega_weight, tolerance, 4 pi
    DevScalars *S;
};

// r_sh from m_bh and the horizon check of custom.f90:481-484 (STOP: the
// stage kernels of every later stage return at once)
__device__ __forceinline__ void grav_horizon(DevScalars *S, double c_code, double xF0) {
    S->r_sh = 2.0 * S->m_bh / (c_code * c_code);
    if (S->r_sh >= xF0) {
        S->grav_stop = 1;
        S->active = 0;
    }
}

// custom.f90:450-484 for p = STAGE: mdot = sum over l, k (serial order) of
// -mflux_x(k,l) xF(0)^2 dcose(k) dz(l), clipped at 0, then the RK combine of
// m_bh.  The 64 lanes form the terms; lane 0 adds them in the serial order.
template <int STAGE> __global__ __launch_bounds__(64) void k_accrete(ArgsT<double> a, double c_code) {
    if (!a.S->active) return;
    __shared__ double term[64];
    const int lane = threadIdx.x;
    const int nk = a.ny * a.nz;
    const double xf0 = a.ax[0].xF[3];
    double mdot = 0.0;
    for (int base = 0; base < nk; base += 64) {
        const int t = base + lane;
        if (t < nk) {
            const int k = 1 + t % a.ny, l = 1 + t / a.ny;
            const double mf = a.F[0][0][cidx(a, 0, k, l)];
            term[lane] = (-mf) * xf0 * xf0 * a.dcose[k + 3] * a.ax[2].d[l + 3];
        }
        __syncthreads();
        if (lane == 0) {
            const int m = nk - base < 64 ? nk - base : 64;
            for (int i = 0; i < m; ++i) mdot = mdot + term[i];
        }
        __syncthreads();
    }
    if (lane != 0) return;
    mdot = fmax2(mdot, 0.0);
    DevScalars *S = a.S;
    const double dt = S->dt;
    if constexpr (STAGE == 1) {
        S->mbh_old = S->m_bh;
        S->m_bh = S->mbh_old + dt * mdot;
    } else if constexpr (STAGE == 2) {
        S->m_bh = (3.0 / 4.0) * S->mbh_old + (1.0 / 4.0) * S->m_bh + (1.0 / 4.0) * dt * mdot;
    } else {
        S->m_bh = (1.0 / 3.0) * S->mbh_old + (2.0 / 3.0) * S->m_bh + (2.0 / 3.0) * dt * mdot;
    }
    grav_horizon(S, c_code, xf0);
}

// UPDATE(0): m_bh as the initial model left it, then r_sh and the check
__global__ void k_grav_start(DevScalars *S, double m_bh, double c_code, double xF0) {
    S->m_bh = m_bh;
    S->mbh_old = m_bh;
    S->grav_stop = 0;
    grav_horizon(S, c_code, xF0);
}

// Start of a solve: the zero guess phi = 0 on cells 0..n+1 for p = 0
// (custom.f90:491-507; the whole box here, the outer layers are never read)
// and the iteration state.  An inactive run (finished or stopped) skips it.
__global__ void k_sor_begin(DevScalars *S) {
    if (!S->active) {
        S->sor_state = 1;
        S->sor_iter = 0;
        return;
    }
    S->sor_state = 0;
    S->sor_n = 0;
    S->sor_ctr = 0;
}

__global__ __launch_bounds__(256) void k_fill0(double *p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// One colour of the sweep (custom.f90:537-584).  Threads walk the cells of
// the colour along x (every second cell), rows in (k, l) order.
template <int COLOUR> __global__ __launch_bounds__(256) void k_sor_sweep(SorArgs a) {
    if (a.S->sor_state || !a.S->active) return;
    const int hx = (a.nx + 1) >> 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)hx * a.ny * a.nz) return;
    const int i = (int)(t % hx);
    const long r = t / hx;
    const int k = 1 + (int)(r % a.ny), l = 1 + (int)(r / a.ny);
    const int j = (((1 + k + l) & 1) == COLOUR ? 1 : 2) + 2 * i;   // (j + k + l) & 1 == COLOUR
    if (j > a.nx) return;
    const long sy = a.NX, sz = a.NX * a.NY;
    const long c = (long)(j + 2) + a.NX * ((long)(k + 2) + a.NY * (long)(l + 2));
    const long q = (long)(j - 1) + (long)a.nx * ((long)(k - 1) + (long)a.ny * (l - 1));
    double *phi = a.phi;
    const double rho = a.rho[c];
    const double diff = rho - a.pa;
    const double factor = diff > 0.0 ? 1.0 : 0.0;
    const double rho_in = factor * rho;
    const long jk = (long)k * (a.nx + 1) + j;
    const double s = a.ajp1[j] * phi[c + 1] + a.ajm1[j] * phi[c - 1] + a.bkp1[jk] * phi[c + sy] +
                     a.bkm1[jk] * phi[c - sy] + a.clp1[q] * phi[c + sz] + a.clm1[q] * phi[c - sz];
    const double rhs = (a.pi4 * rho_in - s) / a.epsc[q];
    const double old = phi[c];
    const double nw = (1.0 - a.omega) * old + a.omega * rhs;
    phi[c] = nw;
    a.err[q] = fabs((nw - old) / old);
}

__device__ __forceinline__ long long block_max_ll(long long v) {
    __shared__ long long red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        long long o = __shfl_down(v, off, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    long long m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = red[w] > m ? red[w] : m;
    __syncthreads();
    return m;
}

// max of non-NaN values >= -1 (the terms are >= 0; -1 marks none)
__device__ __forceinline__ double block_max_d(double v) {
    __shared__ double red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_down(v, off, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = red[w] > m ? red[w] : m;
    __syncthreads();
    return m;
}

// Blocks [0, nerr): the error fold (custom.f90:587-598); the last of them to
// finish combines the partials and decides the loop exit (:638, :644).
// Blocks [nerr, ...): the boundary values (:601-631): phi(0) = phi(1),
// phi(nx+1) = -1; zero gradient at both theta ends; periodic in phi.
__global__ __launch_bounds__(256) void k_sor_post(SorArgs a) {
    if (a.S->sor_state || !a.S->active) return;
    const long N = (long)a.nx * a.ny * a.nz;
    if ((int)blockIdx.x >= a.nerr) {
        long t = (long)(blockIdx.x - a.nerr) * blockDim.x + threadIdx.x;
        const long sy = a.NX, sz = a.NX * a.NY;
        double *phi = a.phi;
        auto ci = [&](int j, int k, int l) { return (long)(j + 2) + a.NX * ((long)(k + 2) + a.NY * (long)(l + 2)); };
        if (t < (long)a.ny * a.nz) {
            const int k = 1 + (int)(t % a.ny), l = 1 + (int)(t / a.ny);
            const long c = ci(1, k, l);
            phi[c - 1] = phi[c];
            phi[ci(a.nx + 1, k, l)] = -1.0;
            return;
        }
        t -= (long)a.ny * a.nz;
        if (t < (long)a.nx * a.nz) {
            const int j = 1 + (int)(t % a.nx), l = 1 + (int)(t / a.nx);
            const long c1 = ci(j, 1, l), cn = ci(j, a.ny, l);
            phi[c1 - sy] = phi[c1];
            phi[cn + sy] = phi[cn];
            return;
        }
        t -= (long)a.nx * a.nz;
        if (t < (long)a.nx * a.ny) {
            const int j = 1 + (int)(t % a.nx), k = 1 + (int)(t / a.nx);
            const long c1 = ci(j, k, 1), cn = ci(j, k, a.nz);
            phi[c1 - sz] = phi[cn];
            phi[cn + sz] = phi[c1];
        }
        return;
    }
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const double e = q < N ? a.err[q] : 0.0;
    const bool nan = q < N && e != e;
    const long long last = block_max_ll(nan ? (long long)q : -1LL);
    const double m = block_max_d(q < N && !nan && q > last ? e : -1.0);
    __shared__ int is_last;
    if (threadIdx.x == 0) {
        a.part_nan[blockIdx.x] = last;
        a.part_max[blockIdx.x] = m;
        __threadfence();
        const unsigned done = atomicAdd(&a.S->sor_ctr, 1u);
        is_last = done == (unsigned)(a.nerr - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    long long L = -1;
    for (int b = threadIdx.x; b < a.nerr; b += blockDim.x) {
        const long long v = a.part_nan[b];
        L = v > L ? v : L;
    }
    L = block_max_ll(L);
    const long long bL = L >= 0 ? L / 256 : 0;
    double M = -1.0;
    for (int b = threadIdx.x; b < a.nerr; b += blockDim.x) {
        if (b < bL) continue;
        const double v = a.part_max[b];
        M = v > M ? v : M;
    }
    M = block_max_d(M);
    if (threadIdx.x != 0) return;
    DevScalars *S = a.S;
    double err;
    if (L == N - 1) err = 1.0e-50;   // the private result is NaN: the combine keeps 1e-50
    else err = M > 1.0e-50 ? M : 1.0e-50;
    S->sor_err = err;
    S->sor_ctr = 0;
    const int n = S->sor_n + 1;
    S->sor_n = n;
    if (err <= a.tol) {
        S->sor_state = 1;
        S->sor_iter = n;
    } else if (n == a.relax_max) {
        S->sor_state = 1;
        S->sor_iter = n + 1;   // the DO loop ran out: n = relax_max + 1 after it
    }
}

}  // namespace cumc3d
