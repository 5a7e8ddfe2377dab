// Self-gravity and black-hole accretion of the collapsar model: the parts of
// CUSTOM_UPDATE (model/collapsar/src/custom.f90:416-658) that run after
// FINDPRESSURE in UPDATE(p).
//
//   k_mdot            mass accretion rate through the inner face from the x
//                     sweep's density flux at face 0 (GETFLUX_X, :222-245)
//   k_mbh<STAGE>      the RK update of m_bh, r_sh and the horizon check
//   k_sor_red         the red half of one SOR sweep pair (and the decision
//                     of the pair before it)
//   k_sor_black       the black half and the blocks' error partials
//   k_sor_decide      the decision of a batch's last pair
//   k_sor_bc          the boundary values of the final potential
//
// FP64 only.  The sweeps are bit-identical to the serial loop: a red cell
// ((-1)**(j+k+l) > 0) reads only black neighbours and vice versa, so the order
// within a sweep does not matter.  During a solve the potential lives split
// by colour (every cell of one colour in a dense array, like rho_in) and
// ping-pongs: a sweep pair reads the red and black values at its
// start and writes both colours anew; red cells read the start values, black
// cells the new red ones.  The boundary values a pair reads are those the
// reference's previous pair set (custom.f90:601-631), formed on the fly from
// the start values: phi(0) = phi(1), phi(nx+1) = -1, zero theta gradient,
// periodic in phi -- and 0 on the first pair after the zero guess of
// UPDATE(0) (:497-505), which had set no boundary values yet.  k_sor_copy
// moves phi into the split arrays at the start and back at the end, and
// k_sor_bc writes the boundary values for the gravity source.
// The error is the OpenMP MAX reduction of :588-598 as the reference's
// one-thread CPU build evaluates it: a private fold of
// MAX(abserror, |dphi/phi_old|) in l, k, j order with flang's MAX
// (a > b ? a : b) -- a NaN term replaces the running value and the next term
// replaces the NaN -- then a combine with the shared 1e-50 that drops a NaN.
// So the result is 1e-50 when the last term is NaN (0/0 where phi stays 0,
// as on the first sweep from the zero guess), else the larger of 1e-50 and
// the maximum of the terms after the last NaN.  k_sor_black leaves per-block
// partials (the last NaN index of the block and the maximum of the terms
// after it) and one block of the next kernel combines them exactly.  The
// host runs sweep pairs in batches and reads the end flag between batches;
// pairs queued after the solve ended return at once.
#pragma once
#include "kernels.hpp"

namespace cumc3d {

// colour-split slots of each colour per black block (1024 cells, 4 per thread)
constexpr int SOR_SLOTS = 512;

// Per-cell SOR data are split by colour so that each sweep reads its cells
// densely: colour c (0 red, 1 black) holds, for row (k, l), the cells
// j = j0 + 2i with j0 = 1 when (1 + k + l) % 2 == c, else 2, at
// c*hx*ny*nz + i + hx*((k-1) + ny*(l-1)), hx = (nx+1)/2 (a padded slot when
// j > nx).  Per cell only rho_in, the density above the atmosphere
// (custom.f90:544-546), refreshed per solve.  poisson_coef's azimuthal
// coefficients (poisson.f90:52-63) are formed in the kernel from a (j, k)
// factor and the cell's phi spacings with the same divisions in the same
// order, so they are the values the reference stores:
//   clp1 = ((2/xc^2/sin^2) / (zp1-zm1)) / (zp1-zc)
//   clm1 = ((2/xc^2/sin^2) / (zp1-zm1)) / (zc-zm1)
//   epsc = (radial + polar terms) - ((2/xc^2/sin^2) / (zp1-zc)) / (zc-zm1)

struct SorArgs {
    int nx, ny, nz;
    long NX, NY;
    // phi during a solve, split by colour like rho_in (index
    // i + hx((k-1) + ny(l-1)) within a colour): the red and black values at
    // the start of the sweep pair and after it
    const double *rin, *bin;
    double *rout, *bout;
    const double *rho_in;               // per cell, split by colour (see above)
    const double *ajp1, *ajm1;          // [j], j = 1..nx
    const double *bkp1, *bkm1;          // [k(nx+1) + j]
    const double *a2, *e2;              // [k(nx+1) + j]: 2/xc^2/sin^2 and epsc's radial + polar terms
    const double *zpm, *zpc, *zcm;      // [l]: zp1-zm1, zp1-zc, zc-zm1
    long long *part_nan;                // per black block: q of its last NaN term, -1 if none
    double *part_max;                   // per black block: max of the terms after it, -1 if none
    int nblk;                           // black blocks: ceil(hx ny / SOR_SLOTS) per plane, times nz
    int relax_max;
    int zero_ghosts;                    // the first pair after the zero guess: boundary values 0
    int combine;                        // red kernel: block 0 first decides the previous pair's exit
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

// custom.f90:450-465: mdot = sum over l, k (serial order) of
// -mflux_x(k,l) xF(0)^2 dcose(k) dz(l), clipped at 0, from the x sweep's
// density flux at face 0.  256 threads form up to 4096 terms at a time into
// LDS (all loads in flight together); one wave then adds them in the serial
// order, reading ahead of the dependent add chain, which is the cost.  Runs
// on a second stream beside the stage's later kernels.
constexpr int ACC_CHUNK = 4096;
__global__ __launch_bounds__(256) void k_mdot(ArgsT<double> a) {
    if (!a.S->active) return;
    __shared__ double term[ACC_CHUNK];
    const int nk = a.ny * a.nz;
    const double xf0 = a.ax[0].xF[3];
    double mdot = 0.0;
    for (int base = 0; base < nk; base += ACC_CHUNK) {
        const int m = nk - base < ACC_CHUNK ? nk - base : ACC_CHUNK;
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
            const int t = base + i;
            const int k = 1 + t % a.ny, l = 1 + t / a.ny;
            const double mf = a.F[0][0][cidx(a, 0, k, l)];
            term[i] = (-mf) * xf0 * xf0 * a.dcose[k + 3] * a.ax[2].d[l + 3];
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            int i = 0;
            for (; i + 8 <= m; i += 8) {
                double r[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) r[u] = term[i + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) mdot = mdot + r[u];
            }
            for (; i < m; ++i) mdot = mdot + term[i];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) a.S->mdot = fmax2(mdot, 0.0);
}

// custom.f90:467-484 for p = STAGE: the RK combine of m_bh with the stage's
// mdot, then r_sh and the horizon check
template <int STAGE> __global__ void k_mbh(DevScalars *S, double c_code, double xF0) {
    if (!S->active) return;
    const double dt = S->dt, mdot = S->mdot;
    if constexpr (STAGE == 1) {
        S->mbh_old = S->m_bh;
        S->m_bh = S->mbh_old + dt * mdot;
    } else if constexpr (STAGE == 2) {
        S->m_bh = (3.0 / 4.0) * S->mbh_old + (1.0 / 4.0) * S->m_bh + (1.0 / 4.0) * dt * mdot;
    } else {
        S->m_bh = (1.0 / 3.0) * S->mbh_old + (2.0 / 3.0) * S->m_bh + (2.0 / 3.0) * dt * mdot;
    }
    grav_horizon(S, c_code, xF0);
}

// UPDATE(0): m_bh as the initial model left it, then r_sh and the check
__global__ void k_grav_start(DevScalars *S, double m_bh, double c_code, double xF0) {
    S->m_bh = m_bh;
    S->mbh_old = m_bh;
    S->grav_stop = 0;
    grav_horizon(S, c_code, xF0);
}

// Start of a solve: the iteration state (k_fill0 writes the zero guess phi = 0
// of p = 0, custom.f90:491-507, over the whole box; the outer layers are
// never read).  An inactive run (finished or stopped) skips the solve.
__global__ void k_sor_begin(DevScalars *S) {
    S->sor_n = 0;
    if (!S->active) {
        S->sor_state = 1;
        S->sor_iter = 0;
        return;
    }
    S->sor_state = 0;
}

__global__ __launch_bounds__(256) void k_fill0(double *p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__device__ __forceinline__ long sor_c(const SorArgs &a, int j, int k, int l) {
    return (long)(j + 2) + a.NX * ((long)(k + 2) + a.NY * (long)(l + 2));
}

// index of cell (j, k, l) within its colour's arrays
__device__ __forceinline__ long sor_i(const SorArgs &a, int j, int k, int l) {
    const long hx = (a.nx + 1) >> 1;
    return (long)((j - 1) >> 1) + hx * ((long)(k - 1) + (long)a.ny * (l - 1));
}

// The neighbour values one cell's update reads (custom.f90:548-550): cells
// of the other colour from `other` (the black start values for a red cell,
// the new red values for a black one), the boundary values from the start
// values of the pair: phi(0) = phi(1) and the theta ends copy the cell
// itself (`own`, its start value), phi(nx+1) = -1, periodic in phi.
__device__ __forceinline__ void sor_neigh(const SorArgs &a, const double *other, int j, int k, int l, double own,
                                          double nb[6]) {
    const bool z0 = a.zero_ghosts;
    nb[0] = j < a.nx ? other[sor_i(a, j + 1, k, l)] : (z0 ? 0.0 : -1.0);
    nb[1] = j > 1 ? other[sor_i(a, j - 1, k, l)] : (z0 ? 0.0 : own);
    nb[2] = k < a.ny ? other[sor_i(a, j, k + 1, l)] : (z0 ? 0.0 : own);
    nb[3] = k > 1 ? other[sor_i(a, j, k - 1, l)] : (z0 ? 0.0 : own);
    if (l < a.nz) nb[4] = other[sor_i(a, j, k, l + 1)];
    else if (z0) nb[4] = 0.0;
    else nb[4] = (((j + k + 1) & 1) ? a.bin : a.rin)[sor_i(a, j, k, 1)];
    if (l > 1) nb[5] = other[sor_i(a, j, k, l - 1)];
    else if (z0) nb[5] = 0.0;
    else nb[5] = (((j + k + a.nz) & 1) ? a.bin : a.rin)[sor_i(a, j, k, a.nz)];
}

// one cell's relaxation (custom.f90:544-551); returns the new value
__device__ __forceinline__ double sor_cell(const SorArgs &a, int j, int k, int l, double rho_in, const double nb[6],
                                           double old) {
    const int jk = k * (a.nx + 1) + j;
    const double A = a.a2[jk], zpc = a.zpc[l], zcm = a.zcm[l];
    const double t = A / a.zpm[l];
    const double clp1 = t / zpc, clm1 = t / zcm;
    const double epsc = a.e2[jk] - A / zpc / zcm;
    const double s = a.ajp1[j] * nb[0] + a.ajm1[j] * nb[1] + a.bkp1[jk] * nb[2] + a.bkm1[jk] * nb[3] +
                     clp1 * nb[4] + clm1 * nb[5];
    const double rhs = (a.pi4 * rho_in - s) / epsc;
    return (1.0 - a.omega) * old + a.omega * rhs;
}

// position of cell i of row (k, l) of a colour in the colour-split rho_in
__device__ __forceinline__ long sor_rec(const SorArgs &a, int colour, int i, int k, int l) {
    const long hx = (a.nx + 1) >> 1;
    return (long)colour * hx * a.ny * a.nz + i + hx * ((long)(k - 1) + (long)a.ny * (l - 1));
}

// Start of a solve: rho_in of every cell, split by colour
// (custom.f90:544-546: MERGE(1, 0, rho - prim_a(irho) > 0) * rho).
__global__ __launch_bounds__(256) void k_sor_rho(SorArgs a, const double *rho, double *rho_in) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)(a.nx * a.ny)) return;
    const int l = 1 + blockIdx.y;
    const int k = 1 + (int)(t / (unsigned)a.nx), j = 1 + (int)(t - (unsigned)(k - 1) * a.nx);
    const double r = rho[sor_c(a, j, k, l)];
    const double diff = r - a.pa;
    const double factor = diff > 0.0 ? 1.0 : 0.0;
    const int colour = (j + k + l) & 1;
    const int j0 = ((1 + k + l) & 1) == colour ? 1 : 2;
    rho_in[sor_rec(a, colour, (j - j0) >> 1, k, l)] = factor * r;
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

// Two partials of consecutive q ranges, the right one folded into the left:
// the last NaN of the right range supersedes everything before it.
__device__ __forceinline__ void fold_partial(long long &L, double &M, long long L2, double M2) {
    if (L2 >= 0) {
        L = L2;
        M = M2;
    } else {
        M = M2 > M ? M2 : M;
    }
}

// The previous pair's abserror from the black blocks' partials and the exit
// decision (custom.f90:638, :644), by one block of 256 threads.  The
// partials are in q order (block index plane-major), so their left-to-right
// fold is (the last NaN q, the maximum of the terms after it): each thread
// folds a contiguous run of them, then the threads' results are folded in
// thread order (lane i takes lane i + off at each level of the tree).
__device__ void sor_decide(const SorArgs &a) {
    DevScalars *S = a.S;
    const int per = (a.nblk + 255) >> 8;
    const int i0 = (int)threadIdx.x * per, i1 = i0 + per < a.nblk ? i0 + per : a.nblk;
    long long L = -1;
    double M = -1.0;
    for (int i = i0; i < i1; ++i) fold_partial(L, M, a.part_nan[i], a.part_max[i]);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const long long L2 = __shfl_down(L, off, 64);
        const double M2 = __shfl_down(M, off, 64);
        if ((threadIdx.x & 63) + off < 64) fold_partial(L, M, L2, M2);
    }
    __shared__ long long wl[4];
    __shared__ double wm[4];
    if ((threadIdx.x & 63) == 0) {
        wl[threadIdx.x >> 6] = L;
        wm[threadIdx.x >> 6] = M;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) fold_partial(L, M, wl[w], wm[w]);
    const long long N = (long long)a.nx * a.ny * a.nz;
    double err;
    if (L == N - 1) err = 1.0e-50;   // the private result is NaN: the combine keeps 1e-50
    else err = M > 1.0e-50 ? M : 1.0e-50;
    S->sor_err = err;
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

// Red cells ((j + k + l) even), every second cell along x: grid (x: the
// plane's red cells / 256 + 1 for the decision block, y: l - 1).  With
// `combine`, block (0, 0) first decides the previous pair's exit; the other
// blocks do not wait for it: a pair begun after the solve ended only writes
// the buffer that is not the result, and the black kernel and every later
// kernel see the decision and return.
__global__ __launch_bounds__(256) void k_sor_red(SorArgs a) {
    if (a.combine && blockIdx.x == 0 && blockIdx.y == 0) {
        if (!a.S->sor_state && a.S->active) sor_decide(a);
        return;
    }
    if (a.S->sor_state || !a.S->active) return;
    const int hx = (a.nx + 1) >> 1;
    const unsigned t = (blockIdx.x - 1) * blockDim.x + threadIdx.x;
    if (t >= (unsigned)(hx * a.ny)) return;
    const int k = 1 + (int)(t / (unsigned)hx), i = (int)(t - (unsigned)(k - 1) * hx), l = 1 + blockIdx.y;
    const int j = (((1 + k + l) & 1) == 0 ? 1 : 2) + 2 * i;
    if (j > a.nx) return;
    const long ic = sor_i(a, j, k, l);
    const double rho_in = a.rho_in[sor_rec(a, 0, i, k, l)];
    const double old = a.rin[ic];
    double nb[6];
    sor_neigh(a, a.bin, j, k, l, old, nb);
    a.rout[ic] = sor_cell(a, j, k, l, rho_in, nb, old);
}

// Every interior cell, in the colour-split order: block b of plane l takes
// the split indices [b S, (b+1) S) of both colours (S = SOR_SLOTS), the cells
// (k, j = 2i+1, 2i+2) of rows k at index i; together they are one contiguous
// range of q.  A thread's slots 0-1 are black cells, which relax from the new
// red values, and its slots 2-3 red ones, so the lanes of a wave share a
// colour and neither branch diverges.  Every cell forms its error term
// |dphi/phi_old| (custom.f90:593) from the pair's start and end values, and
// the block folds them into its partial: the fold is a maximum, so the order
// of the slots does not matter.
__global__ __launch_bounds__(256) void k_sor_black(SorArgs a) {
    if (a.S->sor_state || !a.S->active) return;
    const int hx = (a.nx + 1) >> 1, nh = hx * a.ny, nxy = a.nx * a.ny;
    const int l = 1 + blockIdx.y;
    double e[4];
    long long qq[4];
    long long last = -1;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int colour = u < 2 ? 1 : 0;
        const int s = blockIdx.x * SOR_SLOTS + (u & 1) * 256 + (int)threadIdx.x;
        qq[u] = -1;
        e[u] = -1.0;
        if (s >= nh) continue;
        const int k = 1 + s / hx, i = s - (k - 1) * hx;
        const int j = (((1 + k + l) & 1) == colour ? 1 : 2) + 2 * i;
        if (j > a.nx) continue;   // the padded slot of an odd row
        qq[u] = (long long)(l - 1) * nxy + (long long)(k - 1) * a.nx + (j - 1);
        const long ic = (long)s + (long)nh * (l - 1);   // sor_i(a, j, k, l)
        double old, nw;
        if (colour) {
            const double rho_in = a.rho_in[(long)nh * a.nz + ic];   // sor_rec(a, 1, i, k, l)
            old = a.bin[ic];
            double nb[6];
            sor_neigh(a, a.rout, j, k, l, old, nb);
            nw = sor_cell(a, j, k, l, rho_in, nb, old);
            a.bout[ic] = nw;
        } else {
            old = a.rin[ic];
            nw = a.rout[ic];
        }
        e[u] = fabs((nw - old) / old);
        if (e[u] != e[u]) last = qq[u] > last ? qq[u] : last;
    }
    last = block_max_ll(last);
    double m = -1.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (qq[u] > last && e[u] == e[u]) m = e[u] > m ? e[u] : m;
    m = block_max_d(m);
    if (threadIdx.x == 0) {
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        a.part_nan[b] = last;
        a.part_max[b] = m;
    }
}

// The decision of a solve's last pair when no red kernel follows it.
__global__ __launch_bounds__(256) void k_sor_decide(SorArgs a) {
    if (!a.S->sor_state && a.S->active) sor_decide(a);
}

// phi over the interior between the ghosted box and the colour-split
// arrays: to_split at the start of a solve, back at its end.
__global__ __launch_bounds__(256) void k_sor_copy(SorArgs a, double *phi, double *red, double *black, int to_split) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)(a.nx * a.ny)) return;
    const int l = 1 + blockIdx.y;
    const int k = 1 + (int)(t / (unsigned)a.nx), j = 1 + (int)(t - (unsigned)(k - 1) * a.nx);
    double *s = ((j + k + l) & 1) ? black : red;
    const long c = sor_c(a, j, k, l), ic = sor_i(a, j, k, l);
    if (to_split) s[ic] = phi[c];
    else phi[c] = s[ic];
}

// The boundary values of the final potential (custom.f90:601-631), written
// once a solve has ended: phi(0) = phi(1), phi(nx+1) = -1 over the interior
// (k, l); theta ends copy their neighbours over (j, l); phi is periodic over
// (j, k).
__global__ __launch_bounds__(256) void k_sor_bc(double *phi, int nx, int ny, int nz, long NX, long NY) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long sy = NX, sz = NX * NY;
    auto ci = [&](int j, int k, int l) { return (long)(j + 2) + NX * ((long)(k + 2) + NY * (long)(l + 2)); };
    if (t < (long)ny * nz) {
        const int k = 1 + (int)(t % ny), l = 1 + (int)(t / ny);
        const long c = ci(1, k, l);
        phi[c - 1] = phi[c];
        phi[ci(nx + 1, k, l)] = -1.0;
        return;
    }
    t -= (long)ny * nz;
    if (t < (long)nx * nz) {
        const int j = 1 + (int)(t % nx), l = 1 + (int)(t / nx);
        const long c1 = ci(j, 1, l), cn = ci(j, ny, l);
        phi[c1 - sy] = phi[c1];
        phi[cn + sy] = phi[cn];
        return;
    }
    t -= (long)nx * nz;
    if (t < (long)nx * ny) {
        const int j = 1 + (int)(t % nx), k = 1 + (int)(t / nx);
        const long c1 = ci(j, k, 1), cn = ci(j, k, nz);
        phi[c1 - sz] = phi[cn];
        phi[cn + sz] = phi[c1];
    }
}

}  // namespace cumc3d
