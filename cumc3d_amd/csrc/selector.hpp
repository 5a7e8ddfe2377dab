// Learned reconstruction selector: a 16(15)->32->32->3 MLP that picks, per
// cell and sweep direction, the reconstruction the flux kernels run there
// (class 0 TVD-MC, 1 PPM, 2 WENO-Z).  The reference has no counterpart: its
// reconstruction is one compile-time choice (spatial.f90:33-43).
//
// Inference runs on the matrix cores: v_mfma_f32_16x16x16_bf16 with BF16
// weights and activations and FP32 accumulation, 16 cells per MFMA group.
// Training (on samples the solver generates from its own state) runs in FP32
// on the vector units, one sample per thread, gradients reduced in LDS.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cumc3d {

constexpr int SEL_IN = 16;    // 15 features + one zero pad
constexpr int SEL_H = 32;
constexpr int SEL_OUT = 3;
// FP32 parameter vector: W1[in][h], b1[h], W2[h][h], b2[h], W3[h][out], b3[out]
constexpr int SEL_W1 = 0;
constexpr int SEL_B1 = SEL_W1 + SEL_IN * SEL_H;
constexpr int SEL_W2 = SEL_B1 + SEL_H;
constexpr int SEL_B2 = SEL_W2 + SEL_H * SEL_H;
constexpr int SEL_W3 = SEL_B2 + SEL_H;
constexpr int SEL_B3 = SEL_W3 + SEL_H * SEL_OUT;
constexpr int SEL_NPARAM = SEL_B3 + SEL_OUT;   // 1699
constexpr int SEL_NFRAG = 8;                    // BF16 A fragments of the three layers

typedef short sel_s4 __attribute__((ext_vector_type(4)));
typedef float sel_f4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short sel_bf16(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float sel_from_bf16(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// ---------------------------------------------------------------------------
// features and labels
// ---------------------------------------------------------------------------

// The 16 inputs of cell c along a sweep with stride st: feature 5v+s is
// variable v (rho, p, |v|) at stencil offset s-2 as a bounded relative
// difference from the cell itself (|v| is scaled by |v| + c_s); input 15 is a
// zero pad.  Training samples and inference share this function.
template <class W>
__device__ __forceinline__ void sel_features(const ArgsT<W> &a, long c, long st, float f[SEL_IN]) {
    float rho[5], pr[5], u[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const long q = c + (long)(m - 2) * st;
        rho[m] = (float)a.P[0][q];
        pr[m] = (float)a.P[4][q];
        const float vx = (float)a.P[1][q], vy = (float)a.P[2][q], vz = (float)a.P[3][q];
        u[m] = sqrtf(vx * vx + vy * vy + vz * vz);
    }
    const float c0 = (float)a.cs[c];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        f[m] = (rho[m] - rho[2]) / (fabsf(rho[m]) + fabsf(rho[2]) + 1.0e-30f);
        f[5 + m] = (pr[m] - pr[2]) / (fabsf(pr[m]) + fabsf(pr[2]) + 1.0e-30f);
        f[10 + m] = (u[m] - u[2]) / (u[m] + u[2] + c0 + 1.0e-30f);
    }
    f[15] = 0.0f;
}

// Training label of cell c along D: the largest relative gap between the
// TVD-MC and WENO-Z faces of rho and p; class 2 above thr_weno, 1 above
// thr_ppm, else 0.
template <class W, int D>
__device__ __forceinline__ int sel_label(const ArgsT<W> &a, long c, int i, long st, float thr_ppm, float thr_weno) {
    float wt[NW], ww[NW];
    load_w<float>(a.ax[D].wsel[0] + (long)(i + 3) * NW, wt, 2);
    load_w<float>(a.ax[D].wsel[2] + (long)(i + 3) * NW, ww, 10);
    float err = 0.0f;
#pragma unroll
    for (int v = 0; v < 2; ++v) {
        const W *x = v == 0 ? a.P[0] : a.P[4];
        float s[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) s[m] = (float)x[c + (long)(m - 2) * st];
        float tl, tr, wl, wr;
        recon_tvd<float, 2>(wt[0], wt[1], s[1], s[2], s[3], tl, tr);
        recon_weno<float>(ww, s[0], s[1], s[2], s[3], s[4], wl, wr);
        float scale = fabsf(s[2]) + 1.0e-30f;
        err = fmaxf(err, fmaxf(fabsf(wl - tl), fabsf(wr - tr)) / scale);
    }
    return err > thr_weno ? 2 : (err > thr_ppm ? 1 : 0);
}

// cell of the selector's region along D: 0..n+1 along D (the cells the sweep
// reconstructs), 1..n across
template <int D, class A>
__device__ __forceinline__ long sel_cell(const A &a, long q, int &i) {
    const int ex = (D == 0) ? 2 : 0, ey = (D == 1) ? 2 : 0;
    const long n0 = a.nx + ex, n1 = a.ny + ey;
    int j = (int)(q % n0), k = (int)((q / n0) % n1), l = (int)(q / (n0 * n1));
    j += (D == 0) ? 0 : 1;
    k += (D == 1) ? 0 : 1;
    l += (D == 2) ? 0 : 1;
    i = (D == 0) ? j : (D == 1 ? k : l);
    return cidx(a, j, k, l);
}

template <int D, class A> __device__ __forceinline__ long sel_count(const A &a) {
    return (long)(a.nx + (D == 0 ? 2 : 0)) * (a.ny + (D == 1 ? 2 : 0)) * (a.nz + (D == 2 ? 2 : 0));
}

// ---------------------------------------------------------------------------
// inference on MFMA
// ---------------------------------------------------------------------------
//
// v_mfma_f32_16x16x16_bf16, lane l (0..63), r = l%16, g = l/16:
//   A (16x16): A[r][4g..4g+3]   B (16x16): B[4g..4g+3][r]   D (16x16 f32): D[4g..4g+3][r].
// Each layer computes H'^T = W^T H^T: A = W^T (output units x inputs) is a
// fixed fragment, B = H^T holds 4 inputs of cell r, and D holds 4 outputs of
// cell r -- exactly the B layout of the next layer, so activations never
// leave the lane.  Fragments (4 bf16 per lane each):
//   0,1: layer 1, output tile 0/1;  2..5: layer 2, (output tile, input tile)
//   = (0,0),(0,1),(1,0),(1,1);  6,7: layer 3, input tile 0/1 (outputs 0..2).

// Pack the FP32 parameters into the BF16 A fragments (64 lanes x 8 fragments x 4).
static __global__ void k_sel_pack(const float *__restrict__ w, unsigned short *__restrict__ frag) {
    const int t = threadIdx.x;   // 0..511: lane + 64*fragment
    if (t >= 64 * SEL_NFRAG) return;
    const int lane = t & 63, fr = t >> 6, r = lane & 15, g = lane >> 4;
    for (int e = 0; e < 4; ++e) {
        float v = 0.0f;
        if (fr < 2) {
            int o = fr * 16 + r, k = 4 * g + e;
            v = w[SEL_W1 + k * SEL_H + o];
        } else if (fr < 6) {
            int to = (fr - 2) >> 1, tk = (fr - 2) & 1;
            int o = to * 16 + r, k = tk * 16 + 4 * g + e;
            v = w[SEL_W2 + k * SEL_H + o];
        } else {
            int tk = fr - 6, o = r, k = tk * 16 + 4 * g + e;
            v = o < SEL_OUT ? w[SEL_W3 + k * SEL_OUT + o] : 0.0f;
        }
        frag[(fr * 64 + lane) * 4 + e] = sel_bf16(v);
    }
}

struct SelNet {
    sel_s4 A[SEL_NFRAG];
    float b1[2][4], b2[2][4], b3[4];
};

__device__ __forceinline__ void sel_load_net(const unsigned short *frag, const float *w, SelNet &n) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
#pragma unroll
    for (int fr = 0; fr < SEL_NFRAG; ++fr) {
        const unsigned short *p = frag + (fr * 64 + lane) * 4;
        n.A[fr] = sel_s4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        n.b1[0][e] = w[SEL_B1 + 4 * g + e];
        n.b1[1][e] = w[SEL_B1 + 16 + 4 * g + e];
        n.b2[0][e] = w[SEL_B2 + 4 * g + e];
        n.b2[1][e] = w[SEL_B2 + 16 + 4 * g + e];
        n.b3[e] = (4 * g + e) < SEL_OUT ? w[SEL_B3 + 4 * g + e] : 0.0f;
    }
}

__device__ __forceinline__ sel_s4 sel_act(sel_f4 d, const float b[4]) {
    sel_s4 r;
    r.x = (short)sel_bf16(fmaxf(d.x + b[0], 0.0f));
    r.y = (short)sel_bf16(fmaxf(d.y + b[1], 0.0f));
    r.z = (short)sel_bf16(fmaxf(d.z + b[2], 0.0f));
    r.w = (short)sel_bf16(fmaxf(d.w + b[3], 0.0f));
    return r;
}

// Logits of 16 cells: x = this lane's 4 features (inputs 4g..4g+3 of cell r).
// Returns D3: lanes 0..15 hold logits 0..3 (3 used) of cell r.
__device__ __forceinline__ sel_f4 sel_mlp16_b(const SelNet &n, sel_s4 bx) {
    const sel_f4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    sel_f4 d1a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(n.A[0], bx, z, 0, 0, 0);
    sel_f4 d1b = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(n.A[1], bx, z, 0, 0, 0);
    sel_s4 h1a = sel_act(d1a, n.b1[0]), h1b = sel_act(d1b, n.b1[1]);
    sel_f4 d2a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(n.A[2], h1a, z, 0, 0, 0);
    d2a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(n.A[3], h1b, d2a, 0, 0, 0);
    sel_f4 d2b = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(n.A[4], h1a, z, 0, 0, 0);
    d2b = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(n.A[5], h1b, d2b, 0, 0, 0);
    sel_s4 h2a = sel_act(d2a, n.b2[0]), h2b = sel_act(d2b, n.b2[1]);
    sel_f4 d3 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(n.A[6], h2a, z, 0, 0, 0);
    d3 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(n.A[7], h2b, d3, 0, 0, 0);
    d3.x += n.b3[0];
    d3.y += n.b3[1];
    d3.z += n.b3[2];
    d3.w += n.b3[3];
    return d3;
}

__device__ __forceinline__ sel_f4 sel_mlp16(const SelNet &n, const float x[4]) {
    sel_s4 bx = {(short)sel_bf16(x[0]), (short)sel_bf16(x[1]), (short)sel_bf16(x[2]), (short)sel_bf16(x[3])};
    return sel_mlp16_b(n, bx);
}

__device__ __forceinline__ int sel_argmax(sel_f4 d) {
    // ties go to the higher-order class
    int c = 2;
    float m = d.z;
    if (d.y > m) { c = 1; m = d.y; }
    if (d.x > m) c = 0;
    return c;
}

// Class mask of every cell the sweep along D reconstructs.  Each wave takes
// 64 consecutive cells (one per lane, coalesced stencil loads), stages their
// BF16 features in its own LDS tile, and runs four 16-cell MFMA groups from
// it.  Waves never wait on each other, so only wavefront-scope ordering is
// needed around the tile.  Blocks are dealt to the 8 XCDs round-robin; with a
// block count that is a multiple of 8, the waves of one XCD walk one
// contiguous eighth of the cells, so the rows above and below (which the y
// and z stencils read, and neighbouring chunks share) stay in that XCD's L2.
template <class W, int D>
__global__ __launch_bounds__(256) void k_sel_infer(ArgsT<W> a, const unsigned short *frag, const float *w,
                                                   unsigned char *mask) {
    __shared__ unsigned short tile[4][64][SEL_IN + 4];
    SelNet net;
    sel_load_net(frag, w, net);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const long nq = sel_count<D>(a);
    const long nchunk = (nq + 63) / 64;
    const int wpb = blockDim.x >> 6;
    long first = 0, last = nchunk, wave, nwaves;
    if (gridDim.x % 8 == 0) {
        const int xcd = blockIdx.x % 8;
        first = nchunk * xcd / 8;
        last = nchunk * (xcd + 1) / 8;
        wave = (long)(blockIdx.x / 8) * wpb + wv;
        nwaves = (long)(gridDim.x / 8) * wpb;
    } else {
        wave = (long)blockIdx.x * wpb + wv;
        nwaves = (long)gridDim.x * wpb;
    }
    const long st = D == 0 ? 1L : (D == 1 ? a.NX : a.NX * a.NY);
    for (long ch = first + wave; ch < last; ch += nwaves) {
        const long q = ch * 64 + lane;
        const bool ok = q < nq;
        int i = 0;
        const long c = ok ? sel_cell<D>(a, q, i) : 0;
        float f[SEL_IN];
        if (ok) sel_features(a, c, st, f);
        else
            for (int e = 0; e < SEL_IN; ++e) f[e] = 0.0f;
#pragma unroll
        for (int e = 0; e < SEL_IN; ++e) tile[wv][lane][e] = sel_bf16(f[e]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        unsigned char cls_out = 0;
#pragma unroll
        for (int sub = 0; sub < 4; ++sub) {
            const unsigned short *src = &tile[wv][sub * 16 + r][4 * g];
            sel_s4 bx = {(short)src[0], (short)src[1], (short)src[2], (short)src[3]};
            sel_f4 d = sel_mlp16_b(net, bx);
            // lanes 0..15 hold the logits of cell sub*16 + r; hand them to that lane
            const int cls = sel_argmax(d);
            const int want = __shfl(cls, lane & 15, 64);
            if ((lane >> 4) == sub) cls_out = (unsigned char)want;
        }
        if (ok) mask[c] = cls_out;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Logits of given feature rows (testing the MFMA path against a host model).
static __global__ __launch_bounds__(64) void k_sel_mlp_rows(const unsigned short *frag, const float *w, const float *feat,
                                                     long n, float *logits) {
    SelNet net;
    sel_load_net(frag, w, net);
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (long grp = blockIdx.x; grp * 16 < n; grp += gridDim.x) {
        const long q = grp * 16 + r;
        float x[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) x[e] = q < n ? feat[q * SEL_IN + 4 * g + e] : 0.0f;
        sel_f4 d = sel_mlp16(net, x);
        if (q < n && g == 0) {
            logits[q * 3 + 0] = d.x;
            logits[q * 3 + 1] = d.y;
            logits[q * 3 + 2] = d.z;
        }
    }
}

// ---------------------------------------------------------------------------
// training data and training
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned sel_hash(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Training samples by reservoir sampling over every candidate offered since
// cumc3d_selector_init: candidate s (the candidates of earlier collects plus
// its index in this one, `base + t`) takes slot s while the buffer has room;
// after that it replaces slot r = hash(s) % (s + 1) when r < cap, so each
// candidate offered so far is held with probability cap / seen and a late
// snapshot is represented as well as an early one.  The candidates are every
// stride-th cell (from a hashed offset) of the region along D.
__device__ __forceinline__ unsigned long long sel_slot(unsigned long long s, unsigned seed) {
    const unsigned h = sel_hash((unsigned)s ^ sel_hash((unsigned)(s >> 32) ^ seed));
    const unsigned long long r = ((unsigned long long)sel_hash(h ^ 0x5bd1e995U) << 32) | h;
    return r % (s + 1);
}

template <int D, class A> __device__ __forceinline__ long sel_candidate(const A &a, long t, int stride, unsigned seed) {
    const long q = t * stride + (long)(sel_hash(seed ^ (unsigned)t) % (unsigned)stride);
    return q < sel_count<D>(a) ? q : -1;
}

template <class W, int D>
__device__ __forceinline__ void sel_store_sample(const ArgsT<W> &a, long q, float thr_ppm, float thr_weno, float *feat,
                                                 unsigned char *label, unsigned long long slot) {
    int i;
    const long c = sel_cell<D>(a, q, i);
    const long st = D == 0 ? 1L : (D == 1 ? a.NX : a.NX * a.NY);
    float f[SEL_IN];
    sel_features(a, c, st, f);
    for (int e = 0; e < SEL_IN; ++e) feat[slot * SEL_IN + e] = f[e];
    label[slot] = (unsigned char)sel_label<W, D>(a, c, i, st, thr_ppm, thr_weno);
}

// Pass 1: candidates below cap store their sample; later ones claim their
// slot in owner[] (the largest candidate index of this collect wins, id =
// index within the collect + 1; owner starts at 0).
template <class W, int D>
__global__ __launch_bounds__(256) void k_sel_collect(ArgsT<W> a, int stride, unsigned seed, float thr_ppm,
                                                     float thr_weno, float *feat, unsigned char *label, long cap,
                                                     unsigned long long base, long id0, int *owner) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long q = sel_candidate<D>(a, t, stride, seed);
    if (q < 0) return;
    const unsigned long long s = base + (unsigned long long)t;
    if (s < (unsigned long long)cap) {
        sel_store_sample<W, D>(a, q, thr_ppm, thr_weno, feat, label, s);
        return;
    }
    const unsigned long long r = sel_slot(s, seed);
    if (r < (unsigned long long)cap) atomicMax(&owner[r], (int)(id0 + t + 1));
}

// Pass 2: a candidate past cap that owns its slot replaces the sample there.
template <class W, int D>
__global__ __launch_bounds__(256) void k_sel_place(ArgsT<W> a, int stride, unsigned seed, float thr_ppm,
                                                   float thr_weno, float *feat, unsigned char *label, long cap,
                                                   unsigned long long base, long id0, const int *owner) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long q = sel_candidate<D>(a, t, stride, seed);
    if (q < 0) return;
    const unsigned long long s = base + (unsigned long long)t;
    if (s < (unsigned long long)cap) return;
    const unsigned long long r = sel_slot(s, seed);
    if (r < (unsigned long long)cap && owner[r] == (int)(id0 + t + 1))
        sel_store_sample<W, D>(a, q, thr_ppm, thr_weno, feat, label, r);
}

// counts[0] = samples held, counts[1..3] = their class histogram (counts zeroed before)
static __global__ __launch_bounds__(256) void k_sel_hist(const unsigned char *__restrict__ label, long n, int *counts) {
    __shared__ int h[3];
    if (threadIdx.x < 3) h[threadIdx.x] = 0;
    __syncthreads();
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = label[i];
        if (c < 3) atomicAdd(&h[c], 1);
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicAdd(&counts[1 + threadIdx.x], h[threadIdx.x]);
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[0] = (int)n;
}

static __global__ void k_fill_i32(int *p, long n, int v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

constexpr int SEL_TB = 128;   // samples (threads) per training block

// One minibatch gradient: each block takes SEL_TB samples drawn by hash,
// runs forward and backward in FP32 (weights rounded to BF16 as inference
// sees them), and writes its summed gradient and loss.  Parameter gradients
// are batch dot products over the block's activations held in LDS.
static __global__ __launch_bounds__(SEL_TB) void k_sel_grad(const float *__restrict__ w, const float *__restrict__ feat,
                                                     const unsigned char *__restrict__ label, int n, unsigned seed,
                                                     float cw0, float cw1, float cw2, float *part, float *lossp) {
    __shared__ float sw[SEL_NPARAM];
    __shared__ float sx[SEL_TB][SEL_IN + 1];
    __shared__ float sh1[SEL_TB][SEL_H + 1];
    __shared__ float sh2[SEL_TB][SEL_H + 1];
    __shared__ float sd1[SEL_TB][SEL_H + 1];
    __shared__ float sd2[SEL_TB][SEL_H + 1];
    __shared__ float sdz[SEL_TB][4];
    __shared__ float sl[SEL_TB];
    const int t = threadIdx.x;
    for (int p = t; p < SEL_NPARAM; p += SEL_TB) {
        float v = w[p];
        // the weights inference uses are BF16; the biases stay FP32
        bool isw = (p < SEL_B1) || (p >= SEL_W2 && p < SEL_B2) || (p >= SEL_W3 && p < SEL_B3);
        sw[p] = isw ? sel_from_bf16(sel_bf16(v)) : v;
    }
    const int idx = (int)(sel_hash(seed * 0x9E3779B9U ^ (unsigned)(blockIdx.x * SEL_TB + t)) % (unsigned)n);
    for (int f = 0; f < SEL_IN; ++f) sx[t][f] = sel_from_bf16(sel_bf16(feat[(long)idx * SEL_IN + f]));
    const int y = label[idx];
    __syncthreads();
    // forward
    float h1[SEL_H], h2[SEL_H], z[SEL_OUT];
    for (int o = 0; o < SEL_H; ++o) {
        float s = sw[SEL_B1 + o];
        for (int i = 0; i < SEL_IN; ++i) s += sx[t][i] * sw[SEL_W1 + i * SEL_H + o];
        h1[o] = fmaxf(s, 0.0f);
        sh1[t][o] = h1[o];
    }
    for (int o = 0; o < SEL_H; ++o) {
        float s = sw[SEL_B2 + o];
        for (int i = 0; i < SEL_H; ++i) s += h1[i] * sw[SEL_W2 + i * SEL_H + o];
        h2[o] = fmaxf(s, 0.0f);
        sh2[t][o] = h2[o];
    }
    for (int o = 0; o < SEL_OUT; ++o) {
        float s = sw[SEL_B3 + o];
        for (int i = 0; i < SEL_H; ++i) s += h2[i] * sw[SEL_W3 + i * SEL_OUT + o];
        z[o] = s;
    }
    // weighted softmax cross-entropy
    const float cw = y == 0 ? cw0 : (y == 1 ? cw1 : cw2);
    float m = fmaxf(z[0], fmaxf(z[1], z[2]));
    float e0 = expf(z[0] - m), e1 = expf(z[1] - m), e2 = expf(z[2] - m), se = e0 + e1 + e2;
    float pr[3] = {e0 / se, e1 / se, e2 / se};
    sl[t] = -cw * logf(fmaxf(pr[y], 1.0e-30f));
    float dz[3];
    for (int o = 0; o < 3; ++o) {
        dz[o] = cw * (pr[o] - (o == y ? 1.0f : 0.0f));
        sdz[t][o] = dz[o];
    }
    // backward through the hidden layers
    float d2[SEL_H];
    for (int i = 0; i < SEL_H; ++i) {
        float s = 0.0f;
        for (int o = 0; o < SEL_OUT; ++o) s += sw[SEL_W3 + i * SEL_OUT + o] * dz[o];
        d2[i] = h2[i] > 0.0f ? s : 0.0f;
        sd2[t][i] = d2[i];
    }
    for (int i = 0; i < SEL_H; ++i) {
        float s = 0.0f;
        for (int o = 0; o < SEL_H; ++o) s += sw[SEL_W2 + i * SEL_H + o] * d2[o];
        sd1[t][i] = h1[i] > 0.0f ? s : 0.0f;
    }
    __syncthreads();
    // parameter gradients: sums over the block's samples
    float *out = part + (long)blockIdx.x * SEL_NPARAM;
    for (int p = t; p < SEL_NPARAM; p += SEL_TB) {
        float s = 0.0f;
        if (p < SEL_B1) {
            int i = p / SEL_H, o = p % SEL_H;
            for (int b = 0; b < SEL_TB; ++b) s += sx[b][i] * sd1[b][o];
        } else if (p < SEL_W2) {
            int o = p - SEL_B1;
            for (int b = 0; b < SEL_TB; ++b) s += sd1[b][o];
        } else if (p < SEL_B2) {
            int i = (p - SEL_W2) / SEL_H, o = (p - SEL_W2) % SEL_H;
            for (int b = 0; b < SEL_TB; ++b) s += sh1[b][i] * sd2[b][o];
        } else if (p < SEL_W3) {
            int o = p - SEL_B2;
            for (int b = 0; b < SEL_TB; ++b) s += sd2[b][o];
        } else if (p < SEL_B3) {
            int i = (p - SEL_W3) / SEL_OUT, o = (p - SEL_W3) % SEL_OUT;
            for (int b = 0; b < SEL_TB; ++b) s += sh2[b][i] * sdz[b][o];
        } else {
            int o = p - SEL_B3;
            for (int b = 0; b < SEL_TB; ++b) s += sdz[b][o];
        }
        out[p] = s;
    }
    if (t == 0) {
        float s = 0.0f;
        for (int b = 0; b < SEL_TB; ++b) s += sl[b];
        lossp[blockIdx.x] = s;
    }
}

// Adam step on the mean gradient of nblk blocks; loss_out = mean loss.
static __global__ __launch_bounds__(256) void k_sel_adam(float *w, float *m, float *v, const float *part, const float *lossp,
                                                  int nblk, float lr, int step, float *loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const float inv = 1.0f / (float)(nblk * SEL_TB);
    if (p < SEL_NPARAM) {
        float gsum = 0.0f;
        for (int b = 0; b < nblk; ++b) gsum += part[(long)b * SEL_NPARAM + p];
        const float gr = gsum * inv, b1 = 0.9f, b2 = 0.999f;
        float mm = b1 * m[p] + (1.0f - b1) * gr;
        float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
        m[p] = mm;
        v[p] = vv;
        const float mh = mm / (1.0f - powf(b1, (float)step)), vh = vv / (1.0f - powf(b2, (float)step));
        w[p] -= lr * mh / (sqrtf(vh) + 1.0e-8f);
    }
    if (p == 0 && loss_out) {
        float s = 0.0f;
        for (int b = 0; b < nblk; ++b) s += lossp[b];
        *loss_out = s * inv;
    }
}

}  // namespace cumc3d
