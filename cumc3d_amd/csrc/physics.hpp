// Device-side physics of the MHD update step: reconstruction, EOS, face
// states, Riemann solvers.  Everything is templated on the arithmetic type T
// (double for the FP64 mode, float for the FP32 split) and, where the normal
// direction matters, on the sweep direction D (0=x, 1=y, 2=z) so that the
// component permutation is resolved at compile time and every state vector
// lives in registers.
//
// Operation order follows the reference expression by expression (Fortran
// evaluates a*b*c as (a*b)*c and a+b+c as (a+b)+c); together with
// -ffp-contract=off and IEEE division/square root this makes the FP64 mode
// reproduce the reference bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace cumc3d {

template <class T> __device__ __forceinline__ T fsign1(T b) { return __builtin_copysign(T(1), b); }
// Fortran MIN/MAX as flang evaluates them: MIN(a,b) = a < b ? a : b, so a
// NaN or a tie yields the second argument.
template <class T> __device__ __forceinline__ T fmin2(T a, T b) { return (a < b) ? a : b; }
template <class T> __device__ __forceinline__ T fmax2(T a, T b) { return (a > b) ? a : b; }
// MIN(a, 0.0) as the reference evaluates it: (a < 0) ? a : +0.  The backend
// turns that select into v_min (and drops a following + 0), which returns -0
// for a = -0; a condition the compiler cannot see through keeps the select.
template <class T> __device__ __forceinline__ T fmin_zero(T a) {
    int lt = a < T(0);
    __asm__("" : "+v"(lt));
    return lt ? a : T(0);
}
template <class T> __device__ __forceinline__ T fabsT(T a) { return __builtin_fabs(a); }
__device__ __forceinline__ float fabsT(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ float fsign1(float b) { return __builtin_copysignf(1.0f, b); }
__device__ __forceinline__ double sqrtT(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float sqrtT(float a) { return __builtin_sqrtf(a); }

// Permutation (normal, transverse 1, transverse 2) of BUILDSTATES
// (solver/mhd/spatial.f90:307-331): x -> (x,y,z), y -> (y,z,x), z -> (z,x,y).
template <int D> struct Perm {
    static constexpr int n = D, t1 = (D + 1) % 3, t2 = (D + 2) % 3;
};

// Primitive face state: natural component order for v and B.
template <class T> struct Q {
    T rho, v[3], p, b[3];
};
// Conserved vector / flux in the reference's variable order
// (rho, mx, my, mz, tau, Bx, By, Bz) = indices 1..8 of SETUP_EQNS.
template <class T> struct V8 {
    T a[8];
};

// ---------------------------------------------------------------------------
// EOS hooks: EOSEPSILON_NM and EOSSOUNDSPEED (model/riemann/src/getepsilon.f90:4-17,
// soundspeed.f90:4-21), ideal gas.
// ---------------------------------------------------------------------------
template <class T> __device__ __forceinline__ T eos_eps(T g1, T rho, T p) { return p / rho / g1; }
template <class T> __device__ __forceinline__ T eos_cs(T g1, T p, T rho, T eps) {
    T dpdden = eps * g1;
    T dpdint = rho * g1;
    return sqrtT(dpdden + dpdint * p / (rho * rho));
}

// ---------------------------------------------------------------------------
// Reconstruction of one cell's two faces from its stencil.
// ---------------------------------------------------------------------------

// TVD slope limiters (tvd_module.f90:279-344); lim 0 minmod, 1 van Leer, 2 MC.
template <class T, int LIM>
__device__ __forceinline__ void recon_tvd(T cf, T cb, T vm, T vc, T vp, T &vm_out, T &vp_out) {
    if constexpr (sizeof(T) == 4 && LIM != 1) {
        // Single precision: the same limiters without a divide or a branch.
        // dqf = 2 df / cf and dqb = 2 db / cb carry the signs of df and db
        // (cf, cb > 0).  When dqf dqb > 0, minmod is sign(df) min(|dqf|, |dqb|)
        // and MC's dqf max(0, min((1 + r) / 2, cf, cb r)), r = dqb / dqf, is
        // sign(df) min(|df / cf + db / cb|, 2 |df|, 2 |db|); both are 0
        // otherwise.  (On gfx950 v_rcp_f32 issues at a third of a multiply's
        // rate, and each range-checked divide adds four instructions.)
        const T df = vp - vc, db = vc - vm;
        const T icf = __builtin_amdgcn_rcpf(cf), icb = __builtin_amdgcn_rcpf(cb);
        T m;
        if constexpr (LIM == 0)
            m = T(2) * __builtin_fminf(__builtin_fabsf(df * icf), __builtin_fabsf(db * icb));
        else
            m = __builtin_fminf(__builtin_fabsf(df * icf + db * icb),
                                T(2) * __builtin_fminf(__builtin_fabsf(df), __builtin_fabsf(db)));
        const T slope = (df * db > T(0)) ? __builtin_copysignf(m, df) : T(0);
        vp_out = vc + T(0.5) * slope;
        vm_out = vc - T(0.5) * slope;
        return;
    }
    T dqf = T(2) * (vp - vc) / cf;
    T dqb = T(2) * (vc - vm) / cb;
    T slope;
    if constexpr (LIM == 0) {
        slope = T(0.5) * (fsign1(dqf) + fsign1(dqb)) * fmin2(fabsT(dqf), fabsT(dqb));
    } else if constexpr (LIM == 1) {
        T dq2 = dqf * dqb;
        if (dq2 <= T(0)) slope = T(0);
        else slope = (dq2 * (cf * dqb + cb * dqf) / (dqb * dqb + dqf * dqf + dq2 * (cf + cb - T(2))));
    } else {
        if (dqf == T(0)) {
            slope = T(0);
        } else {
            T r = dqb / dqf;
            slope = fmax2(T(0), fmin2(fmin2(T(0.5) * (T(1) + r), cf), cb * r));
            slope = dqf * slope;
        }
    }
    vp_out = vc + T(0.5) * slope;
    vm_out = vc - T(0.5) * slope;
}

// PPM without steepening/flattening (ppmc_module.f90:359-464); w = 14
// interpolation weights of PPM_INTERPOLANT.
template <class T>
__device__ __forceinline__ void recon_ppm(const T *__restrict__ w, T vm2, T vm1, T vc, T vp1, T vp2, T &vm_out,
                                          T &vp_out) {
    T deltap1 = w[12] * (vp2 - vp1) + w[13] * (vp1 - vc);
    T deltac = w[10] * (vp1 - vc) + w[11] * (vc - vm1);
    T deltam1 = w[8] * (vc - vm1) + w[9] * (vm1 - vm2);
    T dmc = T(0), dmp1 = T(0), dmm1 = T(0);
    if ((vp1 - vc) * (vc - vm1) > T(0))
        dmc = fmin2(fmin2(fabsT(deltac), T(2) * fabsT(vc - vm1)), T(2) * fabsT(vc - vp1)) * fsign1(deltac);
    if ((vp2 - vp1) * (vp1 - vc) > T(0))
        dmp1 = fmin2(fmin2(fabsT(deltap1), T(2) * fabsT(vp1 - vc)), T(2) * fabsT(vp1 - vp2)) * fsign1(deltap1);
    if ((vc - vm1) * (vm1 - vm2) > T(0))
        dmm1 = fmin2(fmin2(fabsT(deltam1), T(2) * fabsT(vm1 - vm2)), T(2) * fabsT(vm1 - vc)) * fsign1(deltam1);
    T vp = vc + w[4] * (vp1 - vc) + w[5] * (vp1 - vc) + w[6] * dmp1 + w[7] * dmc;
    T vm = vm1 + w[0] * (vc - vm1) + w[1] * (vc - vm1) + w[2] * dmc + w[3] * dmm1;
    T vl = vm, vr = vp;
    if ((vr - vc) * (vc - vl) <= T(0)) {
        vp = vc;
        vm = vc;
    } else {
        if ((vr - vl) * (vl - T(3) * vc + T(2) * vr) < T(0)) vm = T(3) * vc - T(2) * vr;
        if ((vp - vm) * (T(3) * vc - T(2) * vl - vr) < T(0)) vp = T(3) * vc - T(2) * vl;
    }
    vm_out = vm;
    vp_out = vp;
}

// WENO-Z, q=2, eps=1e-40 (weno_module.f90:315-408); w = 10 weights of
// WENO_INTERPOLANT.  In single precision eps is 1e-30 (1e-40 is below the
// normal range) and tau5/(eps+beta) is capped at 1e18 so that its square
// stays finite; both change nothing unless a smoothness indicator is within
// 18 decades of zero while tau5 is not.
template <class T> __device__ __forceinline__ T weno_ratio(T tau5, T sp, T b) {
    T q = tau5 / (sp + b);
    if constexpr (sizeof(T) == 4) q = fmin2(q, T(1.0e18));
    return q;
}
template <class T>
__device__ __forceinline__ T weno_face(T P1, T P2, T P3, T B1, T B2, T B3) {
    const T sp = sizeof(T) == 4 ? T(1.0e-30) : T(1.0e-40);
    const T g1 = T(1.0 / 10.0), g2 = T(6.0 / 10.0), g3 = T(3.0 / 10.0);
    T tau5 = fabsT(B1 - B3);
    T q = weno_ratio(tau5, sp, B1);
    T w1 = g1 * (T(1) + q * q);
    q = weno_ratio(tau5, sp, B2);
    T w2 = g2 * (T(1) + q * q);
    q = weno_ratio(tau5, sp, B3);
    T w3 = g3 * (T(1) + q * q);
    return (w1 * P1 + w2 * P2 + w3 * P3) / (w1 + w2 + w3);
}

template <class T>
__device__ __forceinline__ void recon_weno(const T *__restrict__ w, T vm2, T vm1, T vc, T vp1, T vp2, T &vm_out,
                                           T &vp_out) {
    const T r1 = T(13.0 / 12.0), r2 = T(1.0 / 4.0);
    T U01 = w[0] * vc + w[1] * vp1;
    T U12 = w[2] * vp1 + w[3] * vp2;
    T U10 = w[4] * vm1 + w[5] * vc;
    T U21 = w[6] * vm2 + w[7] * vm1;
    T U00 = vm1 + w[8] * (vm1 - vm2);
    T U11 = vp1 + w[9] * (vp1 - vp2);
    T t, s;
    {
        T P1 = (T(6) * vc - T(1) * U10 - T(2) * U00) / T(3);
        T P2 = (-T(1) * U10 + T(2) * vc + T(2) * U01) / T(3);
        T P3 = (T(2) * U01 + T(2) * vp1 - T(1) * U12) / T(3);
        t = T(2) * U10 - T(2) * U00; s = T(4) * vc - T(2) * U10 - T(2) * U00;
        T B1 = r1 * (t * t) + r2 * (s * s);
        t = T(2) * U10 - T(4) * vc + T(2) * U01; s = -T(2) * U10 + T(2) * U01;
        T B2 = r1 * (t * t) + r2 * (s * s);
        t = T(2) * U01 - T(4) * vp1 + T(2) * U12; s = -T(6) * U01 + T(8) * vp1 - T(2) * U12;
        T B3 = r1 * (t * t) + r2 * (s * s);
        vp_out = weno_face(P1, P2, P3, B1, B2, B3);
    }
    {
        T P1 = (T(6) * vc - T(1) * U01 - T(2) * U11) / T(3);
        T P2 = (-T(1) * U01 + T(2) * vc + T(2) * U10) / T(3);
        T P3 = (T(2) * U10 + T(2) * vm1 - T(1) * U21) / T(3);
        t = T(2) * U01 - T(2) * U11; s = T(4) * vc - T(2) * U01 - T(2) * U11;
        T B1 = r1 * (t * t) + r2 * (s * s);
        t = T(2) * U01 - T(4) * vc + T(2) * U10; s = T(2) * U01 - T(2) * U10;
        T B2 = r1 * (t * t) + r2 * (s * s);
        t = T(2) * U10 - T(4) * vm1 + T(2) * U21; s = T(8) * vm1 - T(6) * U10 - T(2) * U21;
        T B3 = r1 * (t * t) + r2 * (s * s);
        vm_out = weno_face(P1, P2, P3, B1, B2, B3);
    }
}

// Scheme ids as in the build descriptor: 0 minmod, 1 van Leer, 2 MC, 3 PPM, 4 WENO-Z.
template <class T, int R>
__device__ __forceinline__ void recon1(const T *__restrict__ w, const T s[5], T &vm_out, T &vp_out) {
    if constexpr (R <= 2) recon_tvd<T, R>(w[0], w[1], s[1], s[2], s[3], vm_out, vp_out);
    else if constexpr (R == 3) recon_ppm<T>(w, s[0], s[1], s[2], s[3], s[4], vm_out, vp_out);
    else recon_weno<T>(w, s[0], s[1], s[2], s[3], s[4], vm_out, vp_out);
}

// Both faces of one cell: 5 hydro primitives and the 2 transverse cell-centred
// fields, then the EOS at each face (TVD_reconx etc., tvd_module.f90:97-113).
// Face "m" is the cell's left face (reference primR(c-1)), face "p" its right
// face (primL(c)).  The normal field component is filled in by the caller.
template <class T> struct Face {
    T rho, v[3], p, bt1, bt2, eps, cs;
};

template <class T, int R>
__device__ __forceinline__ void recon_cell(const T *__restrict__ w, const T st[7][5], T g1, Face<T> &fm,
                                           Face<T> &fp) {
    recon1<T, R>(w, st[0], fm.rho, fp.rho);
    recon1<T, R>(w, st[1], fm.v[0], fp.v[0]);
    recon1<T, R>(w, st[2], fm.v[1], fp.v[1]);
    recon1<T, R>(w, st[3], fm.v[2], fp.v[2]);
    recon1<T, R>(w, st[4], fm.p, fp.p);
    if constexpr (sizeof(T) == 4) {
        // single precision: one reciprocal per face.  eps = (p / rho) / g1 and
        // eps g1 + rho g1 p / rho^2 = (g1 + 1) p / rho, the ideal-gas c_s^2
        const T ig1 = __builtin_amdgcn_rcpf(g1);
        const T prm = fm.p * __builtin_amdgcn_rcpf(fm.rho), prp = fp.p * __builtin_amdgcn_rcpf(fp.rho);
        fm.eps = prm * ig1;
        fp.eps = prp * ig1;
        fm.cs = sqrtT((g1 + T(1)) * prm);
        fp.cs = sqrtT((g1 + T(1)) * prp);
    } else {
        fm.eps = eos_eps(g1, fm.rho, fm.p);
        fp.eps = eos_eps(g1, fp.rho, fp.p);
        fm.cs = eos_cs(g1, fm.p, fm.rho, fm.eps);
        fp.cs = eos_cs(g1, fp.p, fp.rho, fp.eps);
    }
    recon1<T, R>(w, st[5], fm.bt1, fp.bt1);
    recon1<T, R>(w, st[6], fm.bt2, fp.bt2);
}

// Face -> full primitive state with the normal field of the interface.
// bt1/bt2 in Face are the two transverse components in the reconstruction
// order of the sweep: x -> (By,Bz), y -> (Bx,Bz), z -> (Bx,By)
// (tvd_module.f90:108-109, 181-182, 254-255).
template <class T, int D>
__device__ __forceinline__ Q<T> to_q(const Face<T> &f, T bn) {
    Q<T> q;
    q.rho = f.rho;
    q.v[0] = f.v[0]; q.v[1] = f.v[1]; q.v[2] = f.v[2];
    q.p = f.p;
    if constexpr (D == 0) { q.b[0] = bn; q.b[1] = f.bt1; q.b[2] = f.bt2; }
    else if constexpr (D == 1) { q.b[0] = f.bt1; q.b[1] = bn; q.b[2] = f.bt2; }
    else { q.b[0] = f.bt1; q.b[1] = f.bt2; q.b[2] = bn; }
    return q;
}

template <class T> __device__ __forceinline__ T dot3(const T *a, const T *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// BUILDSTATES for one face state (spatial.f90:342-388): u (conserved) and f
// (physical flux along D).  Index layout of u/f: 0 rho, 1-3 momentum (x,y,z),
// 4 tau, 5-7 B (x,y,z).
template <class T, int D>
__device__ __forceinline__ void build_state(const Q<T> &q, T eps, T u[8], T f[8]) {
    constexpr int n = Perm<D>::n, t1 = Perm<D>::t1, t2 = Perm<D>::t2;
    T b2 = dot3(q.b, q.b), v2 = dot3(q.v, q.v), vb = dot3(q.b, q.v);
    u[0] = q.rho;
    u[1] = q.v[0] * q.rho;
    u[2] = q.v[1] * q.rho;
    u[3] = q.v[2] * q.rho;
    u[5] = q.b[0]; u[6] = q.b[1]; u[7] = q.b[2];
    u[4] = q.rho * (T(0.5) * v2 + eps) + T(0.5) * b2;
    T vn = q.v[n];
    for (int i = 0; i < 5; ++i) f[i] = u[i] * vn;
    f[1 + n] = f[1 + n] + q.p + T(0.5) * (b2) - q.b[n] * q.b[n];
    f[1 + t1] = f[1 + t1] - q.b[n] * q.b[t1];
    f[1 + t2] = f[1 + t2] - q.b[n] * q.b[t2];
    f[4] = f[4] + (q.p + T(0.5) * (b2)) * vn - q.b[n] * vb;
    f[5 + n] = T(0);
    f[5 + t1] = (vn * q.b[t1] - q.v[t1] * q.b[n]);
    f[5 + t2] = (vn * q.b[t2] - q.v[t2] * q.b[n]);
}

// rsolver_module.f90:32-93
template <class T>
__device__ __forceinline__ T signalspeed(T cs, T bn, T bt1, T bt2, T rho) {
    T a2 = cs * cs, a4 = a2 * a2;
    T b2n = (bn * bn / rho), b2t1 = (bt1 * bt1 / rho), b2t2 = (bt2 * bt2 / rho);
    T b2 = b2n + b2t1 + b2t2, b4 = b2 * b2;
    return sqrtT(T(0.5) * (a2 + b2 + sqrtT(fmax2((a4 + b4 + T(2) * b2 * a2) - T(4) * a2 * b2n, T(0)))));
}
template <class T> __device__ __forceinline__ T hll_state(T ul, T ur, T fl, T fr, T sl, T sr) {
    return (sr * ur - sl * ul - (fr - fl)) / (sr - sl);
}
template <class T> __device__ __forceinline__ T hll_flux(T yl, T yr, T xl, T xr, T sl, T sr) {
    return (sr * yl - sl * yr + sl * sr * (xr - xl)) / (sr - sl);
}
template <class T> __device__ __forceinline__ T roe_avg(T xl, T xr, T rl, T rr) {
    return (sqrtT(rl) * xl + sqrtT(rr) * xr) / (sqrtT(rl) + sqrtT(rr));
}
template <class T>
__device__ __forceinline__ T sstar_hllc(T pl, T pr, T rl, T rr, T vl, T vr, T sl, T sr, T bl, T br) {
    return (rr * vr * (sr - vr) - rl * vl * (sl - vl) + pl - pr + br * br - bl * bl) /
           (rr * (sr - vr) - rl * (sl - vl));
}

// Riemann solver ids: 0 LF, 1 HLL, 2 HLLC, 3 HLLD.
// Result of one interface.  ds: the HLLD double-star branch was taken;
// uss_t1/uss_t2: the transverse fields of that double-star state, needed to
// reproduce the reference's carried usstar(ibn) (see riemann()).
// none: HLL took none of its branches (a NaN signal speed) and assigned no flux.
template <class T> struct RiemannOut {
    T f[8];
    bool ds;
    bool none;
    T uss_t1, uss_t2;
};

// bn_mode: 0 -> usstar(ibn)=0, 1 -> Bn, 2 -> carried value `carry`.
template <class T, int D, int S>
__device__ __forceinline__ void riemann(const Q<T> &qL, const Q<T> &qR, T epsL, T epsR, T csL, T csR, int bn_mode,
                                        T carry, RiemannOut<T> &out) {
    constexpr int n = Perm<D>::n, t1 = Perm<D>::t1, t2 = Perm<D>::t2;
    T uL[8], uR[8], fL[8], fR[8];
    build_state<T, D>(qL, epsL, uL, fL);
    build_state<T, D>(qR, epsR, uR, fR);
    out.ds = false;
    out.none = false;
    out.uss_t1 = T(0);
    out.uss_t2 = T(0);
    T cfsL = signalspeed(csL, qL.b[n], qL.b[t1], qL.b[t2], qL.rho);
    T cfsR = signalspeed(csR, qR.b[n], qR.b[t1], qR.b[t2], qR.rho);
    if constexpr (S == 0) {   // LFNM, rsolver_module.f90:199-207
        T sL = fabsT(qL.v[n]) + cfsL, sR = fabsT(qR.v[n]) + cfsR, splus = fmax2(sL, sR);
        for (int i = 0; i < 8; ++i) out.f[i] = T(0.5) * (fL[i] + fR[i] - splus * (uR[i] - uL[i]));
        return;
    }
    T ubar = roe_avg(qL.v[n], qR.v[n], qL.rho, qR.rho);
    T cbar = roe_avg(cfsL, cfsR, qL.rho, qR.rho);
    T sL = fmin2(qL.v[n] - cfsL, ubar - cbar);
    T sR = fmax2(qR.v[n] + cfsR, ubar + cbar);
    if constexpr (S == 1) {   // HLLNM, rsolver_module.f90:303-311 (no ELSE branch)
        if (sL >= T(0)) { for (int i = 0; i < 8; ++i) out.f[i] = fL[i]; }
        else if (sL <= T(0) && sR >= T(0)) { for (int i = 0; i < 8; ++i) out.f[i] = hll_flux(fL[i], fR[i], uL[i], uR[i], sL, sR); }
        else if (sR <= T(0)) { for (int i = 0; i < 8; ++i) out.f[i] = fR[i]; }
        else out.none = true;
        return;
    }
    if (sL >= T(0)) { for (int i = 0; i < 8; ++i) out.f[i] = fL[i]; return; }
    if (sR <= T(0)) { for (int i = 0; i < 8; ++i) out.f[i] = fR[i]; return; }
    T b2L = dot3(qL.b, qL.b), b2R = dot3(qR.b, qR.b);
    T pLt = qL.p + T(0.5) * b2L, pRt = qR.p + T(0.5) * b2R;
    if constexpr (S == 2) {   // HLLCNM, rsolver_module.f90:425-499
        T sstar = sstar_hllc(pLt, pRt, qL.rho, qR.rho, qL.v[n], qR.v[n], sL, sR, qL.b[n], qR.b[n]);
        T bh[3], vh[3];
        bh[n] = qL.b[n];
        bh[t1] = hll_state(qL.b[t1], qR.b[t1], fL[5 + t1], fR[5 + t1], sL, sR);
        bh[t2] = hll_state(qL.b[t2], qR.b[t2], fL[5 + t2], fR[5 + t2], sL, sR);
        T rh = hll_state(qL.rho, qR.rho, fL[0], fR[0], sL, sR);
        T mh1 = hll_state(uL[1 + t1], uR[1 + t1], fL[1 + t1], fR[1 + t1], sL, sR);
        T mh2 = hll_state(uL[1 + t2], uR[1 + t2], fL[1 + t2], fR[1 + t2], sL, sR);
        vh[n] = sstar;
        vh[t1] = mh1 / rh;
        vh[t2] = mh2 / rh;
        const bool left = sstar >= T(0);
        const Q<T> &q = left ? qL : qR;
        const T *u = left ? uL : uR;
        const T *f = left ? fL : fR;
        T s = left ? sL : sR, pt = left ? pLt : pRt;
        T us[8];
        us[0] = u[0] * (s - q.v[n]) / (s - sstar);
        T vb = dot3(q.b, q.v);
        T vbhll = dot3(bh, vh);
        T ps = q.rho * (s - q.v[n]) * (sstar - q.v[n]) + pt;
        us[5 + n] = bh[n];
        us[5 + t1] = bh[t1];
        us[5 + t2] = bh[t2];
        us[1 + n] = us[0] * sstar;
        us[1 + t1] = (u[1 + t1] * (s - q.v[n]) - (us[5 + n] * us[5 + t1] - q.b[n] * q.b[t1])) / (s - sstar);
        us[1 + t2] = (u[1 + t2] * (s - q.v[n]) - (us[5 + n] * us[5 + t2] - q.b[n] * q.b[t2])) / (s - sstar);
        us[4] = (u[4] * (s - q.v[n]) + (ps * sstar - pt * q.v[n]) - ((vbhll * us[5 + n] - vb * q.b[n]))) / (s - sstar);
        for (int i = 0; i < 8; ++i) out.f[i] = f[i] + s * (us[i] - u[i]);
        return;
    }
    // HLLDNM, rsolver_module.f90:615-752
    T sM = sstar_hllc(pLt, pRt, qL.rho, qR.rho, qL.v[n], qR.v[n], sL, sR, qL.b[n], qR.b[n]);
    T usL[8], usR[8], psL[3], psR[3];
    for (int i = 0; i < 5; ++i) usL[i] = uL[i] * (sL - qL.v[n]) / (sL - sM);
    for (int i = 0; i < 5; ++i) usR[i] = uR[i] * (sR - qR.v[n]) / (sR - sM);
    T sstarL = sM - fabsT(qL.b[n]) / sqrtT(usL[0]);
    T sstarR = sM + fabsT(qR.b[n]) / sqrtT(usR[0]);
    if (((sstarL - sL) < T(1.0e-4) * (sM - sL)) || ((sstarR - sR) > -T(1.0e-4) * (sR - sM))) {
        usL[5 + n] = qL.b[n];
        usL[5 + t1] = hll_state(qL.b[t1], qR.b[t1], fL[5 + t1], fR[5 + t1], sL, sR);
        usL[5 + t2] = hll_state(qL.b[t2], qR.b[t2], fL[5 + t2], fR[5 + t2], sL, sR);
        usR[5 + n] = qR.b[n];
        usR[5 + t1] = usL[5 + t1];
        usR[5 + t2] = usL[5 + t2];
        sstarL = sM;
        sstarR = sM;
    } else {
        T fac = (qL.rho * (sL - qL.v[n]) * (sL - sM) - qL.b[n] * qL.b[n]);
        usL[5 + n] = qL.b[n];
        usL[5 + t1] = (qL.b[t1] * (qL.rho * (sL - qL.v[n]) * (sL - qL.v[n]) - qL.b[n] * qL.b[n])) / fac;
        usL[5 + t2] = (qL.b[t2] * (qL.rho * (sL - qL.v[n]) * (sL - qL.v[n]) - qL.b[n] * qL.b[n])) / fac;
        fac = (qR.rho * (sR - qR.v[n]) * (sR - sM) - qR.b[n] * qR.b[n]);
        usR[5 + n] = qR.b[n];
        usR[5 + t1] = (qR.b[t1] * (qR.rho * (sR - qR.v[n]) * (sR - qR.v[n]) - qR.b[n] * qR.b[n])) / fac;
        usR[5 + t2] = (qR.b[t2] * (qR.rho * (sR - qR.v[n]) * (sR - qR.v[n]) - qR.b[n] * qR.b[n])) / fac;
    }
    usL[1 + n] = usL[0] * sM;
    usL[1 + t1] = usL[1 + t1] - (usL[5 + n] * usL[5 + t1] - qL.b[n] * qL.b[t1]) / (sL - sM);
    usL[1 + t2] = usL[1 + t2] - (usL[5 + n] * usL[5 + t2] - qL.b[n] * qL.b[t2]) / (sL - sM);
    usR[1 + n] = usR[0] * sM;
    usR[1 + t1] = usR[1 + t1] - (usR[5 + n] * usR[5 + t1] - qR.b[n] * qR.b[t1]) / (sR - sM);
    usR[1 + t2] = usR[1 + t2] - (usR[5 + n] * usR[5 + t2] - qR.b[n] * qR.b[t2]) / (sR - sM);
    psL[n] = sM;
    psL[t1] = usL[1 + t1] / usL[0];
    psL[t2] = usL[1 + t2] / usL[0];
    psR[n] = sM;
    psR[t1] = usR[1 + t1] / usR[0];
    psR[t2] = usR[1 + t2] / usR[0];
    T pLs = qL.rho * (sL - qL.v[n]) * (sM - qL.v[n]) + pLt;
    T pRs = qR.rho * (sR - qR.v[n]) * (sM - qR.v[n]) + pRt;
    T bdotu = dot3(qL.b, qL.v);
    T bdotus = dot3(&usL[5], psL);
    usL[4] = usL[4] + ((pLs * sM - pLt * qL.v[n]) - (bdotus - bdotu) * qL.b[n]) / (sL - sM);
    bdotu = dot3(qR.b, qR.v);
    bdotus = dot3(&usR[5], psR);
    usR[4] = usR[4] + ((pRs * sM - pRt * qR.v[n]) - (bdotus - bdotu) * qR.b[n]) / (sR - sM);
    if (sstarL >= T(0)) {
        for (int i = 0; i < 8; ++i) out.f[i] = fL[i] + sL * (usL[i] - uL[i]);
        return;
    }
    if (sstarR <= T(0)) {
        for (int i = 0; i < 8; ++i) out.f[i] = fR[i] + sR * (usR[i] - uR[i]);
        return;
    }
    // double-star states.  usstar(ibn) is never assigned by the reference
    // (rsolver_module.f90:704-741): its value is selected by bn_mode.
    T ussL[8], ussR[8], pss[3];
    T bnL = (bn_mode == 1) ? qL.b[n] : ((bn_mode == 2) ? carry : T(0));
    T bnR = (bn_mode == 1) ? qR.b[n] : ((bn_mode == 2) ? carry : T(0));
    ussL[5 + n] = bnL;
    ussR[5 + n] = bnR;
    ussL[0] = usL[0];
    ussR[0] = usR[0];
    T f1 = T(1) / (sqrtT(usL[0]) + sqrtT(usR[0]));
    T f2 = fsign1(qL.b[n]);
    pss[n] = sM;
    pss[t1] = f1 * (sqrtT(usL[0]) * psL[t1] + sqrtT(usR[0]) * psR[t1] + (usR[5 + t1] - usL[5 + t1]) * f2);
    pss[t2] = f1 * (sqrtT(usL[0]) * psL[t2] + sqrtT(usR[0]) * psR[t2] + (usR[5 + t2] - usL[5 + t2]) * f2);
    ussL[5 + t1] = f1 * (sqrtT(usL[0]) * usR[5 + t1] + sqrtT(usR[0]) * usL[5 + t1] +
                         sqrtT(usL[0] * usR[0]) * (psR[t1] - psL[t1]) * f2);
    ussL[5 + t2] = f1 * (sqrtT(usL[0]) * usR[5 + t2] + sqrtT(usR[0]) * usL[5 + t2] +
                         sqrtT(usL[0] * usR[0]) * (psR[t2] - psL[t2]) * f2);
    ussL[1 + n] = ussL[0] * pss[n];
    ussL[1 + t1] = ussL[0] * pss[t1];
    ussL[1 + t2] = ussL[0] * pss[t2];
    bdotus = dot3(&usL[5], psL);
    T bdotuss = dot3(&ussL[5], pss);
    ussL[4] = usL[4] - sqrtT(usL[0]) * (bdotus - bdotuss) * f2;
    ussR[5 + t1] = ussL[5 + t1];
    ussR[5 + t2] = ussL[5 + t2];
    ussR[1 + n] = ussR[0] * pss[n];
    ussR[1 + t1] = ussR[0] * pss[t1];
    ussR[1 + t2] = ussR[0] * pss[t2];
    bdotus = dot3(&usR[5], psR);
    bdotuss = dot3(&ussR[5], pss);
    ussR[4] = usR[4] + sqrtT(usR[0]) * (bdotus - bdotuss) * f2;
    out.ds = true;
    out.uss_t1 = ussL[5 + t1];
    out.uss_t2 = ussL[5 + t2];
    if (sM >= T(0)) {
        for (int i = 0; i < 8; ++i) out.f[i] = fL[i] + sL * (usL[i] - uL[i]) + sstarL * (ussL[i] - usL[i]);
    } else {
        for (int i = 0; i < 8; ++i) out.f[i] = fR[i] + sR * (usR[i] - uR[i]) + sstarR * (ussR[i] - usR[i]);
    }
}

}  // namespace cumc3d
