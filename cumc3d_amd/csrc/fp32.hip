// The FP32 split's kernels: every working-type (float) instantiation of the
// stage, selector and layout kernels.  This unit is compiled with
// contraction (v_fma_f32) and the short single-precision divide and square
// root; the FP64 mode lives in capi.hip, compiled without either, so it
// keeps the reference's arithmetic bit for bit.
#include "context.hpp"

namespace cumc3d {
namespace f32 {

void run_stage(cumc3d_ctx *x, int stage) { run_stage_t<float>(x, stage); }

void spatial(cumc3d_ctx *x, double *lrk) { spatial_t<float>(x, lrk); }

void sel_infer(cumc3d_ctx *x) { sel_infer_t<float>(x); }

void sel_collect(cumc3d_ctx *x, int stride, unsigned seed, float t1, float t2) {
    sel_collect_t<float>(x, stride, seed, t1, t2);
}

void finddt(cumc3d_ctx *x) {
    const long nn = (long)x->nx * x->ny * x->nz;
    hipLaunchKernelGGL(k_finddt<float>, dim3(nblk(nn)), dim3(256), 0, x->stream, make_args<float>(x));
}

void split_prim(cumc3d_ctx *x, const double *aos, const double *eps64) {
    hipLaunchKernelGGL(k_split_prim<float>, dim3(nblk(x->G)), dim3(256), 0, x->stream, aos, eps64,
                       make_args<float>(x));
}

void initial_update(cumc3d_ctx *x) {
    const unsigned gi = nblk((long)x->nx * x->ny * x->nz), gg = nblk(x->G);
    hipLaunchKernelGGL(k_init_interior<float>, dim3(gi), dim3(256), 0, x->stream, make_args<float>(x));
    hipLaunchKernelGGL(k_ghost<float>, dim3(nblk(ghost_count(*x))), dim3(256), 0, x->stream, make_args<float>(x));
}

void join_prim(cumc3d_ctx *x, double *tmp) {
    hipLaunchKernelGGL(k_join_prim<float>, dim3(nblk(x->G)), dim3(256), 0, x->stream, tmp, make_args<float>(x));
}

void join_bcell(cumc3d_ctx *x, double *tmp) {
    hipLaunchKernelGGL(k_join3<float>, dim3(nblk(x->G)), dim3(256), 0, x->stream, tmp, (const float *)x->Bc[0],
                       (const float *)x->Bc[1], (const float *)x->Bc[2], x->G);
}

void widen(cumc3d_ctx *x, double *tmp, const void *src) {
    hipLaunchKernelGGL(k_widen<float>, dim3(nblk(x->G)), dim3(256), 0, x->stream, tmp, (const float *)src, x->G);
}

}  // namespace f32
}  // namespace cumc3d
