// The FP32 split's state kernels: the update (flux difference, CT, RK
// combine in double, pressure recovery in double, FINDDT's per-cell step),
// the ghost fill, FINDDT and the layout kernels, in their working-type
// (float) instantiations.  This unit is compiled with contraction and the
// short single-precision divide and square root, but keeps IEEE division in
// double precision: the conserved-state accumulators and the energy-to-pressure
// recovery are double arithmetic as in the FP64 mode.
#include "context.hpp"

namespace cumc3d {
namespace f32 {

void update(cumc3d_ctx *x, const ArgsT<float> &a, int stage, dim3 grid) {
    launch_update<float>(x->d.coordinate_flag, stage, grid, x->stream, a);
}

void ghost(cumc3d_ctx *x, const ArgsT<float> &g) {
    hipLaunchKernelGGL(k_ghost<float>, dim3(nblk(ghost_count(*x))), dim3(256), 0, x->stream, g);
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
    const unsigned gi = nblk((long)x->nx * x->ny * x->nz);
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
