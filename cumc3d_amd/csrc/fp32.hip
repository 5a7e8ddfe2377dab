// The FP32 split's stage sequence, flux sweeps, EMF kernels and selector:
// the working-type (float) instantiations of the stage and selector kernels.
// This unit is compiled with contraction (v_fma_f32), the short
// single-precision divide and square root, and reciprocal divides; the
// update, ghost fill, FINDDT and layout kernels are in fp32_state.hip, whose
// double arithmetic keeps IEEE division, and the FP64 mode lives in
// capi.hip, compiled without any of these, so it keeps the reference's
// arithmetic bit for bit.
#include "context.hpp"

namespace cumc3d {
namespace f32 {

void run_stage(cumc3d_ctx *x, int stage) { run_stage_t<float>(x, stage); }

void spatial(cumc3d_ctx *x, double *lrk) { spatial_t<float>(x, lrk); }

void sel_infer(cumc3d_ctx *x) { sel_infer_t<float>(x); }

void sel_collect(cumc3d_ctx *x, int stride, unsigned seed, float t1, float t2) {
    sel_collect_t<float>(x, stride, seed, t1, t2);
}

}  // namespace f32
}  // namespace cumc3d
