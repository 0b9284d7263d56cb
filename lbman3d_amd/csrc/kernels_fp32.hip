// kernels_fp32.hip — FP32 instantiation of the step kernels (shifted
// populations, FMA contraction on) and the fused single-kernel FP32 step.
#include "step_kernels.hpp"
#include "fused_step.hpp"

namespace lbm {
template void launch_qstep<float>(const Fields<float>&, const BCDesc<float>&, const Phys<float>&, long,
                                  hipStream_t);
template void launch_force_lbm<float, true>(const Fields<float>&, const BCDesc<float>&, const Phys<float>&, long,
                                            int, int, hipStream_t);
template void launch_stream_only<float>(const Fields<float>&, const BCDesc<float>&, long, int, float*,
                                        hipStream_t);
template void launch_diagnostics<float>(const Fields<float>&, const BCDesc<float>&, const Phys<double>&, long,
                                        double*, double*, int, double*, hipStream_t);
template void launch_order_director<float>(const Fields<float>&, long, double*, double*, hipStream_t);
}  // namespace lbm
