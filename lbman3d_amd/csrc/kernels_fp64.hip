// kernels_fp64.hip — FP64 instantiation of the step kernels. Built with
// -ffp-contract=off so that, with the operand order kept in physics.hpp,
// results reproduce the reference CPU arithmetic bit for bit.
#include "step_kernels.hpp"

namespace lbm {
template void launch_qstep<double>(const Fields<double>&, const BCDesc<double>&, const Phys<double>&, long,
                                   hipStream_t);
template void launch_force_lbm<double, false>(const Fields<double>&, const BCDesc<double>&, const Phys<double>&,
                                              long, int, int, hipStream_t);
template void launch_stream_only<double>(const Fields<double>&, const BCDesc<double>&, long, int, double*,
                                         hipStream_t);
template void launch_diagnostics<double>(const Fields<double>&, const BCDesc<double>&, const Phys<double>&, long,
                                         double*, double*, int, double*, hipStream_t);
template void launch_order_director<double>(const Fields<double>&, long, double*, double*, hipStream_t);
}  // namespace lbm
