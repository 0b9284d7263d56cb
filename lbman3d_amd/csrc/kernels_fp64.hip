// kernels_fp64.hip — FP64 instantiation of the step kernels. Built with
// -ffp-contract=off so that, with the operand order kept in physics.hpp,
// results reproduce the reference CPU arithmetic bit for bit.
#include "step_kernels.hpp"
#define LBM_FUSED_BODY_ONLY
#include "fused_step.hpp"

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

// FP64 tile catalogue of the fused step (first pipeline, periodic
// specialisation). Rows of 16 doubles are one 128-byte line.
struct Tile64 {
    int tx, ty, zc;
    const char* name;
};
constexpr Tile64 kTiles64[] = {
    {32, 4, 64, "fp64:32x4x64"}, {32, 4, 128, "fp64:32x4x128"}, {16, 8, 64, "fp64:16x8x64"},
    {16, 8, 32, "fp64:16x8x32"}, {32, 8, 64, "fp64:32x8x64"},   {64, 4, 64, "fp64:64x4x64"},
    {32, 4, 32, "fp64:32x4x32"},
};
constexpr int kNumTiles64 = sizeof(kTiles64) / sizeof(kTiles64[0]);

int fused64_tile_count() { return kNumTiles64; }
const char* fused64_tile_name(int tile) { return (tile >= 0 && tile < kNumTiles64) ? kTiles64[tile].name : "?"; }

template <int TX, int TY, int ZC>
static void launch_tile64(const Fields<double>& f, const BCDesc<double>& bc, const Phys<double>& k, int pull,
                          int aux, hipStream_t s) {
    const int ntx = (bc.n[0] + TX - 1) / TX, nty = (bc.n[1] + TY - 1) / TY, nzc = (bc.n[2] + ZC - 1) / ZC;
    const int grid = (ntx * nty * nzc + 7) / 8 * 8;
    // Grids of whole column tiles take the form with every memory
    // instruction issued unconditionally.
    if (bc.n[0] % TX == 0 && bc.n[1] % TY == 0)
        hipLaunchKernelGGL((fused_step_kernel<TX, TY, ZC, true, double, true>), dim3(grid), dim3(TX * TY), 0, s, f,
                           bc, k, pull, aux, ntx, nty, nzc, 0);
    else
        hipLaunchKernelGGL((fused_step_kernel<TX, TY, ZC, true, double>), dim3(grid), dim3(TX * TY), 0, s, f, bc, k,
                           pull, aux, ntx, nty, nzc, 0);
}

void launch_fused_step_f64(const Fields<double>& f, const BCDesc<double>& bc, const Phys<double>& k, int tile,
                           int pull, int aux, hipStream_t s) {
    switch (tile) {
        case 0: launch_tile64<32, 4, 64>(f, bc, k, pull, aux, s); break;
        case 1: launch_tile64<32, 4, 128>(f, bc, k, pull, aux, s); break;
        case 2: launch_tile64<16, 8, 64>(f, bc, k, pull, aux, s); break;
        case 3: launch_tile64<16, 8, 32>(f, bc, k, pull, aux, s); break;
        case 4: launch_tile64<32, 8, 64>(f, bc, k, pull, aux, s); break;
        case 5: launch_tile64<64, 4, 64>(f, bc, k, pull, aux, s); break;
        default: launch_tile64<32, 4, 32>(f, bc, k, pull, aux, s); break;
    }
}
}  // namespace lbm
