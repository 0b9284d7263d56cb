// fused_step.hpp — placeholder until the fused kernel lands.
#pragma once
#include "launch.hpp"
namespace lbm {
int fused_tile_count() { return 0; }
const char* fused_tile_name(int) { return "none"; }
void launch_fused_step_f32(const Fields<float>&, const BCDesc<float>&, const Phys<float>&, int, int, int,
                           hipStream_t) {}
}  // namespace lbm
