// launch.hpp — host-callable launchers of the step kernels. Instantiated for
// double in kernels_fp64.hip (no FMA contraction, bitwise parity path) and for
// float in kernels_fp32.hip (shifted populations, contraction on).
#pragma once

#include "common.hpp"

namespace lbm {

// Which parts of the force + collide-stream kernel run.
enum : int { FL_FORCE = 1, FL_LBM = 2, FL_STORE_F = 4, FL_STORE_RHO = 8 };

// Q update: Q^n, u^{n-1} -> Q^{n+1}, Sigma^n, Tau^n, F = Ericksen force.
template <typename T>
void launch_qstep(const Fields<T>& f, const BCDesc<T>& bc, const Phys<T>& k, long V, hipStream_t s);

// Body force and/or collide-stream. pull = 0 reads the node's own slots
// (populations uploaded as post-streaming values), pull = 1 gathers from the
// upwind neighbours / walls (post-collision storage).
template <typename T, bool SHIFTED>
void launch_force_lbm(const Fields<T>& f, const BCDesc<T>& bc, const Phys<T>& k, long V, int flags,
                      int pull, hipStream_t s);

// Post-streaming populations f(x, j) of the stored state into out[15][vs].
template <typename T>
void launch_stream_only(const Fields<T>& f, const BCDesc<T>& bc, long V, int pull, T* out,
                        hipStream_t s);

// FP64 reductions: per-block partials of mass, px, py, pz, kinetic, du2, u2,
// umax^2, free energy, nonfinite count; then a single-block final pass.
template <typename T>
void launch_diagnostics(const Fields<T>& f, const BCDesc<T>& bc, const Phys<double>& k, long V,
                        double* uprev, double* partials, int nblocks, double* out, hipStream_t s);

// Order parameter and director per cell in FP64.
template <typename T>
void launch_order_director(const Fields<T>& f, long V, double* order, double* director, hipStream_t s);

// Fused single-kernel FP32 step (Q update + body force + collide-stream).
// tile: index of the tile shape (see fused_tile_count()).
int fused_tile_count();
const char* fused_tile_name(int tile);
void launch_fused_step_f32(const Fields<float>& f, const BCDesc<float>& bc, const Phys<float>& k,
                           int tile, int pull, int store_aux, hipStream_t s);

// Fused single-kernel FP64 step on fully periodic domains: the first z-marching
// pipeline with the unfused kernels' per-cell arithmetic, so its results are
// those of launch_qstep + launch_force_lbm bit for bit (requires
// 15 * vs * 8 < 2^32).
int fused64_tile_count();
const char* fused64_tile_name(int tile);
void launch_fused_step_f64(const Fields<double>& f, const BCDesc<double>& bc, const Phys<double>& k, int tile,
                           int pull, int store_aux, hipStream_t s);

// Plain device-to-device copy kernel (HBM bandwidth calibration).
void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);

}  // namespace lbm
