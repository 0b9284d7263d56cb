// solver_state.hpp — the solver handle behind lbm_handle, shared by the
// single-simulation runtime (capi.cpp) and the batched form (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/lbman3d.h"
#include "common.hpp"
#include "runtime.hpp"

using lbm::BCDesc;
using lbm::Fields;
using lbm::Phys;

struct lbm_solver {
    lbm_config cfg{};
    long V = 0, vs = 0;
    bool fp64 = true;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    Fields<double> d64{};
    Fields<float> d32{};
    BCDesc<double> bc64{};
    BCDesc<float> bc32{};
    Phys<double> k64{};
    Phys<float> k32{};
    uint16_t* slots = nullptr;
    double* uprev = nullptr;
    double* partials = nullptr;
    double* diag_out = nullptr;
    int diag_blocks = 0;
    int pulled = 0;        // 0: g holds post-streaming f; 1: g holds post-collision
    int schedule = 0;      // FP32: 0 auto, 1 two-kernel, 2+k fused tile k
    int tuned = 0;         // schedule picked by lbm_autotune for "auto" (0: not tuned)
    bool fused_ok = false; // fused kernel usable (FP32, 32-bit offsets fit)
    bool fused64_ok = false; // fused FP64 kernel usable (fully periodic, 32-bit offsets fit)
    bool moving = false;

    size_t esize() const { return fp64 ? sizeof(double) : sizeof(float); }
    void* gptr() const { return fp64 ? (void*)d64.g : (void*)d32.g; }
};

namespace lbm {
// One full step of a handle (capi.cpp); `last` materialises rho and F.
int run_step(lbm_solver* s, bool last, KernelRecorder* rec);
// Tunes the automatic FP32 schedule on first use (capi.cpp).
void ensure_tuned(lbm_solver* s);
}  // namespace lbm
