// common.hpp — types shared by the host runtime (capi.cpp) and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbm {

// ---- D3Q15 velocity set (src/lattice_stencil.h:42-93) ------------------------
// Direction order, weights and the opposite / specular maps are the
// reference's, because they fix the population layout of the C ABI.
constexpr int kQ = 15;
__host__ __device__ constexpr int cx(int i) {
    constexpr int v[kQ] = {0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    return v[i];
}
__host__ __device__ constexpr int cy(int i) {
    constexpr int v[kQ] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    return v[i];
}
__host__ __device__ constexpr int cz(int i) {
    constexpr int v[kQ] = {0, 0, 0, 0, 0, 1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
    return v[i];
}
__host__ __device__ constexpr int ce(int i, int a) { return a == 0 ? cx(i) : (a == 1 ? cy(i) : cz(i)); }
__host__ __device__ constexpr double wgt(int i) {
    return i == 0 ? 2.0 / 9 : (i < 7 ? 1.0 / 9 : 1.0 / 72);
}
__host__ __device__ constexpr int opposite(int i) {
    constexpr int v[kQ] = {0, 3, 4, 1, 2, 6, 5, 13, 14, 11, 12, 9, 10, 7, 8};
    return v[i];
}
// Mirror partner across a wall normal to axis a: flips component a of e_i.
__host__ __device__ constexpr int mirror(int i, int a) {
    constexpr int v[3][kQ] = {{0, 3, 2, 1, 4, 5, 6, 8, 7, 10, 9, 12, 11, 14, 13},
                              {0, 1, 4, 3, 2, 5, 6, 10, 9, 8, 7, 14, 13, 12, 11},
                              {0, 1, 2, 3, 4, 6, 5, 11, 12, 13, 14, 7, 8, 9, 10}};
    return v[a][i];
}

// ---- boundary tags ----------------------------------------------------------
enum : int { QB_PERIODIC = 0, QB_NEUMANN = 1, QB_ANCHOR = 2 };
enum : int { UB_PERIODIC = 0, UB_NOSLIP = 1, UB_MOVING = 2, UB_SPECULAR = 3 };

// Pull-slot table: for a node whose position on each axis is classified as
// interior / low face / high face / both (n == 1), each incoming slot j is
// either streamed from the upwind node x - e_j (periodic wrap) or rebuilt
// from one of the node's own post-collision populations by a wall.
// Entry bits: [1:0] kind, [5:2] source direction, [8:6] face.
enum : int { SLOT_STREAM = 0, SLOT_BOUNCE = 1, SLOT_MIRROR = 2 };
constexpr int kClasses = 64;
__host__ __device__ inline int slot_kind(uint16_t e) { return e & 3; }
__host__ __device__ inline int slot_src(uint16_t e) { return (e >> 2) & 15; }
__host__ __device__ inline int slot_face(uint16_t e) { return (e >> 6) & 7; }
__host__ __device__ inline uint16_t make_slot(int kind, int src, int face) {
    return (uint16_t)(kind | (src << 2) | (face << 6));
}
__host__ __device__ inline int node_class(int x, int y, int z, int nx, int ny, int nz) {
    const int a = (x == 0 ? 1 : 0) | (x == nx - 1 ? 2 : 0);
    const int b = (y == 0 ? 1 : 0) | (y == ny - 1 ? 2 : 0);
    const int c = (z == 0 ? 1 : 0) | (z == nz - 1 ? 2 : 0);
    return a | (b << 2) | (c << 4);
}

// Boundary description handed to every kernel by value.
template <typename T>
struct BCDesc {
    int n[3];
    int qkind[6];        // per face LBM_Q_* (XLo, XHi, YLo, YHi, ZLo, ZHi)
    int ukind[6];        // per face LBM_U_*
    T anchor[6][5];      // strong-anchoring Q per face (xx, xy, xz, yy, yz)
    T wall_u[6][3];      // wall velocity per face (MovingWall; zero otherwise)
    int all_periodic_u;  // every velocity face periodic: no wall slots anywhere
    int slots_xy;        // every reachable class has the slot row of its x, y
                         // bits alone (the z bits change nothing)
    T rho_base;          // the density array holds rho - rho_base (rho0 for the
                         // shifted FP32 form, 0 for FP64)
};

// Physical constants in the arithmetic type of the kernel.
template <typename T>
struct Phys {
    T dt, rho0, L, A, B, C, lambda, gamma, alpha, mu;
    T omega, omega_p, omega_f;   // 1 - DT/TAUF, DT/TAUF, 1 - DT/2/TAUF
    int advection;               // 0 centred, 1 upwind
    int force_full;              // 1: active + passive divergence + friction
};

// Device buffers of one simulation (structure of arrays, cell = (z*ny+y)*nx+x).
template <typename T>
struct Fields {
    T* q[5];        // Q^n            (xx, xy, xz, yy, yz)
    T* qn[5];       // Q^{n+1}
    T* sig[5];      // Sigma^n
    T* tau[3];      // Tau^n (xy, xz, yz)
    T* F[3];        // body force
    T* u[3];        // velocity u^{n-1} (unfused kernels update it in place)
    T* un[3];       // velocity u^n written by the fused kernel
    T* rho;         // density, stored as rho - BCDesc::rho_base
    T* g;           // populations, source buffer   [15][vs]
    T* gn;          // populations, destination     [15][vs]
    const uint16_t* slots;  // [kClasses][15] pull-slot table
    long vs;        // population stride (>= V, multiple of 64)
};

}  // namespace lbm
