// fused_step.hpp — the whole coupled step in one FP32 kernel.
//
// Per output cell the step needs Q^{n+1}, Sigma^n and Tau^n at the cell and
// its six neighbours (body force: grad Q^{n+1} and div(Sigma + Tau),
// src/qtensor_solver.tpp:154-217), which in turn need Q^n and u^{n-1} two
// cells away. Instead of writing Q^{n+1}, Sigma, Tau and the Ericksen force
// to HBM and reading them back (the reference's three-kernel data flow), a
// block owns a TX x TY column tile, marches up a chunk of z planes, and keeps
//   * the current input plane (Q^n, u^{n-1}) of the (TX+4) x (TY+4) region
//     in LDS for xy neighbours, with z neighbours in registers;
//   * the Q-update results (Q^{n+1}, Sigma, Tau) of the (TX+2) x (TY+2)
//     region for the last two planes in LDS;
// and finishes each plane with the force, pull gather and BGK collision of
// the interior cells. HBM traffic per cell is then the compulsory state only:
// populations 15 + 15, Q 5 + 5, u 3 + 3 values.
//
// Boundary handling is the unfused kernels' (physics.hpp, gather): every
// column of the extended tile holds the values of the domain column that the
// reference's index rule maps it to (periodic wrap or clamp, per face tag),
// and neighbour reads of the force stage substitute the cell itself where
// the reference clamps.
#pragma once

#include <algorithm>
#include <type_traits>

#include "launch.hpp"
#include "physics.hpp"
#include "shell_region.hpp"
#include "tile_map.hpp"
#include "step_kernels.hpp"

namespace lbm {

// Global access as uniform base + 32-bit byte offset, so the compiler can use
// the scalar-base addressing form (one VGPR per address instead of two).
__device__ __forceinline__ float ldo(const float* base, unsigned boff) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + boff);
}
__device__ __forceinline__ void sto(float* base, unsigned boff, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + boff) = v;
}
__device__ __forceinline__ double ldo(const double* base, unsigned boff) {
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + boff);
}
__device__ __forceinline__ void sto(double* base, unsigned boff, double v) {
    *reinterpret_cast<double*>(reinterpret_cast<char*>(base) + boff) = v;
}
// Store with an optional non-temporal hint (streaming output read back only
// next step, so it need not displace the halo rows held in L2).
template <bool NT>
__device__ __forceinline__ void sto_hint(float* base, unsigned boff, float v) {
    float* p = reinterpret_cast<float*>(reinterpret_cast<char*>(base) + boff);
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Pull gather with 32-bit byte offsets (fused kernels; requires
// 15 * vs * sizeof(T) < 2^32).
template <bool PERIODIC, typename T>
__device__ __forceinline__ void gather32(const Fields<T>& f, const BCDesc<T>& bc, int x, int y, int z,
                                         unsigned c, int pull, T out[kQ]) {
    constexpr unsigned E = sizeof(T);
    const int nx = bc.n[0], ny = bc.n[1], nz = bc.n[2];
    const unsigned vs4 = (unsigned)f.vs * E;
    const T* g = f.g;
    if (!pull) {
#pragma unroll
        for (int j = 0; j < kQ; ++j) out[j] = ldo(g, j * vs4 + c * E);
        return;
    }
    const int cls = node_class(x, y, z, nx, ny, nz);
    if (PERIODIC || cls == 0) {
        const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
        const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
        const int zm = z == 0 ? nz - 1 : z - 1, zp = z == nz - 1 ? 0 : z + 1;
        const int xs[3] = {xp, x, xm}, ys[3] = {yp, y, ym}, zs[3] = {zp, z, zm};
#pragma unroll
        for (int j = 0; j < kQ; ++j) {
            const unsigned src = ((unsigned)zs[cz(j) + 1] * ny + ys[cy(j) + 1]) * nx + xs[cx(j) + 1];
            out[j] = ldo(g, j * vs4 + src * E);
        }
        return;
    }
    const uint16_t* row = f.slots + cls * kQ;
    T rho_prev = T(0);
    bool have_rho = false;
#pragma unroll
    for (int j = 0; j < kQ; ++j) {
        const uint16_t e = row[j];
        const int kind = slot_kind(e);
        if (kind == SLOT_STREAM) {
            int xs = x - cx(j), ys = y - cy(j), zs = z - cz(j);
            xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
            ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
            zs = zs < 0 ? zs + nz : (zs >= nz ? zs - nz : zs);
            out[j] = ldo(g, j * vs4 + (((unsigned)zs * ny + ys) * nx + xs) * E);
        } else if (kind == SLOT_MIRROR) {
            out[j] = ldo(g, slot_src(e) * vs4 + c * E);
        } else {
            const int face = slot_face(e);
            if (!have_rho) { rho_prev = bc.rho_base + ldo(f.rho, c * E); have_rho = true; }
            const T dot = T(cx(j)) * bc.wall_u[face][0] + T(cy(j)) * bc.wall_u[face][1]
                        + T(cz(j)) * bc.wall_u[face][2];
            out[j] = ldo(g, slot_src(e) * vs4 + c * E) + T(6.0) * rho_prev * T(wgt(j)) * dot;
        }
    }
}

// Pull gather of the wall-aware third pipeline, given the slot-table row of
// the cell's x, y class (the rows do not depend on z: BCDesc::slots_xy)
// packed two 16-bit entries per word. No face has a moving wall, so a
// bounce-back slot reads its source population at the cell itself like a
// mirror slot, and each population is one load from a selected (population,
// cell) pair: the values of gather32<false> for those tags.
__device__ __forceinline__ void gather32_row(const Fields<float>& f, const BCDesc<float>& bc, int x, int y, int z,
                                             unsigned c, int pull, const uint32_t* srow, float out[kQ]) {
    const int nx = bc.n[0], ny = bc.n[1], nz = bc.n[2];
    const unsigned vs4 = (unsigned)f.vs * 4u;
    if (!pull) {
#pragma unroll
        for (int j = 0; j < kQ; ++j) out[j] = ldo(f.g, j * vs4 + c * 4u);
        return;
    }
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const int zm = z == 0 ? nz - 1 : z - 1, zp = z == nz - 1 ? 0 : z + 1;
    const int xs[3] = {xp, x, xm}, ys[3] = {yp, y, ym}, zs[3] = {zp, z, zm};
#pragma unroll
    for (int j = 0; j < kQ; ++j) {
        const uint16_t e = (uint16_t)(srow[j >> 1] >> (16 * (j & 1)));
        const unsigned nb = ((unsigned)zs[cz(j) + 1] * ny + ys[cy(j) + 1]) * nx + xs[cx(j) + 1];
        const unsigned off = slot_kind(e) == SLOT_STREAM ? j * vs4 + nb * 4u : slot_src(e) * vs4 + c * 4u;
        out[j] = ldo(f.g, off);
    }
}

template <int TX, int TY>
struct TileGeom {
    static constexpr int NT = TX * TY;               // threads = interior columns
    static constexpr int AX = TX + 2, AY = TY + 2;   // Q-update region
    static constexpr int IX = TX + 4, IY = TY + 4;   // input region
    static constexpr int NA = AX * AY, NI = IX * IY;
    static constexpr int NR = 2 * AX + 2 * (AY - 2); // Q-update ring columns
    static constexpr int NO = 2 * IX + 2 * (IY - 2); // input outer ring columns
    static_assert(NR <= NT && NO <= NT, "tile too small for its halo");
};

// Border column r of a W x H rectangle (top row, bottom row, left, right).
__device__ __forceinline__ void ring_col(int r, int W, int H, int& a, int& b) {
    if (r < W) { a = r; b = 0; }
    else if (r < 2 * W) { a = r - W; b = H - 1; }
    else if (r < 2 * W + (H - 2)) { a = 0; b = 1 + r - 2 * W; }
    else { a = W - 1; b = 1 + r - 2 * W - (H - 2); }
}

struct AxisTags {
    bool qlo, qhi, ulo, uhi;   // periodic flags per side for Q and velocity tags
};

// Q-update (stage A) results kept per plane.
template <typename T>
struct StageAT {
    T qn[5], sg[5], ta[3];
};
using StageA = StageAT<float>;

constexpr int kNumIn = 8;    // Q (5) + u (3)
constexpr int kNumA = 13;    // Q^{n+1} (5) + Sigma (5) + Tau (3)

template <int TX, int TY, typename T>
__device__ __forceinline__ void load_column(const Fields<T>& f, int qcol, int ucol_xy, unsigned qplane,
                                            unsigned uplane, T v[kNumIn]) {
    constexpr unsigned E = sizeof(T);
    const unsigned qo = (qplane + (unsigned)qcol) * E, uo = (uplane + (unsigned)ucol_xy) * E;
#pragma unroll
    for (int m = 0; m < 5; ++m) v[m] = ldo(f.q[m], qo);
#pragma unroll
    for (int m = 0; m < 3; ++m) v[5 + m] = ldo(f.u[m], uo);
}

// Stage A at one column: Beris-Edwards update from the input plane in LDS
// (xy neighbours) and the column's own z neighbours.
template <int TX, int TY, bool NOGHOST, typename T>
__device__ __forceinline__ void stage_a(const BCDesc<T>& bc, const Phys<T>& k,
                                        const T (*sIn)[TileGeom<TX, TY>::NI], int ii, int jj,
                                        const T zm[kNumIn], const T c[kNumIn], const T zp[kNumIn],
                                        const int qpos[3], const int upos[3], StageAT<T>& out, T fe[3]) {
    using G = TileGeom<TX, TY>;
    const int cc = jj * G::IX + ii;
    const int xm = cc - 1, xp = cc + 1, ym = cc - G::IX, yp = cc + G::IX;
    T Q[5];
    Deriv<T> dq[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        Q[m] = c[m];
        dq[m] = q_deriv<T, NOGHOST>(bc, m, qpos, c[m], sIn[m][xm], sIn[m][xp], sIn[m][ym], sIn[m][yp], zm[m], zp[m]);
    }
    T u[3], g[3][3];
#pragma unroll
    for (int A = 0; A < 3; ++A) {
        const T v0 = c[5 + A];
        u[A] = v0;
        g[A][0] = u_grad<T, NOGHOST>(bc, 0, A, upos[0], sIn[5 + A][xm], v0, sIn[5 + A][xp]);
        g[A][1] = u_grad<T, NOGHOST>(bc, 1, A, upos[1], sIn[5 + A][ym], v0, sIn[5 + A][yp]);
        g[A][2] = u_grad<T, NOGHOST>(bc, 2, A, upos[2], zm[5 + A], v0, zp[5 + A]);
    }
    q_pointwise(k, Q, u, dq, g, out.qn, out.sg, out.ta, fe);
}

// Register-lean form of the Q update for the FP32 fused kernels: the same
// formulas as q_pointwise (src/model.h:102-342), evaluated component by
// component so that the 35 stencil derivatives are never live together.
// Operand order inside each formula follows q_pointwise; the Ericksen force
// is accumulated per component (a different summation order, within FP32
// rounding of q_pointwise).
// GH: 1 every face periodic (no ghosts), 0 the general ghost rules, 2 walls
// on x and y faces only (Neumann Q; no-slip or specular velocity; z periodic).
template <int TX, int TY, int GH, bool UPWIND>
__device__ __forceinline__ void stage_a_lean(const BCDesc<float>& bc, const Phys<float>& k,
                                             const float (*sIn)[TileGeom<TX, TY>::NI], int ii, int jj,
                                             const float zm[kNumIn], const float c[kNumIn],
                                             const float zp[kNumIn], const int qpos[3], const int upos[3],
                                             StageA& out, float fe[3]) {
    using G = TileGeom<TX, TY>;
    const int cc = jj * G::IX + ii;
    const int xm = cc - 1, xp = cc + 1, ym = cc - G::IX, yp = cc + G::IX;
    // Velocity gradient -> strain rate E and vorticity W.
    float g[3][3];
#pragma unroll
    for (int A = 0; A < 3; ++A) {
        const float v0 = c[5 + A];
        if constexpr (GH == 2) {
            g[A][0] = u_grad_xy<float>(bc, 0, A, upos[0], sIn[5 + A][xm], v0, sIn[5 + A][xp]);
            g[A][1] = u_grad_xy<float>(bc, 1, A, upos[1], sIn[5 + A][ym], v0, sIn[5 + A][yp]);
            g[A][2] = u_grad<float, true>(bc, 2, A, upos[2], zm[5 + A], v0, zp[5 + A]);
        } else {
            g[A][0] = u_grad<float, GH == 1>(bc, 0, A, upos[0], sIn[5 + A][xm], v0, sIn[5 + A][xp]);
            g[A][1] = u_grad<float, GH == 1>(bc, 1, A, upos[1], sIn[5 + A][ym], v0, sIn[5 + A][yp]);
            g[A][2] = u_grad<float, GH == 1>(bc, 2, A, upos[2], zm[5 + A], v0, zp[5 + A]);
        }
    }
    const float Exx = g[0][0], Eyy = g[1][1];
    const float Exy = 0.5f * (g[0][1] + g[1][0]);
    const float Exz = 0.5f * (g[0][2] + g[2][0]);
    const float Eyz = 0.5f * (g[1][2] + g[2][1]);
    const float Wxy = 0.5f * (g[0][1] - g[1][0]);
    const float Wxz = 0.5f * (g[0][2] - g[2][0]);
    const float Wyz = 0.5f * (g[1][2] - g[2][1]);
    const float Qxx = c[0], Qxy = c[1], Qxz = c[2], Qyy = c[3], Qyz = c[4];
    const float u0 = c[5], u1 = c[6], u2 = c[7];
    const float TrQ2 = 2.0f * (Qxx * Qxx + Qyy * Qyy + Qxx * Qyy + Qxy * Qxy + Qxz * Qxz + Qyz * Qyz);
    const float third = 1.0f / 3.0f;
    const float ld = k.A + k.C * TrQ2;

    float H[5], adv[5];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    float dxx0 = 0.0f, dxx1 = 0.0f, dxx2 = 0.0f;   // grad Qxx, kept for the cross term with Hyy
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const Deriv<float> d =
            GH == 2 ? q_deriv_xy<float>(bc, qpos, c[m], sIn[m][xm], sIn[m][xp], sIn[m][ym], sIn[m][yp], zm[m], zp[m])
                    : q_deriv<float, GH == 1>(bc, m, qpos, c[m], sIn[m][xm], sIn[m][xp], sIn[m][ym], sIn[m][yp], zm[m],
                                               zp[m]);
        float Q2;
        switch (m) {
            case 0: Q2 = Qxx * Qxx + Qxy * Qxy + Qxz * Qxz - third * TrQ2; break;
            case 1: Q2 = Qxx * Qxy + Qxy * Qyy + Qxz * Qyz; break;
            case 2: Q2 = Qxy * Qyz - Qxz * Qyy; break;
            case 3: Q2 = Qxy * Qxy + Qyy * Qyy + Qyz * Qyz - third * TrQ2; break;
            default: Q2 = Qxy * Qxz - Qyz * Qxx; break;
        }
        // With upwind advection the per-axis second differences exist anyway;
        // their sum is the same Laplacian with three fewer additions.
        const float lap = UPWIND ? (d.d2[0] + d.d2[1]) + d.d2[2] : d.lap;
        H[m] = k.L * lap - ld * c[m] - k.B * Q2;
        adv[m] = -(adv_axis(UPWIND, u0, d.d[0], d.d2[0]) + adv_axis(UPWIND, u1, d.d[1], d.d2[1])
                   + adv_axis(UPWIND, u2, d.d[2], d.d2[2]));
        // -H : grad_a Q over the nine index pairs (Qzz = -Qxx - Qyy, Hzz = -Hxx - Hyy).
        fx -= 2.0f * H[m] * d.d[0];
        fy -= 2.0f * H[m] * d.d[1];
        fz -= 2.0f * H[m] * d.d[2];
        if (m == 0) { dxx0 = d.d[0]; dxx1 = d.d[1]; dxx2 = d.d[2]; }
        if (m == 3) {
            fx -= H[0] * d.d[0] + H[3] * dxx0;
            fy -= H[0] * d.d[1] + H[3] * dxx1;
            fz -= H[0] * d.d[2] + H[3] * dxx2;
        }
    }
    fe[0] = fx;
    fe[1] = fy;
    fe[2] = fz;
    const float Hxx = H[0], Hxy = H[1], Hxz = H[2], Hyy = H[3], Hyz = H[4];

    const float two3 = 2.0f / 3.0f;
    const float trQH = (2.0f * Hxx + Hyy) * Qxx + (2.0f * Hxy) * Qxy + (2.0f * Hxz) * Qxz + (Hxx + 2.0f * Hyy) * Qyy
                     + 2.0f * Hyz * Qyz;
    out.ta[0] = (Hxy * Qxx + Hyy * Qxy + Hyz * Qxz) - (Qxy * Hxx + Qyy * Hxy + Qyz * Hxz);
    out.ta[1] = (Hxz * Qxx + Hyz * Qxy + (-Hxx - Hyy) * Qxz) - (Qxz * Hxx + Qyz * Hxy + (-Qxx - Qyy) * Hxz);
    out.ta[2] = (Hxz * Qxy + Hyz * Qyy + (-Hxx - Hyy) * Qyz) - (Qxz * Hxy + Qyz * Hyy + (-Qxx - Qyy) * Hyz);
    const float trQE = (2.0f * Exx + Eyy) * Qxx + (2.0f * Exy) * Qxy + (2.0f * Exz) * Qxz + (Exx + 2.0f * Eyy) * Qyy
                     + 2.0f * Eyz * Qyz;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        float QH, aln2, cor, E;
        switch (m) {
            case 0:
                QH = 2.0f * (Hxx * Qxx + Hxy * Qxy + Hxz * Qxz) - two3 * trQH;
                aln2 = 2.0f * (Exx * Qxx + Exy * Qxy + Exz * Qxz) - two3 * trQE;
                cor = 2.0f * (Qxy * Wxy + Qxz * Wxz);
                E = Exx;
                break;
            case 1:
                QH = Hxy * Qxx + Hyy * Qxy + Hyz * Qxz + Qxy * Hxx + Qyy * Hxy + Qyz * Hxz;
                aln2 = Exy * Qxx + Eyy * Qxy + Eyz * Qxz + Qxy * Exx + Qyy * Exy + Qyz * Exz;
                cor = -Qxx * Wxy + Qxz * Wyz + Qyy * Wxy + Qyz * Wxz;
                E = Exy;
                break;
            case 2:
                QH = Hxz * Qxx + Hyz * Qxy + (-Hxx - Hyy) * Qxz + Qxz * Hxx + Qyz * Hxy + (-Qxx - Qyy) * Hxz;
                aln2 = Exz * Qxx + Eyz * Qxy + (-Exx - Eyy) * Qxz + Qxz * Exx + Qyz * Exy + (-Qxx - Qyy) * Exz;
                cor = -2.0f * Qxx * Wxz - Qxy * Wyz - Qyy * Wxz + Qyz * Wxy;
                E = Exz;
                break;
            case 3:
                QH = 2.0f * (Hxy * Qxy + Hyy * Qyy + Hyz * Qyz) - two3 * trQH;
                aln2 = 2.0f * (Exy * Qxy + Eyy * Qyy + Eyz * Qyz) - two3 * trQE;
                cor = 2.0f * (-Qxy * Wxy + Qyz * Wyz);
                E = Eyy;
                break;
            default:
                QH = Hxz * Qxy + Hyz * Qyy + (-Hxx - Hyy) * Qyz + Qxz * Hxy + Qyz * Hyy + (-Qxx - Qyy) * Hyz;
                aln2 = Exz * Qxy + Eyz * Qyy + (-Exx - Eyy) * Qyz + Qxz * Exy + Qyz * Eyy + (-Qxx - Qyy) * Eyz;
                cor = -Qxx * Wyz - Qxy * Wxz - Qxz * Wxy - 2.0f * Qyy * Wyz;
                E = Eyz;
                break;
        }
        const float aln3 = -2.0f * trQE * c[m];
        out.sg[m] = -two3 * k.lambda * H[m] - k.lambda * QH + 2.0f * k.lambda * trQH * c[m];
        out.qn[m] = c[m] + k.dt * (adv[m] + cor + k.lambda * (two3 * E + aln2 + aln3) + k.gamma * H[m]);
    }
}

// Components of the stage-A results whose z differences the force needs:
// Q^{n+1} xx, xz, yy, yz; Sigma xx, xz, yy, yz; Tau xz, yz.
constexpr int kNumDZ = 10;
__device__ __forceinline__ int dz_slot(int m) {
    switch (m) {
        case 0: return 0;
        case 2: return 1;
        case 3: return 2;
        case 4: return 3;
        case 5: return 4;
        case 7: return 5;
        case 8: return 6;
        case 9: return 7;
        case 11: return 8;
        case 12: return 9;
        default: return -1;
    }
}
template <typename T>
__device__ __forceinline__ T stage_value(const StageAT<T>& s, int m) {
    return m < 5 ? s.qn[m] : (m < 10 ? s.sg[m - 5] : s.ta[m - 10]);
}

// Cell classes on domains with walls. A face is open unless both its Q and
// its velocity tags are periodic. A cell at least two cells away from every
// open face has its whole dependency cone (inputs two cells out, pull
// sources and force stencil one cell out) inside the domain and away from
// every ghost, clamp or wall slot, so the periodic specialisation computes it
// exactly: class 0. The others lie in the shell of an axis, x first: class
// 1 + axis.
//
// Write modes of a launch: 0 every cell, 1 class 0 only, 2 + a the shell of
// axis a (the launches use modes 0 and 1; the shells are written by the
// per-cell shell passes). A mode gates the stores only; the force stage runs
// for every active cell as in an unmasked launch. (With the whole force stage under the
// mask, the build for gfx950 applied the Neumann ghost across a periodic z
// face in the lanes that wrote, at the first and last planes; gating the
// stores alone gives the reference values.)
struct CellMask {
    int mode, col, zlo, zhi, nz;
    __device__ __forceinline__ bool writes(int z) const {
        const int zin = (zlo & (z < 2)) | (zhi & (z >= nz - 2));
        const int cls = col + (col == 0) * 3 * zin;
        return (mode == 0) | ((mode == 1) & (cls == 0)) | ((mode >= 2) & (cls == mode - 1));
    }
};
template <typename T>
__device__ __forceinline__ CellMask cell_mask(const BCDesc<T>& bc, int mode, int x, int y) {
    int open[6];
#pragma unroll
    for (int face = 0; face < 6; ++face)
        open[face] = (bc.qkind[face] != QB_PERIODIC) | (bc.ukind[face] != UB_PERIODIC);
    const int inx = (open[0] & (x < 2)) | (open[1] & (x >= bc.n[0] - 2));
    const int iny = (open[2] & (y < 2)) | (open[3] & (y >= bc.n[1] - 2));
    CellMask m;
    m.mode = mode;
    m.col = inx + (1 - inx) * 2 * iny;
    m.zlo = open[4];
    m.zhi = open[5];
    m.nz = bc.n[2];
    return m;
}

// First pipeline, one tile: z planes [tz*ZC, tz*ZC + ZC) of the TX x TY
// column tile (tx, ty), writing the cells that `mode` selects. LDS buffers
// come from the calling kernel so that a kernel instantiating several
// variants does not allocate LDS per variant.
//
// UNCOND (fully periodic grids whose x and y sides are whole tiles, every
// cell written): the ring and outer-ring loads, the pull gather and the
// stores are issued by every thread and the first two planes of the chunk
// are peeled, so that no memory instruction sits under a branch. The memory
// counter waits the compiler places after such a branch assume its loads or
// stores were skipped and then wait for everything issued before, the
// previous plane's stores included, and at one wave per SIMD (FP64) nothing
// covers that wait.
template <int TX, int TY, int ZC, bool PERIODIC, typename T, bool UNCOND = false>
__device__ __forceinline__ void fused_body(const Fields<T>& f, const BCDesc<T>& bc, const Phys<T>& k,
                                           int pull, int aux, int tx, int ty, int tz,
                                           T (*sIn)[TileGeom<TX, TY>::NI],
                                           T (*sA)[kNumA][TileGeom<TX, TY>::NA],
                                           T (*sRing)[kNumIn][TileGeom<TX, TY>::NR], int mode = 0) {
    constexpr unsigned E = sizeof(T);
    using G = TileGeom<TX, TY>;
    const int nx = bc.n[0], ny = bc.n[1], nz = bc.n[2];
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = tz * ZC, z1 = min(z0 + ZC, nz);
    const unsigned plane = (unsigned)nx * ny;

    // Periodic flags per face; all true in the fully periodic instantiation,
    // which then compiles without any ghost, clamp or wall-slot code.
    AxisTags ax[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ax[a].qlo = PERIODIC || bc.qkind[2 * a] == QB_PERIODIC;
        ax[a].qhi = PERIODIC || bc.qkind[2 * a + 1] == QB_PERIODIC;
        ax[a].ulo = PERIODIC || bc.ukind[2 * a] == UB_PERIODIC;
        ax[a].uhi = PERIODIC || bc.ukind[2 * a + 1] == UB_PERIODIC;
    }

    const int t = threadIdx.x;
    // Interior column of this thread (stage-A region coordinates ia, ib).
    const int ia = t % TX + 1, ib = t / TX + 1;
    const int acell = ib * G::AX + ia;
    const int gx = x0 + ia - 1, gy = y0 + ib - 1;
    const bool active = gx < nx && gy < ny;
    const CellMask wmask = cell_mask(bc, mode, gx, gy);
    // Stage-A ring column (threads t < NR) and input outer-ring column (last NO threads).
    const bool has_ring = t < G::NR;
    int ra = 0, rb = 0;
    if (has_ring) ring_col(t, G::AX, G::AY, ra, rb);
    const int orank = t - (G::NT - G::NO);
    const bool has_outer = orank >= 0;
    int oi = 0, oj = 0;
    if (has_outer) ring_col(orank, G::IX, G::IY, oi, oj);

    // Domain columns (Q tags / velocity tags) of the owned columns.
    const int qx0 = map_col(gx, nx, ax[0].qlo, ax[0].qhi), qy0 = map_col(gy, ny, ax[1].qlo, ax[1].qhi);
    const int ux0 = map_col(gx, nx, ax[0].ulo, ax[0].uhi), uy0 = map_col(gy, ny, ax[1].ulo, ax[1].uhi);
    const int rgx = x0 - 1 + ra, rgy = y0 - 1 + rb;
    const int qxr = map_col(rgx, nx, ax[0].qlo, ax[0].qhi), qyr = map_col(rgy, ny, ax[1].qlo, ax[1].qhi);
    const int uxr = map_col(rgx, nx, ax[0].ulo, ax[0].uhi), uyr = map_col(rgy, ny, ax[1].ulo, ax[1].uhi);
    int oqc = 0, ouc = 0;
    if (has_outer) {
        const int ogx = x0 - 2 + oi, ogy = y0 - 2 + oj;
        oqc = map_col(ogy, ny, ax[1].qlo, ax[1].qhi) * nx + map_col(ogx, nx, ax[0].qlo, ax[0].qhi);
        ouc = map_col(ogy, ny, ax[1].ulo, ax[1].uhi) * nx + map_col(ogx, nx, ax[0].ulo, ax[0].uhi);
    }
    const int qc0 = qy0 * nx + qx0, uc0 = uy0 * nx + ux0;
    const int qcr = qyr * nx + qxr, ucr = uyr * nx + uxr;

    auto qplane = [&](int z) { return (unsigned)map_coord(z, nz, ax[2].qlo, ax[2].qhi) * plane; };
    auto uplane = [&](int z) { return (unsigned)map_coord(z, nz, ax[2].ulo, ax[2].uhi) * plane; };

    // Input history of the interior column in iteration pa: planes pa-1, pa,
    // pa+1 (registers); the next plane, pa + 2, is loaded one iteration ahead
    // of its use. It joins the history at the end of the iteration in the
    // unconditional form (the wait for it then falls after the stores are
    // issued) and at the start of the next iteration otherwise.
    static_assert(!UNCOND || PERIODIC, "unconditional form: fully periodic grids");
    T hin[3][kNumIn], hnext[kNumIn];
    load_column<TX, TY>(f, qc0, uc0, qplane(z0 - 2), uplane(z0 - 2), hin[UNCOND ? 0 : 1]);
    load_column<TX, TY>(f, qc0, uc0, qplane(z0 - 1), uplane(z0 - 1), hin[UNCOND ? 1 : 2]);
    load_column<TX, TY>(f, qc0, uc0, qplane(z0), uplane(z0), UNCOND ? hin[2] : hnext);
    if (has_ring) {
        T v[kNumIn];
        load_column<TX, TY>(f, qcr, ucr, qplane(z0 - 2), uplane(z0 - 2), v);
#pragma unroll
        for (int m = 0; m < kNumIn; ++m) sRing[(z0 - 2 + 6) % 3][m][t] = v[m];
        load_column<TX, TY>(f, qcr, ucr, qplane(z0 - 1), uplane(z0 - 1), v);
#pragma unroll
        for (int m = 0; m < kNumIn; ++m) sRing[(z0 - 1 + 6) % 3][m][t] = v[m];
    }
    // Ring and outer-ring inputs are loaded one plane ahead of their LDS
    // store: ring plane pa + 1 and outer plane pa arrive in iteration pa.
    T rnext[kNumIn], onext[kNumIn];
    // Columns of the ring and outer-ring loads; in the unconditional form a
    // thread without such a column loads its own (the values go unused).
    const int rq = (UNCOND && !has_ring) ? qc0 : qcr, ru = (UNCOND && !has_ring) ? uc0 : ucr;
    const int oq = (UNCOND && !has_outer) ? qc0 : oqc, ou = (UNCOND && !has_outer) ? uc0 : ouc;
    if (UNCOND || has_ring) load_column<TX, TY>(f, rq, ru, qplane(z0), uplane(z0), rnext);
    auto load_outer = [&](int z, T* v) {
        const unsigned qo = (qplane(z) + (unsigned)oq) * E, uo = (uplane(z) + (unsigned)ou) * E;
#pragma unroll
        for (int m = 0; m < 5; ++m) v[m] = ldo(f.q[m], qo);
#pragma unroll
        for (int m = 0; m < 3; ++m) v[5 + m] = ldo(f.u[m], uo);
    };
    if (UNCOND || has_outer) load_outer(z0 - 1, onext);

    // z-difference values of the interior column's stage A at plane pa - 2
    // (the lower z neighbour of the plane finished in step 3), and the
    // Ericksen force of the interior column for planes pa - 1 and pa.
    T dz_back[kNumDZ];
#pragma unroll
    for (int m = 0; m < kNumDZ; ++m) dz_back[m] = T(0);
    T fe_prev[3] = {0, 0, 0}, fe_cur[3] = {0, 0, 0};

    // One iteration: WB (compile time) false skips the gather and the force
    // stage, for the two peeled planes of the unconditional form.
    auto iteration = [&](const int pa, auto WB) {
        const int pb = pa - 1;
        const bool do_b = decltype(WB)::value && (UNCOND || (pb >= z0 && active));
        if constexpr (!UNCOND) {
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) {
                hin[0][m] = hin[1][m];
                hin[1][m] = hin[2][m];
                hin[2][m] = hnext[m];
            }
        }
        // ---- step 1: input plane pa into LDS; next-plane loads ----
        // At pa == z1 the loaded plane is never used; clamping it to z1 + 1
        // keeps every request within one plane past the chunk.
        const int pn = min(pa + 2, z1 + 1);
        load_column<TX, TY>(f, qc0, uc0, qplane(pn), uplane(pn), hnext);
        // The pull gather of plane pb = pa - 1 reads the previous state only,
        // so it is issued here and completes under the barrier and the Q
        // update. (The moving-wall term reads rho of the own cell, which only
        // this thread writes, after the collision below.)
        const unsigned cpb = (unsigned)pb * plane + (unsigned)gy * nx + gx;
        T pop[kQ];
        if (do_b) gather32<PERIODIC>(f, bc, gx, gy, pb, cpb, pull, pop);
        {
            const int cc = (ib + 1) * G::IX + ia + 1;
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) sIn[m][cc] = hin[1][m];
        }
        const int r_m = (pa - 1 + 6) % 3, r_c = (pa + 6) % 3, r_p = (pa + 1 + 6) % 3;
        if (has_ring) {
            const int cc = (rb + 1) * G::IX + ra + 1;
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) {
                sRing[r_p][m][t] = rnext[m];
                sIn[m][cc] = sRing[r_c][m][t];
            }
            if (!UNCOND && pa < z1) load_column<TX, TY>(f, rq, ru, qplane(pa + 2), uplane(pa + 2), rnext);
        }
        if (UNCOND) load_column<TX, TY>(f, rq, ru, qplane(pn), uplane(pn), rnext);
        if (has_outer) {
            const int cc = oj * G::IX + oi;
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) sIn[m][cc] = onext[m];
            if (!UNCOND && pa < z1) load_outer(pa + 1, onext);
        }
        if (UNCOND) load_outer(min(pa + 1, z1), onext);
        __syncthreads();

        // ---- step 2: Q update (stage A) at plane pa ----
        const int par = (pa + 2) & 1;
        const int zq = map_coord(pa, nz, ax[2].qlo, ax[2].qhi), zu = map_coord(pa, nz, ax[2].ulo, ax[2].uhi);
        T dz_fwd[kNumDZ];
        fe_prev[0] = fe_cur[0];
        fe_prev[1] = fe_cur[1];
        fe_prev[2] = fe_cur[2];
        // Interior column (slot 0, inputs in registers) and, for the first
        // NR threads, one ring column (slot 1, inputs in LDS); one copy of
        // the Q-update code serves both.
#pragma unroll 1
        for (int slot = 0; slot < (has_ring ? 2 : 1); ++slot) {
            T zm[kNumIn], c[kNumIn], zp[kNumIn];
            int qpos[3], upos[3], ci, cj, acol;
            if (slot == 0) {
#pragma unroll
                for (int m = 0; m < kNumIn; ++m) {
                    zm[m] = hin[0][m];
                    c[m] = hin[1][m];
                    zp[m] = hin[2][m];
                }
                qpos[0] = qx0; qpos[1] = qy0; upos[0] = ux0; upos[1] = uy0;
                ci = ia; cj = ib; acol = acell;
            } else {
#pragma unroll
                for (int m = 0; m < kNumIn; ++m) {
                    zm[m] = sRing[r_m][m][t];
                    c[m] = sRing[r_c][m][t];
                    zp[m] = sRing[r_p][m][t];
                }
                qpos[0] = qxr; qpos[1] = qyr; upos[0] = uxr; upos[1] = uyr;
                ci = ra; cj = rb; acol = rb * G::AX + ra;
            }
            qpos[2] = zq;
            upos[2] = zu;
            StageAT<T> a;
            T fe[3];
            stage_a<TX, TY, PERIODIC>(bc, k, sIn, ci + 1, cj + 1, zm, c, zp, qpos, upos, a, fe);
#pragma unroll
            for (int m = 0; m < kNumA; ++m) sA[par][m][acol] = stage_value(a, m);
            if (slot == 0) {
#pragma unroll
                for (int m = 0; m < kNumA; ++m)
                    if (dz_slot(m) >= 0) dz_fwd[dz_slot(m)] = stage_value(a, m);
                fe_cur[0] = fe[0];
                fe_cur[1] = fe[1];
                fe_cur[2] = fe[2];
            }
        }
        __syncthreads();

        // ---- step 3: force + collide-stream at plane pb = pa - 1 ----
        const T* A0 = &sA[(pb + 2) & 1][0][0];
        auto A = [&](int m, int col) { return A0[m * G::NA + col]; };
        if (do_b) {
            // Neighbour columns with the reference's fetch rule: where a
            // non-periodic face clamps, the cell itself.
            const int lxm = (qx0 == 0 && !ax[0].qlo) ? acell : acell - 1;
            const int lxp = (qx0 == nx - 1 && !ax[0].qhi) ? acell : acell + 1;
            const int lym = (qy0 == 0 && !ax[1].qlo) ? acell : acell - G::AX;
            const int lyp = (qy0 == ny - 1 && !ax[1].qhi) ? acell : acell + G::AX;
            const bool zm_self = (pb == 0 && !ax[2].qlo);
            const bool zp_self = (pb == nz - 1 && !ax[2].qhi);
            auto zmv = [&](int m) -> T {
                const int sl = dz_slot(m);
                return (zm_self || sl < 0) ? A(m, acell) : dz_back[sl];
            };
            auto zpv = [&](int m) -> T {
                const int sl = dz_slot(m);
                return (zp_self || sl < 0) ? A(m, acell) : dz_fwd[sl];
            };
            const int pos[3] = {qx0, qy0, pb};
            T qc[5], dQ[5][3];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                qc[m] = A(m, acell);
                const Deriv<T> d = q_deriv<T, PERIODIC>(bc, m, pos, qc[m], A(m, lxm), A(m, lxp), A(m, lym), A(m, lyp),
                                               zmv(m), zpv(m));
                dQ[m][0] = d.d[0];
                dQ[m][1] = d.d[1];
                dQ[m][2] = d.d[2];
            }
            T dS[8][3];
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                dS[m][0] = (A(5 + m, lxp) - A(5 + m, lxm)) / T(2.0);
                dS[m][1] = (A(5 + m, lyp) - A(5 + m, lym)) / T(2.0);
                dS[m][2] = (zpv(5 + m) - zmv(5 + m)) / T(2.0);
            }
            const T uo[3] = {hin[0][5], hin[0][6], hin[0][7]};   // u^{n-1} at (gx, gy, pb)
            T F[3];
            if (k.force_full) {
                body_force(k, dQ, dS, uo, fe_prev, F);
            } else {
                F[0] = fe_prev[0];
                F[1] = fe_prev[1];
                F[2] = fe_prev[2];
            }
            // Density in its stored form: rho - rho0 for the shifted FP32
            // populations, rho for FP64.
            T drho, u[3];
            if constexpr (sizeof(T) == 4) collide_shifted(k, pop, F, drho, u);
            else collide_plain(k, pop, F, drho, u);
            const unsigned vs4 = (unsigned)f.vs * E, c4 = cpb * E;
            if (UNCOND || wmask.writes(pb)) {
#pragma unroll
                for (int j = 0; j < kQ; ++j) sto(f.gn, j * vs4 + c4, pop[j]);
#pragma unroll
                for (int m = 0; m < 3; ++m) sto(f.un[m], c4, u[m]);
#pragma unroll
                for (int m = 0; m < 5; ++m) sto(f.qn[m], c4, qc[m]);
                if (aux) {
                    sto(f.rho, c4, drho);
                    sto(f.F[0], c4, F[0]);
                    sto(f.F[1], c4, F[1]);
                    sto(f.F[2], c4, F[2]);
                }
            }
        }
        // Plane pb becomes the lower z neighbour of the next plane.
#pragma unroll
        for (int m = 0; m < kNumA; ++m)
            if (dz_slot(m) >= 0) dz_back[dz_slot(m)] = A(m, acell);
        if constexpr (UNCOND) {
            // History advance: planes pa, pa + 1, pa + 2 for the next iteration.
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) {
                hin[0][m] = hin[1][m];
                hin[1][m] = hin[2][m];
                hin[2][m] = hnext[m];
            }
        }
    };
    if constexpr (UNCOND) {
        iteration(z0 - 1, std::false_type{});
        iteration(z0, std::false_type{});
        for (int pa = z0 + 1; pa <= z1; ++pa) iteration(pa, std::true_type{});
    } else {
        for (int pa = z0 - 1; pa <= z1; ++pa) iteration(pa, std::true_type{});
    }
}

template <int TX, int TY, int ZC, bool PERIODIC, typename T = float, bool UNCOND = false>
__global__ void __launch_bounds__(TX * TY)
fused_step_kernel(Fields<T> f, BCDesc<T> bc, Phys<T> k, int pull, int aux, int ntx, int nty,
                  int nzc, int mode) {
    using G = TileGeom<TX, TY>;
    __shared__ T sIn[kNumIn][G::NI];          // input plane pa, extended region
    __shared__ T sA[2][kNumA][G::NA];         // stage-A results, planes pa and pa - 1
    __shared__ T sRing[3][kNumIn][G::NR];     // input history of the stage-A ring columns
    // XCD-aware block order: hardware deals consecutive block ids round-robin
    // over the 8 XCDs; renumber so each XCD gets a contiguous run of tiles
    // (neighbouring tiles share halos in that XCD's L2).
    const int nblk = ntx * nty * nzc;
    const int per_xcd = (nblk + 7) / 8;
    const int L = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (L >= nblk) return;
    const int tx = L % ntx, ty = (L / ntx) % nty, tz = L / (ntx * nty);
    fused_body<TX, TY, ZC, PERIODIC, T, UNCOND>(f, bc, k, pull, aux, tx, ty, tz, sIn, sA, sRing, mode);
}

// ---- wall shells of a single simulation, one thread per cell ------------------
#ifndef LBM_FUSED_BODY_ONLY
// Q update over the shell and its one-cell halo (w = 3). Every cell stores
// Sigma and Tau for the force pass; only shell cells store Q^{n+1} and the
// Ericksen force (the interior launch owns the rest).
static __global__ void __launch_bounds__(kBlock)
shell_qstep_kernel(Fields<float> f, BCDesc<float> bc, Phys<float> k, ShellRegion r) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= region_cells(r)) return;
    int q[3];
    region_cell(r, bc.n, i, q);
    const Pos p{q[0], q[1], q[2]};
    const long c = cell_of(p.x, p.y, p.z, bc.n[0], bc.n[1]);
    qstep_cell(f, bc, k, c, p, !cell_mask(bc, 1, p.x, p.y).writes(p.z));
}

// Body force and collide-stream of the shell cells (w = 2); f.q points at
// Q^{n+1}, and the velocity goes to un.
static __global__ void __launch_bounds__(kBlock)
shell_force_lbm_kernel(Fields<float> f, BCDesc<float> bc, Phys<float> k, ShellRegion r, int flags, int pull) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= region_cells(r)) return;
    int q[3];
    region_cell(r, bc.n, i, q);
    const Pos p{q[0], q[1], q[2]};
    force_lbm_cell<float, true, true>(f, bc, k, cell_of(p.x, p.y, p.z, bc.n[0], bc.n[1]), p, flags, pull);
}
#endif  // LBM_FUSED_BODY_ONLY

// Second pipeline: the same computation with one stage-A plane in LDS (the
// force stage reads plane pb while the fresh plane pa waits in registers and
// is published after a barrier), and with every global load issued one z
// plane ahead of its use: the pull gather of plane pb and the input planes
// pa + 2 are in flight while the Q update of plane pa computes. Only the
// periodic specialisation is built. On a domain with walls (MASKED) its
// values for the shell cells are wrong and the shell passes that follow
// overwrite them; rho and F, of which rho is also read by the shell passes,
// are written for class-0 cells only.
template <int TX, int TY, int ZC, bool MASKED, int WPE, bool UPWIND, bool NT>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
fused_step_kernel_pf(Fields<float> f, BCDesc<float> bc, Phys<float> k, int pull, int aux, int ntx, int nty,
                     int nzc) {
    using G = TileGeom<TX, TY>;
    constexpr bool PERIODIC = true;
    __shared__ float sIn[kNumIn][G::NI];           // input plane pa, extended region
    __shared__ float sA[kNumA][G::NA];             // stage-A results of plane pa - 1
    __shared__ float sRing[4][kNumIn][G::NR];      // ring-column inputs, planes mod 4

    const int nblk = ntx * nty * nzc;
    const int per_xcd = (nblk + 7) / 8;
    const int L = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (L >= nblk) return;
    const int tx = L % ntx, ty = (L / ntx) % nty, tz = L / (ntx * nty);
    const int nx = bc.n[0], ny = bc.n[1], nz = bc.n[2];
    const int x0 = tx * TX, y0 = ty * TY;
    const int z0 = tz * ZC, z1 = min(z0 + ZC, nz);
    const unsigned plane = (unsigned)nx * ny;

    AxisTags ax[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ax[a].qlo = PERIODIC || bc.qkind[2 * a] == QB_PERIODIC;
        ax[a].qhi = PERIODIC || bc.qkind[2 * a + 1] == QB_PERIODIC;
        ax[a].ulo = PERIODIC || bc.ukind[2 * a] == UB_PERIODIC;
        ax[a].uhi = PERIODIC || bc.ukind[2 * a + 1] == UB_PERIODIC;
    }

    const int t = threadIdx.x;
    const int ia = t % TX + 1, ib = t / TX + 1;
    const int acell = ib * G::AX + ia;
    const int gx = x0 + ia - 1, gy = y0 + ib - 1;
    const bool active = gx < nx && gy < ny;
    const bool has_ring = t < G::NR;
    int ra = 0, rb = 0;
    if (has_ring) ring_col(t, G::AX, G::AY, ra, rb);
    const int orank = t - (G::NT - G::NO);
    const bool has_outer = orank >= 0;
    int oi = 0, oj = 0;
    if (has_outer) ring_col(orank, G::IX, G::IY, oi, oj);

    const int qx0 = map_col(gx, nx, ax[0].qlo, ax[0].qhi), qy0 = map_col(gy, ny, ax[1].qlo, ax[1].qhi);
    const int ux0 = map_col(gx, nx, ax[0].ulo, ax[0].uhi), uy0 = map_col(gy, ny, ax[1].ulo, ax[1].uhi);
    const int rgx = x0 - 1 + ra, rgy = y0 - 1 + rb;
    const int qxr = map_col(rgx, nx, ax[0].qlo, ax[0].qhi), qyr = map_col(rgy, ny, ax[1].qlo, ax[1].qhi);
    const int uxr = map_col(rgx, nx, ax[0].ulo, ax[0].uhi), uyr = map_col(rgy, ny, ax[1].ulo, ax[1].uhi);
    int oqc = 0, ouc = 0;
    if (has_outer) {
        const int ogx = x0 - 2 + oi, ogy = y0 - 2 + oj;
        oqc = map_col(ogy, ny, ax[1].qlo, ax[1].qhi) * nx + map_col(ogx, nx, ax[0].qlo, ax[0].qhi);
        ouc = map_col(ogy, ny, ax[1].ulo, ax[1].uhi) * nx + map_col(ogx, nx, ax[0].ulo, ax[0].uhi);
    }
    const int qc0 = qy0 * nx + qx0, uc0 = uy0 * nx + ux0;
    const int qcr = qyr * nx + qxr, ucr = uyr * nx + uxr;
    auto qplane = [&](int z) { return (unsigned)map_coord(z, nz, ax[2].qlo, ax[2].qhi) * plane; };
    auto uplane = [&](int z) { return (unsigned)map_coord(z, nz, ax[2].ulo, ax[2].uhi) * plane; };
    const int cc_int = (ib + 1) * G::IX + ia + 1;
    const int cc_ring = (rb + 1) * G::IX + ra + 1;
    const int cc_out = oj * G::IX + oi;
    const int rcell = rb * G::AX + ra;

    // Prologue: interior planes z0-2 .. z0, ring planes z0-2 .. z0, outer plane z0-1.
    float hin[3][kNumIn];
    load_column<TX, TY>(f, qc0, uc0, qplane(z0 - 2), uplane(z0 - 2), hin[0]);
    load_column<TX, TY>(f, qc0, uc0, qplane(z0 - 1), uplane(z0 - 1), hin[1]);
    load_column<TX, TY>(f, qc0, uc0, qplane(z0), uplane(z0), hin[2]);
    if (has_ring) {
#pragma unroll 1
        for (int z = z0 - 2; z <= z0; ++z) {
            float v[kNumIn];
            load_column<TX, TY>(f, qcr, ucr, qplane(z), uplane(z), v);
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) sRing[z & 3][m][t] = v[m];
        }
    }
    float ov[kNumIn];
    if (has_outer) load_column<TX, TY>(f, oqc, ouc, qplane(z0 - 1), uplane(z0 - 1), ov);

    float dz_back[kNumDZ];
#pragma unroll
    for (int m = 0; m < kNumDZ; ++m) dz_back[m] = 0.0f;
    float fe_prev[3] = {0, 0, 0}, fe_cur[3] = {0, 0, 0};

    for (int pa = z0 - 1; pa <= z1; ++pa) {
        const int pb = pa - 1;
        const bool do_b = pb >= z0 && active;
        // ---- step 1: issue next-plane loads and the pull gather; stage plane pa ----
        float hnext[kNumIn], rnext[kNumIn], onext[kNumIn], pop[kQ];
        // At pa == z1 the loaded plane is never used; clamping it to z1 + 1
        // keeps every request within one plane past the chunk.
        const int pn = min(pa + 2, z1 + 1);
        load_column<TX, TY>(f, qc0, uc0, qplane(pn), uplane(pn), hnext);
        if (has_ring) load_column<TX, TY>(f, qcr, ucr, qplane(pn), uplane(pn), rnext);
        if (has_outer) load_column<TX, TY>(f, oqc, ouc, qplane(pa + 1), uplane(pa + 1), onext);
        const unsigned cb = (unsigned)pb * plane + (unsigned)gy * nx + gx;
        if (do_b) gather32<PERIODIC>(f, bc, gx, gy, pb, cb, pull, pop);
#pragma unroll
        for (int m = 0; m < kNumIn; ++m) sIn[m][cc_int] = hin[1][m];
        if (has_ring) {
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) sIn[m][cc_ring] = sRing[pa & 3][m][t];
        }
        if (has_outer) {
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) sIn[m][cc_out] = ov[m];
        }
        __syncthreads();

        // ---- step 2: Q update at plane pa (results stay in registers) ----
        const int zq = map_coord(pa, nz, ax[2].qlo, ax[2].qhi), zu = map_coord(pa, nz, ax[2].ulo, ax[2].uhi);
        fe_prev[0] = fe_cur[0];
        fe_prev[1] = fe_cur[1];
        fe_prev[2] = fe_cur[2];
        StageA ai, ar;
        {
            const int qpos[3] = {qx0, qy0, zq};
            const int upos[3] = {ux0, uy0, zu};
            stage_a_lean<TX, TY, PERIODIC, UPWIND>(bc, k, sIn, ia + 1, ib + 1, hin[0], hin[1], hin[2], qpos, upos, ai, fe_cur);
        }
        if (has_ring) {
            const int qpos[3] = {qxr, qyr, zq};
            const int upos[3] = {uxr, uyr, zu};
            float zm[kNumIn], c[kNumIn], zp[kNumIn], fer[3];
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) {
                zm[m] = sRing[(pa - 1) & 3][m][t];
                c[m] = sRing[pa & 3][m][t];
                zp[m] = sRing[(pa + 1) & 3][m][t];
            }
            stage_a_lean<TX, TY, PERIODIC, UPWIND>(bc, k, sIn, ra + 1, rb + 1, zm, c, zp, qpos, upos, ar, fer);
        }

        // ---- step 3: force + collide-stream at plane pb (sA holds plane pb) ----
        if (do_b) {
            const int lxm = (qx0 == 0 && !ax[0].qlo) ? acell : acell - 1;
            const int lxp = (qx0 == nx - 1 && !ax[0].qhi) ? acell : acell + 1;
            const int lym = (qy0 == 0 && !ax[1].qlo) ? acell : acell - G::AX;
            const int lyp = (qy0 == ny - 1 && !ax[1].qhi) ? acell : acell + G::AX;
            const bool zm_self = (pb == 0 && !ax[2].qlo);
            const bool zp_self = (pb == nz - 1 && !ax[2].qhi);
            auto zmv = [&](int m) -> float {
                const int sl = dz_slot(m);
                return (zm_self || sl < 0) ? sA[m][acell] : dz_back[sl];
            };
            auto zpv = [&](int m) -> float {
                const int sl = dz_slot(m);
                return (zp_self || sl < 0) ? sA[m][acell] : stage_value(ai, m);
            };
            const int pos[3] = {qx0, qy0, pb};
            float qc[5], dQ[5][3];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                qc[m] = sA[m][acell];
                const Deriv<float> d = q_deriv<float, PERIODIC>(bc, m, pos, qc[m], sA[m][lxm], sA[m][lxp], sA[m][lym], sA[m][lyp],
                                               zmv(m), zpv(m));
                dQ[m][0] = d.d[0];
                dQ[m][1] = d.d[1];
                dQ[m][2] = d.d[2];
            }
            float dS[8][3];
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                dS[m][0] = (sA[5 + m][lxp] - sA[5 + m][lxm]) / 2.0f;
                dS[m][1] = (sA[5 + m][lyp] - sA[5 + m][lym]) / 2.0f;
                dS[m][2] = (zpv(5 + m) - zmv(5 + m)) / 2.0f;
            }
            const float uo[3] = {hin[0][5], hin[0][6], hin[0][7]};   // u^{n-1} at (gx, gy, pb)
            float F[3];
            if (k.force_full) {
                body_force(k, dQ, dS, uo, fe_prev, F);
            } else {
                F[0] = fe_prev[0];
                F[1] = fe_prev[1];
                F[2] = fe_prev[2];
            }
            float drho, u[3];
            collide_shifted(k, pop, F, drho, u);
            const unsigned vs4 = (unsigned)f.vs * 4u, c4 = cb * 4u;
#pragma unroll
            for (int j = 0; j < kQ; ++j) sto_hint<NT>(f.gn, j * vs4 + c4, pop[j]);
#pragma unroll
            for (int m = 0; m < 3; ++m) sto_hint<NT>(f.un[m], c4, u[m]);
#pragma unroll
            for (int m = 0; m < 5; ++m) sto_hint<NT>(f.qn[m], c4, qc[m]);
            // rho is also an input (moving-wall bounce of the shell cells),
            // so on a domain with walls only the class-0 cells write it.
            if (aux && (!MASKED || cell_mask(bc, 1, gx, gy).writes(pb))) {
                sto(f.rho, c4, drho);
                sto(f.F[0], c4, F[0]);
                sto(f.F[1], c4, F[1]);
                sto(f.F[2], c4, F[2]);
            }
        }
        // Plane pb becomes the lower z neighbour of the next plane.
#pragma unroll
        for (int m = 0; m < kNumA; ++m)
            if (dz_slot(m) >= 0) dz_back[dz_slot(m)] = sA[m][acell];
        __syncthreads();

        // ---- step 4: publish stage A of plane pa; advance histories ----
#pragma unroll
        for (int m = 0; m < kNumA; ++m) sA[m][acell] = stage_value(ai, m);
        if (has_ring) {
#pragma unroll
            for (int m = 0; m < kNumA; ++m) sA[m][rcell] = stage_value(ar, m);
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) sRing[(pa + 2) & 3][m][t] = rnext[m];
        }
#pragma unroll
        for (int m = 0; m < kNumIn; ++m) {
            hin[0][m] = hin[1][m];
            hin[1][m] = hin[2][m];
            hin[2][m] = hnext[m];
        }
        if (has_outer) {
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) ov[m] = onext[m];
        }
    }
}

// Divergence form of the body force. With P = -alpha Q^{n+1} + (Sigma +/- Tau)
// in the sign pattern of body_force (src/qtensor_solver.tpp:154-217,
// src/boundary_handler.h:335-392), F_a = d_b P_ab - mu u_a + F_E,a, and a
// neighbour along axis b contributes only the column (P_xb, P_yb, P_zb):
//   X set (x neighbours): -aQxx + Sxx, -aQxy + Sxy - Txy, -aQxz + Sxz - Txz;
//   Y set (y neighbours): -aQxy + Sxy + Txy, -aQyy + Syy, -aQyz + Syz - Tyz;
//   Z set (z neighbours): -aQxz + Sxz + Txz, -aQyz + Syz + Tyz, a(Qxx + Qyy) - Sxx - Syy.
// (Szz = -Sxx - Syy and Qzz = -Qxx - Qyy.)
__device__ __forceinline__ void p_sets(float alpha, const StageA& a, float pxy[6], float pz[3]) {
    const float* q = a.qn;
    const float* sg = a.sg;
    const float* ta = a.ta;
    pxy[0] = sg[0] - alpha * q[0];
    pxy[1] = (sg[1] - ta[0]) - alpha * q[1];
    pxy[2] = (sg[2] - ta[1]) - alpha * q[2];
    pxy[3] = (sg[1] + ta[0]) - alpha * q[1];
    pxy[4] = sg[3] - alpha * q[3];
    pxy[5] = (sg[4] - ta[2]) - alpha * q[4];
    pz[0] = (sg[2] + ta[1]) - alpha * q[2];
    pz[1] = (sg[4] + ta[2]) - alpha * q[4];
    pz[2] = alpha * (q[0] + q[3]) - (sg[0] + sg[3]);
}

// Boundary tags the third pipeline's WALLS mode computes exactly: z faces
// periodic (and a slot table whose rows do not depend on the z position:
// BCDesc::slots_xy), x and y faces with a periodic or Neumann Q tag and a periodic,
// no-slip or specular velocity tag, periodic for both or for neither
// (anchoring changes the Q ghost of the force stage, a moving wall's slot
// reads the density, which this pipeline stores only with aux, and equal
// periodicity lets Q and velocity share the mapped columns).
inline bool p2_walls_ok(const BCDesc<float>& bc) {
    if (!bc.slots_xy) return false;
    for (int face = 4; face < 6; ++face)
        if (bc.qkind[face] != QB_PERIODIC || bc.ukind[face] != UB_PERIODIC) return false;
    for (int face = 0; face < 4; ++face) {
        if (bc.qkind[face] == QB_ANCHOR || bc.ukind[face] == UB_MOVING) return false;
        if ((bc.qkind[face] == QB_PERIODIC) != (bc.ukind[face] == UB_PERIODIC)) return false;
    }
    return true;
}

// Third pipeline (periodic specialisation, like the second): the force
// stage reads only the divergence-form sets above (6 values per column in
// LDS instead of 13; the z set of the own column stays in registers), and the
// input plane and the sets are double-buffered in LDS, so one barrier per
// plane separates the writes of a plane from its reads. Per iteration pa:
//   before the barrier: publish the sets of plane pa - 1, stage input plane pa,
//     store ring inputs of plane pa + 1, issue the loads of planes pa + 1/pa + 2
//     and the pull gather of plane pa - 1;
//   after it: Q update at plane pa, then force + collide-stream at plane pa - 1.
// A wave can run at most one iteration ahead of the slowest (the next
// barrier), which the two LDS buffers and the four ring-input slots cover.
// One call computes z planes [z0, z1) (0 <= z0 < z1 <= nz) of the TX x TY
// column tile (tx, ty), with LDS buffers from the calling kernel; a kernel
// that runs several tiles per block puts a barrier between them.
//
// WALLS: the x and y faces may be walls (Q tag periodic or Neumann, velocity
// tag periodic, no-slip or specular; the z faces periodic, which the caller
// checks with p2_walls_ok). Every cell is then computed exactly, with the
// unfused kernels' wall handling: input columns mapped by the face tags,
// ghost values in the Q update, the slot table in the pull gather, and the
// force stage's x and y neighbours clamped to the cell itself where the
// reference clamps. For a Neumann face the clamped neighbour is the ghost of
// Q^{n+1} as well as of Sigma and Tau, so the divergence-form sets clamp as
// a whole.
template <int TX, int TY, bool MASKED, bool UPWIND, bool NT, bool UNCOND = false, bool WALLS = false>
__device__ __forceinline__ void p2_range(const Fields<float>& f, const BCDesc<float>& bc, const Phys<float>& k,
                                        int pull, int aux, int tx, int ty, int z0, int z1,
                                        float (*sIn)[kNumIn][TileGeom<TX, TY>::NI],
                                        float (*sP)[6][TileGeom<TX, TY>::NA],
                                        float (*sRing)[kNumIn][TileGeom<TX, TY>::NR]) {
    using G = TileGeom<TX, TY>;
    const int nx = bc.n[0], ny = bc.n[1], nz = bc.n[2];
    const int x0 = tx * TX, y0 = ty * TY;
    const unsigned plane = (unsigned)nx * ny;

    const int t = threadIdx.x;
    const int ia = t % TX + 1, ib = t / TX + 1;
    const int acell = ib * G::AX + ia;
    const int gx = x0 + ia - 1, gy = y0 + ib - 1;
    const bool active = gx < nx && gy < ny;
    const bool has_ring = t < G::NR;
    int ra = 0, rb = 0;
    if (has_ring) ring_col(t, G::AX, G::AY, ra, rb);
    const int orank = t - (G::NT - G::NO);
    const bool has_outer = orank >= 0;
    int oi = 0, oj = 0;
    if (has_outer) ring_col(orank, G::IX, G::IY, oi, oj);

    // Periodic flags of the x and y faces (all periodic unless WALLS; Q and
    // velocity agree, see p2_walls_ok).
    bool plo[2] = {true, true}, phi[2] = {true, true};
    if constexpr (WALLS) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            plo[a] = bc.qkind[2 * a] == QB_PERIODIC;
            phi[a] = bc.qkind[2 * a + 1] == QB_PERIODIC;
        }
    }
    const int qx0 = map_col(gx, nx, plo[0], phi[0]), qy0 = map_col(gy, ny, plo[1], phi[1]);
    const int qxr = map_col(x0 - 1 + ra, nx, plo[0], phi[0]), qyr = map_col(y0 - 1 + rb, ny, plo[1], phi[1]);
    int oc = 0;
    if (has_outer) oc = map_col(y0 - 2 + oj, ny, plo[1], phi[1]) * nx + map_col(x0 - 2 + oi, nx, plo[0], phi[0]);
    const int c0 = qy0 * nx + qx0, cr = qyr * nx + qxr;
    auto zplane = [&](int z) { return (unsigned)map_coord(z, nz, true, true) * plane; };
    const int cc_int = (ib + 1) * G::IX + ia + 1;
    const int cc_ring = (rb + 1) * G::IX + ra + 1;
    const int cc_out = oj * G::IX + oi;
    const int rcell = rb * G::AX + ra;
    const int qpos_i[3] = {qx0, qy0, 0}, qpos_r[3] = {qxr, qyr, 0};
    // In-plane parts of the periodic pull sources (x - cx, y - cy) of this
    // column, [cy + 1][cx + 1]: each plane's gather adds the plane offset of
    // z - cz, which is uniform over the block.
    unsigned ioff[3][3];
    {
        const int xs[3] = {qx0 == nx - 1 ? 0 : qx0 + 1, qx0, qx0 == 0 ? nx - 1 : qx0 - 1};
        const int ys[3] = {qy0 == ny - 1 ? 0 : qy0 + 1, qy0, qy0 == 0 ? ny - 1 : qy0 - 1};
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) ioff[b][a] = (unsigned)ys[b] * nx + xs[a];
    }
    // Force-stage x and y neighbours in the set arrays: the cell itself where
    // a face clamps (Neumann).
    int lxm = acell - 1, lxp = acell + 1, lym = acell - G::AX, lyp = acell + G::AX;
    // Slot-table rows of the 16 x, y classes (no z bits; see
    // BCDesc::slots_xy), two entries per word, in LDS; the class of this
    // column; and
    // whether any column of the wavefront lies on an x or y face (else the
    // periodic gather serves the whole wavefront). The first pull gather
    // follows two barriers, so the rows are in place by then; a caller that
    // runs several ranges per block puts a barrier between them.
    int cxy = 0;
    bool wave_walls = false;
    const uint32_t* slot_rows = nullptr;
    if constexpr (WALLS) {
        if (qx0 == 0 && !plo[0]) lxm = acell;
        if (qx0 == nx - 1 && !phi[0]) lxp = acell;
        if (qy0 == 0 && !plo[1]) lym = acell;
        if (qy0 == ny - 1 && !phi[1]) lyp = acell;
        __shared__ uint32_t sSlot[16][8];
        if (t < 16 * 8) {
            const int cl = t >> 3, w = t & 7;
            const uint16_t* row = f.slots + cl * kQ;
            const uint32_t lo = row[2 * w], hi = 2 * w + 1 < kQ ? row[2 * w + 1] : 0u;
            sSlot[cl][w] = lo | (hi << 16);
        }
        cxy = node_class(min(gx, nx - 1), min(gy, ny - 1), 1, nx, ny, 3);
        wave_walls = __builtin_amdgcn_ballot_w64(cxy != 0) != 0;
        slot_rows = &sSlot[0][0];
    }

    // Prologue: interior planes z0-2 .. z0 (registers), ring planes z0-2 and
    // z0-1 (LDS) and z0 (registers, stored in the first iteration), outer
    // plane z0-1.
    float hin[3][kNumIn], rcur[kNumIn], ov[kNumIn];
    load_column<TX, TY>(f, c0, c0, zplane(z0 - 2), zplane(z0 - 2), hin[0]);
    load_column<TX, TY>(f, c0, c0, zplane(z0 - 1), zplane(z0 - 1), hin[1]);
    load_column<TX, TY>(f, c0, c0, zplane(z0), zplane(z0), hin[2]);
    if (has_ring) {
#pragma unroll 1
        for (int z = z0 - 2; z <= z0 - 1; ++z) {
            float v[kNumIn];
            load_column<TX, TY>(f, cr, cr, zplane(z), zplane(z), v);
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) sRing[z & 3][m][t] = v[m];
        }
        load_column<TX, TY>(f, cr, cr, zplane(z0), zplane(z0), rcur);
    }
    if (has_outer) load_column<TX, TY>(f, oc, oc, zplane(z0 - 1), zplane(z0 - 1), ov);

    // Carried across iterations (interior column): sets of plane pa - 1 to
    // publish, z sets of planes pa - 2 and pa - 1, Q^{n+1} and Ericksen force
    // of plane pa - 1; ring column: sets of plane pa - 1.
    float pxy_prev[6] = {0, 0, 0, 0, 0, 0}, pr_prev[6] = {0, 0, 0, 0, 0, 0};
    float zm2[3] = {0, 0, 0}, zm1[3] = {0, 0, 0};
    float qn_prev[5] = {0, 0, 0, 0, 0}, fe_prev[3] = {0, 0, 0};

    // One iteration: WB is false for the first two (pa = z0 - 1 and z0), whose
    // planes pb = pa - 1 lie below the chunk; the loop below runs the others,
    // which finish planes z0 .. z1 - 1.
    auto iteration = [&](const int pa, auto WB) {
        const int pb = pa - 1;
        const int bi = pa & 1;
        // UNCOND (full tiles only: every column active): the gather, the
        // ring and outer loads and the stores are issued by every thread, so
        // that no memory instruction sits under a branch. (The memory
        // counter waits the compiler inserts after a branch assume its
        // memory instructions were skipped, and then wait for everything
        // issued before, including the previous plane's stores.)
        const bool do_b = decltype(WB)::value && (UNCOND || (pb >= z0 && active));
        // ---- before the barrier: loads in flight, publish plane pa - 1, stage plane pa ----
        // The loads of a plane are issued at raised wave priority, so that a
        // wave reaching this point overtakes the other wave of its SIMD (in
        // its arithmetic) and its loads start early (256^3: 0.667 ->
        // 0.653 ms; the level, 1 to 3, and where it drops back, after the
        // loads or after the barrier, make no difference).
        __builtin_amdgcn_s_setprio(1);
        float hnext[kNumIn], rnext[kNumIn], onext[kNumIn], pop[kQ];
        const int pn = min(pa + 2, z1 + 1);   // unused at pa == z1, as in fused_step_kernel_pf
        const unsigned cb = (unsigned)pb * plane + (unsigned)gy * nx + gx;
        load_column<TX, TY>(f, c0, c0, zplane(pn), zplane(pn), hnext);
        if (UNCOND) {
            const int rc = has_ring ? cr : c0, ocol = has_outer ? oc : c0;
            load_column<TX, TY>(f, rc, rc, zplane(pn), zplane(pn), rnext);
            load_column<TX, TY>(f, ocol, ocol, zplane(pa + 1), zplane(pa + 1), onext);
        } else {
            if (has_ring) load_column<TX, TY>(f, cr, cr, zplane(pn), zplane(pn), rnext);
            if (has_outer) load_column<TX, TY>(f, oc, oc, zplane(pa + 1), zplane(pa + 1), onext);
        }
        if (do_b) {
            if constexpr (WALLS) {
                if (wave_walls) gather32_row(f, bc, gx, gy, pb, cb, pull, slot_rows + 8 * cxy, pop);
                else gather32<true>(f, bc, gx, gy, pb, cb, pull, pop);
            } else {
                const unsigned vs4 = (unsigned)f.vs * 4u;
                if (pull) {
                    const unsigned zoff[3] = {zplane(pb + 1), zplane(pb), zplane(pb - 1)};
#pragma unroll
                    for (int j = 0; j < kQ; ++j)
                        pop[j] = ldo(f.g, j * vs4 + (zoff[cz(j) + 1] + ioff[cy(j) + 1][cx(j) + 1]) * 4u);
                } else {
#pragma unroll
                    for (int j = 0; j < kQ; ++j) pop[j] = ldo(f.g, j * vs4 + cb * 4u);
                }
            }
        }
        __builtin_amdgcn_s_setprio(0);
#pragma unroll
        for (int m = 0; m < 6; ++m) sP[bi ^ 1][m][acell] = pxy_prev[m];
#pragma unroll
        for (int m = 0; m < kNumIn; ++m) sIn[bi][m][cc_int] = hin[1][m];
        if (has_ring) {
#pragma unroll
            for (int m = 0; m < 6; ++m) sP[bi ^ 1][m][rcell] = pr_prev[m];
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) {
                sIn[bi][m][cc_ring] = sRing[pa & 3][m][t];
                sRing[(pa + 1) & 3][m][t] = rcur[m];
            }
        }
        if (has_outer) {
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) sIn[bi][m][cc_out] = ov[m];
        }
        __syncthreads();

        // ---- Q update at plane pa ----
        const int zq = map_coord(pa, nz, true, true);
        StageA ai;
        float fe_new[3], z_new[3], pxy_new[6];
        {
            const int qpos[3] = {qpos_i[0], qpos_i[1], zq};
            stage_a_lean<TX, TY, WALLS ? 2 : 1, UPWIND>(bc, k, sIn[bi], ia + 1, ib + 1, hin[0], hin[1], hin[2], qpos,
                                                        qpos, ai, fe_new);
            p_sets(k.alpha, ai, pxy_new, z_new);
        }
        float pr_new[6] = {0, 0, 0, 0, 0, 0};
        if (has_ring) {
            const int qpos[3] = {qpos_r[0], qpos_r[1], zq};
            float zm[kNumIn], c[kNumIn], zp[kNumIn], fer[3], zr[3];
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) {
                zm[m] = sRing[(pa - 1) & 3][m][t];
                c[m] = sRing[pa & 3][m][t];
                zp[m] = sRing[(pa + 1) & 3][m][t];
            }
            StageA ar;
            stage_a_lean<TX, TY, WALLS ? 2 : 1, UPWIND>(bc, k, sIn[bi], ra + 1, rb + 1, zm, c, zp, qpos, qpos, ar, fer);
            p_sets(k.alpha, ar, pr_new, zr);
        }

        // ---- force + collide-stream at plane pb (sP[bi ^ 1] holds plane pb) ----
        if (do_b) {
            const float* P = &sP[bi ^ 1][0][0];
            float F[3];
            if (k.force_full) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float dx = P[a * G::NA + lxp] - P[a * G::NA + lxm];
                    const float dy = P[(3 + a) * G::NA + lyp] - P[(3 + a) * G::NA + lym];
                    const float dz = z_new[a] - zm2[a];
                    F[a] = fe_prev[a] + (0.5f * ((dx + dy) + dz) - k.mu * hin[0][5 + a]);
                }
            } else {
                F[0] = fe_prev[0];
                F[1] = fe_prev[1];
                F[2] = fe_prev[2];
            }
            float drho, u[3];
            collide_shifted(k, pop, F, drho, u);
            const unsigned vs4 = (unsigned)f.vs * 4u, c4 = cb * 4u;
#pragma unroll
            for (int j = 0; j < kQ; ++j) sto_hint<NT>(f.gn, j * vs4 + c4, pop[j]);
#pragma unroll
            for (int m = 0; m < 3; ++m) sto_hint<NT>(f.un[m], c4, u[m]);
#pragma unroll
            for (int m = 0; m < 5; ++m) sto_hint<NT>(f.qn[m], c4, qn_prev[m]);
            if (aux && (!MASKED || cell_mask(bc, 1, gx, gy).writes(pb))) {
                sto(f.rho, c4, drho);
                sto(f.F[0], c4, F[0]);
                sto(f.F[1], c4, F[1]);
                sto(f.F[2], c4, F[2]);
            }
        }

        // ---- advance the carried values ----
#pragma unroll
        for (int m = 0; m < 6; ++m) pxy_prev[m] = pxy_new[m];
        if (UNCOND || has_ring) {
#pragma unroll
            for (int m = 0; m < 6; ++m) pr_prev[m] = pr_new[m];
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) rcur[m] = rnext[m];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            zm2[a] = zm1[a];
            zm1[a] = z_new[a];
            fe_prev[a] = fe_new[a];
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) qn_prev[m] = ai.qn[m];
#pragma unroll
        for (int m = 0; m < kNumIn; ++m) {
            hin[0][m] = hin[1][m];
            hin[1][m] = hin[2][m];
            hin[2][m] = hnext[m];
        }
        if (UNCOND || has_outer) {
#pragma unroll
            for (int m = 0; m < kNumIn; ++m) ov[m] = onext[m];
        }
    };
    iteration(z0 - 1, std::false_type{});
    iteration(z0, std::false_type{});
    for (int pa = z0 + 1; pa <= z1; ++pa) iteration(pa, std::true_type{});
}

// Chunk tz of ZC planes of the column tile (tx, ty).
template <int TX, int TY, int ZC, bool MASKED, bool UPWIND, bool NT, bool UNCOND = false, bool WALLS = false>
__device__ __forceinline__ void p2_tile(const Fields<float>& f, const BCDesc<float>& bc, const Phys<float>& k,
                                       int pull, int aux, int tx, int ty, int tz,
                                       float (*sIn)[kNumIn][TileGeom<TX, TY>::NI],
                                       float (*sP)[6][TileGeom<TX, TY>::NA],
                                       float (*sRing)[kNumIn][TileGeom<TX, TY>::NR]) {
    const int z0 = tz * ZC;
    p2_range<TX, TY, MASKED, UPWIND, NT, UNCOND, WALLS>(f, bc, k, pull, aux, tx, ty, z0, min(z0 + ZC, bc.n[2]), sIn,
                                                       sP, sRing);
}

template <int TX, int TY, int ZC, bool MASKED, int WPE, bool UPWIND, bool NT, bool ZSLOW, bool UNCOND = false,
          bool WALLS = false>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
fused_step_kernel_p2(Fields<float> f, BCDesc<float> bc, Phys<float> k, int pull, int aux, int ntx, int nty,
                     int nzc) {
    using G = TileGeom<TX, TY>;
    __shared__ float sIn[2][kNumIn][G::NI];        // input plane, by parity
    __shared__ float sP[2][6][G::NA];              // X and Y sets, by plane parity
    __shared__ float sRing[4][kNumIn][G::NR];      // ring-column inputs, planes mod 4

    const int nblk = ntx * nty * nzc;
    const int per_xcd = (nblk + 7) / 8;
    const int L = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (L >= nblk) return;
    // Two block orders. ZSLOW: columns along x, then rows along y, then z
    // chunks, so the blocks of one XCD cover a band of whole rows and every
    // plane they touch is one contiguous stretch of each array. Otherwise the
    // z chunks of one column are neighbours (they share halo planes), then
    // columns along x, then rows along y.
    int tz, tx, ty;
    if constexpr (ZSLOW) {
        tx = L % ntx; ty = (L / ntx) % nty; tz = L / (ntx * nty);
    } else {
        tz = L % nzc; tx = (L / nzc) % ntx; ty = L / (nzc * ntx);
    }
    if constexpr (WALLS) {
        // A tile with a cell within two cells of a wall runs the wall-aware
        // form. The others are whole tiles of cells whose inputs, halo and
        // pull sources lie inside the domain away from every wall, which the
        // periodic form computes exactly (with UNCOND, every memory
        // instruction unconditional).
        const int x0 = tx * TX, y0 = ty * TY, nx = bc.n[0], ny = bc.n[1];
        const bool near = (bc.qkind[0] != QB_PERIODIC && x0 < 2) || (bc.qkind[1] != QB_PERIODIC && x0 + TX > nx - 2) ||
                          (bc.qkind[2] != QB_PERIODIC && y0 < 2) || (bc.qkind[3] != QB_PERIODIC && y0 + TY > ny - 2);
        if (near) p2_tile<TX, TY, ZC, false, UPWIND, NT, false, true>(f, bc, k, pull, aux, tx, ty, tz, sIn, sP, sRing);
        else p2_tile<TX, TY, ZC, false, UPWIND, NT, UNCOND, false>(f, bc, k, pull, aux, tx, ty, tz, sIn, sP, sRing);
    } else {
        p2_tile<TX, TY, ZC, MASKED, UPWIND, NT, UNCOND, false>(f, bc, k, pull, aux, tx, ty, tz, sIn, sP, sRing);
    }
}

#ifndef LBM_FUSED_BODY_ONLY
// ---- tile catalogue -----------------------------------------------------------
// kind 0: first pipeline (stage-A planes double-buffered in LDS);
// kind 1: prefetch pipeline (single stage-A plane, loads one plane ahead);
// kind 2: prefetch pipeline with non-temporal output stores;
// kind 3: third pipeline (divergence-form sets, one barrier per plane);
// kind 4: the same with z chunks slowest in the block order;
// kinds 5, 6: kinds 3, 4 with no memory instruction under a branch (grids
// of whole tiles).
struct TileShape {
    int kind, tx, ty, zc;
    const char* name;
};
constexpr TileShape kTiles[] = {
    {0, 32, 8, 32, "32x8x32"},        {0, 32, 8, 64, "32x8x64"},        {0, 64, 4, 32, "64x4x32"},
    {1, 64, 4, 64, "pf:64x4x64"},     {1, 32, 8, 64, "pf:32x8x64"},     {1, 64, 4, 32, "pf:64x4x32"},
    {1, 32, 8, 32, "pf:32x8x32"},     {2, 64, 4, 64, "pf-nt:64x4x64"},  {2, 32, 8, 64, "pf-nt:32x8x64"},
    // 512-thread tiles: the ring of halo columns whose Q update is recomputed
    // is a smaller share of the interior (64x8: 148 of 512, 32x16: 100 of 512)
    {1, 64, 8, 64, "pf:64x8x64"},     {1, 32, 16, 64, "pf:32x16x64"},   {1, 64, 8, 32, "pf:64x8x32"},
    {1, 32, 16, 32, "pf:32x16x32"},
    // longer z chunks: fewer halo planes per chunk, and at 256^3 every block
    // is resident at once (512 blocks of 4 waves, or 256 of 8 waves)
    {2, 64, 4, 128, "pf-nt:64x4x128"}, {2, 32, 8, 128, "pf-nt:32x8x128"}, {2, 64, 8, 128, "pf-nt:64x8x128"},
    // short z chunks: enough blocks to fill the CUs on thin or small grids
    // (a 256 x 256 x 32 slab, 128^3)
    {2, 64, 4, 16, "pf-nt:64x4x16"},   {2, 32, 8, 16, "pf-nt:32x8x16"},
    // third pipeline: divergence-form sets, one barrier per plane
    {3, 64, 4, 128, "p2:64x4x128"},    {3, 32, 8, 128, "p2:32x8x128"},   {3, 64, 4, 64, "p2:64x4x64"},
    {3, 64, 4, 16, "p2:64x4x16"},      {3, 32, 8, 16, "p2:32x8x16"},
    // third pipeline, z chunks slowest in the block order
    {4, 64, 4, 128, "p2z:64x4x128"},   {4, 64, 4, 64, "p2z:64x4x64"},    {4, 64, 4, 32, "p2z:64x4x32"},
    {4, 32, 8, 128, "p2z:32x8x128"},
    // third pipeline, no memory instruction under a branch: the compiler's
    // memory-counter waits after a branch assume its loads and stores were
    // skipped, and then wait for everything issued before it, including the
    // previous plane's stores (3% faster at 256^3)
    {5, 64, 4, 128, "p2u:64x4x128"},   {6, 64, 4, 128, "p2uz:64x4x128"}, {6, 64, 4, 64, "p2uz:64x4x64"},
    {5, 64, 4, 16, "p2u:64x4x16"},
};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

int fused_tile_count() { return kNumTiles; }
const char* fused_tile_name(int tile) { return (tile >= 0 && tile < kNumTiles) ? kTiles[tile].name : "?"; }

// Shell launches of a domain with walls: after the interior launch, the
// unfused step kernels' per-cell code recomputes the cells within two cells
// of an open face, with full wall handling, in two passes over the shell
// (one thread per cell, so the few shell cells of a large or small grid are
// spread over the whole GPU). Both passes read the previous state, the new
// Q^{n+1} of the interior launch next to the shell, and Sigma/Tau that the
// first pass writes.
static void launch_shells(const Fields<float>& f, const BCDesc<float>& bc, const Phys<float>& k, int pull, int aux,
                          hipStream_t s) {
    const ShellRegion r3 = make_shell_region(bc, 3), r2 = make_shell_region(bc, 2);
    if (region_cells(r2) == 0) return;
    hipLaunchKernelGGL(shell_qstep_kernel, dim3(grid_for(region_cells(r3))), dim3(kBlock), 0, s, f, bc, k, r3);
    Fields<float> fq = f;
    for (int m = 0; m < 5; ++m) fq.q[m] = f.qn[m];
    const int flags = FL_FORCE | FL_LBM | (aux ? (FL_STORE_F | FL_STORE_RHO) : 0);
    hipLaunchKernelGGL(shell_force_lbm_kernel, dim3(grid_for(region_cells(r2))), dim3(kBlock), 0, s, fq, bc, k, r2,
                       flags, pull);
}

template <int KIND, int TX, int TY, int ZC>
static void launch_tile(const Fields<float>& f, const BCDesc<float>& bc, const Phys<float>& k, int pull, int aux,
                        hipStream_t s) {
    const int ntx = (bc.n[0] + TX - 1) / TX, nty = (bc.n[1] + TY - 1) / TY, nzc = (bc.n[2] + ZC - 1) / ZC;
    const int nblk = ntx * nty * nzc;
    const int grid = (nblk + 7) / 8 * 8;
    bool periodic = true;
    for (int face = 0; face < 6; ++face)
        periodic = periodic && bc.qkind[face] == QB_PERIODIC && bc.ukind[face] == UB_PERIODIC;
    const bool up = k.advection != 0;
#define LBM_FUSED_PF(M, U, N)                                                                              \
    hipLaunchKernelGGL((fused_step_kernel_pf<TX, TY, ZC, M, 2, U, N>), dim3(grid), dim3(TX * TY), 0, s, f, bc, k, \
                       pull, aux, ntx, nty, nzc)
    // Periodic domains run one launch. On domains with walls the periodic
    // specialisation of the chosen pipeline runs over every tile; its values
    // are exact for the cells two or more cells away from every open face.
    // The two shell passes then overwrite the cells within two cells of an
    // open face. rho, which the force pass reads for moving walls, is
    // written by the owner of each cell alone.
    const int mode = periodic ? 0 : 1;
    if constexpr (KIND == 0) {
        hipLaunchKernelGGL((fused_step_kernel<TX, TY, ZC, true>), dim3(grid), dim3(TX * TY), 0, s, f, bc, k, pull,
                           aux, ntx, nty, nzc, mode);
    } else if constexpr (KIND == 1) {
        if (periodic) {
            if (up) LBM_FUSED_PF(false, true, false); else LBM_FUSED_PF(false, false, false);
        } else {
            if (up) LBM_FUSED_PF(true, true, false); else LBM_FUSED_PF(true, false, false);
        }
    } else if constexpr (KIND == 2) {
        if (periodic) {
            if (up) LBM_FUSED_PF(false, true, true); else LBM_FUSED_PF(false, false, true);
        } else {
            if (up) LBM_FUSED_PF(true, true, true); else LBM_FUSED_PF(true, false, true);
        }
    } else {
        // kinds 5, 6: every memory instruction unconditional; only for
        // periodic grids whose x and y sides are multiples of the tile
        // (otherwise the conditional form of the same kind runs), and for
        // the tiles away from the walls of a domain with walls in x and y.
        constexpr bool uc = KIND >= 5;
        constexpr bool zs = KIND == 4 || KIND == 6;
        const bool full = periodic && bc.n[0] % TX == 0 && bc.n[1] % TY == 0;
#define LBM_FUSED_P2(M, U, UC, W)                                                                                  \
    hipLaunchKernelGGL((fused_step_kernel_p2<TX, TY, ZC, M, 2, U, true, zs, UC, W>), dim3(grid), dim3(TX * TY), 0, s, \
                       f, bc, k, pull, aux, ntx, nty, nzc)
        if (uc && full) {
            if (up) LBM_FUSED_P2(false, true, uc, false); else LBM_FUSED_P2(false, false, uc, false);
        } else if (periodic) {
            if (up) LBM_FUSED_P2(false, true, false, false); else LBM_FUSED_P2(false, false, false, false);
        } else if (p2_walls_ok(bc)) {
            // walls in x and y only: every cell exact in the one launch (tiles
            // away from the walls on the periodic form, unconditional in the
            // kinds that are)
            if (up) LBM_FUSED_P2(false, true, uc, true); else LBM_FUSED_P2(false, false, uc, true);
            return;
        } else {
            if (up) LBM_FUSED_P2(true, true, false, false); else LBM_FUSED_P2(true, false, false, false);
        }
#undef LBM_FUSED_P2
    }
#undef LBM_FUSED_PF
    if (!periodic) launch_shells(f, bc, k, pull, aux, s);
}

void launch_fused_step_f32(const Fields<float>& f, const BCDesc<float>& bc, const Phys<float>& k, int tile,
                           int pull, int aux, hipStream_t s) {
    switch (tile) {
        case 0: launch_tile<0, 32, 8, 32>(f, bc, k, pull, aux, s); break;
        case 1: launch_tile<0, 32, 8, 64>(f, bc, k, pull, aux, s); break;
        case 2: launch_tile<0, 64, 4, 32>(f, bc, k, pull, aux, s); break;
        case 3: launch_tile<1, 64, 4, 64>(f, bc, k, pull, aux, s); break;
        case 4: launch_tile<1, 32, 8, 64>(f, bc, k, pull, aux, s); break;
        case 5: launch_tile<1, 64, 4, 32>(f, bc, k, pull, aux, s); break;
        case 6: launch_tile<1, 32, 8, 32>(f, bc, k, pull, aux, s); break;
        case 7: launch_tile<2, 64, 4, 64>(f, bc, k, pull, aux, s); break;
        case 8: launch_tile<2, 32, 8, 64>(f, bc, k, pull, aux, s); break;
        case 9: launch_tile<1, 64, 8, 64>(f, bc, k, pull, aux, s); break;
        case 10: launch_tile<1, 32, 16, 64>(f, bc, k, pull, aux, s); break;
        case 11: launch_tile<1, 64, 8, 32>(f, bc, k, pull, aux, s); break;
        case 12: launch_tile<1, 32, 16, 32>(f, bc, k, pull, aux, s); break;
        case 13: launch_tile<2, 64, 4, 128>(f, bc, k, pull, aux, s); break;
        case 14: launch_tile<2, 32, 8, 128>(f, bc, k, pull, aux, s); break;
        case 15: launch_tile<2, 64, 8, 128>(f, bc, k, pull, aux, s); break;
        case 16: launch_tile<2, 64, 4, 16>(f, bc, k, pull, aux, s); break;
        case 17: launch_tile<2, 32, 8, 16>(f, bc, k, pull, aux, s); break;
        case 18: launch_tile<3, 64, 4, 128>(f, bc, k, pull, aux, s); break;
        case 19: launch_tile<3, 32, 8, 128>(f, bc, k, pull, aux, s); break;
        case 20: launch_tile<3, 64, 4, 64>(f, bc, k, pull, aux, s); break;
        case 21: launch_tile<3, 64, 4, 16>(f, bc, k, pull, aux, s); break;
        case 22: launch_tile<3, 32, 8, 16>(f, bc, k, pull, aux, s); break;
        case 23: launch_tile<4, 64, 4, 128>(f, bc, k, pull, aux, s); break;
        case 24: launch_tile<4, 64, 4, 64>(f, bc, k, pull, aux, s); break;
        case 25: launch_tile<4, 64, 4, 32>(f, bc, k, pull, aux, s); break;
        case 26: launch_tile<4, 32, 8, 128>(f, bc, k, pull, aux, s); break;
        case 27: launch_tile<5, 64, 4, 128>(f, bc, k, pull, aux, s); break;
        case 28: launch_tile<6, 64, 4, 128>(f, bc, k, pull, aux, s); break;
        case 29: launch_tile<6, 64, 4, 64>(f, bc, k, pull, aux, s); break;
        default: launch_tile<5, 64, 4, 16>(f, bc, k, pull, aux, s); break;
    }
}

// ---- HBM calibration: plain 16-byte-per-lane copy ------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) copy_kernel(const v4f* __restrict__ src, v4f* __restrict__ dst, long n4) {
    // Four independent 16-byte loads in flight per lane, then the stores.
    const long base = ((long)blockIdx.x * 4) * blockDim.x + threadIdx.x;
    v4f v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const long i = base + (long)r * blockDim.x;
        if (i < n4) v[r] = __builtin_nontemporal_load(&src[i]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const long i = base + (long)r * blockDim.x;
        if (i < n4) __builtin_nontemporal_store(v[r], &dst[i]);
    }
}
void launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const long n4 = (long)(bytes / 16);
    const long blocks = (n4 + 4L * 256 - 1) / (4L * 256);
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (const v4f*)src, (v4f*)dst, n4);
}

#endif  // LBM_FUSED_BODY_ONLY

}  // namespace lbm
