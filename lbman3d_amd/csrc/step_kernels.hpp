// step_kernels.hpp — the unfused step kernels, templated on the arithmetic
// type. One thread per cell, cells linear in x-fastest order so every field
// access of a wavefront is a contiguous, coalesced run.
#pragma once

#include "common.hpp"
#include "launch.hpp"
#include "physics.hpp"

namespace lbm {

constexpr int kBlock = 256;

struct Pos {
    int x, y, z;
};
__device__ __forceinline__ Pos cell_pos(long c, int nx, int ny) {
    Pos p;
    const long xy = (long)nx * ny;
    p.z = (int)(c / xy);
    const long r = c - (long)p.z * xy;
    p.y = (int)(r / nx);
    p.x = (int)(r - (long)p.y * nx);
    return p;
}
__device__ __forceinline__ long cell_of(int x, int y, int z, int nx, int ny) {
    return ((long)z * ny + y) * nx + x;
}

// Gather the 15 incoming populations of node (x, y, z) from the
// post-collision buffer g: upwind neighbour, or the node's own reflected
// population at a wall (push semantics of src/lbm_solver.tpp:144-192 and
// HandleBoundaryPoint, src/boundary_handler.h:403-426, restated as pull
// through the host-built slot table). rho_prev is the density this node had
// when those populations were produced (moving walls).
template <typename T, bool SHIFTED>
__device__ __forceinline__ void gather(const Fields<T>& f, const BCDesc<T>& bc, int x, int y, int z, long c,
                                       int pull, T out[kQ]) {
    const int nx = bc.n[0], ny = bc.n[1], nz = bc.n[2];
    const T* g = f.g;
    const long vs = f.vs;
    if (!pull) {
#pragma unroll
        for (int j = 0; j < kQ; ++j) out[j] = g[j * vs + c];
        return;
    }
    const int cls = node_class(x, y, z, nx, ny, nz);
    if (cls == 0) {
#pragma unroll
        for (int j = 0; j < kQ; ++j) out[j] = g[j * vs + c - cx(j) - (long)cy(j) * nx - (long)cz(j) * nx * ny];
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
            out[j] = g[j * vs + cell_of(xs, ys, zs, nx, ny)];
        } else if (kind == SLOT_MIRROR) {
            out[j] = g[slot_src(e) * vs + c];
        } else {
            const int face = slot_face(e);
            if (!have_rho) { rho_prev = bc.rho_base + f.rho[c]; have_rho = true; }
            const T dot = T(cx(j)) * bc.wall_u[face][0] + T(cy(j)) * bc.wall_u[face][1]
                        + T(cz(j)) * bc.wall_u[face][2];
            out[j] = g[slot_src(e) * vs + c] + T(6.0) * rho_prev * T(wgt(j)) * dot;
        }
    }
}

// ---- Q update ----------------------------------------------------------------
// QTensorSolver::StepAndSetupBodyForce (src/qtensor_solver.tpp:50-147).
// One cell. `own` false stores Sigma and Tau only (a halo cell of a region
// pass whose Q^{n+1} and Ericksen force belong to another launch).
template <typename T>
__device__ __forceinline__ void qstep_cell(const Fields<T>& f, const BCDesc<T>& bc, const Phys<T>& k, long c, Pos p,
                                           bool own) {
    const int nx = bc.n[0], ny = bc.n[1], nz = bc.n[2];
    const int pos[3] = {p.x, p.y, p.z};

    // Q stencil neighbours (Q tags), velocity neighbours (velocity tags).
    const int qxm = nb_index(p.x, -1, nx, bc.qkind[0] == QB_PERIODIC, bc.qkind[1] == QB_PERIODIC);
    const int qxp = nb_index(p.x, +1, nx, bc.qkind[0] == QB_PERIODIC, bc.qkind[1] == QB_PERIODIC);
    const int qym = nb_index(p.y, -1, ny, bc.qkind[2] == QB_PERIODIC, bc.qkind[3] == QB_PERIODIC);
    const int qyp = nb_index(p.y, +1, ny, bc.qkind[2] == QB_PERIODIC, bc.qkind[3] == QB_PERIODIC);
    const int qzm = nb_index(p.z, -1, nz, bc.qkind[4] == QB_PERIODIC, bc.qkind[5] == QB_PERIODIC);
    const int qzp = nb_index(p.z, +1, nz, bc.qkind[4] == QB_PERIODIC, bc.qkind[5] == QB_PERIODIC);
    const long q_nb[6] = {cell_of(qxm, p.y, p.z, nx, ny), cell_of(qxp, p.y, p.z, nx, ny),
                          cell_of(p.x, qym, p.z, nx, ny), cell_of(p.x, qyp, p.z, nx, ny),
                          cell_of(p.x, p.y, qzm, nx, ny), cell_of(p.x, p.y, qzp, nx, ny)};
    T Q[5];
    Deriv<T> dq[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const T* q = f.q[m];
        Q[m] = q[c];
        dq[m] = q_deriv(bc, m, pos, Q[m], q[q_nb[0]], q[q_nb[1]], q[q_nb[2]], q[q_nb[3]], q[q_nb[4]], q[q_nb[5]]);
    }
    const int uxm = nb_index(p.x, -1, nx, bc.ukind[0] == UB_PERIODIC, bc.ukind[1] == UB_PERIODIC);
    const int uxp = nb_index(p.x, +1, nx, bc.ukind[0] == UB_PERIODIC, bc.ukind[1] == UB_PERIODIC);
    const int uym = nb_index(p.y, -1, ny, bc.ukind[2] == UB_PERIODIC, bc.ukind[3] == UB_PERIODIC);
    const int uyp = nb_index(p.y, +1, ny, bc.ukind[2] == UB_PERIODIC, bc.ukind[3] == UB_PERIODIC);
    const int uzm = nb_index(p.z, -1, nz, bc.ukind[4] == UB_PERIODIC, bc.ukind[5] == UB_PERIODIC);
    const int uzp = nb_index(p.z, +1, nz, bc.ukind[4] == UB_PERIODIC, bc.ukind[5] == UB_PERIODIC);
    const long u_nb[6] = {cell_of(uxm, p.y, p.z, nx, ny), cell_of(uxp, p.y, p.z, nx, ny),
                          cell_of(p.x, uym, p.z, nx, ny), cell_of(p.x, uyp, p.z, nx, ny),
                          cell_of(p.x, p.y, uzm, nx, ny), cell_of(p.x, p.y, uzp, nx, ny)};
    T u[3], g[3][3];
#pragma unroll
    for (int A = 0; A < 3; ++A) {
        const T* v = f.u[A];
        const T v0 = v[c];
        u[A] = v0;
#pragma unroll
        for (int B = 0; B < 3; ++B) g[A][B] = u_grad(bc, B, A, pos[B], v[u_nb[2 * B]], v0, v[u_nb[2 * B + 1]]);
    }
    T qn[5], sg[5], ta[3], fe[3];
    q_pointwise(k, Q, u, dq, g, qn, sg, ta, fe);
#pragma unroll
    for (int m = 0; m < 5; ++m) f.sig[m][c] = sg[m];
#pragma unroll
    for (int m = 0; m < 3; ++m) f.tau[m][c] = ta[m];
    if (own) {
#pragma unroll
        for (int m = 0; m < 5; ++m) f.qn[m][c] = qn[m];
#pragma unroll
        for (int m = 0; m < 3; ++m) f.F[m][c] = fe[m];
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) qstep_kernel(Fields<T> f, BCDesc<T> bc, Phys<T> k, long V) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= V) return;
    qstep_cell(f, bc, k, c, cell_pos(c, bc.n[0], bc.n[1]), true);
}

// ---- body force and collide-stream ----------------------------------------
// SetActiveStressAndComputeBodyForce (src/qtensor_solver.tpp:154-217) and
// LatticeBoltzmannStep (src/lbm_solver.tpp:89-224) at one node. Fusing the two
// is the merge the reference documents as valid (src/qtensor_solver.tpp:219-247):
// the force at a node feeds only that node's collision.
// One cell. UN: the velocity goes to un (fused-step buffers, where u still
// holds u^{n-1} for other cells) instead of replacing u in place.
template <typename T, bool SHIFTED, bool UN>
__device__ __forceinline__ void force_lbm_cell(const Fields<T>& f, const BCDesc<T>& bc, const Phys<T>& k, long c,
                                               Pos p, int flags, int pull) {
    const int nx = bc.n[0], ny = bc.n[1], nz = bc.n[2];
    T F[3];
    if (flags & FL_FORCE) {
        const T fe[3] = {f.F[0][c], f.F[1][c], f.F[2][c]};
        if (k.force_full) {
            const int pos[3] = {p.x, p.y, p.z};
            const int xm = nb_index(p.x, -1, nx, bc.qkind[0] == QB_PERIODIC, bc.qkind[1] == QB_PERIODIC);
            const int xp = nb_index(p.x, +1, nx, bc.qkind[0] == QB_PERIODIC, bc.qkind[1] == QB_PERIODIC);
            const int ym = nb_index(p.y, -1, ny, bc.qkind[2] == QB_PERIODIC, bc.qkind[3] == QB_PERIODIC);
            const int yp = nb_index(p.y, +1, ny, bc.qkind[2] == QB_PERIODIC, bc.qkind[3] == QB_PERIODIC);
            const int zm = nb_index(p.z, -1, nz, bc.qkind[4] == QB_PERIODIC, bc.qkind[5] == QB_PERIODIC);
            const int zp = nb_index(p.z, +1, nz, bc.qkind[4] == QB_PERIODIC, bc.qkind[5] == QB_PERIODIC);
            const long nb[6] = {cell_of(xm, p.y, p.z, nx, ny), cell_of(xp, p.y, p.z, nx, ny),
                                cell_of(p.x, ym, p.z, nx, ny), cell_of(p.x, yp, p.z, nx, ny),
                                cell_of(p.x, p.y, zm, nx, ny), cell_of(p.x, p.y, zp, nx, ny)};
            T dQ[5][3];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const T* q = f.q[m];
                const Deriv<T> d = q_deriv(bc, m, pos, q[c], q[nb[0]], q[nb[1]], q[nb[2]], q[nb[3]], q[nb[4]], q[nb[5]]);
                dQ[m][0] = d.d[0];
                dQ[m][1] = d.d[1];
                dQ[m][2] = d.d[2];
            }
            T dS[8][3];
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const T* s = m < 5 ? f.sig[m] : f.tau[m - 5];
#pragma unroll
                for (int a = 0; a < 3; ++a) dS[m][a] = (s[nb[2 * a + 1]] - s[nb[2 * a]]) / T(2.0);
            }
            const T u[3] = {f.u[0][c], f.u[1][c], f.u[2][c]};
            body_force(k, dQ, dS, u, fe, F);
        } else {
            F[0] = fe[0];
            F[1] = fe[1];
            F[2] = fe[2];
        }
        if (flags & FL_STORE_F) {
            f.F[0][c] = F[0];
            f.F[1][c] = F[1];
            f.F[2][c] = F[2];
        }
    } else {
        F[0] = f.F[0][c];
        F[1] = f.F[1][c];
        F[2] = f.F[2][c];
    }
    if (!(flags & FL_LBM)) return;

    T pop[kQ];
    gather<T, SHIFTED>(f, bc, p.x, p.y, p.z, c, pull, pop);
    T rho_st, u[3];   // the density in its stored form (rho - BCDesc::rho_base)
    if constexpr (SHIFTED) collide_shifted(k, pop, F, rho_st, u);
    else collide_plain(k, pop, F, rho_st, u);
    if (flags & FL_STORE_RHO) f.rho[c] = rho_st;
    T* const* uo = UN ? f.un : f.u;
    uo[0][c] = u[0];
    uo[1][c] = u[1];
    uo[2][c] = u[2];
#pragma unroll
    for (int j = 0; j < kQ; ++j) f.gn[j * f.vs + c] = pop[j];
}

template <typename T, bool SHIFTED>
__global__ void __launch_bounds__(kBlock)
force_lbm_kernel(Fields<T> f, BCDesc<T> bc, Phys<T> k, long V, int flags, int pull) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= V) return;
    force_lbm_cell<T, SHIFTED, false>(f, bc, k, c, cell_pos(c, bc.n[0], bc.n[1]), flags, pull);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
stream_only_kernel(Fields<T> f, BCDesc<T> bc, long V, int pull, T* out) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= V) return;
    const Pos p = cell_pos(c, bc.n[0], bc.n[1]);
    T pop[kQ];
    gather<T, false>(f, bc, p.x, p.y, p.z, c, pull, pop);
#pragma unroll
    for (int j = 0; j < kQ; ++j) out[j * f.vs + c] = pop[j];
}

// ---- FP64 diagnostics ---------------------------------------------------------
// SimIO::Log sums (src/sim_io.cc:78-109) and TotalNematicFreeEnergy
// (src/analysis_fields.h:39-72, Laplacian with the Neumann clamp, neighbour
// order xp, xm, yp, ym, zp, zm; density from src/analysis_fields.cc:38-57).
constexpr int kDiag = 10;
template <typename T>
__global__ void __launch_bounds__(kBlock)
diag_partial_kernel(Fields<T> f, BCDesc<T> bc, Phys<double> k, long V, double* uprev, double* partials) {
    __shared__ double red[kBlock / 64][kDiag];
    double acc[kDiag] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int nx = bc.n[0], ny = bc.n[1], nz = bc.n[2];
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < V; c += (long)gridDim.x * blockDim.x) {
        const double r = (sizeof(T) == 4 ? k.rho0 : 0.0) + (double)f.rho[c];
        const double ux = (double)f.u[0][c], uy = (double)f.u[1][c], uz = (double)f.u[2][c];
        acc[0] += r;
        acc[1] += r * ux;
        acc[2] += r * uy;
        acc[3] += r * uz;
        const double u2 = ux * ux + uy * uy + uz * uz;
        acc[4] += 0.5 * r * u2;
        const double dx = ux - uprev[c], dy = uy - uprev[V + c], dz = uz - uprev[2 * V + c];
        acc[5] += dx * dx + dy * dy + dz * dz;
        acc[6] += u2;
        acc[7] = u2 > acc[7] ? u2 : acc[7];
        uprev[c] = ux;
        uprev[V + c] = uy;
        uprev[2 * V + c] = uz;
        if (!isfinite(r) || !isfinite(ux) || !isfinite(uy) || !isfinite(uz)) acc[9] += 1.0;

        const Pos p = cell_pos(c, nx, ny);
        const int xm = nb_index(p.x, -1, nx, bc.qkind[0] == QB_PERIODIC, bc.qkind[1] == QB_PERIODIC);
        const int xp = nb_index(p.x, +1, nx, bc.qkind[0] == QB_PERIODIC, bc.qkind[1] == QB_PERIODIC);
        const int ym = nb_index(p.y, -1, ny, bc.qkind[2] == QB_PERIODIC, bc.qkind[3] == QB_PERIODIC);
        const int yp = nb_index(p.y, +1, ny, bc.qkind[2] == QB_PERIODIC, bc.qkind[3] == QB_PERIODIC);
        const int zm = nb_index(p.z, -1, nz, bc.qkind[4] == QB_PERIODIC, bc.qkind[5] == QB_PERIODIC);
        const int zp = nb_index(p.z, +1, nz, bc.qkind[4] == QB_PERIODIC, bc.qkind[5] == QB_PERIODIC);
        double Q[5], lap[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const T* v = f.q[m];
            Q[m] = (double)v[c];
            lap[m] = (double)v[cell_of(xp, p.y, p.z, nx, ny)] + (double)v[cell_of(xm, p.y, p.z, nx, ny)]
                   + (double)v[cell_of(p.x, yp, p.z, nx, ny)] + (double)v[cell_of(p.x, ym, p.z, nx, ny)]
                   + (double)v[cell_of(p.x, p.y, zp, nx, ny)] + (double)v[cell_of(p.x, p.y, zm, nx, ny)]
                   - 6.0 * Q[m];
        }
        const double Qxx = Q[0], Qxy = Q[1], Qxz = Q[2], Qyy = Q[3], Qyz = Q[4];
        const double TrQ2 = 2.0 * (Qxx * Qxx + Qxx * Qyy + Qyy * Qyy + Qxy * Qxy + Qxz * Qxz + Qyz * Qyz);
        const double third = 1.0 / 3.0;
        const double Q2xx = Qxx * Qxx + Qxy * Qxy + Qxz * Qxz - third * TrQ2;
        const double Q2xy = Qxx * Qxy + Qxy * Qyy + Qxz * Qyz;
        const double Q2xz = Qxy * Qyz - Qxz * Qyy;
        const double Q2yy = Qxy * Qxy + Qyy * Qyy + Qyz * Qyz - third * TrQ2;
        const double Q2yz = Qxy * Qxz - Qyz * Qxx;
        const double TrQ3 = 2.0 * Qxx * Q2xx + 2.0 * Qyy * Q2yy + Qxx * Q2yy + Qyy * Q2xx
                          + 2.0 * (Qxy * Q2xy + Qxz * Q2xz + Qyz * Q2yz);
        const double QlapQ = 2.0 * Qxx * lap[0] + 2.0 * Qyy * lap[3] + Qxx * lap[3] + Qyy * lap[0]
                           + 2.0 * (Qxy * lap[1] + Qxz * lap[2] + Qyz * lap[4]);
        acc[8] += 0.5 * k.A * TrQ2 + (k.B / 3.0) * TrQ3 + 0.25 * k.C * TrQ2 * TrQ2 - 0.5 * k.L * QlapQ;
    }
    // Wavefront reduction (64 lanes), then across the block's wavefronts.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < kDiag; ++m) {
        double v = acc[m];
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_down(v, off, 64);
            v = (m == 7) ? (o > v ? o : v) : v + o;
        }
        if (lane == 0) red[wave][m] = v;
    }
    __syncthreads();
    if (threadIdx.x < kDiag) {
        const int m = threadIdx.x;
        double v = red[0][m];
        for (int w = 1; w < kBlock / 64; ++w) v = (m == 7) ? (red[w][m] > v ? red[w][m] : v) : v + red[w][m];
        partials[(long)blockIdx.x * kDiag + m] = v;
    }
}

template <int Unused = 0>
__global__ void __launch_bounds__(64) diag_final_kernel(const double* partials, int nblocks, double* out) {
    const int m = threadIdx.x;
    if (m >= kDiag) return;
    double v = 0.0;
    for (int b = 0; b < nblocks; ++b) {
        const double x = partials[(long)b * kDiag + m];
        v = (m == 7) ? (x > v ? x : v) : v + x;
    }
    out[m] = v;
}

// QtensorToOrderDirector (src/analysis_fields.cc:59-93) in FP64.
template <typename T>
__global__ void __launch_bounds__(kBlock)
order_director_kernel(Fields<T> f, long V, double* order, double* director) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= V) return;
    const double Qxx = f.q[0][c], Qxy = f.q[1][c], Qxz = f.q[2][c], Qyy = f.q[3][c], Qyz = f.q[4][c];
    const double pp = Qxx * Qxx + Qxx * Qyy + Qyy * Qyy + Qxy * Qxy + Qxz * Qxz + Qyz * Qyz;
    const double qd = -(Qxx + Qyy) * (Qxx * Qyy - Qxy * Qxy) - Qyy * Qxz * Qxz - Qxx * Qyz * Qyz + 2 * Qxy * Qxz * Qyz;
    const double r = 2.0 * sqrt(pp / 3.0);
    double arg = 4 * qd / (r * r * r);
    arg = arg < -1.0 ? -1.0 : (arg > 1.0 ? 1.0 : arg);
    const double S = r * cos(1.0 / 3.0 * acos(arg));
    double nxv = Qxz * (Qyy - S) - Qxy * Qyz;
    double nyv = Qyz * (Qxx - S) - Qxy * Qxz;
    double nzv = Qxy * Qxy - (Qxx - S) * (Qyy - S);
    const double inv = 1.0 / sqrt(nxv * nxv + nyv * nyv + nzv * nzv);
    order[c] = S;
    director[3 * c + 0] = nxv * inv;
    director[3 * c + 1] = nyv * inv;
    director[3 * c + 2] = nzv * inv;
}

// ---- launchers ---------------------------------------------------------------
inline unsigned grid_for(long V) { return (unsigned)((V + kBlock - 1) / kBlock); }

template <typename T>
void launch_qstep(const Fields<T>& f, const BCDesc<T>& bc, const Phys<T>& k, long V, hipStream_t s) {
    hipLaunchKernelGGL(qstep_kernel<T>, dim3(grid_for(V)), dim3(kBlock), 0, s, f, bc, k, V);
}

template <typename T, bool SHIFTED>
void launch_force_lbm(const Fields<T>& f, const BCDesc<T>& bc, const Phys<T>& k, long V, int flags, int pull,
                      hipStream_t s) {
    hipLaunchKernelGGL((force_lbm_kernel<T, SHIFTED>), dim3(grid_for(V)), dim3(kBlock), 0, s, f, bc, k, V, flags,
                       pull);
}

template <typename T>
void launch_stream_only(const Fields<T>& f, const BCDesc<T>& bc, long V, int pull, T* out, hipStream_t s) {
    hipLaunchKernelGGL(stream_only_kernel<T>, dim3(grid_for(V)), dim3(kBlock), 0, s, f, bc, V, pull, out);
}

template <typename T>
void launch_diagnostics(const Fields<T>& f, const BCDesc<T>& bc, const Phys<double>& k, long V, double* uprev,
                        double* partials, int nblocks, double* out, hipStream_t s) {
    hipLaunchKernelGGL(diag_partial_kernel<T>, dim3(nblocks), dim3(kBlock), 0, s, f, bc, k, V, uprev, partials);
    hipLaunchKernelGGL(diag_final_kernel<0>, dim3(1), dim3(64), 0, s, partials, nblocks, out);
}

template <typename T>
void launch_order_director(const Fields<T>& f, long V, double* order, double* director, hipStream_t s) {
    hipLaunchKernelGGL(order_director_kernel<T>, dim3(grid_for(V)), dim3(kBlock), 0, s, f, V, order, director);
}

}  // namespace lbm
