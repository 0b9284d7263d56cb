// physics.hpp — pointwise physics of the lbman3d step as device templates.
//
// Every expression keeps the operand order of the reference formula it
// implements (file:line at each function). Compiled without FMA contraction
// (kernels_fp64.hip) the FP64 instantiation therefore reproduces the
// reference CPU arithmetic bit for bit; the FP32 instantiation
// (kernels_fp32.hip) is compiled with contraction and uses the shifted
// population algebra of collide_shifted below.
#pragma once

#include "common.hpp"

namespace lbm {

template <typename T>
struct Deriv {
    T d[3];    // centred first differences along x, y, z
    T lap;     // 7-point Laplacian as one six-neighbour sum
    T d2[3];   // per-axis second differences
};

// Neighbour index along one axis: a periodic side wraps, any other side
// clamps (src/boundary_handler.h:76-89 for gradients, src/offsets.h:190-203
// for the stress divergence and free energy; both reduce to this rule).
__device__ __forceinline__ int nb_index(int i, int s, int n, bool lo_per, bool hi_per) {
    const int t = i + s;
    if (t < 0) return lo_per ? t + n : 0;
    if (t >= n) return hi_per ? t - n : n - 1;
    return t;
}

// Ghost-substituted (minus, plus) pair of Q component c along axis a
// (QAxisGhostPair / QGhost, src/boundary_handler.h:257-282): only the low
// face's tag selects the periodic short-cut.
template <typename T>
__device__ __forceinline__ void q_pair(const BCDesc<T>& bc, int a, int c, int i, T qm, T q0, T qp,
                                       T& m, T& p) {
    m = qm;
    p = qp;
    if (bc.qkind[2 * a] == QB_PERIODIC) return;
    if (i == 0) m = (bc.qkind[2 * a] == QB_ANCHOR) ? T(2.0) * bc.anchor[2 * a][c] - q0 : q0;
    if (i == bc.n[a] - 1) p = (bc.qkind[2 * a + 1] == QB_ANCHOR) ? T(2.0) * bc.anchor[2 * a + 1][c] - q0 : q0;
}

// QGradientAndLaplacian for one component from its seven stencil values
// (src/boundary_handler.h:287-320).
// NOGHOST: every face periodic, so the fetched neighbours are used as they are.
template <typename T, bool NOGHOST = false>
__device__ __forceinline__ Deriv<T> q_deriv(const BCDesc<T>& bc, int c, const int pos[3], T q0,
                                            T xm, T xp, T ym, T yp, T zm, T zp) {
    T mx = xm, px = xp, my = ym, py = yp, mz = zm, pz = zp;
    if (!NOGHOST) {
        q_pair(bc, 0, c, pos[0], xm, q0, xp, mx, px);
        q_pair(bc, 1, c, pos[1], ym, q0, yp, my, py);
        q_pair(bc, 2, c, pos[2], zm, q0, zp, mz, pz);
    }
    Deriv<T> r;
    r.d[0] = (px - mx) / T(2.0);
    r.d[1] = (py - my) / T(2.0);
    r.d[2] = (pz - mz) / T(2.0);
    r.lap = (mx + px + my + py + mz + pz) - T(6.0) * q0;
    r.d2[0] = (mx + px) - T(2.0) * q0;
    r.d2[1] = (my + py) - T(2.0) * q0;
    r.d2[2] = (mz + pz) - T(2.0) * q0;
    return r;
}

// Velocity ghost pair for component comp along axis a
// (VelocityGhost / VelocityAxisGhostPair, src/boundary_handler.h:121-146).
template <typename T>
__device__ __forceinline__ T u_ghost(const BCDesc<T>& bc, int face, int comp, bool normal, T vb) {
    const int k = bc.ukind[face];
    const T wv = (k == UB_MOVING) ? bc.wall_u[face][comp] : T(0.0);
    if (k == UB_SPECULAR) return normal ? (T(2.0) * wv - vb) : vb;
    return T(2.0) * wv - vb;
}
template <typename T, bool NOGHOST = false>
__device__ __forceinline__ T u_grad(const BCDesc<T>& bc, int a, int comp, int i, T vm, T v0, T vp) {
    T m = vm, p = vp;
    if (!NOGHOST && bc.ukind[2 * a] != UB_PERIODIC) {
        const bool normal = (comp == a);
        if (i == 0) m = u_ghost(bc, 2 * a, comp, normal, v0);
        if (i == bc.n[a] - 1) p = u_ghost(bc, 2 * a + 1, comp, normal, v0);
    }
    return (p - m) / T(2.0);
}

// The same derivatives for domains whose x and y faces are periodic or walls
// with a Neumann Q tag and a periodic, no-slip or specular velocity tag, and
// whose z faces are periodic (the fused step's wall-aware form): q_pair and
// u_ghost reduced to those tags, with the z axis taken as periodic.
template <typename T>
__device__ __forceinline__ void q_pair_neumann(const BCDesc<T>& bc, int a, int i, T qm, T q0, T qp, T& m, T& p) {
    const bool ghost = bc.qkind[2 * a] != QB_PERIODIC;
    m = (ghost && i == 0) ? q0 : qm;
    p = (ghost && i == bc.n[a] - 1) ? q0 : qp;
}
template <typename T>
__device__ __forceinline__ Deriv<T> q_deriv_xy(const BCDesc<T>& bc, const int pos[3], T q0, T xm, T xp, T ym, T yp,
                                               T zm, T zp) {
    T mx, px, my, py;
    q_pair_neumann(bc, 0, pos[0], xm, q0, xp, mx, px);
    q_pair_neumann(bc, 1, pos[1], ym, q0, yp, my, py);
    Deriv<T> r;
    r.d[0] = (px - mx) / T(2.0);
    r.d[1] = (py - my) / T(2.0);
    r.d[2] = (zp - zm) / T(2.0);
    r.lap = (mx + px + my + py + zm + zp) - T(6.0) * q0;
    r.d2[0] = (mx + px) - T(2.0) * q0;
    r.d2[1] = (my + py) - T(2.0) * q0;
    r.d2[2] = (zm + zp) - T(2.0) * q0;
    return r;
}
// Velocity gradient along x or y: no-slip ghost -v, specular ghost -v for
// the normal component and v for the tangential ones.
template <typename T>
__device__ __forceinline__ T u_grad_xy(const BCDesc<T>& bc, int a, int comp, int i, T vm, T v0, T vp) {
    T m = vm, p = vp;
    if (bc.ukind[2 * a] != UB_PERIODIC) {
        const bool tangential = comp != a;
        if (i == 0) m = (bc.ukind[2 * a] == UB_SPECULAR && tangential) ? v0 : -v0;
        if (i == bc.n[a] - 1) p = (bc.ukind[2 * a + 1] == UB_SPECULAR && tangential) ? v0 : -v0;
    }
    return (p - m) / T(2.0);
}

// One axis of u.grad(Q) (AdvectiveAxisTerm, src/model.h:60-68).
template <typename T>
__device__ __forceinline__ T adv_axis(int upwind, T u, T dq, T d2q) {
    if (upwind) {
        const T au = u < T(0.0) ? -u : u;
        return u * dq - T(0.5) * au * d2q;
    }
    return u * dq;
}

// Beris-Edwards pointwise update, stresses and Ericksen force
// (PointwiseStepAndSetupBodyForce, src/model.h:102-342).
// g[A][B] = d u_A / d x_B with g[2][2] unused.
template <typename T>
__device__ __forceinline__ void q_pointwise(const Phys<T>& k, const T Q[5], const T u[3],
                                            const Deriv<T> dq[5], const T g[3][3],
                                            T qn[5], T sg[5], T ta[3], T fe[3]) {
    const T Qxx = Q[0], Qxy = Q[1], Qxz = Q[2], Qyy = Q[3], Qyz = Q[4];
    const T TrQ2 = T(2.0) * (Qxx * Qxx + Qyy * Qyy + Qxx * Qyy + Qxy * Qxy + Qxz * Qxz + Qyz * Qyz);
    const T third = T(1.0 / 3.0);
    const T Q2xx = Qxx * Qxx + Qxy * Qxy + Qxz * Qxz - third * TrQ2;
    const T Q2xy = Qxx * Qxy + Qxy * Qyy + Qxz * Qyz;
    const T Q2xz = Qxy * Qyz - Qxz * Qyy;
    const T Q2yy = Qxy * Qxy + Qyy * Qyy + Qyz * Qyz - third * TrQ2;
    const T Q2yz = Qxy * Qxz - Qyz * Qxx;
    const T Q2[5] = {Q2xx, Q2xy, Q2xz, Q2yy, Q2yz};

    const T Exx = g[0][0];
    const T Exy = T(0.5) * (g[0][1] + g[1][0]);
    const T Exz = T(0.5) * (g[0][2] + g[2][0]);
    const T Eyy = g[1][1];
    const T Eyz = T(0.5) * (g[1][2] + g[2][1]);
    const T Wxy = T(0.5) * (g[0][1] - g[1][0]);
    const T Wxz = T(0.5) * (g[0][2] - g[2][0]);
    const T Wyz = T(0.5) * (g[1][2] - g[2][1]);
    const T E[5] = {Exx, Exy, Exz, Eyy, Eyz};

    T advc[5];
#pragma unroll
    for (int c = 0; c < 5; ++c)
        advc[c] = -(adv_axis(k.advection, u[0], dq[c].d[0], dq[c].d2[0])
                    + adv_axis(k.advection, u[1], dq[c].d[1], dq[c].d2[1])
                    + adv_axis(k.advection, u[2], dq[c].d[2], dq[c].d2[2]));

    const T cor[5] = {T(2.0) * (Qxy * Wxy + Qxz * Wxz),
                      -Qxx * Wxy + Qxz * Wyz + Qyy * Wxy + Qyz * Wxz,
                      T(-2.0) * Qxx * Wxz - Qxy * Wyz - Qyy * Wxz + Qyz * Wxy,
                      T(2.0) * (-Qxy * Wxy + Qyz * Wyz),
                      -Qxx * Wyz - Qxy * Wxz - Qxz * Wxy - T(2.0) * Qyy * Wyz};

    const T two3 = T(2.0 / 3.0);
    const T trQE = (T(2.0) * Exx + Eyy) * Qxx + (T(2.0) * Exy) * Qxy + (T(2.0) * Exz) * Qxz
                 + (Exx + T(2.0) * Eyy) * Qyy + T(2.0) * Eyz * Qyz;
    const T aln2[5] = {
        T(2.0) * (Exx * Qxx + Exy * Qxy + Exz * Qxz) - two3 * trQE,
        Exy * Qxx + Eyy * Qxy + Eyz * Qxz + Qxy * Exx + Qyy * Exy + Qyz * Exz,
        Exz * Qxx + Eyz * Qxy + (-Exx - Eyy) * Qxz + Qxz * Exx + Qyz * Exy + (-Qxx - Qyy) * Exz,
        T(2.0) * (Exy * Qxy + Eyy * Qyy + Eyz * Qyz) - two3 * trQE,
        Exz * Qxy + Eyz * Qyy + (-Exx - Eyy) * Qyz + Qxz * Exy + Qyz * Eyy + (-Qxx - Qyy) * Eyz};

    const T ld = k.A + k.C * TrQ2;
    T H[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) H[c] = k.L * dq[c].lap - ld * Q[c] - k.B * Q2[c];
    const T Hxx = H[0], Hxy = H[1], Hxz = H[2], Hyy = H[3], Hyz = H[4];

#pragma unroll
    for (int a = 0; a < 3; ++a)
        fe[a] = T(-2.0) * (Hxx * dq[0].d[a] + Hxy * dq[1].d[a] + Hxz * dq[2].d[a] + Hyy * dq[3].d[a]
                           + Hyz * dq[4].d[a])
              - Hxx * dq[3].d[a] - Hyy * dq[0].d[a];

    const T trQH = (T(2.0) * Hxx + Hyy) * Qxx + (T(2.0) * Hxy) * Qxy + (T(2.0) * Hxz) * Qxz
                 + (Hxx + T(2.0) * Hyy) * Qyy + T(2.0) * Hyz * Qyz;
    const T QH[5] = {
        T(2.0) * (Hxx * Qxx + Hxy * Qxy + Hxz * Qxz) - two3 * trQH,
        Hxy * Qxx + Hyy * Qxy + Hyz * Qxz + Qxy * Hxx + Qyy * Hxy + Qyz * Hxz,
        Hxz * Qxx + Hyz * Qxy + (-Hxx - Hyy) * Qxz + Qxz * Hxx + Qyz * Hxy + (-Qxx - Qyy) * Hxz,
        T(2.0) * (Hxy * Qxy + Hyy * Qyy + Hyz * Qyz) - two3 * trQH,
        Hxz * Qxy + Hyz * Qyy + (-Hxx - Hyy) * Qyz + Qxz * Hxy + Qyz * Hyy + (-Qxx - Qyy) * Hyz};
    ta[0] = (Hxy * Qxx + Hyy * Qxy + Hyz * Qxz) - (Qxy * Hxx + Qyy * Hxy + Qyz * Hxz);
    ta[1] = (Hxz * Qxx + Hyz * Qxy + (-Hxx - Hyy) * Qxz) - (Qxz * Hxx + Qyz * Hxy + (-Qxx - Qyy) * Hxz);
    ta[2] = (Hxz * Qxy + Hyz * Qyy + (-Hxx - Hyy) * Qyz) - (Qxz * Hxy + Qyz * Hyy + (-Qxx - Qyy) * Hyz);

#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const T aln3 = T(-2.0) * trQE * Q[c];
        sg[c] = -two3 * k.lambda * H[c] - k.lambda * QH[c] + T(2.0) * k.lambda * trQH * Q[c];
        qn[c] = Q[c] + k.dt * (advc[c] + cor[c] + k.lambda * (two3 * E[c] + aln2[c] + aln3) + k.gamma * H[c]);
    }
}

// Active force + passive stress divergence - friction, added to the Ericksen
// force (PointwiseSetActiveStressAndComputeBodyForce, src/model.h:344-361;
// PassiveStressDivergence, src/boundary_handler.h:335-392). dS[m][a] is the
// centred difference of stress m (Sigma xx, xy, xz, yy, yz, Tau xy, xz, yz)
// along axis a; dQ[c][a] the centred differences of Q^{n+1}.
template <typename T>
__device__ __forceinline__ void body_force(const Phys<T>& k, const T dQ[5][3], const T dS[8][3],
                                           const T u[3], const T fe[3], T F[3]) {
    const T dxSxx = dS[0][0], dxSxy = dS[1][0], dxSxz = dS[2][0], dxTxy = dS[5][0], dxTxz = dS[6][0];
    const T dySxy = dS[1][1], dySyy = dS[3][1], dySyz = dS[4][1], dyTxy = dS[5][1], dyTyz = dS[7][1];
    const T dzSxz = dS[2][2], dzSyz = dS[4][2], dzTxz = dS[6][2], dzTyz = dS[7][2];
    const T dzSzz = -(dS[0][2] + dS[3][2]);
    const T pdx = dxSxx + (dySxy + dyTxy) + (dzSxz + dzTxz);
    const T pdy = (dxSxy - dxTxy) + dySyy + (dzSyz + dzTyz);
    const T pdz = (dxSxz - dxTxz) + (dySyz - dyTyz) + dzSzz;
    const T fx = -k.alpha * (dQ[0][0] + dQ[1][1] + dQ[2][2]) + pdx - k.mu * u[0];
    const T fy = -k.alpha * (dQ[1][0] + dQ[3][1] + dQ[4][2]) + pdy - k.mu * u[1];
    const T fz = -k.alpha * (dQ[2][0] + dQ[4][1] - dQ[0][2] - dQ[3][2]) + pdz - k.mu * u[2];
    F[0] = fe[0] + fx;
    F[1] = fe[1] + fy;
    F[2] = fe[2] + fz;
}

// Plain populations: moments with the half-force correction, then Feq + Guo
// forcing + BGK for every direction (ComputeMoments / ComputeFeqAndForcing /
// PointwiseBGKCollide, src/physics_helpers.h:69-129). f[] in, f[] out.
template <typename T>
__device__ __forceinline__ void collide_plain(const Phys<T>& k, T f[kQ], const T F[3], T& rho_out,
                                              T u_out[3]) {
    T rho = T(0.0), mx = T(0.0), my = T(0.0), mz = T(0.0);
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
        rho += f[i];
        mx += T(cx(i)) * f[i];
        my += T(cy(i)) * f[i];
        mz += T(cz(i)) * f[i];
    }
    mx += T(0.5) * F[0] * k.dt;
    my += T(0.5) * F[1] * k.dt;
    mz += T(0.5) * F[2] * k.dt;
    mx /= rho;
    my /= rho;
    mz /= rho;
    rho_out = rho;
    u_out[0] = mx;
    u_out[1] = my;
    u_out[2] = mz;
    const T uF = mx * F[0] + my * F[1] + mz * F[2];
    const T u2 = mx * mx + my * my + mz * mz;
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
        const T e0 = T(cx(i)), e1 = T(cy(i)), e2 = T(cz(i));
        const T wi = T(wgt(i));
        const T ue = mx * e0 + my * e1 + mz * e2;
        const T eF = e0 * F[0] + e1 * F[1] + e2 * F[2];
        const T feq = wi * rho * (T(1.0) + T(3.0) * ue + T(4.5) * ue * ue - T(1.5) * u2);
        const T frc = k.omega_f * wi * (T(3.0) * eF - T(3.0) * uF + T(9.0) * ue * eF);
        f[i] = k.omega * f[i] + k.omega_p * feq + k.dt * frc;
    }
}

// Shifted populations d_i = f_i - w_i*rho0 (the FP32 storage form). The same
// update as collide_plain, rewritten so that only small quantities are
// rounded: rho = rho0 + sum d_i, sum e_i f_i = sum e_i d_i, and since
// omega + omega' = 1,
//   d*_i = omega d_i + omega' w_i (drho + rho (3 ue + 4.5 ue^2 - 1.5 u^2)) + DT F_i.
// The directions are taken in opposite pairs (e_o = -e_i, w_o = w_i): with
// ue = e_i.u and eF = e_i.F the update splits into a part even in e_i,
//   S = w (omega' (drho - 1.5 rho u^2 + 4.5 rho ue^2) + DT omega_f (-3 u.F + 9 ue eF)),
// and an odd part A = 3 w (omega' rho ue + DT omega_f eF), so that
// d*_i = omega d_i + S + A and d*_o = omega d_o + S - A.
// drho_out is rho - rho0, the form the density array stores next to shifted
// populations (rho itself, rounded to FP32 near 1, would keep only about
// 6e-8 of the deviation).
template <typename T>
__device__ __forceinline__ void collide_shifted(const Phys<T>& k, T d[kQ], const T F[3], T& drho_out,
                                                T u_out[3]) {
    // Pair differences p = d_i - d_o carry the momentum; e_i of the first of
    // each pair: 1 (1,0,0), 2 (0,1,0), 5 (0,0,1), 7 (1,1,1), 8 (-1,1,1),
    // 9 (-1,-1,1), 10 (1,-1,1).
    const T p1 = d[1] - d[3], p2 = d[2] - d[4], p5 = d[5] - d[6];
    const T p7 = d[7] - d[13], p8 = d[8] - d[14], p9 = d[9] - d[11], p10 = d[10] - d[12];
    const T drho = ((d[0] + (d[1] + d[3])) + ((d[2] + d[4]) + (d[5] + d[6])))
                 + (((d[7] + d[13]) + (d[8] + d[14])) + ((d[9] + d[11]) + (d[10] + d[12])));
    T mx = (p1 + p7) + (p10 - (p8 + p9));
    T my = (p2 + p7) + (p8 - (p9 + p10));
    T mz = (p5 + p7) + ((p8 + p9) + p10);
    const T rho = k.rho0 + drho;
    const T hdt = T(0.5) * k.dt;
    T inv;
    if constexpr (sizeof(T) == 4) inv = __builtin_amdgcn_rcpf(rho);   // v_rcp_f32, 1 ulp
    else inv = T(1.0) / rho;
    mx = (mx + hdt * F[0]) * inv;
    my = (my + hdt * F[1]) * inv;
    mz = (mz + hdt * F[2]) * inv;
    drho_out = drho;
    u_out[0] = mx;
    u_out[1] = my;
    u_out[2] = mz;
    const T uF = mx * F[0] + my * F[1] + mz * F[2];
    const T u2 = mx * mx + my * my + mz * mz;
    const T cf = k.dt * k.omega_f;
    // Direction-independent part of S / w, and the factors of the two weight classes.
    const T s0 = k.omega_p * (drho - T(1.5) * rho * u2) - T(3.0) * cf * uF;
    const T r45 = T(4.5) * k.omega_p * rho, c9 = T(9.0) * cf;
    const T a1 = T(3.0) * k.omega_p * rho, a2 = T(3.0) * cf;
    d[0] = k.omega * d[0] + T(wgt(0)) * s0;
    auto pair = [&](int i, int o, T ue, T eF) {
        const T w = T(wgt(i));
        const T S = w * (s0 + ue * (r45 * ue + c9 * eF));
        const T A = w * (a1 * ue + a2 * eF);
        d[i] = k.omega * d[i] + (S + A);
        d[o] = k.omega * d[o] + (S - A);
    };
    pair(1, 3, mx, F[0]);
    pair(2, 4, my, F[1]);
    pair(5, 6, mz, F[2]);
    pair(7, 13, mx + my + mz, F[0] + F[1] + F[2]);
    pair(8, 14, my + mz - mx, F[1] + F[2] - F[0]);
    pair(9, 11, mz - mx - my, F[2] - F[0] - F[1]);
    pair(10, 12, mx - my + mz, F[0] - F[1] + F[2]);
}

}  // namespace lbm
