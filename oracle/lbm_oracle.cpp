// lbm_oracle.cpp — FP64 CPU oracle for the lbman3d solver step.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP product is
// compared against; it is loaded only by tests/, __graft_entry__.smoke() and
// bench.py's CPU-baseline leg, never by the product library.
//
// It restates, with runtime dimensions, constants and boundary tags, the
// reference's CPU algorithm for one time step, keeping the reference's
// structure so its FP64 results can match bit for bit:
//   * populations in the host layout cell*15 + i (src/local_grid.h:66),
//   * push streaming with per-source-node wall handling (src/lbm_solver.tpp:89-224),
//   * the two Q-tensor loops (src/qtensor_solver.tpp:50-217),
//   * ActiveNematicSim::Step order (src/active_nematic.h:84-108).
// Every arithmetic expression keeps the reference's evaluation order, because
// without FMA contraction (-ffp-contract=off, as the reference's own CPU build
// compiles, CMakeLists.txt:50-62) that order fixes the result bits.
//
// Parity pins (see tests/): every function here is checked bit for bit
// against the reference's own CPU solver compiled from its sources, through
// the committed fixtures tests/golden/*.npz and a live run of oracle/_ref
// where the reference is present (tests/test_oracle_golden.py), and against
// the closed forms the reference's integration tests use (Poiseuille profile,
// equilibrium scalar order, conservation; tests/test_invariants.py), and its
// pointwise entry points against the closed forms of the reference's unit
// tests (lattice tables, equilibrium and Guo moments, BGK fixed point,
// polynomial stencils; tests/test_oracle_kat.py).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../include/lbman3d.h"

namespace {

// D3Q15 velocity set and symmetry maps (src/lattice_stencil.h:42-93).
constexpr int kQ = 15;
constexpr int EX[kQ] = {0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int EY[kQ] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int EZ[kQ] = {0, 0, 0, 0, 0, 1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
constexpr double W[kQ] = {2.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                          1.0 / 72, 1.0 / 72, 1.0 / 72, 1.0 / 72, 1.0 / 72, 1.0 / 72,
                          1.0 / 72, 1.0 / 72};
constexpr int OPP[kQ] = {0, 3, 4, 1, 2, 6, 5, 13, 14, 11, 12, 9, 10, 7, 8};
constexpr int SPEC[3][kQ] = {
    {0, 3, 2, 1, 4, 5, 6, 8, 7, 10, 9, 12, 11, 14, 13},   // x-walls
    {0, 1, 4, 3, 2, 5, 6, 10, 9, 8, 7, 14, 13, 12, 11},   // y-walls
    {0, 1, 2, 3, 4, 6, 5, 11, 12, 13, 14, 7, 8, 9, 10}};  // z-walls

struct Consts {
    double dt, rho0, tauf, L, A, B, C, lambda, gamma, alpha, mu;
    double omega, omega_prime, omega_forcing;
};

Consts make_consts(const lbm_params& p) {
    Consts c{};
    c.dt = p.dt; c.rho0 = p.density; c.tauf = p.tauf; c.L = p.L; c.A = p.A; c.B = p.B;
    c.C = p.C; c.lambda = p.lambda; c.gamma = p.gamma; c.alpha = p.alpha; c.mu = p.mu;
    // src/params.h:52-54
    c.omega = 1.0 - c.dt / c.tauf;
    c.omega_prime = c.dt / c.tauf;
    c.omega_forcing = 1.0 - c.dt / 2.0 / c.tauf;
    return c;
}

// Strong-anchoring target Q for one face (src/boundary_handler.h:232-251).
void anchored_q(const lbm_wall& w, double out[5]) {
    if (w.qbc != LBM_Q_ANCHORING) { for (int k = 0; k < 5; ++k) out[k] = 0.0; return; }
    const double phi = w.anchor_phi, th = w.anchor_theta, S = w.anchor_s;
    const double sphi2 = 0.5 * (1.0 - std::cos(2.0 * phi));
    const double cphi2 = 1.0 - sphi2;
    const double sphicphi = 0.5 * std::sin(2.0 * phi);
    const double sth2 = 0.5 * (1.0 - std::cos(2.0 * th));
    const double sthcth = 0.5 * std::sin(2.0 * th);
    out[0] = S * (cphi2 * sth2 - 1.0 / 3.0);
    out[1] = S * (sphicphi * sth2);
    out[2] = S * (std::cos(phi) * sthcth);
    out[3] = S * (sphi2 * sth2 - 1.0 / 3.0);
    out[4] = S * (std::sin(phi) * sthcth);
}

struct Sim {
    lbm_config cfg{};
    Consts k{};
    int n[3]{};
    long V = 0;
    double anchor[6][5]{};
    std::vector<double> f, f_new;            // AoS cell*15+i
    std::vector<double> rho, u[3], F[3];
    std::vector<double> q[5], q_new[5], sig[5], tau[3];
    std::vector<double> u_prev[3];           // Log's "past" velocity (src/analysis_fields.h:81)
    int threads = 1;

    long cell(int x, int y, int z) const { return ((long)z * n[1] + y) * n[0] + x; }
    int qlo(int a) const { return cfg.walls[2 * a].qbc; }
    int qhi(int a) const { return cfg.walls[2 * a + 1].qbc; }
    int ulo(int a) const { return cfg.walls[2 * a].ubc; }
    int uhi(int a) const { return cfg.walls[2 * a + 1].ubc; }
};

// ---------------------------------------------------------------------------
// Index helpers
// ---------------------------------------------------------------------------

// Fetch index for a gradient/Laplacian neighbour: periodic side wraps, any
// other side clamps (SafeFetchAxisOffset, src/boundary_handler.h:76-89).
// The same rule keyed on the Q tag is the Neumann clamp of the stress
// divergence and free energy (QAxisOffset, src/offsets.h:190-203).
inline int fetch_index(int i, int s, int n, bool lo_periodic, bool hi_periodic) {
    const int t = i + s;
    if (t < 0) return lo_periodic ? ((t % n) + n) % n : 0;
    if (t >= n) return hi_periodic ? ((t % n) + n) % n : n - 1;
    return t;
}

struct Pair { double m, p; };

// Ghost-substituted neighbour pair for Q component c along axis a
// (QAxisGhostPair/QGhost, src/boundary_handler.h:257-282). Only the low
// face's tag decides the periodic short-cut, as in the reference.
inline Pair q_pair(const Sim& s, int a, int comp, int i, double qm, double q0, double qp) {
    if (s.qlo(a) == LBM_Q_PERIODIC) return {qm, qp};
    Pair r{qm, qp};
    if (i == 0) r.m = (s.qlo(a) == LBM_Q_ANCHORING) ? 2.0 * s.anchor[2 * a][comp] - q0 : q0;
    if (i == s.n[a] - 1) r.p = (s.qhi(a) == LBM_Q_ANCHORING) ? 2.0 * s.anchor[2 * a + 1][comp] - q0 : q0;
    return r;
}

// Velocity ghost (VelocityGhost / VelocityAxisGhostPair, src/boundary_handler.h:121-146).
// comp: velocity component, a: differentiation axis.
inline double u_ghost(const lbm_wall& w, int comp, bool is_normal, double vb) {
    const double wv = (w.ubc == LBM_U_MOVING) ? w.wall_u[comp] : 0.0;
    if (w.ubc == LBM_U_SPECULAR) return is_normal ? (2.0 * wv - vb) : vb;
    return 2.0 * wv - vb;
}
inline Pair u_pair(const Sim& s, int a, int comp, int i, double vm, double v0, double vp) {
    if (s.ulo(a) == LBM_U_PERIODIC) return {vm, vp};
    Pair r{vm, vp};
    const bool normal = (comp == a);
    if (i == 0) r.m = u_ghost(s.cfg.walls[2 * a], comp, normal, v0);
    if (i == s.n[a] - 1) r.p = u_ghost(s.cfg.walls[2 * a + 1], comp, normal, v0);
    return r;
}

// dx, dy, dz, lap, d2x, d2y, d2z (QDerivs, src/qtensor_types.h:201).
struct Derivs { double d[3], lap, d2[3]; };

// Gradient and 7-point Laplacian of one Q component (QGradientAndLaplacian,
// src/boundary_handler.h:287-320).
Derivs q_derivs(const Sim& s, const double* qc, int comp, int x, int y, int z) {
    const int c[3] = {x, y, z};
    int im[3], ip[3];
    for (int a = 0; a < 3; ++a) {
        im[a] = fetch_index(c[a], -1, s.n[a], s.qlo(a) == LBM_Q_PERIODIC, s.qhi(a) == LBM_Q_PERIODIC);
        ip[a] = fetch_index(c[a], +1, s.n[a], s.qlo(a) == LBM_Q_PERIODIC, s.qhi(a) == LBM_Q_PERIODIC);
    }
    const double q0 = qc[s.cell(x, y, z)];
    const Pair px = q_pair(s, 0, comp, x, qc[s.cell(im[0], y, z)], q0, qc[s.cell(ip[0], y, z)]);
    const Pair py = q_pair(s, 1, comp, y, qc[s.cell(x, im[1], z)], q0, qc[s.cell(x, ip[1], z)]);
    const Pair pz = q_pair(s, 2, comp, z, qc[s.cell(x, y, im[2])], q0, qc[s.cell(x, y, ip[2])]);
    Derivs r;
    r.d[0] = (px.p - px.m) / 2.0;
    r.d[1] = (py.p - py.m) / 2.0;
    r.d[2] = (pz.p - pz.m) / 2.0;
    r.lap = (px.m + px.p + py.m + py.p + pz.m + pz.p) - 6.0 * q0;
    r.d2[0] = (px.m + px.p) - 2.0 * q0;
    r.d2[1] = (py.m + py.p) - 2.0 * q0;
    r.d2[2] = (pz.m + pz.p) - 2.0 * q0;
    return r;
}

// Velocity gradient g[A][B] = d u_A / d x_B, uz_z not formed
// (VelocityGradientTensor, src/boundary_handler.h:175-209).
void velocity_gradient(const Sim& s, int x, int y, int z, double g[3][3]) {
    const int c[3] = {x, y, z};
    int im[3], ip[3];
    for (int a = 0; a < 3; ++a) {
        im[a] = fetch_index(c[a], -1, s.n[a], s.ulo(a) == LBM_U_PERIODIC, s.uhi(a) == LBM_U_PERIODIC);
        ip[a] = fetch_index(c[a], +1, s.n[a], s.ulo(a) == LBM_U_PERIODIC, s.uhi(a) == LBM_U_PERIODIC);
    }
    const long c0 = s.cell(x, y, z);
    for (int A = 0; A < 3; ++A) {
        const double* v = s.u[A].data();
        const double v0 = v[c0];
        for (int B = 0; B < 3; ++B) {
            if (A == 2 && B == 2) { g[2][2] = 0.0; continue; }
            long cm, cp;
            if (B == 0) { cm = s.cell(im[0], y, z); cp = s.cell(ip[0], y, z); }
            else if (B == 1) { cm = s.cell(x, im[1], z); cp = s.cell(x, ip[1], z); }
            else { cm = s.cell(x, y, im[2]); cp = s.cell(x, y, ip[2]); }
            const Pair pr = u_pair(s, B, A, c[B], v[cm], v0, v[cp]);
            g[A][B] = (pr.p - pr.m) / 2.0;
        }
    }
}

// One axis of u.grad(Q) (AdvectiveAxisTerm, src/model.h:60-68).
inline double adv_term(int scheme, double u, double dq, double d2q) {
    if (scheme == LBM_ADV_UPWIND) {
        const double au = u < 0.0 ? -u : u;
        return u * dq - 0.5 * au * d2q;
    }
    return u * dq;
}
inline double adv(int scheme, const double u[3], const Derivs& d) {
    return adv_term(scheme, u[0], d.d[0], d.d2[0])
         + adv_term(scheme, u[1], d.d[1], d.d2[1])
         + adv_term(scheme, u[2], d.d[2], d.d2[2]);
}

// Pointwise Beris-Edwards update plus stresses and Ericksen force
// (PointwiseStepAndSetupBodyForce, src/model.h:102-342). Component order of
// the 5-vectors is xx, xy, xz, yy, yz.
void pointwise_q(const Consts& k, int scheme, const double Q[5], const double u[3],
                 const Derivs dq[5], const double g[3][3],
                 double qn[5], double sg[5], double ta[3], double fe[3]) {
    const double Qxx = Q[0], Qxy = Q[1], Qxz = Q[2], Qyy = Q[3], Qyz = Q[4];
    const double TrQ2 = 2.0 * (Qxx * Qxx + Qyy * Qyy + Qxx * Qyy + Qxy * Qxy + Qxz * Qxz + Qyz * Qyz);
    const double third = 1.0 / 3.0;
    const double Q2[5] = {
        Qxx * Qxx + Qxy * Qxy + Qxz * Qxz - third * TrQ2,
        Qxx * Qxy + Qxy * Qyy + Qxz * Qyz,
        Qxy * Qyz - Qxz * Qyy,
        Qxy * Qxy + Qyy * Qyy + Qyz * Qyz - third * TrQ2,
        Qxy * Qxz - Qyz * Qxx};

    const double Exx = g[0][0], Eyy = g[1][1];
    const double Exy = 0.5 * (g[0][1] + g[1][0]);
    const double Exz = 0.5 * (g[0][2] + g[2][0]);
    const double Eyz = 0.5 * (g[1][2] + g[2][1]);
    const double Wxy = 0.5 * (g[0][1] - g[1][0]);
    const double Wxz = 0.5 * (g[0][2] - g[2][0]);
    const double Wyz = 0.5 * (g[1][2] - g[2][1]);
    const double E[5] = {Exx, Exy, Exz, Eyy, Eyz};

    double advc[5];
    for (int c = 0; c < 5; ++c) advc[c] = -adv(scheme, u, dq[c]);

    // Co-rotation (Omega Q - Q Omega).
    const double cor[5] = {
        2.0 * (Qxy * Wxy + Qxz * Wxz),
        -Qxx * Wxy + Qxz * Wyz + Qyy * Wxy + Qyz * Wxz,
        -2.0 * Qxx * Wxz - Qxy * Wyz - Qyy * Wxz + Qyz * Wxy,
        2.0 * (-Qxy * Wxy + Qyz * Wyz),
        -Qxx * Wyz - Qxy * Wxz - Qxz * Wxy - 2.0 * Qyy * Wyz};

    const double two3 = 2.0 / 3.0;
    const double trQE = (2.0 * Exx + Eyy) * Qxx + (2.0 * Exy) * Qxy + (2.0 * Exz) * Qxz
                      + (Exx + 2.0 * Eyy) * Qyy + 2.0 * Eyz * Qyz;
    const double aln2[5] = {
        2.0 * (Exx * Qxx + Exy * Qxy + Exz * Qxz) - two3 * trQE,
        Exy * Qxx + Eyy * Qxy + Eyz * Qxz + Qxy * Exx + Qyy * Exy + Qyz * Exz,
        Exz * Qxx + Eyz * Qxy + (-Exx - Eyy) * Qxz + Qxz * Exx + Qyz * Exy + (-Qxx - Qyy) * Exz,
        2.0 * (Exy * Qxy + Eyy * Qyy + Eyz * Qyz) - two3 * trQE,
        Exz * Qxy + Eyz * Qyy + (-Exx - Eyy) * Qyz + Qxz * Exy + Qyz * Eyy + (-Qxx - Qyy) * Eyz};
    double aln3[5];
    for (int c = 0; c < 5; ++c) aln3[c] = -2.0 * trQE * Q[c];

    // Molecular field.
    const double ld = k.A + k.C * TrQ2;
    double H[5];
    for (int c = 0; c < 5; ++c) H[c] = k.L * dq[c].lap - ld * Q[c] - k.B * Q2[c];
    const double Hxx = H[0], Hxy = H[1], Hxz = H[2], Hyy = H[3], Hyz = H[4];

    // Ericksen force -H : d_a Q over all nine index pairs.
    for (int a = 0; a < 3; ++a) {
        fe[a] = -2.0 * (Hxx * dq[0].d[a] + Hxy * dq[1].d[a] + Hxz * dq[2].d[a] + Hyy * dq[3].d[a]
                        + Hyz * dq[4].d[a])
              - Hxx * dq[3].d[a] - Hyy * dq[0].d[a];
    }

    const double trQH = (2.0 * Hxx + Hyy) * Qxx + (2.0 * Hxy) * Qxy + (2.0 * Hxz) * Qxz
                      + (Hxx + 2.0 * Hyy) * Qyy + 2.0 * Hyz * Qyz;
    const double QH[5] = {
        2.0 * (Hxx * Qxx + Hxy * Qxy + Hxz * Qxz) - two3 * trQH,
        Hxy * Qxx + Hyy * Qxy + Hyz * Qxz + Qxy * Hxx + Qyy * Hxy + Qyz * Hxz,
        Hxz * Qxx + Hyz * Qxy + (-Hxx - Hyy) * Qxz + Qxz * Hxx + Qyz * Hxy + (-Qxx - Qyy) * Hxz,
        2.0 * (Hxy * Qxy + Hyy * Qyy + Hyz * Qyz) - two3 * trQH,
        Hxz * Qxy + Hyz * Qyy + (-Hxx - Hyy) * Qyz + Qxz * Hxy + Qyz * Hyy + (-Qxx - Qyy) * Hyz};
    ta[0] = (Hxy * Qxx + Hyy * Qxy + Hyz * Qxz) - (Qxy * Hxx + Qyy * Hxy + Qyz * Hxz);
    ta[1] = (Hxz * Qxx + Hyz * Qxy + (-Hxx - Hyy) * Qxz) - (Qxz * Hxx + Qyz * Hxy + (-Qxx - Qyy) * Hxz);
    ta[2] = (Hxz * Qxy + Hyz * Qyy + (-Hxx - Hyy) * Qyz) - (Qxz * Hxy + Qyz * Hyy + (-Qxx - Qyy) * Hyz);

    for (int c = 0; c < 5; ++c) {
        sg[c] = -two3 * k.lambda * H[c] - k.lambda * QH[c] + 2.0 * k.lambda * trQH * Q[c];
        qn[c] = Q[c] + k.dt * (advc[c] + cor[c] + k.lambda * (two3 * E[c] + aln2[c] + aln3[c])
                               + k.gamma * H[c]);
    }
}

// Divergence of Sigma + Tau under the row convention with the Neumann clamp
// keyed on the Q tags (PassiveStressDivergence, src/boundary_handler.h:335-392).
void stress_divergence(const Sim& s, int x, int y, int z, double out[3]) {
    const int c[3] = {x, y, z};
    int im[3], ip[3];
    for (int a = 0; a < 3; ++a) {
        im[a] = fetch_index(c[a], -1, s.n[a], s.qlo(a) == LBM_Q_PERIODIC, s.qhi(a) == LBM_Q_PERIODIC);
        ip[a] = fetch_index(c[a], +1, s.n[a], s.qlo(a) == LBM_Q_PERIODIC, s.qhi(a) == LBM_Q_PERIODIC);
    }
    const long xm = s.cell(im[0], y, z), xp = s.cell(ip[0], y, z);
    const long ym = s.cell(x, im[1], z), yp = s.cell(x, ip[1], z);
    const long zm = s.cell(x, y, im[2]), zp = s.cell(x, y, ip[2]);
    const std::vector<double>* S = s.sig;
    const std::vector<double>* T = s.tau;
    auto D = [](const std::vector<double>& v, long p, long m) { return (v[p] - v[m]) / 2.0; };
    const double dxSxx = D(S[0], xp, xm), dxSxy = D(S[1], xp, xm), dxSxz = D(S[2], xp, xm);
    const double dxTxy = D(T[0], xp, xm), dxTxz = D(T[1], xp, xm);
    const double dySxy = D(S[1], yp, ym), dySyy = D(S[3], yp, ym), dySyz = D(S[4], yp, ym);
    const double dyTxy = D(T[0], yp, ym), dyTyz = D(T[2], yp, ym);
    const double dzSxz = D(S[2], zp, zm), dzSyz = D(S[4], zp, zm);
    const double dzTxz = D(T[1], zp, zm), dzTyz = D(T[2], zp, zm);
    const double dzSzz = -(D(S[0], zp, zm) + D(S[3], zp, zm));
    out[0] = dxSxx + (dySxy + dyTxy) + (dzSxz + dzTxz);
    out[1] = (dxSxy - dxTxy) + dySyy + (dzSyz + dzTyz);
    out[2] = (dxSxz - dxTxz) + (dySyz - dyTyz) + dzSzz;
}

// ---------------------------------------------------------------------------
// Loops
// ---------------------------------------------------------------------------

// Loop 1 (QTensorSolver::StepAndSetupBodyForce, src/qtensor_solver.tpp:50-147),
// followed by the Q/Q_new swap.
void q_loop1(Sim& s) {
    const int scheme = s.cfg.advection;
    #pragma omp parallel for schedule(static) num_threads(s.threads)
    for (int z = 0; z < s.n[2]; ++z)
        for (int y = 0; y < s.n[1]; ++y)
            for (int x = 0; x < s.n[0]; ++x) {
                const long c = s.cell(x, y, z);
                double Q[5], u[3], g[3][3], qn[5], sg[5], ta[3], fe[3];
                Derivs dq[5];
                for (int k = 0; k < 5; ++k) { Q[k] = s.q[k][c]; }
                for (int a = 0; a < 3; ++a) u[a] = s.u[a][c];
                for (int k = 0; k < 5; ++k) dq[k] = q_derivs(s, s.q[k].data(), k, x, y, z);
                velocity_gradient(s, x, y, z, g);
                pointwise_q(s.k, scheme, Q, u, dq, g, qn, sg, ta, fe);
                for (int a = 0; a < 3; ++a) s.F[a][c] = fe[a];
                for (int k = 0; k < 5; ++k) { s.sig[k][c] = sg[k]; s.q_new[k][c] = qn[k]; }
                for (int k = 0; k < 3; ++k) s.tau[k][c] = ta[k];
            }
    for (int k = 0; k < 5; ++k) std::swap(s.q[k], s.q_new[k]);
}

// Loop 2 (QTensorSolver::SetActiveStressAndComputeBodyForce,
// src/qtensor_solver.tpp:154-217, with PointwiseSetActiveStressAndComputeBodyForce,
// src/model.h:344-361). Reads the already-swapped (new) Q.
void q_loop2(Sim& s) {
    const double alpha = s.k.alpha, mu = s.k.mu;
    #pragma omp parallel for schedule(static) num_threads(s.threads)
    for (int z = 0; z < s.n[2]; ++z)
        for (int y = 0; y < s.n[1]; ++y)
            for (int x = 0; x < s.n[0]; ++x) {
                const long c = s.cell(x, y, z);
                Derivs dq[5];
                for (int k = 0; k < 5; ++k) dq[k] = q_derivs(s, s.q[k].data(), k, x, y, z);
                double pd[3];
                stress_divergence(s, x, y, z, pd);
                const double ux = s.u[0][c], uy = s.u[1][c], uz = s.u[2][c];
                const double fx = -alpha * (dq[0].d[0] + dq[1].d[1] + dq[2].d[2]) + pd[0] - mu * ux;
                const double fy = -alpha * (dq[1].d[0] + dq[3].d[1] + dq[4].d[2]) + pd[1] - mu * uy;
                const double fz = -alpha * (dq[2].d[0] + dq[4].d[1] - dq[0].d[2] - dq[3].d[2]) + pd[2]
                                - mu * uz;
                s.F[0][c] += fx;
                s.F[1][c] += fy;
                s.F[2][c] += fz;
            }
}

// Feq and the Guo forcing term for one direction (src/physics_helpers.h:69-90).
inline void feq_forcing(const Consts& k, double rho, const double u[3], double u2, double uF,
                        const double F[3], int i, double* feq, double* forcing) {
    const double e0 = EX[i], e1 = EY[i], e2 = EZ[i];
    const double ue = u[0] * e0 + u[1] * e1 + u[2] * e2;
    const double eF = e0 * F[0] + e1 * F[1] + e2 * F[2];
    *feq = W[i] * rho * (1.0 + 3.0 * ue + 4.5 * ue * ue - 1.5 * u2);
    *forcing = k.omega_forcing * W[i] * (3.0 * eF - 3.0 * uF + 9.0 * ue * eF);
}

// LatticeBoltzmannStep (src/lbm_solver.tpp:89-224): moments, BGK with Guo
// forcing, push streaming, wall handling at the source node, buffer swap.
void lbm_loop(Sim& s) {
    const Consts& k = s.k;
    const int nx = s.n[0], ny = s.n[1], nz = s.n[2];
    #pragma omp parallel for schedule(static) num_threads(s.threads)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const long c = s.cell(x, y, z);
                const double F[3] = {s.F[0][c], s.F[1][c], s.F[2][c]};
                // ComputeMoments (src/physics_helpers.h:92-121)
                double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
                const double* fc = &s.f[c * kQ];
                for (int i = 0; i < kQ; ++i) {
                    const double fi = fc[i];
                    rho += fi;
                    mx += EX[i] * fi;
                    my += EY[i] * fi;
                    mz += EZ[i] * fi;
                }
                mx += 0.5 * F[0] * k.dt;
                my += 0.5 * F[1] * k.dt;
                mz += 0.5 * F[2] * k.dt;
                mx /= rho;
                my /= rho;
                mz /= rho;
                const double u[3] = {mx, my, mz};
                s.rho[c] = rho;
                s.u[0][c] = mx;
                s.u[1][c] = my;
                s.u[2][c] = mz;
                const double uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
                const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                const int pos[3] = {x, y, z};
                for (int i = 0; i < kQ; ++i) {
                    double feq, forcing;
                    feq_forcing(k, rho, u, u2, uF, F, i, &feq, &forcing);
                    const double fs = k.omega * fc[i] + k.omega_prime * feq + k.dt * forcing;
                    const int e[3] = {EX[i], EY[i], EZ[i]};
                    int d[3];
                    int cross[3] = {0, 0, 0};   // -1: low face, +1: high face
                    for (int a = 0; a < 3; ++a) {
                        const int t = pos[a] + e[a];
                        if (t < 0) {
                            if (s.ulo(a) == LBM_U_PERIODIC) d[a] = ((t % s.n[a]) + s.n[a]) % s.n[a];
                            else { d[a] = t; cross[a] = -1; }
                        } else if (t >= s.n[a]) {
                            if (s.uhi(a) == LBM_U_PERIODIC) d[a] = ((t % s.n[a]) + s.n[a]) % s.n[a];
                            else { d[a] = t; cross[a] = +1; }
                        } else {
                            d[a] = t;
                        }
                    }
                    if (!cross[0] && !cross[1] && !cross[2]) {
                        s.f_new[s.cell(d[0], d[1], d[2]) * kQ + i] = fs;
                        continue;
                    }
                    // HandleBoundaryPoint (src/boundary_handler.h:403-426), x then y then z.
                    for (int a = 0; a < 3; ++a) {
                        if (!cross[a]) continue;
                        const lbm_wall& w = s.cfg.walls[2 * a + (cross[a] > 0 ? 1 : 0)];
                        if (w.ubc == LBM_U_PERIODIC) continue;
                        if (w.ubc == LBM_U_SPECULAR) {
                            s.f_new[c * kQ + SPEC[a][i]] = fs;
                        } else {
                            const double Uw[3] = {w.ubc == LBM_U_MOVING ? w.wall_u[0] : 0.0,
                                                  w.ubc == LBM_U_MOVING ? w.wall_u[1] : 0.0,
                                                  w.ubc == LBM_U_MOVING ? w.wall_u[2] : 0.0};
                            const int m = OPP[i];
                            const double em[3] = {(double)EX[m], (double)EY[m], (double)EZ[m]};
                            s.f_new[c * kQ + m] =
                                fs + 6.0 * rho * W[m] * (em[0] * Uw[0] + em[1] * Uw[1] + em[2] * Uw[2]);
                        }
                    }
                }
            }
    std::swap(s.f, s.f_new);
}

// Seven-point Laplacian with the Neumann clamp, neighbour order xp, xm, yp,
// ym, zp, zm (TotalNematicFreeEnergy, src/analysis_fields.h:39-72).
double free_energy(const Sim& s, int threads) {
    const Consts& k = s.k;
    double total = 0.0;
    #pragma omp parallel for schedule(static) reduction(+ : total) num_threads(threads)
    for (int z = 0; z < s.n[2]; ++z)
        for (int y = 0; y < s.n[1]; ++y)
            for (int x = 0; x < s.n[0]; ++x) {
                const int c3[3] = {x, y, z};
                int im[3], ip[3];
                for (int a = 0; a < 3; ++a) {
                    im[a] = fetch_index(c3[a], -1, s.n[a], s.qlo(a) == LBM_Q_PERIODIC, s.qhi(a) == LBM_Q_PERIODIC);
                    ip[a] = fetch_index(c3[a], +1, s.n[a], s.qlo(a) == LBM_Q_PERIODIC, s.qhi(a) == LBM_Q_PERIODIC);
                }
                const long c = s.cell(x, y, z);
                double Q[5], lap[5];
                for (int m = 0; m < 5; ++m) {
                    const double* v = s.q[m].data();
                    Q[m] = v[c];
                    lap[m] = v[s.cell(ip[0], y, z)] + v[s.cell(im[0], y, z)] + v[s.cell(x, ip[1], z)]
                           + v[s.cell(x, im[1], z)] + v[s.cell(x, y, ip[2])] + v[s.cell(x, y, im[2])]
                           - 6.0 * Q[m];
                }
                const double Qxx = Q[0], Qxy = Q[1], Qxz = Q[2], Qyy = Q[3], Qyz = Q[4];
                // NematicFreeEnergyDensity (src/analysis_fields.cc:38-57)
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
                total += 0.5 * k.A * TrQ2 + (k.B / 3.0) * TrQ3 + 0.25 * k.C * TrQ2 * TrQ2 - 0.5 * k.L * QlapQ;
            }
    return total;
}

}  // namespace

// ===========================================================================
// C interface used by the tests and the CPU baseline (ctypes)
// ===========================================================================
extern "C" {

void* orc_create(const lbm_config* cfg, int threads) {
    Sim* s = new Sim();
    s->cfg = *cfg;
    s->k = make_consts(cfg->p);
    s->n[0] = cfg->nx; s->n[1] = cfg->ny; s->n[2] = cfg->nz;
    s->V = (long)cfg->nx * cfg->ny * cfg->nz;
    s->threads = threads > 0 ? threads : 1;
    for (int w = 0; w < 6; ++w) anchored_q(cfg->walls[w], s->anchor[w]);
    const long V = s->V;
    // FluidFields / QTensorFields constructors (src/fluid_fields.cc:6-17,
    // src/qtensor_fields.cc:5-26): f, f_new, F, u, Q, Sigma, Tau zero; rho = kDensity.
    s->f.assign(V * kQ, 0.0);
    s->f_new.assign(V * kQ, 0.0);
    s->rho.assign(V, cfg->p.density);
    for (int a = 0; a < 3; ++a) { s->u[a].assign(V, 0.0); s->F[a].assign(V, 0.0); s->u_prev[a].assign(V, 0.0); }
    for (int k = 0; k < 5; ++k) { s->q[k].assign(V, 0.0); s->q_new[k].assign(V, 0.0); s->sig[k].assign(V, 0.0); }
    for (int k = 0; k < 3; ++k) s->tau[k].assign(V, 0.0);
    return s;
}

void orc_destroy(void* h) { delete static_cast<Sim*>(h); }

void orc_set_threads(void* h, int threads) { static_cast<Sim*>(h)->threads = threads > 0 ? threads : 1; }

// Overwrite macroscopic fields (NULL leaves a field unchanged).
void orc_set_fields(void* h, const double* rho, const double* ux, const double* uy,
                    const double* uz, const double* q) {
    Sim& s = *static_cast<Sim*>(h);
    const long V = s.V;
    if (rho) std::memcpy(s.rho.data(), rho, V * sizeof(double));
    if (ux) std::memcpy(s.u[0].data(), ux, V * sizeof(double));
    if (uy) std::memcpy(s.u[1].data(), uy, V * sizeof(double));
    if (uz) std::memcpy(s.u[2].data(), uz, V * sizeof(double));
    if (q) for (int k = 0; k < 5; ++k) std::memcpy(s.q[k].data(), q + k * V, V * sizeof(double));
}

// LbmSolver::Initialize (src/lbm_solver.tpp:59-87): f = Feq(rho, u).
void orc_init_feq(void* h) {
    Sim& s = *static_cast<Sim*>(h);
    for (long c = 0; c < s.V; ++c) {
        const double rho = s.rho[c];
        const double u[3] = {s.u[0][c], s.u[1][c], s.u[2][c]};
        const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        for (int i = 0; i < kQ; ++i) {
            const double ue = u[0] * (double)EX[i] + u[1] * (double)EY[i] + u[2] * (double)EZ[i];
            s.f[c * kQ + i] = W[i] * rho * (1.0 + 3.0 * ue + 4.5 * ue * ue - 1.5 * u2);
        }
    }
}

void orc_set_force(void* h, const double* fx, const double* fy, const double* fz) {
    Sim& s = *static_cast<Sim*>(h);
    const double* src[3] = {fx, fy, fz};
    for (int a = 0; a < 3; ++a) if (src[a]) std::memcpy(s.F[a].data(), src[a], s.V * sizeof(double));
}

// Populations in the device layout i*V + cell.
void orc_set_populations(void* h, const double* f) {
    Sim& s = *static_cast<Sim*>(h);
    for (long c = 0; c < s.V; ++c)
        for (int i = 0; i < kQ; ++i) s.f[c * kQ + i] = f[i * s.V + c];
}
void orc_get_populations(void* h, double* f) {
    Sim& s = *static_cast<Sim*>(h);
    for (long c = 0; c < s.V; ++c)
        for (int i = 0; i < kQ; ++i) f[i * s.V + c] = s.f[c * kQ + i];
}

void orc_get_fields(void* h, double* rho, double* ux, double* uy, double* uz, double* q) {
    Sim& s = *static_cast<Sim*>(h);
    const long V = s.V;
    if (rho) std::memcpy(rho, s.rho.data(), V * sizeof(double));
    if (ux) std::memcpy(ux, s.u[0].data(), V * sizeof(double));
    if (uy) std::memcpy(uy, s.u[1].data(), V * sizeof(double));
    if (uz) std::memcpy(uz, s.u[2].data(), V * sizeof(double));
    if (q) for (int k = 0; k < 5; ++k) std::memcpy(q + k * V, s.q[k].data(), V * sizeof(double));
}
void orc_get_force(void* h, double* fx, double* fy, double* fz) {
    Sim& s = *static_cast<Sim*>(h);
    double* dst[3] = {fx, fy, fz};
    for (int a = 0; a < 3; ++a) if (dst[a]) std::memcpy(dst[a], s.F[a].data(), s.V * sizeof(double));
}
// Sigma (5 blocks) and Tau (3 blocks) from the last Q step.
void orc_get_stress(void* h, double* sigma, double* tau) {
    Sim& s = *static_cast<Sim*>(h);
    for (int k = 0; k < 5; ++k) std::memcpy(sigma + k * s.V, s.sig[k].data(), s.V * sizeof(double));
    for (int k = 0; k < 3; ++k) std::memcpy(tau + k * s.V, s.tau[k].data(), s.V * sizeof(double));
}

void orc_qtensor_step(void* h) {
    Sim& s = *static_cast<Sim*>(h);
    q_loop1(s);
    if (s.cfg.force_model == LBM_FORCE_FULL) q_loop2(s);
}
void orc_lbm_step(void* h) { lbm_loop(*static_cast<Sim*>(h)); }
void orc_step(void* h, int n) {
    for (int t = 0; t < n; ++t) { orc_qtensor_step(h); orc_lbm_step(h); }
}

double orc_free_energy(void* h) {
    Sim& s = *static_cast<Sim*>(h);
    return free_energy(s, s.threads);
}

// SimIO::Log reductions (src/sim_io.cc:78-109): out = {mass, px, py, pz, ke,
// du2, u2, umax}. Updates the stored previous velocity like the reference.
void orc_log_sums(void* h, double* out) {
    Sim& s = *static_cast<Sim*>(h);
    double mass = 0, px = 0, py = 0, pz = 0, ke = 0, e1 = 0, e2 = 0, umax2 = 0;
    for (long c = 0; c < s.V; ++c) {
        const double r = s.rho[c], ux = s.u[0][c], uy = s.u[1][c], uz = s.u[2][c];
        mass += r;
        px += r * ux; py += r * uy; pz += r * uz;
        const double u2 = ux * ux + uy * uy + uz * uz;
        ke += 0.5 * r * u2;
        if (u2 > umax2) umax2 = u2;
        const double dx = ux - s.u_prev[0][c], dy = uy - s.u_prev[1][c], dz = uz - s.u_prev[2][c];
        e1 += dx * dx + dy * dy + dz * dz;
        e2 += u2;
        s.u_prev[0][c] = ux; s.u_prev[1][c] = uy; s.u_prev[2][c] = uz;
    }
    out[0] = mass; out[1] = px; out[2] = py; out[3] = pz; out[4] = ke;
    out[5] = e1; out[6] = e2; out[7] = std::sqrt(umax2);
}

// QtensorToOrderDirector (src/analysis_fields.cc:59-93).
void orc_order_director(void* h, double* order, double* director) {
    Sim& s = *static_cast<Sim*>(h);
    for (long c = 0; c < s.V; ++c) {
        const double Qxx = s.q[0][c], Qxy = s.q[1][c], Qxz = s.q[2][c], Qyy = s.q[3][c], Qyz = s.q[4][c];
        const double p = Qxx * Qxx + Qxx * Qyy + Qyy * Qyy + Qxy * Qxy + Qxz * Qxz + Qyz * Qyz;
        const double qd = -(Qxx + Qyy) * (Qxx * Qyy - Qxy * Qxy) - Qyy * Qxz * Qxz - Qxx * Qyz * Qyz
                        + 2 * Qxy * Qxz * Qyz;
        const double r = 2.0 * std::sqrt(p / 3.0);
        const double S = r * std::cos(1.0 / 3.0 * std::acos(std::clamp(4 * qd / (r * r * r), -1.0, 1.0)));
        double nx = Qxz * (Qyy - S) - Qxy * Qyz;
        double ny = Qyz * (Qxx - S) - Qxy * Qxz;
        double nz = Qxy * Qxy - (Qxx - S) * (Qyy - S);
        const double inv = 1.0 / std::sqrt(nx * nx + ny * ny + nz * nz);
        nx *= inv; ny *= inv; nz *= inv;
        director[3 * c + 0] = nx; director[3 * c + 1] = ny; director[3 * c + 2] = nz;
        order[c] = S;
    }
}

// ---- pointwise entry points for the closed-form unit tests -----------------

// Feq + forcing + BGK for all 15 directions at one node.
void orc_pointwise_collide(const lbm_params* p, double rho, const double* u, const double* F,
                           const double* f, double* feq, double* forcing, double* fstar) {
    const Consts k = make_consts(*p);
    const double uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
    const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    for (int i = 0; i < kQ; ++i) {
        feq_forcing(k, rho, u, u2, uF, F, i, &feq[i], &forcing[i]);
        fstar[i] = k.omega * f[i] + k.omega_prime * feq[i] + k.dt * forcing[i];
    }
}

// Moments of one node's populations with the half-force correction.
void orc_moments(const lbm_params* p, const double* f, const double* F, double* rho_u) {
    double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    for (int i = 0; i < kQ; ++i) {
        rho += f[i]; mx += EX[i] * f[i]; my += EY[i] * f[i]; mz += EZ[i] * f[i];
    }
    mx += 0.5 * F[0] * p->dt; my += 0.5 * F[1] * p->dt; mz += 0.5 * F[2] * p->dt;
    rho_u[0] = rho; rho_u[1] = mx / rho; rho_u[2] = my / rho; rho_u[3] = mz / rho;
}

// Pointwise Q step given explicit derivatives: derivs holds 5 x
// (dx, dy, dz, lap, d2x, d2y, d2z); grad is the 3x3 velocity gradient
// row-major (uz_z ignored).
void orc_pointwise_q(const lbm_params* p, int scheme, const double* Q, const double* u,
                     const double* derivs, const double* grad,
                     double* qn, double* sg, double* ta, double* fe) {
    const Consts k = make_consts(*p);
    Derivs dq[5];
    for (int c = 0; c < 5; ++c) {
        const double* d = derivs + 7 * c;
        dq[c].d[0] = d[0]; dq[c].d[1] = d[1]; dq[c].d[2] = d[2]; dq[c].lap = d[3];
        dq[c].d2[0] = d[4]; dq[c].d2[1] = d[5]; dq[c].d2[2] = d[6];
    }
    double g[3][3];
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) g[a][b] = grad[3 * a + b];
    pointwise_q(k, scheme, Q, u, dq, g, qn, sg, ta, fe);
}

// Q stencil derivatives of the handle's current Q at one node, 5 x 7 values.
void orc_q_derivs(void* h, int x, int y, int z, double* out) {
    Sim& s = *static_cast<Sim*>(h);
    for (int c = 0; c < 5; ++c) {
        const Derivs d = q_derivs(s, s.q[c].data(), c, x, y, z);
        double* o = out + 7 * c;
        o[0] = d.d[0]; o[1] = d.d[1]; o[2] = d.d[2]; o[3] = d.lap; o[4] = d.d2[0]; o[5] = d.d2[1]; o[6] = d.d2[2];
    }
}
// Velocity gradient of the handle's current u at one node (row-major 3x3).
void orc_velocity_gradient(void* h, int x, int y, int z, double* out) {
    Sim& s = *static_cast<Sim*>(h);
    double g[3][3];
    velocity_gradient(s, x, y, z, g);
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) out[3 * a + b] = g[a][b];
}
// Stress divergence from explicitly supplied Sigma (5*V) and Tau (3*V).
void orc_stress_divergence(void* h, const double* sigma, const double* tau, int x, int y, int z, double* out) {
    Sim& s = *static_cast<Sim*>(h);
    for (int k = 0; k < 5; ++k) std::memcpy(s.sig[k].data(), sigma + k * s.V, s.V * sizeof(double));
    for (int k = 0; k < 3; ++k) std::memcpy(s.tau[k].data(), tau + k * s.V, s.V * sizeof(double));
    stress_divergence(s, x, y, z, out);
}
void orc_anchored_q(const lbm_wall* w, double* out) { anchored_q(*w, out); }

// Velocity set tables for table tests.
void orc_lattice(int* ex, int* ey, int* ez, double* w, int* opp, int* spec) {
    for (int i = 0; i < kQ; ++i) {
        ex[i] = EX[i]; ey[i] = EY[i]; ez[i] = EZ[i]; w[i] = W[i]; opp[i] = OPP[i];
        for (int a = 0; a < 3; ++a) spec[a * kQ + i] = SPEC[a][i];
    }
}

}  // extern "C"
