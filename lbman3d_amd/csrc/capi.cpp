// capi.cpp — host runtime behind the C ABI of include/lbman3d.h: handle
// lifetime, HBM allocation, the pull-slot table, host<->device transfers,
// kernel sequencing per step, diagnostics and timing.
//
// Replaces DeviceFields / DeviceSolver<BC> (src/device_fields.h:23-80,
// src/device_solver.h:9-18, src/cuda/device_fields.cu, src/cuda/device_solver.cu)
// with runtime boundary descriptors instead of one template instantiation per
// build.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbman3d.h"
#include "common.hpp"
#include "launch.hpp"
#include "shell_region.hpp"
#include "tile_map.hpp"
#include "runtime.hpp"
#include "solver_state.hpp"

using namespace lbm;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace lbm {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace lbm

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(LBM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) + " (" +   \
                                         __FILE__ + ":" + std::to_string(__LINE__) + ")");        \
    } while (0)

namespace {
// Runs a cleanup when the scope ends, so that the early returns of HIPCHK
// release what a function has acquired so far.
template <typename F>
struct ScopeExit {
    F f;
    ~ScopeExit() { f(); }
};
template <typename F>
ScopeExit<F> scope_exit(F f) {
    return ScopeExit<F>{f};
}
}  // namespace

// ---------------------------------------------------------------------------
// configuration helpers
// ---------------------------------------------------------------------------
namespace lbm {

void anchored_q_host(const lbm_wall& w, double out[5]) {
    // Strong-anchoring target (src/boundary_handler.h:232-251).
    if (w.qbc != LBM_Q_ANCHORING) {
        for (int k = 0; k < 5; ++k) out[k] = 0.0;
        return;
    }
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

int validate_config(const lbm_config& c) {
    if (c.nx < 1 || c.ny < 1 || c.nz < 1)
        return fail(LBM_ERR_INVALID, "grid dimensions must be >= 1");
    const double V = (double)c.nx * c.ny * c.nz;
    if (V * kQ > 4.0e12) return fail(LBM_ERR_INVALID, "grid too large");
    if (c.advection != LBM_ADV_CENTRED && c.advection != LBM_ADV_UPWIND)
        return fail(LBM_ERR_INVALID, "advection must be LBM_ADV_CENTRED or LBM_ADV_UPWIND");
    if (c.precision != LBM_FP64 && c.precision != LBM_FP32_SHIFTED)
        return fail(LBM_ERR_INVALID, "precision must be LBM_FP64 or LBM_FP32_SHIFTED");
    if (c.force_model != LBM_FORCE_FULL && c.force_model != LBM_FORCE_ERICKSEN_ONLY)
        return fail(LBM_ERR_INVALID, "force_model must be LBM_FORCE_FULL or LBM_FORCE_ERICKSEN_ONLY");
    for (int f = 0; f < 6; ++f) {
        if (c.walls[f].qbc < LBM_Q_PERIODIC || c.walls[f].qbc > LBM_Q_ANCHORING)
            return fail(LBM_ERR_INVALID, "invalid Q boundary tag on face " + std::to_string(f));
        if (c.walls[f].ubc < LBM_U_PERIODIC || c.walls[f].ubc > LBM_U_SPECULAR)
            return fail(LBM_ERR_INVALID, "invalid velocity boundary tag on face " + std::to_string(f));
    }
    if (!(c.p.tauf > 0.0) || !(c.p.dt > 0.0)) return fail(LBM_ERR_INVALID, "TAUF and DT must be positive");
    return LBM_OK;
}

// Build the pull-slot table by replaying, for every class of node position,
// the push writes of LatticeBoltzmannStep (src/lbm_solver.tpp:144-192): a
// slot is streamed from its upwind node, or written by the last wall
// reflection of the node itself (directions in order 0..14, axes x, y, z).
// A slot with both a streaming source and a reflection, or with neither,
// has an order-dependent or stale value in the reference and is rejected.
int build_slot_table(const lbm_config& c, std::vector<uint16_t>& table) {
    table.assign(kClasses * kQ, 0);
    const int n[3] = {c.nx, c.ny, c.nz};
    for (int cls = 0; cls < kClasses; ++cls) {
        int lo[3], hi[3];
        bool possible = true;
        for (int a = 0; a < 3; ++a) {
            const int ca = (cls >> (2 * a)) & 3;
            lo[a] = ca & 1;
            hi[a] = (ca >> 1) & 1;
            if (n[a] == 1 && ca != 3) possible = false;
            if (n[a] == 2 && (ca == 0 || ca == 3)) possible = false;
            if (n[a] >= 3 && ca == 3) possible = false;
        }
        if (!possible) continue;
        int writer[kQ];
        for (int j = 0; j < kQ; ++j) writer[j] = -1;
        for (int i = 0; i < kQ; ++i) {
            int cross[3] = {0, 0, 0};
            bool any = false;
            for (int a = 0; a < 3; ++a) {
                const int e = ce(i, a);
                if (e < 0 && lo[a] && c.walls[2 * a].ubc != LBM_U_PERIODIC) cross[a] = -1;
                if (e > 0 && hi[a] && c.walls[2 * a + 1].ubc != LBM_U_PERIODIC) cross[a] = +1;
                any = any || cross[a] != 0;
            }
            if (!any) continue;
            for (int a = 0; a < 3; ++a) {
                if (!cross[a]) continue;
                const int face = 2 * a + (cross[a] > 0 ? 1 : 0);
                const int ubc = c.walls[face].ubc;
                if (ubc == LBM_U_SPECULAR) writer[mirror(i, a)] = make_slot(SLOT_MIRROR, i, face);
                else writer[opposite(i)] = make_slot(SLOT_BOUNCE, i, face);
            }
        }
        for (int j = 0; j < kQ; ++j) {
            bool stream = true;
            for (int a = 0; a < 3; ++a) {
                const int e = ce(j, a);
                if (e > 0 && lo[a] && c.walls[2 * a + 1].ubc != LBM_U_PERIODIC) stream = false;
                if (e < 0 && hi[a] && c.walls[2 * a].ubc != LBM_U_PERIODIC) stream = false;
            }
            if (stream && writer[j] >= 0)
                return fail(LBM_ERR_UNSUPPORTED,
                            "boundary configuration writes population slot " + std::to_string(j) +
                                " both by streaming and by a wall (periodic on one side of an axis only); "
                                "the reference result is order dependent");
            if (!stream && writer[j] < 0)
                return fail(LBM_ERR_UNSUPPORTED,
                            "boundary configuration leaves population slot " + std::to_string(j) +
                                " unwritten at an edge where reflecting walls meet (the reference keeps a stale "
                                "value there)");
            table[cls * kQ + j] = stream ? make_slot(SLOT_STREAM, 0, 0) : (uint16_t)writer[j];
        }
    }
    return LBM_OK;
}

// Whether the pull-slot row of every class a node of this grid can have
// equals the row of the same x, y class with the z bits cleared.
bool slots_independent_of_z(const lbm_config& c, const std::vector<uint16_t>& table) {
    auto bits = [](int i, int n) { return (i == 0 ? 1 : 0) | (i == n - 1 ? 2 : 0); };
    for (int z : {0, c.nz - 1, std::min(1, c.nz - 1)})
        for (int y : {0, c.ny - 1, std::min(1, c.ny - 1)})
            for (int x : {0, c.nx - 1, std::min(1, c.nx - 1)}) {
                const int cls = bits(x, c.nx) | (bits(y, c.ny) << 2) | (bits(z, c.nz) << 4);
                for (int j = 0; j < kQ; ++j)
                    if (table[cls * kQ + j] != table[(cls & 15) * kQ + j]) return false;
            }
    return true;
}

template <typename T>
BCDesc<T> make_bc(const lbm_config& c) {
    BCDesc<T> b{};
    b.n[0] = c.nx;
    b.n[1] = c.ny;
    b.n[2] = c.nz;
    b.all_periodic_u = 1;
    // FP32 runs store shifted populations and rho - rho0; FP64 stores rho.
    b.rho_base = sizeof(T) == 4 ? (T)c.p.density : T(0);
    for (int f = 0; f < 6; ++f) {
        b.qkind[f] = c.walls[f].qbc;
        b.ukind[f] = c.walls[f].ubc;
        if (c.walls[f].ubc != LBM_U_PERIODIC) b.all_periodic_u = 0;
        double a[5];
        anchored_q_host(c.walls[f], a);
        for (int k = 0; k < 5; ++k) b.anchor[f][k] = (T)a[k];
        for (int k = 0; k < 3; ++k) b.wall_u[f][k] = (T)(c.walls[f].ubc == LBM_U_MOVING ? c.walls[f].wall_u[k] : 0.0);
    }
    return b;
}
template BCDesc<double> make_bc<double>(const lbm_config&);
template BCDesc<float> make_bc<float>(const lbm_config&);

template <typename T>
Phys<T> make_phys(const lbm_config& c) {
    const lbm_params& p = c.p;
    Phys<T> k{};
    // src/params.h:52-54
    const double omega = 1.0 - p.dt / p.tauf;
    const double omega_p = p.dt / p.tauf;
    const double omega_f = 1.0 - p.dt / 2.0 / p.tauf;
    k.dt = (T)p.dt;
    k.rho0 = (T)p.density;
    k.L = (T)p.L;
    k.A = (T)p.A;
    k.B = (T)p.B;
    k.C = (T)p.C;
    k.lambda = (T)p.lambda;
    k.gamma = (T)p.gamma;
    k.alpha = (T)p.alpha;
    k.mu = (T)p.mu;
    k.omega = (T)omega;
    k.omega_p = (T)omega_p;
    k.omega_f = (T)omega_f;
    k.advection = c.advection;
    k.force_full = c.force_model == LBM_FORCE_FULL ? 1 : 0;
    return k;
}
template Phys<double> make_phys<double>(const lbm_config&);
template Phys<float> make_phys<float>(const lbm_config&);

bool has_moving_wall(const lbm_config& c) {
    for (int f = 0; f < 6; ++f)
        if (c.walls[f].ubc == LBM_U_MOVING) return true;
    return false;
}

// Feq(rho, u) for one node in the reference's operand order
// (Feq, src/physics_helpers.h:69-71; LbmSolver::Initialize, src/lbm_solver.tpp:59-87).
void feq_host(double rho, double ux, double uy, double uz, double out[kQ]) {
    const double u2 = ux * ux + uy * uy + uz * uz;
    for (int i = 0; i < kQ; ++i) {
        const double ue = ux * (double)cx(i) + uy * (double)cy(i) + uz * (double)cz(i);
        out[i] = wgt(i) * rho * (1.0 + 3.0 * ue + 4.5 * ue * ue - 1.5 * u2);
    }
}

}  // namespace lbm


namespace {

// All field arrays share one allocation. Each array starts kStagger values
// past a 64-value boundary of the previous one, so that the same cell of the
// 28 scalar arrays and of the 15 population planes falls into different HBM
// channels. (With one hipMalloc per array, the arrays of a 256^3 grid sit
// 64 MiB apart and the 46 streams of the fused step ran 4-5% slower.)
constexpr long kStagger = 64;

template <typename T>
int alloc_fields(lbm_solver* s, Fields<T>& f) {
    const long V = s->V, vs = s->vs;
    const size_t fb = V * sizeof(T);
    const size_t sstride = (size_t)((V + 63) / 64 * 64 + kStagger) * sizeof(T);
    const size_t gstride = (size_t)(kQ * vs + kStagger) * sizeof(T);
    const size_t total = 28 * sstride + 2 * gstride;
    char* pool = nullptr;
    HIPCHK(hipMalloc((void**)&pool, total));
    s->allocs.push_back(pool);
    HIPCHK(hipMemsetAsync(pool, 0, total, s->stream));
    size_t off = 0;
    auto get = [&](size_t bytes, void** p) -> int {
        const size_t step = bytes == fb ? sstride : gstride;
        if (off + step > total) return fail(LBM_ERR_STATE, "field pool exhausted");
        *p = pool + off;
        off += step;
        return LBM_OK;
    };
    T* p = nullptr;
    for (int m = 0; m < 5; ++m) {
        if (int r = get(fb, (void**)&p)) return r;
        f.q[m] = p;
        if (int r = get(fb, (void**)&p)) return r;
        f.qn[m] = p;
        if (int r = get(fb, (void**)&p)) return r;
        f.sig[m] = p;
    }
    for (int m = 0; m < 3; ++m) {
        if (int r = get(fb, (void**)&p)) return r;
        f.tau[m] = p;
        if (int r = get(fb, (void**)&p)) return r;
        f.F[m] = p;
        if (int r = get(fb, (void**)&p)) return r;
        f.u[m] = p;
        if (int r = get(fb, (void**)&p)) return r;
        f.un[m] = p;
    }
    if (int r = get(fb, (void**)&p)) return r;
    f.rho = p;
    if (int r = get(kQ * vs * sizeof(T), (void**)&p)) return r;
    f.g = p;
    if (int r = get(kQ * vs * sizeof(T), (void**)&p)) return r;
    f.gn = p;
    f.vs = vs;
    return LBM_OK;
}

// What the density array is stored relative to (rho0 for FP32, 0 for FP64).
double rho_base(const lbm_solver* s) { return s->fp64 ? 0.0 : s->cfg.p.density; }

int fill_value(lbm_solver* s, void* dst, double v) {
    std::vector<char> host(s->V * s->esize());
    if (s->fp64) std::fill((double*)host.data(), (double*)host.data() + s->V, v);
    else std::fill((float*)host.data(), (float*)host.data() + s->V, (float)v);
    HIPCHK(hipMemcpyAsync(dst, host.data(), host.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

// Copy an FP64 host block of n values into a device field of the handle's
// type; an FP32 field stores src - base (the density keeps rho - rho0).
int put_field(lbm_solver* s, void* dst, const double* src, long n, double base = 0.0) {
    if (s->fp64) {
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    } else {
        std::vector<float> tmp(n);
        for (long i = 0; i < n; ++i) tmp[i] = (float)(src[i] - base);
        HIPCHK(hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return LBM_OK;
}
int get_field(lbm_solver* s, double* dst, const void* src, long n, double base = 0.0) {
    if (s->fp64) {
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    } else {
        std::vector<float> tmp(n);
        HIPCHK(hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        for (long i = 0; i < n; ++i) dst[i] = base + (double)tmp[i];
    }
    return LBM_OK;
}

void* field_ptr(lbm_solver* s, int which, int comp) {
    // which: 0 rho, 1 u, 2 q, 3 F
    if (s->fp64) {
        switch (which) {
            case 0: return s->d64.rho;
            case 1: return s->d64.u[comp];
            case 2: return s->d64.q[comp];
            default: return s->d64.F[comp];
        }
    }
    switch (which) {
        case 0: return s->d32.rho;
        case 1: return s->d32.u[comp];
        case 2: return s->d32.q[comp];
        default: return s->d32.F[comp];
    }
}

template <typename T>
void swap_q(Fields<T>& f) {
    for (int m = 0; m < 5; ++m) std::swap(f.q[m], f.qn[m]);
}
template <typename T>
void swap_g(Fields<T>& f) {
    std::swap(f.g, f.gn);
}
template <typename T>
void swap_u(Fields<T>& f) {
    for (int m = 0; m < 3; ++m) std::swap(f.u[m], f.un[m]);
}

}  // namespace

// Kernel sequence of one step, shared by stepping, timing and profiling.
// rec (optional) records an event pair around every launch.
namespace lbm {

int run_qtensor(lbm_solver* s, KernelRecorder* rec) {
    hipStream_t st = s->stream;
    if (s->fp64) {
        if (rec) rec->begin("qstep_f64");
        launch_qstep(s->d64, s->bc64, s->k64, s->V, st);
        if (rec) rec->end();
        swap_q(s->d64);
        if (s->k64.force_full) {
            if (rec) rec->begin("force_f64");
            launch_force_lbm<double, false>(s->d64, s->bc64, s->k64, s->V, FL_FORCE | FL_STORE_F, 0, st);
            if (rec) rec->end();
        }
    } else {
        if (rec) rec->begin("qstep_f32");
        launch_qstep(s->d32, s->bc32, s->k32, s->V, st);
        if (rec) rec->end();
        swap_q(s->d32);
        if (s->k32.force_full) {
            if (rec) rec->begin("force_f32");
            launch_force_lbm<float, true>(s->d32, s->bc32, s->k32, s->V, FL_FORCE | FL_STORE_F, 0, st);
            if (rec) rec->end();
        }
    }
    HIPCHK(hipGetLastError());
    return LBM_OK;
}

int run_lbm(lbm_solver* s, KernelRecorder* rec) {
    hipStream_t st = s->stream;
    if (s->fp64) {
        if (rec) rec->begin("lbm_f64");
        launch_force_lbm<double, false>(s->d64, s->bc64, s->k64, s->V, FL_LBM | FL_STORE_RHO, s->pulled, st);
        if (rec) rec->end();
        swap_g(s->d64);
    } else {
        if (rec) rec->begin("lbm_f32");
        launch_force_lbm<float, true>(s->d32, s->bc32, s->k32, s->V, FL_LBM | FL_STORE_RHO, s->pulled, st);
        if (rec) rec->end();
        swap_g(s->d32);
    }
    s->pulled = 1;
    HIPCHK(hipGetLastError());
    return LBM_OK;
}

// Fused tile index for the next step, or -1 for the two-kernel split.
int active_tile(const lbm_solver* s) {
    if (s->fp64) {
        if (!s->fused64_ok) return -1;
        const int sched = s->schedule ? s->schedule : s->tuned;
        return sched >= 2 ? sched - 2 : -1;
    }
    if (!s->fused_ok) return -1;
    int sched = s->schedule ? s->schedule : (s->tuned ? s->tuned : 2);
    return sched >= 2 ? sched - 2 : -1;
}

// One full step. last: this is the final step of a batch of steps, so the
// auxiliary outputs (rho, F) are materialised for downloads.
int run_step(lbm_solver* s, bool last, KernelRecorder* rec) {
    hipStream_t st = s->stream;
    if (s->fp64 && active_tile(s) >= 0) {
        if (rec) rec->begin("fused_step_f64");
        launch_fused_step_f64(s->d64, s->bc64, s->k64, active_tile(s), s->pulled, (last || s->moving) ? 1 : 0, st);
        if (rec) rec->end();
        swap_q(s->d64);
        swap_g(s->d64);
        swap_u(s->d64);
    } else if (s->fp64) {
        if (rec) rec->begin("qstep_f64");
        launch_qstep(s->d64, s->bc64, s->k64, s->V, st);
        if (rec) rec->end();
        swap_q(s->d64);
        if (rec) rec->begin("force_lbm_f64");
        launch_force_lbm<double, false>(s->d64, s->bc64, s->k64, s->V, FL_FORCE | FL_LBM | FL_STORE_F | FL_STORE_RHO,
                                        s->pulled, st);
        if (rec) rec->end();
        swap_g(s->d64);
    } else {
        const int tile = active_tile(s);
        const bool aux = last || s->moving;
        if (tile >= 0 && tile < fused_tile_count()) {
            if (rec) rec->begin("fused_step_f32");
            launch_fused_step_f32(s->d32, s->bc32, s->k32, tile, s->pulled, aux ? 1 : 0, st);
            if (rec) rec->end();
            swap_q(s->d32);
            swap_g(s->d32);
            swap_u(s->d32);
        } else {
            if (rec) rec->begin("qstep_f32");
            launch_qstep(s->d32, s->bc32, s->k32, s->V, st);
            if (rec) rec->end();
            swap_q(s->d32);
            const int flags = FL_FORCE | FL_LBM | (aux ? (FL_STORE_F | FL_STORE_RHO) : 0);
            if (rec) rec->begin("force_lbm_f32");
            launch_force_lbm<float, true>(s->d32, s->bc32, s->k32, s->V, flags, s->pulled, st);
            if (rec) rec->end();
            swap_g(s->d32);
        }
    }
    s->pulled = 1;
    HIPCHK(hipGetLastError());
    return LBM_OK;
}

}  // namespace lbm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* lbm_last_error(void) { return g_last_error.c_str(); }

int lbm_default_config(lbm_config* cfg, int32_t nx, int32_t ny, int32_t nz) {
    if (!cfg) return fail(LBM_ERR_INVALID, "null config");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->nx = nx;
    cfg->ny = ny;
    cfg->nz = nz;
    cfg->advection = LBM_ADV_CENTRED;
    cfg->precision = LBM_FP64;
    cfg->force_model = LBM_FORCE_FULL;
    for (int f = 0; f < 6; ++f) {
        cfg->walls[f].qbc = LBM_Q_PERIODIC;
        cfg->walls[f].ubc = LBM_U_PERIODIC;
    }
    // namespace Params defaults (src/params.h:29-112).
    cfg->p.dt = 1.0;
    cfg->p.density = 1.0;
    cfg->p.tauf = 2.5;
    cfg->p.L = 0.06;
    cfg->p.A = 0.0;
    cfg->p.B = -0.3;
    cfg->p.C = 0.3;
    cfg->p.lambda = 0.3;
    cfg->p.gamma = 0.34;
    cfg->p.alpha = 0.023;
    cfg->p.mu = 0.0;
    return LBM_OK;
}

int lbm_set_preset(lbm_config* cfg, const char* name) {
    if (!cfg || !name) return fail(LBM_ERR_INVALID, "null argument");
    const std::string n(name);
    auto set = [&](int f, int q, int u) {
        cfg->walls[f].qbc = q;
        cfg->walls[f].ubc = u;
    };
    if (n == "FullyPeriodic" || n == "PeriodicBC") {
        for (int f = 0; f < 6; ++f) set(f, LBM_Q_PERIODIC, LBM_U_PERIODIC);
    } else if (n == "Channel") {
        for (int f = 0; f < 4; ++f) set(f, LBM_Q_NEUMANN, LBM_U_NOSLIP);
        set(4, LBM_Q_PERIODIC, LBM_U_PERIODIC);
        set(5, LBM_Q_PERIODIC, LBM_U_PERIODIC);
    } else if (n == "SlitFreeSlip" || n == "Slit") {
        set(0, LBM_Q_NEUMANN, LBM_U_SPECULAR);
        set(1, LBM_Q_NEUMANN, LBM_U_SPECULAR);
        for (int f = 2; f < 6; ++f) set(f, LBM_Q_PERIODIC, LBM_U_PERIODIC);
    } else if (n == "SlitNoSlip") {
        set(0, LBM_Q_NEUMANN, LBM_U_NOSLIP);
        set(1, LBM_Q_NEUMANN, LBM_U_NOSLIP);
        for (int f = 2; f < 6; ++f) set(f, LBM_Q_PERIODIC, LBM_U_PERIODIC);
    } else {
        return fail(LBM_ERR_INVALID, "unknown preset " + n);
    }
    return LBM_OK;
}

int lbm_create(const lbm_config* cfg, lbm_handle* out) {
    if (!cfg || !out) return fail(LBM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (int r = validate_config(*cfg)) return r;
    std::vector<uint16_t> table;
    if (int r = build_slot_table(*cfg, table)) return r;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(LBM_ERR_HIP, "no HIP device");

    lbm_solver* s = new lbm_solver();
    s->cfg = *cfg;
    s->V = (long)cfg->nx * cfg->ny * cfg->nz;
    s->vs = (s->V + 63) / 64 * 64 + kStagger;   // population planes staggered like the arrays
    s->fp64 = cfg->precision == LBM_FP64;
    s->moving = has_moving_wall(*cfg);
    // The fused kernel addresses populations with 32-bit byte offsets.
    s->fused_ok = !s->fp64 && fused_tile_count() > 0 && (double)kQ * s->vs * sizeof(float) < 4294967296.0;
    {
        bool periodic = true;
        for (int face = 0; face < 6; ++face)
            periodic = periodic && cfg->walls[face].qbc == LBM_Q_PERIODIC && cfg->walls[face].ubc == LBM_U_PERIODIC;
        s->fused64_ok = s->fp64 && periodic && fused64_tile_count() > 0 &&
                        (double)kQ * s->vs * sizeof(double) < 4294967296.0;
    }
    auto cleanup = [&](int r) {
        lbm_destroy(s);
        return r;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(LBM_ERR_HIP, "hipStreamCreate failed"));
    int r = s->fp64 ? alloc_fields(s, s->d64) : alloc_fields(s, s->d32);
    if (r) return cleanup(r);
    void* p = nullptr;
    if (hipMalloc(&p, table.size() * sizeof(uint16_t)) != hipSuccess)
        return cleanup(fail(LBM_ERR_HIP, "hipMalloc(slot table) failed"));
    s->allocs.push_back(p);
    s->slots = (uint16_t*)p;
    if (hipMemcpy(p, table.data(), table.size() * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(LBM_ERR_HIP, "slot table upload failed"));
    s->d64.slots = s->slots;
    s->d32.slots = s->slots;
    if (hipMalloc(&p, 3 * s->V * sizeof(double)) != hipSuccess) return cleanup(fail(LBM_ERR_HIP, "hipMalloc failed"));
    s->allocs.push_back(p);
    s->uprev = (double*)p;
    if (hipMemset(p, 0, 3 * s->V * sizeof(double)) != hipSuccess) return cleanup(fail(LBM_ERR_HIP, "hipMemset failed"));
    s->diag_blocks = (int)std::min<long>(1024, (s->V + 255) / 256);
    if (hipMalloc(&p, (s->diag_blocks * 10 + 16) * sizeof(double)) != hipSuccess)
        return cleanup(fail(LBM_ERR_HIP, "hipMalloc failed"));
    s->allocs.push_back(p);
    s->partials = (double*)p;
    s->diag_out = s->partials + s->diag_blocks * 10;
    s->bc64 = make_bc<double>(*cfg);
    s->bc32 = make_bc<float>(*cfg);
    s->bc32.slots_xy = s->bc64.slots_xy = slots_independent_of_z(*cfg, table) ? 1 : 0;
    s->k64 = make_phys<double>(*cfg);
    s->k32 = make_phys<float>(*cfg);
    // FluidFields constructor state (src/fluid_fields.cc:6-17): rho = kDensity,
    // u = F = 0, populations Feq(rho0, 0).
    if ((r = fill_value(s, field_ptr(s, 0, 0), cfg->p.density - rho_base(s)))) return cleanup(r);
    std::vector<double> rho(s->V, cfg->p.density);
    if ((r = lbm_upload_state(s, rho.data(), nullptr, nullptr, nullptr, nullptr))) return cleanup(r);
    *out = s;
    return LBM_OK;
}

int lbm_destroy(lbm_handle s) {
    if (!s) return LBM_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return LBM_OK;
}

int lbm_get_config(lbm_handle h, lbm_config* out) {
    if (!h || !out) return fail(LBM_ERR_INVALID, "null argument");
    *out = h->cfg;
    return LBM_OK;
}

int lbm_device_info(char* buf, size_t len) {
    if (!buf || len == 0) return fail(LBM_ERR_INVALID, "null buffer");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, len,
                  "HIP device %d: %s (%s), %d CUs, %.1f GiB HBM, %d KiB LDS/block, wavefront %d, L2 %d KiB",
                  dev, prop.name, prop.gcnArchName, prop.multiProcessorCount,
                  prop.totalGlobalMem / 1073741824.0, (int)(prop.sharedMemPerBlock / 1024), prop.warpSize,
                  prop.l2CacheSize / 1024);
    return LBM_OK;
}

int lbm_upload_state(lbm_handle s, const double* rho, const double* ux, const double* uy, const double* uz,
                     const double* q) {
    if (!s) return fail(LBM_ERR_INVALID, "null handle");
    const long V = s->V;
    int r;
    if (rho && (r = put_field(s, field_ptr(s, 0, 0), rho, V, rho_base(s)))) return r;
    const double* uu[3] = {ux, uy, uz};
    for (int a = 0; a < 3; ++a)
        if (uu[a] && (r = put_field(s, field_ptr(s, 1, a), uu[a], V))) return r;
    if (q)
        for (int m = 0; m < 5; ++m)
            if ((r = put_field(s, field_ptr(s, 2, m), q + m * V, V))) return r;
    // Populations = Feq(rho, u) of the (now current) device state.
    std::vector<double> hr(V), hu[3];
    if ((r = get_field(s, hr.data(), field_ptr(s, 0, 0), V, rho_base(s)))) return r;
    for (int a = 0; a < 3; ++a) {
        hu[a].resize(V);
        if ((r = get_field(s, hu[a].data(), field_ptr(s, 1, a), V))) return r;
    }
    // For FP32 the device copy of rho/u is rounded; Feq is formed from the
    // FP64 inputs where given so the shifted populations carry full precision.
    const double* src_r = rho ? rho : hr.data();
    const double* src_u[3] = {ux ? ux : hu[0].data(), uy ? uy : hu[1].data(), uz ? uz : hu[2].data()};
    if (s->fp64) {
        std::vector<double> g(kQ * s->vs, 0.0);
        double fe[kQ];
        for (long c = 0; c < V; ++c) {
            feq_host(src_r[c], src_u[0][c], src_u[1][c], src_u[2][c], fe);
            for (int i = 0; i < kQ; ++i) g[i * s->vs + c] = fe[i];
        }
        HIPCHK(hipMemcpyAsync(s->d64.g, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    } else {
        std::vector<float> g(kQ * s->vs, 0.0f);
        double fe[kQ];
        const double r0 = s->cfg.p.density;
        for (long c = 0; c < V; ++c) {
            feq_host(src_r[c], src_u[0][c], src_u[1][c], src_u[2][c], fe);
            for (int i = 0; i < kQ; ++i) g[i * s->vs + c] = (float)(fe[i] - wgt(i) * r0);
        }
        HIPCHK(hipMemcpyAsync(s->d32.g, g.data(), g.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    s->pulled = 0;
    return LBM_OK;
}

int lbm_upload_populations(lbm_handle s, const double* f) {
    if (!s || !f) return fail(LBM_ERR_INVALID, "null argument");
    const long V = s->V;
    if (s->fp64) {
        std::vector<double> g(kQ * s->vs, 0.0);
        for (int i = 0; i < kQ; ++i) std::memcpy(&g[i * s->vs], f + i * V, V * sizeof(double));
        HIPCHK(hipMemcpyAsync(s->d64.g, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
    } else {
        std::vector<float> g(kQ * s->vs, 0.0f);
        const double r0 = s->cfg.p.density;
        for (int i = 0; i < kQ; ++i)
            for (long c = 0; c < V; ++c) g[i * s->vs + c] = (float)(f[i * V + c] - wgt(i) * r0);
        HIPCHK(hipMemcpyAsync(s->d32.g, g.data(), g.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->pulled = 0;
    return LBM_OK;
}

int lbm_upload_force(lbm_handle s, const double* fx, const double* fy, const double* fz) {
    if (!s) return fail(LBM_ERR_INVALID, "null handle");
    const double* ff[3] = {fx, fy, fz};
    for (int a = 0; a < 3; ++a)
        if (ff[a])
            if (int r = put_field(s, field_ptr(s, 3, a), ff[a], s->V)) return r;
    return LBM_OK;
}

int lbm_download_state(lbm_handle s, double* rho, double* ux, double* uy, double* uz, double* q) {
    if (!s) return fail(LBM_ERR_INVALID, "null handle");
    const long V = s->V;
    int r;
    if (rho && (r = get_field(s, rho, field_ptr(s, 0, 0), V, rho_base(s)))) return r;
    double* uu[3] = {ux, uy, uz};
    for (int a = 0; a < 3; ++a)
        if (uu[a] && (r = get_field(s, uu[a], field_ptr(s, 1, a), V))) return r;
    if (q)
        for (int m = 0; m < 5; ++m)
            if ((r = get_field(s, q + m * V, field_ptr(s, 2, m), V))) return r;
    return LBM_OK;
}

int lbm_download_force(lbm_handle s, double* fx, double* fy, double* fz) {
    if (!s) return fail(LBM_ERR_INVALID, "null handle");
    double* ff[3] = {fx, fy, fz};
    for (int a = 0; a < 3; ++a)
        if (ff[a])
            if (int r = get_field(s, ff[a], field_ptr(s, 3, a), s->V)) return r;
    return LBM_OK;
}

int lbm_download_populations(lbm_handle s, double* f) {
    if (!s || !f) return fail(LBM_ERR_INVALID, "null argument");
    const long V = s->V;
    void* tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, kQ * s->vs * s->esize()));
    if (s->fp64) launch_stream_only(s->d64, s->bc64, V, s->pulled, (double*)tmp, s->stream);
    else launch_stream_only(s->d32, s->bc32, V, s->pulled, (float*)tmp, s->stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(LBM_ERR_HIP, std::string("stream_only launch: ") + hipGetErrorString(e));
    }
    int r = LBM_OK;
    if (s->fp64) {
        std::vector<double> g(kQ * s->vs);
        if (hipMemcpyAsync(g.data(), tmp, g.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
            hipStreamSynchronize(s->stream) != hipSuccess)
            r = fail(LBM_ERR_HIP, "population download failed");
        else
            for (int i = 0; i < kQ; ++i) std::memcpy(f + i * V, &g[i * s->vs], V * sizeof(double));
    } else {
        std::vector<float> g(kQ * s->vs);
        if (hipMemcpyAsync(g.data(), tmp, g.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
            hipStreamSynchronize(s->stream) != hipSuccess)
            r = fail(LBM_ERR_HIP, "population download failed");
        else {
            const double r0 = s->cfg.p.density;
            for (int i = 0; i < kQ; ++i)
                for (long c = 0; c < V; ++c) f[i * V + c] = (double)g[i * s->vs + c] + wgt(i) * r0;
        }
    }
    (void)hipFree(tmp);
    return r;
}

int lbm_qtensor_step(lbm_handle s) {
    if (!s) return fail(LBM_ERR_INVALID, "null handle");
    return run_qtensor(s, nullptr);
}

int lbm_lbm_step(lbm_handle s) {
    if (!s) return fail(LBM_ERR_INVALID, "null handle");
    return run_lbm(s, nullptr);
}

// The automatic FP32 schedule is tuned on first use: the best schedule
// depends on the grid (small grids give the fused z-marching kernel too few
// tiles to fill 256 CUs). Tuning runs on a snapshot and leaves the state as
// it was; if it cannot run, the default fused tile is used.
}  // extern "C"
namespace lbm {
void ensure_tuned(lbm_solver* s) {
    if (s->schedule != 0 || s->tuned != 0) return;
    if (s->fp64 && !s->fused64_ok) {
        s->tuned = 1;
        return;
    }
    const std::string saved_error = g_last_error;
    if (lbm_autotune(s, 3, nullptr, 0, nullptr) != LBM_OK) s->tuned = s->fp64 ? 1 : (s->fused_ok ? 2 : 1);
    g_last_error = saved_error;
}
}  // namespace lbm
extern "C" {

int lbm_step(lbm_handle s, int32_t n) {
    if (!s) return fail(LBM_ERR_INVALID, "null handle");
    if (n < 0) return fail(LBM_ERR_INVALID, "negative step count");
    if (n > 0) ensure_tuned(s);
    for (int t = 0; t < n; ++t)
        if (int r = run_step(s, t == n - 1, nullptr)) return r;
    return LBM_OK;
}

int lbm_synchronize(lbm_handle s) {
    if (!s) return fail(LBM_ERR_INVALID, "null handle");
    HIPCHK(hipStreamSynchronize(s->stream));
    return LBM_OK;
}

int lbm_diagnostics(lbm_handle s, lbm_diag* out) {
    if (!s || !out) return fail(LBM_ERR_INVALID, "null argument");
    if (s->fp64) launch_diagnostics(s->d64, s->bc64, s->k64, s->V, s->uprev, s->partials, s->diag_blocks, s->diag_out, s->stream);
    else launch_diagnostics(s->d32, s->bc32, s->k64, s->V, s->uprev, s->partials, s->diag_blocks, s->diag_out, s->stream);
    HIPCHK(hipGetLastError());
    double h[10];
    HIPCHK(hipMemcpyAsync(h, s->diag_out, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    out->mass = h[0];
    out->px = h[1];
    out->py = h[2];
    out->pz = h[3];
    out->kinetic = h[4];
    out->du2 = h[5];
    out->u2 = h[6];
    out->umax = std::sqrt(h[7]);
    out->free_energy = h[8];
    out->nonfinite = (h[9] > 0.0 || !std::isfinite(h[0]) || !std::isfinite(h[1]) || !std::isfinite(h[2]) ||
                      !std::isfinite(h[3])) ? 1 : 0;
    out->_pad = 0;
    return LBM_OK;
}

int lbm_order_director(lbm_handle s, double* order, double* director) {
    if (!s || !order || !director) return fail(LBM_ERR_INVALID, "null argument");
    void* tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, 4 * s->V * sizeof(double)));
    double* o = (double*)tmp;
    double* d = o + s->V;
    if (s->fp64) launch_order_director(s->d64, s->V, o, d, s->stream);
    else launch_order_director(s->d32, s->V, o, d, s->stream);
    int r = LBM_OK;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(order, o, s->V * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipMemcpyAsync(director, d, 3 * s->V * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        r = fail(LBM_ERR_HIP, "order/director failed");
    (void)hipFree(tmp);
    return r;
}

// ---- order/director without device allocation --------------------------------
// lbm_order_director_noalloc stages the order and director fields in one
// file-scope device array instead of a per-call hipMalloc. The array's name is
// a host-side stand-in: host code uses only the address hipGetSymbolAddress
// returns for the current device (looked up once per device, on that device's
// first call), never the name itself. Every call on one GPU shares that GPU's
// array and overwrites all of it that it reads back (nothing needs clearing);
// calls that overlap on one GPU (two host threads, or two streams) race.
namespace {
constexpr long kNoAllocMaxCells = 1L << 19;   // 524,288 cells; 16 MiB of staging
__device__ __align__(256) double scratch[4 * kNoAllocMaxCells];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(double** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(LBM_ERR_UNSUPPORTED, "device index beyond the staging table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(LBM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return LBM_OK;
}
}  // namespace

int lbm_order_director_noalloc(lbm_handle s, double* order, double* director) {
    if (!s || !order || !director) return fail(LBM_ERR_INVALID, "null argument");
    if (s->V > kNoAllocMaxCells)
        return fail(LBM_ERR_INVALID, "grid of " + std::to_string(s->V) + " cells exceeds the " +
                                         std::to_string(kNoAllocMaxCells) +
                                         "-cell staging limit of lbm_order_director_noalloc");
    double* base = nullptr;
    if (int r = scratch_for_current_device(&base)) return r;
    double* o = base;
    double* d = base + s->V;
    if (s->fp64) launch_order_director(s->d64, s->V, o, d, s->stream);
    else launch_order_director(s->d32, s->V, o, d, s->stream);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(order, o, s->V * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipMemcpyAsync(director, d, 3 * s->V * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return fail(LBM_ERR_HIP, "order/director failed");
    return LBM_OK;
}

int lbm_time_steps(lbm_handle s, int32_t n, double* ms) {
    if (!s || !ms || n < 1) return fail(LBM_ERR_INVALID, "bad argument");
    ensure_tuned(s);
    hipEvent_t a = nullptr, b = nullptr;
    auto release = scope_exit([&] {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    });
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, s->stream));
    int r = LBM_OK;
    for (int t = 0; t < n && r == LBM_OK; ++t) r = run_step(s, t == n - 1, nullptr);
    HIPCHK(hipEventRecord(b, s->stream));
    HIPCHK(hipEventSynchronize(b));
    float f = 0.0f;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    *ms = f;
    return r;
}

int lbm_profile_kernels(lbm_handle s, int32_t n, double* ms_out, int32_t max, char* names, size_t names_len,
                        int32_t* count) {
    if (!s || !ms_out || !count || n < 1) return fail(LBM_ERR_INVALID, "bad argument");
    KernelRecorder rec(s->stream);
    int r = LBM_OK;
    for (int t = 0; t < n && r == LBM_OK; ++t) r = run_step(s, t == n - 1, &rec);
    if (r) return r;
    if (int e = rec.finish()) return e;
    std::string joined;
    int k = 0;
    for (auto& kv : rec.totals()) {
        if (k < max) ms_out[k] = kv.second.first / std::max(1, kv.second.second);
        if (!joined.empty()) joined += ";";
        joined += kv.first;
        ++k;
    }
    *count = std::min(k, (int)max);
    if (names && names_len) {
        std::strncpy(names, joined.c_str(), names_len - 1);
        names[names_len - 1] = 0;
    }
    return LBM_OK;
}

int lbm_set_schedule(lbm_handle s, int32_t schedule) {
    if (!s) return fail(LBM_ERR_INVALID, "null handle");
    if (schedule < 0 || schedule > 1 + (s->fp64 ? fused64_tile_count() : fused_tile_count()))
        return fail(LBM_ERR_INVALID, "schedule out of range");
    s->schedule = schedule;
    return LBM_OK;
}
int lbm_get_schedule(lbm_handle s, int32_t* schedule) {
    if (!s || !schedule) return fail(LBM_ERR_INVALID, "null argument");
    *schedule = s->schedule;
    return LBM_OK;
}
int lbm_tile_count(void) { return fused_tile_count(); }
int lbm_tile_name(int32_t schedule, char* buf, size_t len) {
    if (!buf || !len) return fail(LBM_ERR_INVALID, "null buffer");
    std::string n;
    if (schedule == 0) n = "auto";
    else if (schedule == 1) n = "two-kernel";
    else if (schedule - 2 < fused_tile_count()) n = std::string("fused:") + fused_tile_name(schedule - 2);
    else return fail(LBM_ERR_INVALID, "schedule out of range");
    std::strncpy(buf, n.c_str(), len - 1);
    buf[len - 1] = 0;
    return LBM_OK;
}

int lbm_schedule_name(lbm_handle s, int32_t schedule, char* buf, size_t len) {
    if (!s || !buf || !len) return fail(LBM_ERR_INVALID, "null argument");
    if (!s->fp64) return lbm_tile_name(schedule, buf, len);
    std::string n;
    if (schedule == 0) {
        if (!s->fused64_ok || s->tuned == 1) n = "auto:two-kernel";
        else if (s->tuned >= 2) n = std::string("auto:fused:") + fused64_tile_name(s->tuned - 2);
        else n = "auto";
    }
    else if (schedule == 1) n = "two-kernel";
    else if (schedule - 2 < fused64_tile_count()) n = std::string("fused:") + fused64_tile_name(schedule - 2);
    else return fail(LBM_ERR_INVALID, "schedule out of range");
    std::strncpy(buf, n.c_str(), len - 1);
    buf[len - 1] = 0;
    return LBM_OK;
}

}  // extern "C"
namespace {
// Schedule tuning on a snapshot of the evolving state (fields of type T), so
// that it leaves the simulation as it was; the best schedule is kept in
// s->tuned for the automatic schedule.
template <typename T>
int autotune_impl(lbm_solver* s, const std::vector<int>& cands, int32_t n, double* ms_out, int32_t max,
                  int32_t* count) {
    auto fields = [s]() -> Fields<T>& {
        if constexpr (sizeof(T) == 4) return s->d32;
        else return s->d64;
    };
    Fields<T>& f = fields();
    const size_t fb = s->V * sizeof(T), gb = kQ * s->vs * sizeof(T);
    std::vector<std::pair<void*, size_t>> parts;
    for (int m = 0; m < 5; ++m) parts.push_back({f.q[m], fb});
    for (int m = 0; m < 3; ++m) parts.push_back({f.u[m], fb});
    for (int m = 0; m < 3; ++m) parts.push_back({f.F[m], fb});
    parts.push_back({f.rho, fb});
    size_t total = gb;
    for (auto& p : parts) total += p.second;
    void* snap = nullptr;
    HIPCHK(hipMalloc(&snap, total));
    {
        char* cur = (char*)snap;
        int r = LBM_OK;
        if (hipMemcpyAsync(cur, f.g, gb, hipMemcpyDeviceToDevice, s->stream) != hipSuccess) r = LBM_ERR_HIP;
        cur += gb;
        for (auto& p : parts) {
            if (hipMemcpyAsync(cur, p.first, p.second, hipMemcpyDeviceToDevice, s->stream) != hipSuccess) r = LBM_ERR_HIP;
            cur += p.second;
        }
        if (r) {
            (void)hipFree(snap);
            return fail(LBM_ERR_HIP, "snapshot failed");
        }
    }
    const int pulled0 = s->pulled, sched0 = s->schedule;
    // One trial: m timed steps of schedule c from the snapshot, which is then
    // put back (into whatever buffers are current: the steps permute the
    // buffer pointers).
    auto trial = [&](int c, int m, double& ms) -> int {
        s->schedule = c;
        s->pulled = pulled0;
        if (int e = run_step(s, false, nullptr)) return e;
        if (int e = lbm_time_steps(s, m, &ms)) return e;
        ms /= m;
        Fields<T> now = fields();
        char* cur = (char*)snap;
        if (hipMemcpyAsync(now.g, cur, gb, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
            return fail(LBM_ERR_HIP, "restore failed");
        cur += gb;
        void* dst[12] = {now.q[0], now.q[1], now.q[2], now.q[3], now.q[4], now.u[0], now.u[1], now.u[2], now.F[0], now.F[1], now.F[2], now.rho};
        for (int i = 0; i < 12; ++i) {
            if (hipMemcpyAsync(dst[i], cur, parts[i].second, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
                return fail(LBM_ERR_HIP, "restore failed");
            cur += parts[i].second;
        }
        return LBM_OK;
    };
    // First pass: n steps of every candidate. Second pass: the three fastest
    // again over 8n steps each, which settles near ties (schedules within 1%
    // of each other) and decides on the longer measurement.
    std::vector<double> ms_all(cands.size(), 1e300);
    int r = LBM_OK;
    for (size_t i = 0; i < cands.size() && !r; ++i) r = trial(cands[i], n, ms_all[i]);
    int best_s = 1;
    if (!r) {
        std::vector<size_t> order(cands.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ms_all[a] < ms_all[b]; });
        double best = 1e300;
        for (size_t j = 0; j < std::min<size_t>(3, order.size()) && !r; ++j) {
            const size_t i = order[j];
            double ms = 0.0;
            if ((r = trial(cands[i], 8 * n, ms))) break;
            ms_all[i] = ms;
            if (ms < best) { best = ms; best_s = cands[i]; }
        }
    }
    int k = 0;
    for (size_t i = 0; i < cands.size() && !r; ++i) {
        if (k < max && ms_out) ms_out[k] = ms_all[i];
        ++k;
    }
    (void)hipStreamSynchronize(s->stream);
    (void)hipFree(snap);
    s->pulled = pulled0;
    s->schedule = sched0;
    if (r) return r;
    s->tuned = best_s;
    if (count) *count = std::min(k, (int)max);
    return LBM_OK;
}
}  // namespace
extern "C" {

int lbm_autotune(lbm_handle s, int32_t n, double* ms_out, int32_t max, int32_t* count) {
    if (!s || n < 1) return fail(LBM_ERR_INVALID, "bad argument");
    if (count) *count = 0;
    std::vector<int> cands = {1};
    if (s->fp64) {
        if (s->fused64_ok)
            for (int t = 0; t < fused64_tile_count(); ++t) cands.push_back(2 + t);
        return autotune_impl<double>(s, cands, n, ms_out, max, count);
    }
    if (s->fused_ok)
        for (int t = 0; t < fused_tile_count(); ++t) cands.push_back(2 + t);
    return autotune_impl<float>(s, cands, n, ms_out, max, count);
}

int lbm_bandwidth_probe(size_t bytes, int32_t reps, double* gbs, double* ms_per_copy) {
    if (!gbs || reps < 1 || bytes < 16) return fail(LBM_ERR_INVALID, "bad argument");
    bytes = bytes / 16 * 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = scope_exit([&] {
        if (st) (void)hipStreamSynchronize(st);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
    });
    HIPCHK(hipMalloc(&a, bytes));
    HIPCHK(hipMalloc(&b, bytes));
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemsetAsync(a, 0, bytes, st));
    launch_copy(a, b, bytes, st);  // warm-up
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch_copy(r & 1 ? b : a, r & 1 ? a : b, bytes, st);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    if (ms_per_copy) *ms_per_copy = ms / reps;
    return LBM_OK;
}

int lbm_debug_slot_table(const lbm_config* cfg, uint16_t* table) {
    if (!cfg || !table) return fail(LBM_ERR_INVALID, "null argument");
    if (int r = validate_config(*cfg)) return r;
    std::vector<uint16_t> t;
    if (int r = build_slot_table(*cfg, t)) return r;
    std::memcpy(table, t.data(), t.size() * sizeof(uint16_t));
    return LBM_OK;
}

int lbm_debug_shell_region(const lbm_config* cfg, int32_t w, int32_t* xyz, int64_t max, int64_t* count) {
    if (!cfg || !count || w < 1) return fail(LBM_ERR_INVALID, "bad argument");
    if (int r = validate_config(*cfg)) return r;
    const BCDesc<float> bc = make_bc<float>(*cfg);
    const ShellRegion reg = make_shell_region(bc, w);
    const long n = region_cells(reg);
    *count = n;
    if (!xyz) return LBM_OK;
    if (n > max) return fail(LBM_ERR_INVALID, "xyz holds fewer than count cells");
    for (long i = 0; i < n; ++i) region_cell(reg, bc.n, i, &xyz[3 * i]);
    return LBM_OK;
}

int lbm_debug_tile_map(int32_t kind, int32_t t, int32_t n, int32_t lo_per, int32_t hi_per, int32_t* out) {
    if (!out || n < 1 || (kind != 0 && kind != 1)) return fail(LBM_ERR_INVALID, "bad argument");
    *out = kind == 0 ? map_coord(t, n, lo_per != 0, hi_per != 0) : map_col(t, n, lo_per != 0, hi_per != 0);
    return LBM_OK;
}

int lbm_seed_director_field(uint64_t seed, int32_t nx, int32_t ny, int32_t nz, double s0, double noise,
                            double* q) {
    if (!q || nx < 1 || ny < 1 || nz < 1) return fail(LBM_ERR_INVALID, "bad argument");
    const long V = (long)nx * ny * nz;
    for (long c = 0; c < V; ++c) {
        const double u0 = uniform01(seed, (uint64_t)c, 0);
        const double u1 = uniform01(seed, (uint64_t)c, 1);
        const double u2 = uniform01(seed, (uint64_t)c, 2);
        const double zc = 2.0 * u0 - 1.0;
        const double ph = 2.0 * M_PI * u1;
        const double rr = std::sqrt(std::max(0.0, 1.0 - zc * zc));
        const double n0 = rr * std::cos(ph), n1 = rr * std::sin(ph), n2 = zc;
        const double S = s0 + noise * (2.0 * u2 - 1.0);
        q[0 * V + c] = S * (n0 * n0 - 1.0 / 3.0);
        q[1 * V + c] = S * (n0 * n1);
        q[2 * V + c] = S * (n0 * n2);
        q[3 * V + c] = S * (n1 * n1 - 1.0 / 3.0);
        q[4 * V + c] = S * (n1 * n2);
    }
    return LBM_OK;
}

}  // extern "C"
