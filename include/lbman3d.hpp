// lbman3d.hpp — C++20 host interface over the C ABI of lbman3d.h, with the
// reference's compile-time boundary surface.
//
// The reference selects walls with types (WallSpec<QBC, UBC> per face of a
// BCConfig, src/boundary.h:31-129) and drives the GPU through
// DeviceFields / DeviceSolver<BC> (src/device_fields.h:23-80,
// src/device_solver.h:9-18). This header keeps those names and shapes and
// lowers the types to the runtime lbm_config the library takes:
//   * the tags Periodic, Neumann, NoSlip, SpecularReflection,
//     Anchoring<S, Theta, Phi>, MovingWall<Ux, Uy, Uz>, WallSpec, BCConfig
//     and the presets FullyPeriodicConfig, ChannelConfig, SlitConfig,
//     SlitNoSlipConfig;
//   * qbc_traits / ubc_traits map a tag type to its code. Anchoring- and
//     MovingWall-shaped types from any namespace (static members s/theta/phi
//     or Ux/Uy/Uz) are recognised by shape; the four empty tags of another
//     namespace are mapped by specialising the traits (one line each);
//   * DeviceFields owns one device simulation (RAII over lbm_handle);
//     DeviceSolver<BC> steps it. Failures throw lbman3d::Error, a
//     std::runtime_error, as the reference's CUDA_CHECK does
//     (src/cuda/cuda_utils.h:7-12).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

#include "lbman3d.h"

namespace lbman3d {

// ---- boundary tags ---------------------------------------------------------
struct Periodic {};
struct Neumann {};
struct NoSlip {};
struct SpecularReflection {};

template <double S = 1.0, double Theta = 0.0, double Phi = 0.0>
struct Anchoring {
    static constexpr double s = S;
    static constexpr double theta = Theta;
    static constexpr double phi = Phi;
};

template <double Ux_ = 0.0, double Uy_ = 0.0, double Uz_ = 0.0>
struct MovingWall {
    static constexpr double Ux = Ux_;
    static constexpr double Uy = Uy_;
    static constexpr double Uz = Uz_;
};

template <typename QBC_, typename UBC_>
struct WallSpec {
    using QBC = QBC_;
    using UBC = UBC_;
};

template <typename XLo_ = WallSpec<Neumann, NoSlip>, typename XHi_ = WallSpec<Neumann, NoSlip>,
          typename YLo_ = WallSpec<Neumann, NoSlip>, typename YHi_ = WallSpec<Neumann, NoSlip>,
          typename ZLo_ = WallSpec<Neumann, NoSlip>, typename ZHi_ = WallSpec<Neumann, NoSlip>>
struct BCConfig {
    using XLo = XLo_;
    using XHi = XHi_;
    using YLo = YLo_;
    using YHi = YHi_;
    using ZLo = ZLo_;
    using ZHi = ZHi_;
    static constexpr std::string_view name = "Custom";
};

using PeriodicFace = WallSpec<Periodic, Periodic>;

struct FullyPeriodicConfig : BCConfig<PeriodicFace, PeriodicFace, PeriodicFace, PeriodicFace, PeriodicFace,
                                      PeriodicFace> {
    static constexpr std::string_view name = "FullyPeriodic";
};
struct ChannelConfig : BCConfig<WallSpec<Neumann, NoSlip>, WallSpec<Neumann, NoSlip>, WallSpec<Neumann, NoSlip>,
                                WallSpec<Neumann, NoSlip>, PeriodicFace, PeriodicFace> {
    static constexpr std::string_view name = "Channel";
};
struct SlitConfig : BCConfig<WallSpec<Neumann, SpecularReflection>, WallSpec<Neumann, SpecularReflection>,
                             PeriodicFace, PeriodicFace, PeriodicFace, PeriodicFace> {
    static constexpr std::string_view name = "SlitFreeSlip";
};
struct SlitNoSlipConfig : BCConfig<WallSpec<Neumann, NoSlip>, WallSpec<Neumann, NoSlip>, PeriodicFace,
                                   PeriodicFace, PeriodicFace, PeriodicFace> {
    static constexpr std::string_view name = "SlitNoSlip";
};

// ---- tag -> code -----------------------------------------------------------
// Unknown tags have no definition, so using one fails to compile.
template <typename T, typename = void>
struct qbc_traits;
template <typename T, typename = void>
struct ubc_traits;

template <>
struct qbc_traits<Periodic> {
    static void fill(lbm_wall& w) { w.qbc = LBM_Q_PERIODIC; }
};
template <>
struct qbc_traits<Neumann> {
    static void fill(lbm_wall& w) { w.qbc = LBM_Q_NEUMANN; }
};
template <typename T>
struct qbc_traits<T, std::void_t<decltype(T::s), decltype(T::theta), decltype(T::phi)>> {
    static void fill(lbm_wall& w) {
        w.qbc = LBM_Q_ANCHORING;
        w.anchor_s = T::s;
        w.anchor_theta = T::theta;
        w.anchor_phi = T::phi;
    }
};

template <>
struct ubc_traits<Periodic> {
    static void fill(lbm_wall& w) { w.ubc = LBM_U_PERIODIC; }
};
template <>
struct ubc_traits<NoSlip> {
    static void fill(lbm_wall& w) { w.ubc = LBM_U_NOSLIP; }
};
template <>
struct ubc_traits<SpecularReflection> {
    static void fill(lbm_wall& w) { w.ubc = LBM_U_SPECULAR; }
};
template <typename T>
struct ubc_traits<T, std::void_t<decltype(T::Ux), decltype(T::Uy), decltype(T::Uz)>> {
    static void fill(lbm_wall& w) {
        w.ubc = LBM_U_MOVING;
        w.wall_u[0] = T::Ux;
        w.wall_u[1] = T::Uy;
        w.wall_u[2] = T::Uz;
    }
};

template <typename WS>
lbm_wall wall() {
    lbm_wall w{};
    qbc_traits<typename WS::QBC>::fill(w);
    ubc_traits<typename WS::UBC>::fill(w);
    return w;
}

// ---- errors ----------------------------------------------------------------
class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc) {
    if (rc != LBM_OK) throw Error(rc, std::string("lbman3d: ") + lbm_last_error());
}

// ---- configuration -----------------------------------------------------------
// The params.h defaults (src/params.h:14-112).
inline lbm_params default_params() {
    lbm_config c{};
    check(lbm_default_config(&c, 1, 1, 1));
    return c.p;
}

template <typename BC>
lbm_config make_config(int nx, int ny, int nz, const lbm_params& p = default_params(),
                       int advection = LBM_ADV_CENTRED, int precision = LBM_FP64,
                       int force_model = LBM_FORCE_FULL) {
    lbm_config c{};
    c.nx = nx;
    c.ny = ny;
    c.nz = nz;
    c.advection = advection;
    c.precision = precision;
    c.force_model = force_model;
    c.walls[LBM_XLO] = wall<typename BC::XLo>();
    c.walls[LBM_XHI] = wall<typename BC::XHi>();
    c.walls[LBM_YLO] = wall<typename BC::YLo>();
    c.walls[LBM_YHI] = wall<typename BC::YHi>();
    c.walls[LBM_ZLO] = wall<typename BC::ZLo>();
    c.walls[LBM_ZHI] = wall<typename BC::ZHi>();
    c.p = p;
    return c;
}

// ---- host fields -------------------------------------------------------------
// FP64 SoA host state of V cells (the reference's FluidFields / QTensorFields
// arrays, cell = (z*ny + y)*nx + x). rho starts at kDensity, u and the force
// at zero (src/fluid_fields.cc:6-17).
struct HostFields {
    long V = 0;
    std::vector<double> rho, ux, uy, uz, fx, fy, fz, q;   // q: Qxx, Qxy, Qxz, Qyy, Qyz blocks of V

    HostFields() = default;
    explicit HostFields(long cells, double density = 1.0)
        : V(cells), rho(cells, density), ux(cells, 0.0), uy(cells, 0.0), uz(cells, 0.0), fx(cells, 0.0),
          fy(cells, 0.0), fz(cells, 0.0), q(5 * cells, 0.0) {}
    double* qxx() { return q.data(); }
    double* qxy() { return q.data() + V; }
    double* qxz() { return q.data() + 2 * V; }
    double* qyy() { return q.data() + 3 * V; }
    double* qyz() { return q.data() + 4 * V; }
};

// ---- device fields (DeviceFields) -------------------------------------------
class DeviceFields {
public:
    explicit DeviceFields(const lbm_config& cfg) : cfg_(cfg) { check(lbm_create(&cfg_, &h_)); }
    ~DeviceFields() {
        if (h_) lbm_destroy(h_);
    }
    DeviceFields(const DeviceFields&) = delete;
    DeviceFields& operator=(const DeviceFields&) = delete;
    DeviceFields(DeviceFields&& o) noexcept : cfg_(o.cfg_), h_(std::exchange(o.h_, nullptr)) {}
    DeviceFields& operator=(DeviceFields&& o) noexcept {
        if (this != &o) {
            if (h_) lbm_destroy(h_);
            cfg_ = o.cfg_;
            h_ = std::exchange(o.h_, nullptr);
        }
        return *this;
    }

    // Host state to the device: populations Feq(rho, u) (LbmSolver::Initialize)
    // plus the body force and Q (DeviceFields::Initialize,
    // src/cuda/device_fields.cu:79-115).
    void Initialize(const HostFields& h) {
        expect(h);
        check(lbm_upload_state(h_, h.rho.data(), h.ux.data(), h.uy.data(), h.uz.data(), h.q.data()));
        check(lbm_upload_force(h_, h.fx.data(), h.fy.data(), h.fz.data()));
    }
    // rho, u and Q back to the host (DeviceFields::CopyToHost,
    // src/cuda/device_fields.cu:117-127).
    void CopyToHost(HostFields& h) const {
        expect(h);
        check(lbm_download_state(h_, h.rho.data(), h.ux.data(), h.uy.data(), h.uz.data(), h.q.data()));
    }
    void CopyForceToHost(HostFields& h) const {
        expect(h);
        check(lbm_download_force(h_, h.fx.data(), h.fy.data(), h.fz.data()));
    }
    lbm_diag Diagnostics() const {
        lbm_diag d{};
        check(lbm_diagnostics(h_, &d));
        return d;
    }

    lbm_handle handle() const { return h_; }
    const lbm_config& config() const { return cfg_; }
    long cells() const { return (long)cfg_.nx * cfg_.ny * cfg_.nz; }

private:
    void expect(const HostFields& h) const {
        if (h.V != cells()) throw Error(LBM_ERR_INVALID, "lbman3d: host fields have the wrong cell count");
    }
    lbm_config cfg_{};
    lbm_handle h_ = nullptr;
};

// ---- device solver (DeviceSolver<BC>) ---------------------------------------
// BC is checked against the walls the fields were created with, so a solver
// instantiated for one boundary class cannot step fields of another.
template <typename BC>
struct DeviceSolver {
    void Initialize(DeviceFields& df) {
        const lbm_config want = make_config<BC>(1, 1, 1);
        for (int f = 0; f < 6; ++f) {
            const lbm_wall& a = want.walls[f];
            const lbm_wall& b = df.config().walls[f];
            if (a.qbc != b.qbc || a.ubc != b.ubc || a.anchor_s != b.anchor_s || a.anchor_theta != b.anchor_theta ||
                a.anchor_phi != b.anchor_phi || a.wall_u[0] != b.wall_u[0] || a.wall_u[1] != b.wall_u[1] ||
                a.wall_u[2] != b.wall_u[2])
                throw Error(LBM_ERR_INVALID, "lbman3d: DeviceSolver boundary class differs from the fields' walls");
        }
    }
    void QTensorStep(DeviceFields& df) { check(lbm_qtensor_step(df.handle())); }
    void LBMStep(DeviceFields& df) { check(lbm_lbm_step(df.handle())); }
    // QTensorStep then LBMStep, n times (ActiveNematicSim::Step).
    void Step(DeviceFields& df, int n = 1) { check(lbm_step(df.handle(), n)); }
};

}  // namespace lbman3d
