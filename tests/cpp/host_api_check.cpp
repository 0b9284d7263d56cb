// host_api_check.cpp — exercises the C++ host interface (include/lbman3d.hpp).
//
//   host_api_check host          host-only checks (no GPU): the presets lower
//                                to the configs lbm_set_preset builds, and
//                                boundary tags of another namespace (shaped
//                                like the reference's, src/boundary.h:31-76)
//                                map through the traits.
//   host_api_check run <out>     GPU: 10 steps of a 12x10x8 ChannelConfig case
//                                from the seeded director field (seed 7) via
//                                DeviceFields / DeviceSolver<ChannelConfig>;
//                                writes rho, ux, uy, uz, Q (FP64) to <out>.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "lbman3d.hpp"

namespace other {   // tags declared like the reference's, in another namespace
struct Periodic {};
struct Neumann {};
struct NoSlip {};
struct SpecularReflection {};
template <double S = 1.0, double Theta = 0.0, double Phi = 0.0>
struct Anchoring {
    static constexpr double s = S, theta = Theta, phi = Phi;
};
template <double Ux_ = 0.0, double Uy_ = 0.0, double Uz_ = 0.0>
struct MovingWall {
    static constexpr double Ux = Ux_, Uy = Uy_, Uz = Uz_;
};
template <typename Q, typename U>
struct WallSpec {
    using QBC = Q;
    using UBC = U;
};
struct LidConfig {
    using XLo = WallSpec<Periodic, Periodic>;
    using XHi = WallSpec<Periodic, Periodic>;
    using YLo = WallSpec<Periodic, Periodic>;
    using YHi = WallSpec<Periodic, Periodic>;
    using ZLo = WallSpec<Neumann, NoSlip>;
    using ZHi = WallSpec<Anchoring<0.4, 0.3, 1.1>, MovingWall<0.02, -0.01, 0.0>>;
};
}  // namespace other

// The one-line mappings a binding adds for the empty tags.
template <> struct lbman3d::qbc_traits<other::Periodic> : lbman3d::qbc_traits<lbman3d::Periodic> {};
template <> struct lbman3d::qbc_traits<other::Neumann> : lbman3d::qbc_traits<lbman3d::Neumann> {};
template <> struct lbman3d::ubc_traits<other::Periodic> : lbman3d::ubc_traits<lbman3d::Periodic> {};
template <> struct lbman3d::ubc_traits<other::NoSlip> : lbman3d::ubc_traits<lbman3d::NoSlip> {};
template <> struct lbman3d::ubc_traits<other::SpecularReflection> : lbman3d::ubc_traits<lbman3d::SpecularReflection> {};

static int failures = 0;
#define EXPECT(c)                                                      \
    do {                                                               \
        if (!(c)) {                                                    \
            std::fprintf(stderr, "FAILED %s:%d %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                \
        }                                                              \
    } while (0)

template <typename BC>
static void same_as_preset(const char* name) {
    lbm_config want{};
    lbman3d::check(lbm_default_config(&want, 8, 9, 10));
    lbman3d::check(lbm_set_preset(&want, name));
    const lbm_config got = lbman3d::make_config<BC>(8, 9, 10);
    EXPECT(std::memcmp(&want, &got, sizeof(lbm_config)) == 0);
}

static int host_checks() {
    same_as_preset<lbman3d::FullyPeriodicConfig>("FullyPeriodic");
    same_as_preset<lbman3d::ChannelConfig>("Channel");
    same_as_preset<lbman3d::SlitConfig>("SlitFreeSlip");
    same_as_preset<lbman3d::SlitNoSlipConfig>("SlitNoSlip");
    const lbm_config c = lbman3d::make_config<other::LidConfig>(4, 4, 4);
    EXPECT(c.walls[LBM_XLO].qbc == LBM_Q_PERIODIC && c.walls[LBM_XLO].ubc == LBM_U_PERIODIC);
    EXPECT(c.walls[LBM_ZLO].qbc == LBM_Q_NEUMANN && c.walls[LBM_ZLO].ubc == LBM_U_NOSLIP);
    EXPECT(c.walls[LBM_ZHI].qbc == LBM_Q_ANCHORING && c.walls[LBM_ZHI].ubc == LBM_U_MOVING);
    EXPECT(c.walls[LBM_ZHI].anchor_s == 0.4 && c.walls[LBM_ZHI].anchor_theta == 0.3 &&
           c.walls[LBM_ZHI].anchor_phi == 1.1);
    EXPECT(c.walls[LBM_ZHI].wall_u[0] == 0.02 && c.walls[LBM_ZHI].wall_u[1] == -0.01 &&
           c.walls[LBM_ZHI].wall_u[2] == 0.0);
    const lbm_params p = lbman3d::default_params();
    EXPECT(p.tauf == 2.5 && p.L == 0.06 && p.alpha == 0.023 && p.lambda == 0.3);
    bool threw = false;
    try {
        lbman3d::check(lbm_set_preset(nullptr, "Channel"));
    } catch (const lbman3d::Error& e) {
        threw = e.code() == LBM_ERR_INVALID;
    }
    EXPECT(threw);
    std::printf("host checks: %d failures\n", failures);
    return failures ? 1 : 0;
}

static int run(const char* out) {
    const int nx = 12, ny = 10, nz = 8;
    const long V = (long)nx * ny * nz;
    lbman3d::DeviceFields df(lbman3d::make_config<lbman3d::ChannelConfig>(nx, ny, nz));
    lbman3d::HostFields h(V);
    lbman3d::check(lbm_seed_director_field(7, nx, ny, nz, 0.5, 0.01, h.q.data()));
    df.Initialize(h);
    lbman3d::DeviceSolver<lbman3d::ChannelConfig> solver;
    solver.Initialize(df);
    for (int i = 0; i < 10; ++i) {
        solver.QTensorStep(df);
        solver.LBMStep(df);
    }
    df.CopyToHost(h);
    bool threw = false;
    try {
        lbman3d::DeviceSolver<lbman3d::FullyPeriodicConfig> wrong;
        wrong.Initialize(df);
    } catch (const lbman3d::Error&) {
        threw = true;
    }
    EXPECT(threw);
    FILE* fp = std::fopen(out, "wb");
    if (!fp) return 2;
    std::fwrite(h.rho.data(), sizeof(double), V, fp);
    std::fwrite(h.ux.data(), sizeof(double), V, fp);
    std::fwrite(h.uy.data(), sizeof(double), V, fp);
    std::fwrite(h.uz.data(), sizeof(double), V, fp);
    std::fwrite(h.q.data(), sizeof(double), 5 * V, fp);
    std::fclose(fp);
    const lbm_diag d = df.Diagnostics();
    std::printf("run: mass %.17g umax %.6g, %d failures\n", d.mass, d.umax, failures);
    return failures ? 1 : 0;
}

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::string(argv[1]) == "host") return host_checks();
        if (argc >= 3 && std::string(argv[1]) == "run") return run(argv[2]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    std::fprintf(stderr, "usage: host_api_check host | run <out>\n");
    return 2;
}
