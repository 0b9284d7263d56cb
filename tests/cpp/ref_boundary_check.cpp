// ref_boundary_check.cpp — the binding of include/lbman3d.hpp compiled against
// the reference's own boundary, parameter and scheme headers.
//
// Built only where the reference sources are present, with
//   g++ -std=c++20 -I<reference>/src -I<repo>/include ref_boundary_check.cpp
// so that "boundary.h", "sim_config.h" and "params.h" below are the
// reference's files (src/boundary.h:31-129, src/sim_config.h:19-99,
// src/params.h:14-112), not stand-ins. The five trait lines and
// ReferenceParams() are the binding a maintainer adds, as written in the
// integration guide. Host-only: no GPU call. Exits non-zero on a mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "boundary.h"      // the reference's tags and presets (global namespace)
#include "params.h"        // namespace Params
#include "sim_config.h"    // SlitConfig, SlitNoSlipConfig, SimBC, kQAdvection

#include "lbman3d.hpp"

// ---- the binding ------------------------------------------------------------
template <> struct lbman3d::qbc_traits<::Periodic> : lbman3d::qbc_traits<lbman3d::Periodic> {};
template <> struct lbman3d::qbc_traits<::Neumann>  : lbman3d::qbc_traits<lbman3d::Neumann> {};
template <> struct lbman3d::ubc_traits<::Periodic> : lbman3d::ubc_traits<lbman3d::Periodic> {};
template <> struct lbman3d::ubc_traits<::NoSlip>   : lbman3d::ubc_traits<lbman3d::NoSlip> {};
template <> struct lbman3d::ubc_traits<::SpecularReflection>
    : lbman3d::ubc_traits<lbman3d::SpecularReflection> {};

inline lbm_params ReferenceParams() {
    lbm_params p{};
    p.dt = Params::DT; p.density = Params::kDensity; p.tauf = Params::TAUF;
    p.L = Params::L; p.A = Params::A; p.B = Params::B; p.C = Params::C;
    p.lambda = Params::LAMBDA; p.gamma = Params::GAMMA;
    p.alpha = Params::ALPHA; p.mu = Params::MU;
    return p;
}

inline int ReferenceAdvection() { return kQAdvection == Advection::Upwind ? LBM_ADV_UPWIND : LBM_ADV_CENTRED; }

// ---- checks -----------------------------------------------------------------
static int failures = 0;
#define EXPECT(c)                                                               \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::fprintf(stderr, "FAILED %s:%d %s\n", __FILE__, __LINE__, #c);  \
            ++failures;                                                         \
        }                                                                       \
    } while (0)

// A reference preset lowers to exactly the config that lbm_set_preset builds
// from its name (compared byte for byte, constants included).
template <typename BC>
static void same_as_preset(const char* preset) {
    lbm_config want{};
    lbman3d::check(lbm_default_config(&want, 20, 100, 100));
    lbman3d::check(lbm_set_preset(&want, preset));
    const lbm_config got = lbman3d::make_config<BC>(20, 100, 100);
    EXPECT(std::memcmp(&want, &got, sizeof(lbm_config)) == 0);
    EXPECT(BC::name == std::string_view(preset));
}

// A custom reference BCConfig with every tag kind, templated walls included.
using LidAndAnchors = ::BCConfig<::WallSpec<::Anchoring<0.5, 0.7853981633974483, 0.0>, ::NoSlip>,
                                 ::WallSpec<::Neumann, ::SpecularReflection>,
                                 ::WallSpec<::Periodic, ::Periodic>, ::WallSpec<::Periodic, ::Periodic>,
                                 ::WallSpec<::Neumann, ::NoSlip>,
                                 ::WallSpec<::Anchoring<0.4, 0.3, 1.1>, ::MovingWall<0.02, -0.01, 0.0>>>;

int main() {
    same_as_preset<::FullyPeriodicConfig>("FullyPeriodic");
    same_as_preset<::PeriodicBC>("FullyPeriodic");
    same_as_preset<::ChannelConfig>("Channel");
    same_as_preset<::SlitConfig>("SlitFreeSlip");
    same_as_preset<::SlitNoSlipConfig>("SlitNoSlip");
    same_as_preset<::SimBC>(std::string(::SimBC::name).c_str());

    // The default-constructed reference BCConfig: Neumann/NoSlip on all faces.
    const lbm_config d = lbman3d::make_config<::BCConfig<>>(4, 5, 6);
    for (int f = 0; f < 6; ++f) EXPECT(d.walls[f].qbc == LBM_Q_NEUMANN && d.walls[f].ubc == LBM_U_NOSLIP);

    const lbm_config c = lbman3d::make_config<LidAndAnchors>(4, 5, 6);
    EXPECT(c.walls[LBM_XLO].qbc == LBM_Q_ANCHORING && c.walls[LBM_XLO].ubc == LBM_U_NOSLIP);
    EXPECT(c.walls[LBM_XLO].anchor_s == 0.5 && c.walls[LBM_XLO].anchor_theta == 0.7853981633974483 &&
           c.walls[LBM_XLO].anchor_phi == 0.0);
    EXPECT(c.walls[LBM_XHI].qbc == LBM_Q_NEUMANN && c.walls[LBM_XHI].ubc == LBM_U_SPECULAR);
    EXPECT(c.walls[LBM_YLO].qbc == LBM_Q_PERIODIC && c.walls[LBM_YLO].ubc == LBM_U_PERIODIC);
    EXPECT(c.walls[LBM_YHI].qbc == LBM_Q_PERIODIC && c.walls[LBM_YHI].ubc == LBM_U_PERIODIC);
    EXPECT(c.walls[LBM_ZLO].qbc == LBM_Q_NEUMANN && c.walls[LBM_ZLO].ubc == LBM_U_NOSLIP);
    EXPECT(c.walls[LBM_ZHI].qbc == LBM_Q_ANCHORING && c.walls[LBM_ZHI].ubc == LBM_U_MOVING);
    EXPECT(c.walls[LBM_ZHI].anchor_s == 0.4 && c.walls[LBM_ZHI].anchor_theta == 0.3 &&
           c.walls[LBM_ZHI].anchor_phi == 1.1);
    EXPECT(c.walls[LBM_ZHI].wall_u[0] == 0.02 && c.walls[LBM_ZHI].wall_u[1] == -0.01 &&
           c.walls[LBM_ZHI].wall_u[2] == 0.0);
    // The reference's own type traits agree with the lowering.
    static_assert(is_anchoring_v<LidAndAnchors::XLo::QBC> && is_moving_wall_v<LidAndAnchors::ZHi::UBC>);
    static_assert(periodicity_by_axis<::SlitConfig>[0] == 0 && periodicity_by_axis<::SlitConfig>[1] == 1);

    // params.h lowers to the library defaults (the same namespace Params values).
    lbm_config def{};
    lbman3d::check(lbm_default_config(&def, 1, 1, 1));
    const lbm_params p = ReferenceParams();
    EXPECT(std::memcmp(&p, &def.p, sizeof(lbm_params)) == 0);
    EXPECT(ReferenceAdvection() == LBM_ADV_CENTRED);   // sim_config.h:99 selects Centred

    // The config is accepted as it is by the library's host-side validation.
    uint16_t table[64 * 15];
    EXPECT(lbm_debug_slot_table(&c, table) == LBM_OK);

    if (failures) std::fprintf(stderr, "%d check(s) failed\n", failures);
    else std::printf("reference boundary binding: all checks passed\n");
    return failures ? 1 : 0;
}
