// ref_harness.cpp — drives the reference's own CPU solver core (compiled from
// its sources under /root/reference/src by oracle/build_ref.sh) on a state read
// from a file, and writes the resulting state. TEST INFRASTRUCTURE ONLY: its
// outputs pin the oracle (oracle/lbm_oracle.cpp) and become the committed
// fixtures under tests/golden/. Physical constants come from the params.h
// that oracle/build_ref.sh puts first on the include path, exactly as the
// reference's tests shadow it (tests/CMakeLists.txt:3-9).
//
// usage: ref_<scenario> <bc> <mode> <force_model> <steps> <in.bin> <out.bin> [warmup]
//   bc:    FullyPeriodic | Channel | SlitFreeSlip | SlitNoSlip | PoiseuilleY |
//          SpecularZ | AnchoredZ | MovingLidZ | SpecularXNoSlipY
//   mode:  full (Q step, then LBM step) | lbm (LBM only) | q (Q only)
//   force_model: 0 = full body force, 1 = Ericksen only (empty
//          SetActiveStressAndComputeBodyForce override)
// in.bin : int32 nx, ny, nz, has_fields, has_force; then Q (5V doubles),
//          [rho, ux, uy, uz (4V)], [fx, fy, fz (3V)]
// out.bin: rho, ux, uy, uz (4V), Q (5V), F (3V), f (15V, i*V + cell),
//          free energy (1), SimIO::Log sums (8); "-" writes nothing (timing)
// warmup:  full steps run before the timed loop (timing runs; the protocol
//          of src/benchmark.cc:9-50: kWarmupSteps = 10 untimed steps, then
//          the loop timed per step with a QTensor / LBM split)

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numbers>
#include <string>
#include <vector>

#include <params.h>
#include <sim_config.h>
#include "boundary.h"
#include "lbm_solver.h"
#include "qtensor_solver.h"
#include "analysis_fields.h"

namespace harness {

struct PoiseuilleY {
    using XLo = WallSpec<Periodic, Periodic>;
    using XHi = WallSpec<Periodic, Periodic>;
    using YLo = WallSpec<Neumann, NoSlip>;
    using YHi = WallSpec<Neumann, NoSlip>;
    using ZLo = WallSpec<Periodic, Periodic>;
    using ZHi = WallSpec<Periodic, Periodic>;
    static constexpr std::string_view name = "PoiseuilleY";
};
struct SpecularZ {
    using XLo = WallSpec<Periodic, Periodic>;
    using XHi = WallSpec<Periodic, Periodic>;
    using YLo = WallSpec<Periodic, Periodic>;
    using YHi = WallSpec<Periodic, Periodic>;
    using ZLo = WallSpec<Neumann, SpecularReflection>;
    using ZHi = WallSpec<Neumann, SpecularReflection>;
    static constexpr std::string_view name = "SpecularZ";
};
struct AnchoredZ {
    using XLo = WallSpec<Periodic, Periodic>;
    using XHi = WallSpec<Periodic, Periodic>;
    using YLo = WallSpec<Periodic, Periodic>;
    using YHi = WallSpec<Periodic, Periodic>;
    using ZLo = WallSpec<Anchoring<0.5, std::numbers::pi / 4.0, 0.0>, NoSlip>;
    using ZHi = WallSpec<Anchoring<0.5, std::numbers::pi / 4.0, 0.0>, NoSlip>;
    static constexpr std::string_view name = "AnchoredZ";
};
struct MovingLidZ {
    using XLo = WallSpec<Periodic, Periodic>;
    using XHi = WallSpec<Periodic, Periodic>;
    using YLo = WallSpec<Periodic, Periodic>;
    using YHi = WallSpec<Periodic, Periodic>;
    using ZLo = WallSpec<Neumann, NoSlip>;
    using ZHi = WallSpec<Anchoring<0.4, 0.3, 1.1>, MovingWall<0.02, -0.01, 0.0>>;
    static constexpr std::string_view name = "MovingLidZ";
};
struct SpecularXNoSlipY {
    using XLo = WallSpec<Neumann, SpecularReflection>;
    using XHi = WallSpec<Neumann, SpecularReflection>;
    using YLo = WallSpec<Neumann, NoSlip>;
    using YHi = WallSpec<Neumann, NoSlip>;
    using ZLo = WallSpec<Periodic, Periodic>;
    using ZHi = WallSpec<Periodic, Periodic>;
    static constexpr std::string_view name = "SpecularXNoSlipY";
};

std::vector<double> read_block(FILE* fp, size_t n) {
    std::vector<double> v(n);
    if (std::fread(v.data(), sizeof(double), n, fp) != n) { std::fprintf(stderr, "short read\n"); std::exit(3); }
    return v;
}

template <typename BC>
int run(const std::string& mode, int force_model, int steps, const char* in, const char* out, int warmup) {
    FILE* fp = std::fopen(in, "rb");
    if (!fp) { std::perror(in); return 2; }
    int hdr[5];
    if (std::fread(hdr, sizeof(int), 5, fp) != 5) return 3;
    if (hdr[0] != Params::nx || hdr[1] != Params::ny || hdr[2] != Params::nz) {
        std::fprintf(stderr, "grid %dx%dx%d does not match compiled params %dx%dx%d\n", hdr[0], hdr[1], hdr[2],
                     Params::nx, Params::ny, Params::nz);
        return 4;
    }
    LocalGrid g = LocalGrid::SingleRank();
    const size_t V = g.Volume();
    FluidFields ff(g);
    QTensorFields qf(g);
    std::vector<double> q = read_block(fp, 5 * V);
    if (hdr[3]) {
        auto r = read_block(fp, 4 * V);
        std::memcpy(ff.rho.data(), r.data(), V * sizeof(double));
        std::memcpy(ff.ux.data(), r.data() + V, V * sizeof(double));
        std::memcpy(ff.uy.data(), r.data() + 2 * V, V * sizeof(double));
        std::memcpy(ff.uz.data(), r.data() + 3 * V, V * sizeof(double));
    }
    if (hdr[4]) {
        auto r = read_block(fp, 3 * V);
        std::memcpy(ff.fx.data(), r.data(), V * sizeof(double));
        std::memcpy(ff.fy.data(), r.data() + V, V * sizeof(double));
        std::memcpy(ff.fz.data(), r.data() + 2 * V, V * sizeof(double));
    }
    std::fclose(fp);
    std::vector<double>* qc[5] = {&qf.qxx, &qf.qxy, &qf.qxz, &qf.qyy, &qf.qyz};
    for (int k = 0; k < 5; ++k) std::memcpy(qc[k]->data(), q.data() + k * V, V * sizeof(double));

    LbmSolver<BC> lbm;
    QTensorSolver<BC> qs;
    lbm.Initialize(ff);

    using clock = std::chrono::steady_clock;
    auto q_step = [&] {
        if (mode == "full" || mode == "q") {
            qs.StepAndSetupBodyForce(qf, ff);
            if (force_model == 0) qs.SetActiveStressAndComputeBodyForce(ff, qf);
        }
    };
    auto l_step = [&] {
        if (mode == "full" || mode == "lbm") lbm.LatticeBoltzmannStep(ff);
    };
    for (int t = 0; t < warmup; ++t) {
        q_step();
        l_step();
    }
    double ms_q = 0.0, ms_l = 0.0;
    for (int t = 0; t < steps; ++t) {
        const auto t0 = clock::now();
        q_step();
        const auto t1 = clock::now();
        l_step();
        const auto t2 = clock::now();
        ms_q += std::chrono::duration<double, std::milli>(t1 - t0).count();
        ms_l += std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
    const double ms = ms_q + ms_l;
    std::fprintf(stderr,
                 "ref: %d steps of %zu cells in %.3f ms (%.3f MLUPS, %d threads; %d warm-up steps; "
                 "QTensor %.3f ms/step, LBM %.3f ms/step)\n",
                 steps, V, ms, steps > 0 ? (double)V * steps / (ms * 1e3) : 0.0, (int)Params::kNumOMPThreads, warmup,
                 steps > 0 ? ms_q / steps : 0.0, steps > 0 ? ms_l / steps : 0.0);

    if (std::string(out) == "-") return 0;   // timing run: no output file
    const double fe = TotalNematicFreeEnergy<BC>(qf);
    // SimIO::Log sums without the I/O (src/sim_io.cc:78-109), previous velocity zero.
    double sums[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double umax2 = 0.0;
    for (size_t c = 0; c < V; ++c) {
        const double r = ff.rho[c], ux = ff.ux[c], uy = ff.uy[c], uz = ff.uz[c];
        sums[0] += r; sums[1] += r * ux; sums[2] += r * uy; sums[3] += r * uz;
        const double u2 = ux * ux + uy * uy + uz * uz;
        sums[4] += 0.5 * r * u2;
        if (u2 > umax2) umax2 = u2;
        sums[5] += ux * ux + uy * uy + uz * uz;
        sums[6] += u2;
    }
    sums[7] = std::sqrt(umax2);

    FILE* fo = std::fopen(out, "wb");
    if (!fo) { std::perror(out); return 2; }
    auto put = [&](const std::vector<double>& v) { std::fwrite(v.data(), sizeof(double), V, fo); };
    put(ff.rho); put(ff.ux); put(ff.uy); put(ff.uz);
    for (int k = 0; k < 5; ++k) put(*qc[k]);
    put(ff.fx); put(ff.fy); put(ff.fz);
    std::vector<double> fi(V);
    for (int i = 0; i < Lattice::ndir; ++i) {
        for (size_t c = 0; c < V; ++c) fi[c] = ff.f[c * Lattice::ndir + i];
        put(fi);
    }
    std::fwrite(&fe, sizeof(double), 1, fo);
    std::fwrite(sums, sizeof(double), 8, fo);
    std::fclose(fo);
    return 0;
}

}  // namespace harness

int main(int argc, char** argv) {
    if (argc != 7 && argc != 8) {
        std::fprintf(stderr, "usage: %s <bc> <full|lbm|q> <force_model> <steps> <in.bin> <out.bin> [warmup]\n",
                     argv[0]);
        return 1;
    }
    const std::string bc = argv[1], mode = argv[2];
    const int fm = std::atoi(argv[3]), steps = std::atoi(argv[4]);
    const int warmup = argc == 8 ? std::atoi(argv[7]) : 0;
    const char *in = argv[5], *out = argv[6];
    using namespace harness;
    if (bc == "FullyPeriodic") return run<FullyPeriodicConfig>(mode, fm, steps, in, out, warmup);
    if (bc == "Channel") return run<ChannelConfig>(mode, fm, steps, in, out, warmup);
    if (bc == "SlitFreeSlip") return run<SlitConfig>(mode, fm, steps, in, out, warmup);
    if (bc == "SlitNoSlip") return run<SlitNoSlipConfig>(mode, fm, steps, in, out, warmup);
    if (bc == "PoiseuilleY") return run<PoiseuilleY>(mode, fm, steps, in, out, warmup);
    if (bc == "SpecularZ") return run<SpecularZ>(mode, fm, steps, in, out, warmup);
    if (bc == "AnchoredZ") return run<AnchoredZ>(mode, fm, steps, in, out, warmup);
    if (bc == "MovingLidZ") return run<MovingLidZ>(mode, fm, steps, in, out, warmup);
    if (bc == "SpecularXNoSlipY") return run<SpecularXNoSlipY>(mode, fm, steps, in, out, warmup);
    std::fprintf(stderr, "unknown bc %s\n", bc.c_str());
    return 1;
}
