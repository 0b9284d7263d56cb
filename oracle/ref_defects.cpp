// ref_defects.cpp — drives the reference's own disclination finder
// (QtensorToOrderDirector, src/analysis_fields.cc; DefectFinder<BC>,
// src/analysis/defect_finder.tpp) compiled from its sources under
// /root/reference/src by oracle/build_ref.sh. TEST INFRASTRUCTURE ONLY: its
// outputs pin the host graph code (lbman3d_amd/defects.py), the numpy
// restatement of the winding test (oracle/defects_ref.py) and, through the
// committed fixture tests/golden/defects_ref.npz, the device kernel
// (lbm_defect_faces). The grid comes from the params.h that
// oracle/build_ref.sh puts first on the include path.
//
// usage: ref_defects <bc> <in.bin> <out.bin>
//   bc:     FullyPeriodic | Channel
//   in.bin: int32 nx, ny, nz; Q (5V doubles: Qxx, Qxy, Qxz, Qyy, Qyz, cell = (z*ny + y)*nx + x)
//   out.bin: order (V doubles), director (3V doubles, [cell][3]), def_x, def_y,
//            def_z (bytes), int64 path count, then per path int64 point count
//            and 3 doubles per point (DefectFields::disclinations, in the
//            order IsolateDisclinationsFromGraph produced them)

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <params.h>
#include <sim_config.h>
#include "analysis_fields.h"
#include "boundary.h"
#include "qtensor_fields.h"
#include "analysis/defect_finder.h"

namespace {

template <typename BC>
int run(const char* in, const char* out) {
    FILE* fp = std::fopen(in, "rb");
    if (!fp) { std::perror(in); return 2; }
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, fp) != 3) return 3;
    if (hdr[0] != Params::nx || hdr[1] != Params::ny || hdr[2] != Params::nz) {
        std::fprintf(stderr, "grid %dx%dx%d does not match compiled params %dx%dx%d\n", hdr[0], hdr[1], hdr[2],
                     Params::nx, Params::ny, Params::nz);
        return 4;
    }
    LocalGrid g = LocalGrid::SingleRank();
    const size_t V = g.Volume();
    QTensorFields qf(g);
    std::vector<double>* qc[5] = {&qf.qxx, &qf.qxy, &qf.qxz, &qf.qyy, &qf.qyz};
    for (int k = 0; k < 5; ++k)
        if (std::fread(qc[k]->data(), sizeof(double), V, fp) != V) return 3;
    std::fclose(fp);

    AnalysisFields af(g);
    QtensorToOrderDirector(qf, af);
    DefectFields df;
    DefectFinder<BC> finder;
    finder.FindDefects(af, df);

    FILE* fo = std::fopen(out, "wb");
    if (!fo) { std::perror(out); return 2; }
    std::fwrite(af.order_.data(), sizeof(double), V, fo);
    std::fwrite(af.director_.data(), sizeof(double), 3 * V, fo);
    std::fwrite(df.def_x.data(), 1, df.def_x.size(), fo);
    std::fwrite(df.def_y.data(), 1, df.def_y.size(), fo);
    std::fwrite(df.def_z.data(), 1, df.def_z.size(), fo);
    const int64_t np = (int64_t)df.disclinations.size();
    std::fwrite(&np, sizeof(np), 1, fo);
    for (const Disclination& d : df.disclinations) {
        const int64_t n = (int64_t)(d.points.size() / 3);
        std::fwrite(&n, sizeof(n), 1, fo);
        std::fwrite(d.points.data(), sizeof(double), d.points.size(), fo);
    }
    std::fclose(fo);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <FullyPeriodic|Channel> <in.bin> <out.bin>\n", argv[0]);
        return 1;
    }
    const std::string bc = argv[1];
    if (bc == "FullyPeriodic") return run<FullyPeriodicConfig>(argv[2], argv[3]);
    if (bc == "Channel") return run<ChannelConfig>(argv[2], argv[3]);
    std::fprintf(stderr, "unknown boundary config %s\n", bc.c_str());
    return 1;
}
