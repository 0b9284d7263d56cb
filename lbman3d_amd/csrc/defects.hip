// defects.hip — plaquettes pierced by a disclination, on the device.
//
// A half-integer disclination line pierces a plaquette of four neighbouring
// cells when the director, carried continuously around the plaquette (each
// next director sign-flipped if it points against the previous one), comes
// back antiparallel to where it started
// (DefectFinder::ComputeWindingNumbers, src/analysis/defect_finder.tpp:12-162;
// FlipN, src/analysis/defect_fields.cc:15-24: flip when n1 . n2 <= 0).
// The director is the one QtensorToOrderDirector gives, computed in FP64 by
// order_director_kernel. Built without FMA contraction so each dot product
// is the reference's sum of three rounded products.
#include "step_kernels.hpp"
#include "solver_state.hpp"

using namespace lbm;

namespace {

__device__ __forceinline__ double dot3(const double* __restrict__ d, long a, long b) {
    return d[3 * a] * d[3 * b] + d[3 * a + 1] * d[3 * b + 1] + d[3 * a + 2] * d[3 * b + 2];
}

// Loop c1 -> c2 -> c3 -> c4 -> c1. Negating a vector negates its dot
// products exactly, so the flips are tracked as signs.
__device__ __forceinline__ uint8_t pierced(const double* __restrict__ d, long c1, long c2, long c3, long c4) {
    const double s2 = dot3(d, c1, c2) <= 0.0 ? -1.0 : 1.0;
    const double s3 = s2 * dot3(d, c2, c3) <= 0.0 ? -1.0 : 1.0;
    const double s4 = s3 * dot3(d, c3, c4) <= 0.0 ? -1.0 : 1.0;
    return s4 * dot3(d, c4, c1) <= 0.0 ? 1 : 0;
}

// family 0: plaquettes normal to x, loop (y,z) (y+1,z) (y+1,z+1) (y,z+1);
// family 1: normal to y, loop (x,z) (x+1,z) (x+1,z+1) (x,z+1);
// family 2: normal to z, loop (x,y) (x+1,y) (x+1,y+1) (x,y+1).
__global__ void __launch_bounds__(256)
defect_faces_kernel(const double* __restrict__ dir, int family, int nx, int ny, int nz, uint8_t* __restrict__ out) {
    const int ex = family == 0 ? nx : nx - 1;
    const int ey = family == 1 ? ny : ny - 1;
    const int ez = family == 2 ? nz : nz - 1;
    const long n = (long)ex * ey * ez;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % ex), y = (int)((i / ex) % ey), z = (int)(i / ((long)ex * ey));
    auto cell = [&](int a, int b, int c) { return ((long)c * ny + b) * nx + a; };
    uint8_t v;
    if (family == 0)
        v = pierced(dir, cell(x, y, z), cell(x, y + 1, z), cell(x, y + 1, z + 1), cell(x, y, z + 1));
    else if (family == 1)
        v = pierced(dir, cell(x, y, z), cell(x + 1, y, z), cell(x + 1, y, z + 1), cell(x, y, z + 1));
    else
        v = pierced(dir, cell(x, y, z), cell(x + 1, y, z), cell(x + 1, y + 1, z), cell(x, y + 1, z));
    out[i] = v;
}

}  // namespace

extern "C" int lbm_defect_faces(lbm_handle s, uint8_t* def_x, uint8_t* def_y, uint8_t* def_z) {
    if (!s || !def_x || !def_y || !def_z) return fail(LBM_ERR_INVALID, "null argument");
    const int nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    if (nx < 2 || ny < 2 || nz < 2) return fail(LBM_ERR_INVALID, "defect faces need at least 2 cells per axis");
    const long V = s->V;
    const long nf[3] = {(long)nx * (ny - 1) * (nz - 1), (long)(nx - 1) * ny * (nz - 1), (long)(nx - 1) * (ny - 1) * nz};
    void* tmp = nullptr;
    if (hipMalloc(&tmp, 4 * V * sizeof(double) + (nf[0] + nf[1] + nf[2])) != hipSuccess)
        return fail(LBM_ERR_HIP, "hipMalloc failed");
    double* order = (double*)tmp;
    double* dir = order + V;
    uint8_t* faces = (uint8_t*)(dir + 3 * V);
    if (s->fp64) launch_order_director(s->d64, V, order, dir, s->stream);
    else launch_order_director(s->d32, V, order, dir, s->stream);
    uint8_t* host[3] = {def_x, def_y, def_z};
    long o = 0;
    for (int fam = 0; fam < 3; ++fam) {
        hipLaunchKernelGGL(defect_faces_kernel, dim3((unsigned)((nf[fam] + 255) / 256)), dim3(256), 0, s->stream, dir,
                           fam, nx, ny, nz, faces + o);
        o += nf[fam];
    }
    int r = LBM_OK;
    if (hipGetLastError() != hipSuccess) r = fail(LBM_ERR_HIP, "defect kernel launch failed");
    o = 0;
    for (int fam = 0; fam < 3 && r == LBM_OK; ++fam) {
        if (hipMemcpyAsync(host[fam], faces + o, nf[fam], hipMemcpyDeviceToHost, s->stream) != hipSuccess)
            r = fail(LBM_ERR_HIP, "defect face download failed");
        o += nf[fam];
    }
    if (r == LBM_OK && hipStreamSynchronize(s->stream) != hipSuccess) r = fail(LBM_ERR_HIP, "synchronize failed");
    (void)hipFree(tmp);
    return r;
}
