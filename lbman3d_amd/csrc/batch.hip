// batch.hip — many independent FP32 simulations advanced by one persistent
// kernel launch per step, optionally replayed from a hipGraph.
//
// The reference can run exactly one compile-time configuration per binary
// (src/params.h, src/device_solver.cu:123-125); here every simulation of a
// batch has its own grid, boundary tags and constants. Each simulation is an
// ordinary FP32 handle (own HBM buffers, slot table, diagnostics); the batch
// adds a device array of per-simulation descriptors and a work list of
// (simulation, tile) items that a persistent grid of blocks walks each step.
// A descriptor holds both ping-pong views of the simulation's buffers, so two
// consecutive steps differ only in a parity argument and can be captured once
// as a graph and replayed.
//
// Simulations with walls in x and y only (z periodic, no anchoring, no
// moving wall; the slit and channel presets) run the third pipeline's
// wall-aware form, which computes every cell in the one launch. Other
// simulations with walls take the split of the single-simulation path: the
// periodic specialisation over every tile is exact for the cells two or more
// cells from every open face, and two one-thread-per-cell passes with the
// full wall handling then overwrite their wall shells.
#define LBM_FUSED_BODY_ONLY
#include "step_kernels.hpp"
#include "fused_step.hpp"
#include "solver_state.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

using namespace lbm;

namespace {

struct SimDesc {
    Fields<float> f[2];   // [parity]: parity 1 swaps Q/Q', u/u', g/g'
    BCDesc<float> bc;
    Phys<float> k;
    int periodic;
    int ntx, nty, nzc;
    int tile_begin;       // first work item of this simulation
    int pull;             // 0 until the first step after an upload
    int aux_always;       // moving walls read rho every step
    int walls;            // third pipeline with its wall handling (no shell passes)
    ShellRegion r3, r2;   // shell and halo (Q pass), shell alone (force pass); empty if periodic or walls
    int sq_begin;         // first item of this simulation in the shell Q pass
    int sf_begin;         // first item of this simulation in the shell force pass
};

// Last simulation whose range starts at or before item w (ranges ascending;
// empty ranges share their start with the next simulation).
template <typename Get>
__device__ __forceinline__ int find_sim(int nsims, int w, Get begin_of) {
    int lo = 0, hi = nsims - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin_of(mid) <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Tile shapes of the batched kernel: kind 0 runs the first fused pipeline
// over every tile (cells of the wall shells excluded, the shell passes write
// them), kind 1 the third pipeline (one barrier per plane; simulations with
// walls in x and y only are computed whole by its wall-aware form, others
// with walls have their shells rewritten by the shell passes, as on a single
// simulation).
struct BatchTile {
    int kind, tx, ty, zc;
    const char* name;
};
constexpr BatchTile kBatchTiles[] = {
    {0, 32, 8, 32, "32x8x32"},       {0, 16, 16, 32, "16x16x32"},     {0, 32, 8, 16, "32x8x16"},
    {0, 16, 16, 16, "16x16x16"},     {0, 16, 16, 8, "16x16x8"},       {0, 32, 8, 8, "32x8x8"},
    {1, 32, 8, 64, "p2:32x8x64"},    {1, 32, 8, 32, "p2:32x8x32"},    {1, 32, 8, 16, "p2:32x8x16"},
    {1, 64, 4, 32, "p2:64x4x32"},
};
constexpr int kNumBatchTiles = sizeof(kBatchTiles) / sizeof(kBatchTiles[0]);

template <int TX, int TY, int ZC>
__global__ void __launch_bounds__(TX * TY)
batch_step_kernel(const SimDesc* __restrict__ descs, int nsims, int total, int parity, int aux) {
    using G = TileGeom<TX, TY>;
    __shared__ float sIn[kNumIn][G::NI];
    __shared__ float sA[2][kNumA][G::NA];
    __shared__ float sRing[3][kNumIn][G::NR];
    // Persistent blocks: every block walks the work list with a grid stride;
    // the exit condition (w >= total) is reached by every block.
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        const SimDesc& d = descs[find_sim(nsims, w, [&](int i) { return descs[i].tile_begin; })];
        const int local = w - d.tile_begin;
        const int tx = local % d.ntx, ty = (local / d.ntx) % d.nty, tz = local / (d.ntx * d.nty);
        const int a = aux | d.aux_always;
        fused_body<TX, TY, ZC, true>(d.f[parity], d.bc, d.k, d.pull, a, tx, ty, tz, sIn, sA, sRing,
                                     d.periodic ? 0 : 1);
        __syncthreads();
    }
}

// One work item of the third-pipeline kernel: planes [z0, z1) of column
// tile txy (= tx + ty * ntx) of simulation sim.
struct BatchItem {
    int sim, txy, z0, z1;
};

// The work list on the third pipeline, as an explicit item list that the
// host orders longest first (see build_items). A simulation's advection
// scheme picks the instantiation per work item (uniform over the block).
template <int TX, int TY, int ZC>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(2, 2)))
batch_step_kernel_p2(const SimDesc* __restrict__ descs, const BatchItem* __restrict__ items, int total,
                     int parity, int aux, unsigned* __restrict__ queue) {
    using G = TileGeom<TX, TY>;
    __shared__ float sIn[2][kNumIn][G::NI];
    __shared__ float sP[2][6][G::NA];
    __shared__ float sRing[4][kNumIn][G::NR];
    __shared__ int s_item;
    // Work queue: blocks take the next item as they finish one, which
    // balances tiles of uneven length. queue[0] (next item) only grows, so
    // every block leaves once it passes total; the last block to leave
    // (queue[1] counts them) resets both for the next launch on the stream.
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        const int w = s_item;
        if (w >= total) break;
        const BatchItem it = items[w];
        const SimDesc& d = descs[it.sim];
        const int tx = it.txy % d.ntx, ty = it.txy / d.ntx;
        const int a = aux | d.aux_always;
        if (d.walls) {
            if (d.k.advection != 0)
                p2_range<TX, TY, false, true, true, false, true>(d.f[parity], d.bc, d.k, d.pull, a, tx, ty, it.z0,
                                                                 it.z1, sIn, sP, sRing);
            else
                p2_range<TX, TY, false, false, true, false, true>(d.f[parity], d.bc, d.k, d.pull, a, tx, ty, it.z0,
                                                                  it.z1, sIn, sP, sRing);
        } else if (d.k.advection != 0) {
            p2_range<TX, TY, true, true, true>(d.f[parity], d.bc, d.k, d.pull, a, tx, ty, it.z0, it.z1, sIn, sP,
                                               sRing);
        } else {
            p2_range<TX, TY, true, false, true>(d.f[parity], d.bc, d.k, d.pull, a, tx, ty, it.z0, it.z1, sIn, sP,
                                                sRing);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

// Wall shells of every simulation with walls, one thread per cell, in the
// two passes of the single-simulation path: the Q update over each shell and
// its one-cell halo, then body force and collide-stream of the shell cells.
__global__ void __launch_bounds__(kBlock)
batch_shell_q_kernel(const SimDesc* __restrict__ descs, int nsims, int total, int parity) {
    const int w = blockIdx.x * kBlock + threadIdx.x;
    if (w >= total) return;
    const SimDesc& d = descs[find_sim(nsims, w, [&](int i) { return descs[i].sq_begin; })];
    int q[3];
    region_cell(d.r3, d.bc.n, w - d.sq_begin, q);
    const Pos p{q[0], q[1], q[2]};
    const long c = cell_of(p.x, p.y, p.z, d.bc.n[0], d.bc.n[1]);
    qstep_cell(d.f[parity], d.bc, d.k, c, p, !cell_mask(d.bc, 1, p.x, p.y).writes(p.z));
}

__global__ void __launch_bounds__(kBlock)
batch_shell_force_kernel(const SimDesc* __restrict__ descs, int nsims, int total, int parity, int aux) {
    const int w = blockIdx.x * kBlock + threadIdx.x;
    if (w >= total) return;
    const SimDesc& d = descs[find_sim(nsims, w, [&](int i) { return descs[i].sf_begin; })];
    int q[3];
    region_cell(d.r2, d.bc.n, w - d.sf_begin, q);
    const Pos p{q[0], q[1], q[2]};
    Fields<float> fq = d.f[parity];
    for (int m = 0; m < 5; ++m) fq.q[m] = fq.qn[m];   // the force reads Q^{n+1}
    const int flags = FL_FORCE | FL_LBM | ((aux | d.aux_always) ? (FL_STORE_F | FL_STORE_RHO) : 0);
    force_lbm_cell<float, true, true>(fq, d.bc, d.k, cell_of(p.x, p.y, p.z, d.bc.n[0], d.bc.n[1]), p, flags,
                                      d.pull);
}

template <int KIND, int TX, int TY, int ZC>
int occupancy_blocks() {
    int nb = 0;
    hipError_t e;
    if constexpr (KIND == 0)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_step_kernel<TX, TY, ZC>, TX * TY, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_step_kernel_p2<TX, TY, ZC>, TX * TY, 0);
    return e == hipSuccess ? std::max(1, nb) : 1;
}

template <int KIND, int TX, int TY, int ZC>
void launch_batch(const SimDesc* d, int nsims, int total, const BatchItem* items, int parity, int aux, int grid,
                  unsigned* queue, hipStream_t s) {
    if constexpr (KIND == 0)
        hipLaunchKernelGGL((batch_step_kernel<TX, TY, ZC>), dim3(grid), dim3(TX * TY), 0, s, d, nsims, total,
                           parity, aux);
    else
        hipLaunchKernelGGL((batch_step_kernel_p2<TX, TY, ZC>), dim3(grid), dim3(TX * TY), 0, s, d, items, total,
                           parity, aux, queue);
}

// Every batch tile shape, in catalogue order: fn(TileTag<KIND, TX, TY, ZC>{}).
static_assert(kNumBatchTiles == 10, "the dispatch below lists every catalogue entry");
template <int K, int X, int Y, int Z>
struct TileTag {
    static constexpr int kind = K, tx = X, ty = Y, zc = Z;
};
template <typename F>
auto with_batch_tile(int tile, F&& fn) {
    switch (tile) {
        case 0: return fn(TileTag<0, 32, 8, 32>{});
        case 1: return fn(TileTag<0, 16, 16, 32>{});
        case 2: return fn(TileTag<0, 32, 8, 16>{});
        case 3: return fn(TileTag<0, 16, 16, 16>{});
        case 4: return fn(TileTag<0, 16, 16, 8>{});
        case 5: return fn(TileTag<0, 32, 8, 8>{});
        case 6: return fn(TileTag<1, 32, 8, 64>{});
        case 7: return fn(TileTag<1, 32, 8, 32>{});
        case 8: return fn(TileTag<1, 32, 8, 16>{});
        default: return fn(TileTag<1, 64, 4, 32>{});
    }
}

}  // namespace

struct lbm_batch {
    std::vector<lbm_handle> sims;
    std::vector<SimDesc> desc;
    SimDesc* d_desc = nullptr;
    unsigned* d_queue = nullptr;   // work-queue counters of the third-pipeline kernel (reset by the kernel)
    BatchItem* d_items = nullptr;  // its item list (third-pipeline tiles only)
    int items_cap = 0;
    std::vector<BatchItem> items;  // host copy (the upload reads it asynchronously)
    hipStream_t stream = nullptr;
    int total = 0;
    int sq_total = 0;      // cells of the shell Q pass (simulations with walls)
    int sf_total = 0;      // cells of the shell force pass
    int parity = 0;        // how far the buffers have ping-ponged since the handles were synced
    int tile = 0;
    int grid = 0;
    int grid_limit = 0;    // cap on the persistent grid (0: as many blocks as fit at once)
    bool desc_dirty = true;
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // two steps starting at parity 0 / 1
};

namespace {

void drop_graphs(lbm_batch* b) {
    for (auto& g : b->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

int tiles_of(const lbm_batch* b, int i, int& ntx, int& nty, int& nzc) {
    const BatchTile& t = kBatchTiles[b->tile];
    const lbm_config& c = b->sims[i]->cfg;
    ntx = (c.nx + t.tx - 1) / t.tx;
    nty = (c.ny + t.ty - 1) / t.ty;
    nzc = (c.nz + t.zc - 1) / t.zc;
    return ntx * nty * nzc;
}

int grid_for_tile(int tile, int total, int limit) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
    }
    const int occ = with_batch_tile(tile, [](auto t) {
        using T = decltype(t);
        return occupancy_blocks<T::kind, T::tx, T::ty, T::zc>();
    });
    const int grid = std::max(1, std::min(total, cus * occ));
    return limit > 0 ? std::min(grid, limit) : grid;
}

// Item list of the third-pipeline kernel: every z chunk of every column
// tile, longest first, so that the blocks of the persistent grid end
// together; the items of the last round (grid of them) are split into two
// halves so that the final round is short. Sets total and grid.
int build_items(lbm_batch* b) {
    const BatchTile& t = kBatchTiles[b->tile];
    std::vector<BatchItem>& items = b->items;
    items.clear();
    for (size_t i = 0; i < b->sims.size(); ++i) {
        const SimDesc& d = b->desc[i];
        const int nz = d.bc.n[2];
        for (int txy = 0; txy < d.ntx * d.nty; ++txy)
            for (int z0 = 0; z0 < nz; z0 += t.zc) items.push_back({(int)i, txy, z0, std::min(z0 + t.zc, nz)});
    }
    auto longer = [](const BatchItem& a, const BatchItem& c) { return a.z1 - a.z0 > c.z1 - c.z0; };
    std::stable_sort(items.begin(), items.end(), longer);
    const int grid = grid_for_tile(b->tile, (int)items.size(), b->grid_limit);
    if ((int)items.size() > 2 * grid) {
        std::vector<BatchItem> tail(items.end() - grid, items.end());
        items.resize(items.size() - grid);
        for (const BatchItem& it : tail) {
            const int mid = it.z0 + (it.z1 - it.z0) / 2;
            if (mid > it.z0) {
                items.push_back({it.sim, it.txy, it.z0, mid});
                items.push_back({it.sim, it.txy, mid, it.z1});
            } else {
                items.push_back(it);
            }
        }
        std::stable_sort(items.begin(), items.end(), longer);
    }
    if ((int)items.size() > b->items_cap) {
        if (b->d_items) (void)hipFree(b->d_items);
        b->d_items = nullptr;
        b->items_cap = 0;
        if (hipMalloc(&b->d_items, items.size() * sizeof(BatchItem)) != hipSuccess)
            return fail(LBM_ERR_HIP, "hipMalloc(batch items) failed");
        b->items_cap = (int)items.size();
    }
    if (hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice,
                       b->stream) != hipSuccess)
        return fail(LBM_ERR_HIP, "batch item upload failed");
    b->total = (int)items.size();
    b->grid = grid_for_tile(b->tile, b->total, b->grid_limit);
    return LBM_OK;
}

// Build the descriptors from the handles' current buffers (parity 0 view).
int rebuild(lbm_batch* b) {
    b->desc.resize(b->sims.size());
    int begin = 0, sq_begin = 0, sf_begin = 0;
    for (size_t i = 0; i < b->sims.size(); ++i) {
        lbm_solver* s = b->sims[i];
        SimDesc& d = b->desc[i];
        d.f[0] = s->d32;
        d.f[1] = s->d32;
        for (int m = 0; m < 5; ++m) std::swap(d.f[1].q[m], d.f[1].qn[m]);
        for (int m = 0; m < 3; ++m) std::swap(d.f[1].u[m], d.f[1].un[m]);
        std::swap(d.f[1].g, d.f[1].gn);
        d.bc = s->bc32;
        d.k = s->k32;
        d.periodic = 1;
        for (int f = 0; f < 6; ++f)
            if (s->cfg.walls[f].qbc != LBM_Q_PERIODIC || s->cfg.walls[f].ubc != LBM_U_PERIODIC) d.periodic = 0;
        d.tile_begin = begin;
        begin += tiles_of(b, (int)i, d.ntx, d.nty, d.nzc);
        d.pull = s->pulled;
        d.aux_always = s->moving ? 1 : 0;
        d.walls = (!d.periodic && kBatchTiles[b->tile].kind == 1 && p2_walls_ok(d.bc)) ? 1 : 0;
        d.r3 = d.walls ? ShellRegion{} : make_shell_region(d.bc, 3);
        d.r2 = d.walls ? ShellRegion{} : make_shell_region(d.bc, 2);
        d.sq_begin = sq_begin;
        d.sf_begin = sf_begin;
        sq_begin += (int)region_cells(d.r3);
        sf_begin += (int)region_cells(d.r2);
    }
    b->total = begin;
    b->grid = grid_for_tile(b->tile, b->total, b->grid_limit);
    if (kBatchTiles[b->tile].kind == 1)
        if (int r = build_items(b)) return r;
    b->sq_total = sq_begin;
    b->sf_total = sf_begin;
    if (hipMemcpyAsync(b->d_desc, b->desc.data(), b->desc.size() * sizeof(SimDesc), hipMemcpyHostToDevice,
                       b->stream) != hipSuccess)
        return fail(LBM_ERR_HIP, "descriptor upload failed");
    if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(LBM_ERR_HIP, "descriptor upload failed");
    b->desc_dirty = false;
    drop_graphs(b);
    return LBM_OK;
}

// Bring every handle's buffer pointers in line with the batch parity so
// per-simulation calls (upload, download, diagnostics) see the current state.
int sync_handles(lbm_batch* b) {
    if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(LBM_ERR_HIP, "batch synchronize failed");
    if (b->parity) {
        for (lbm_solver* s : b->sims) {
            for (int m = 0; m < 5; ++m) std::swap(s->d32.q[m], s->d32.qn[m]);
            for (int m = 0; m < 3; ++m) std::swap(s->d32.u[m], s->d32.un[m]);
            std::swap(s->d32.g, s->d32.gn);
        }
        b->parity = 0;
        b->desc_dirty = true;
    }
    return LBM_OK;
}

void launch_step(lbm_batch* b, int aux) {
    const int n = (int)b->sims.size();
    with_batch_tile(b->tile, [&](auto t) {
        using T = decltype(t);
        launch_batch<T::kind, T::tx, T::ty, T::zc>(b->d_desc, n, b->total, b->d_items, b->parity, aux, b->grid,
                                                   b->d_queue, b->stream);
    });
    if (b->sf_total > 0) {
        hipLaunchKernelGGL(batch_shell_q_kernel, dim3(grid_for(b->sq_total)), dim3(kBlock), 0, b->stream, b->d_desc,
                           n, b->sq_total, b->parity);
        hipLaunchKernelGGL(batch_shell_force_kernel, dim3(grid_for(b->sf_total)), dim3(kBlock), 0, b->stream,
                           b->d_desc, n, b->sf_total, b->parity, aux);
    }
}

// One step outside any graph; afterwards every simulation pulls.
int plain_step(lbm_batch* b, int aux) {
    if (b->desc_dirty)
        if (int r = rebuild(b)) return r;
    launch_step(b, aux);
    if (hipGetLastError() != hipSuccess) return fail(LBM_ERR_HIP, "batch step launch failed");
    b->parity ^= 1;
    bool any_fresh = false;
    for (auto& d : b->desc) any_fresh = any_fresh || d.pull == 0;
    if (any_fresh) {
        for (lbm_solver* s : b->sims) s->pulled = 1;
        for (auto& d : b->desc) d.pull = 1;
        if (hipMemcpyAsync(b->d_desc, b->desc.data(), b->desc.size() * sizeof(SimDesc), hipMemcpyHostToDevice,
                           b->stream) != hipSuccess)
            return fail(LBM_ERR_HIP, "descriptor upload failed");
        drop_graphs(b);
    }
    return LBM_OK;
}

int graph_pair(lbm_batch* b) {
    hipGraphExec_t& ge = b->graph[b->parity];
    if (!ge) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail(LBM_ERR_HIP, "graph capture failed to start");
        launch_step(b, 0);
        b->parity ^= 1;
        launch_step(b, 0);
        b->parity ^= 1;
        if (hipStreamEndCapture(b->stream, &g) != hipSuccess) return fail(LBM_ERR_HIP, "graph capture failed");
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            ge = nullptr;
            return fail(LBM_ERR_HIP, "graph instantiate failed");
        }
    }
    if (hipGraphLaunch(ge, b->stream) != hipSuccess) return fail(LBM_ERR_HIP, "graph launch failed");
    return LBM_OK;   // two steps: parity unchanged
}

int batch_steps(lbm_batch* b, int n, int use_graph) {
    if (n <= 0) return LBM_OK;
    int done = 0;
    if (b->desc_dirty)
        if (int r = rebuild(b)) return r;
    bool fresh = false;
    for (auto& d : b->desc) fresh = fresh || d.pull == 0;
    if (fresh) {
        if (int r = plain_step(b, n == 1)) return r;
        done = 1;
    }
    if (use_graph) {
        while (n - done > 2) {
            if (int r = graph_pair(b)) return r;
            done += 2;
        }
    }
    while (done < n) {
        if (int r = plain_step(b, done == n - 1)) return r;
        ++done;
    }
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_create(const lbm_config* cfgs, int32_t count, lbm_batch_handle* out) {
    if (!cfgs || !out || count < 1) return fail(LBM_ERR_INVALID, "bad argument");
    *out = nullptr;
    for (int i = 0; i < count; ++i)
        if (cfgs[i].precision != LBM_FP32_SHIFTED)
            return fail(LBM_ERR_UNSUPPORTED, "batched simulations must use LBM_FP32_SHIFTED");
    lbm_batch* b = new lbm_batch();
    auto bail = [&](int r) {
        lbm_batch_destroy(b);
        return r;
    };
    for (int i = 0; i < count; ++i) {
        lbm_handle h = nullptr;
        if (int r = lbm_create(&cfgs[i], &h)) return bail(r);
        b->sims.push_back(h);
        if (!h->fused_ok) return bail(fail(LBM_ERR_UNSUPPORTED, "simulation too large for the batched kernel"));
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(LBM_ERR_HIP, "hipStreamCreate failed"));
    if (hipMalloc(&b->d_desc, count * sizeof(SimDesc)) != hipSuccess)
        return bail(fail(LBM_ERR_HIP, "hipMalloc(descriptors) failed"));
    if (hipMalloc(&b->d_queue, 2 * sizeof(unsigned)) != hipSuccess ||
        hipMemsetAsync(b->d_queue, 0, 2 * sizeof(unsigned), b->stream) != hipSuccess)
        return bail(fail(LBM_ERR_HIP, "hipMalloc(work queue) failed"));
    if (int r = rebuild(b)) return bail(r);
    *out = b;
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch_handle b) {
    if (!b) return LBM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (lbm_handle h : b->sims) lbm_destroy(h);
    if (b->d_desc) (void)hipFree(b->d_desc);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_queue) (void)hipFree(b->d_queue);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LBM_OK;
}

int lbm_batch_upload_state(lbm_batch_handle b, int32_t sim, const double* rho, const double* ux, const double* uy,
                           const double* uz, const double* q) {
    if (!b || sim < 0 || sim >= (int)b->sims.size()) return fail(LBM_ERR_INVALID, "bad simulation index");
    if (int r = sync_handles(b)) return r;
    if (int r = lbm_upload_state(b->sims[sim], rho, ux, uy, uz, q)) return r;
    b->desc_dirty = true;
    return LBM_OK;
}

int lbm_batch_download_state(lbm_batch_handle b, int32_t sim, double* rho, double* ux, double* uy, double* uz,
                             double* q) {
    if (!b || sim < 0 || sim >= (int)b->sims.size()) return fail(LBM_ERR_INVALID, "bad simulation index");
    if (int r = sync_handles(b)) return r;
    return lbm_download_state(b->sims[sim], rho, ux, uy, uz, q);
}

int lbm_batch_diagnostics(lbm_batch_handle b, int32_t sim, lbm_diag* out) {
    if (!b || sim < 0 || sim >= (int)b->sims.size()) return fail(LBM_ERR_INVALID, "bad simulation index");
    if (int r = sync_handles(b)) return r;
    return lbm_diagnostics(b->sims[sim], out);
}

int lbm_batch_step(lbm_batch_handle b, int32_t n, int32_t use_graph) {
    if (!b || n < 0) return fail(LBM_ERR_INVALID, "bad argument");
    return batch_steps(b, n, use_graph);
}

int lbm_batch_time_steps(lbm_batch_handle b, int32_t n, int32_t use_graph, double* ms) {
    if (!b || !ms || n < 1) return fail(LBM_ERR_INVALID, "bad argument");
    if (b->desc_dirty)
        if (int r = rebuild(b)) return r;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(LBM_ERR_HIP, "hipEventCreate failed");
    }
    (void)hipEventRecord(e0, b->stream);
    int r = batch_steps(b, n, use_graph);
    (void)hipEventRecord(e1, b->stream);
    (void)hipEventSynchronize(e1);
    float f = 0.0f;
    (void)hipEventElapsedTime(&f, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = f;
    return r;
}

int lbm_batch_tile_count(void) { return kNumBatchTiles; }

int lbm_batch_tile_name(int32_t tile, char* buf, size_t len) {
    if (!buf || !len) return fail(LBM_ERR_INVALID, "null buffer");
    if (tile < 0 || tile >= kNumBatchTiles) return fail(LBM_ERR_INVALID, "tile out of range");
    std::strncpy(buf, kBatchTiles[tile].name, len - 1);
    buf[len - 1] = 0;
    return LBM_OK;
}

int lbm_batch_set_tile(lbm_batch_handle b, int32_t tile) {
    if (!b || tile < 0 || tile >= kNumBatchTiles) return fail(LBM_ERR_INVALID, "tile out of range");
    if (int r = sync_handles(b)) return r;
    b->tile = tile;
    b->desc_dirty = true;
    return LBM_OK;
}

int lbm_batch_autotune(lbm_batch_handle b, int32_t n, double* ms_out, int32_t max, int32_t* count) {
    if (!b || n < 1) return fail(LBM_ERR_INVALID, "bad argument");
    if (count) *count = 0;
    if (int r = sync_handles(b)) return r;
    // Tune on scratch copies of every simulation's state, then restore.
    std::vector<std::vector<double>> saved(b->sims.size());
    for (size_t i = 0; i < b->sims.size(); ++i) {
        lbm_solver* s = b->sims[i];
        const long V = s->V;
        saved[i].resize(9 * V + 15 * V);
        double* p = saved[i].data();
        if (int r = lbm_download_state(s, p, p + V, p + 2 * V, p + 3 * V, p + 4 * V)) return r;
        if (int r = lbm_download_populations(s, p + 9 * V)) return r;
    }
    double best = 1e300;
    int best_t = 0;
    for (int t = 0; t < kNumBatchTiles; ++t) {
        if (int r = lbm_batch_set_tile(b, t)) return r;
        double ms = 0.0;
        if (int r = batch_steps(b, 2, 0)) return r;
        if (int r = lbm_batch_time_steps(b, n, 1, &ms)) return r;
        if (ms_out && t < max) ms_out[t] = ms / n;
        if (ms < best) {
            best = ms;
            best_t = t;
        }
    }
    if (int r = lbm_batch_set_tile(b, best_t)) return r;
    for (size_t i = 0; i < b->sims.size(); ++i) {
        lbm_solver* s = b->sims[i];
        const long V = s->V;
        double* p = saved[i].data();
        if (int r = lbm_upload_state(s, p, p + V, p + 2 * V, p + 3 * V, p + 4 * V)) return r;
        if (int r = lbm_upload_populations(s, p + 9 * V)) return r;
        s->pulled = 0;
    }
    b->desc_dirty = true;
    if (count) *count = std::min(kNumBatchTiles, (int)max);
    return LBM_OK;
}

int lbm_batch_set_grid_limit(lbm_batch_handle b, int32_t max_blocks) {
    if (!b || max_blocks < 0) return fail(LBM_ERR_INVALID, "bad argument");
    if (int r = sync_handles(b)) return r;
    b->grid_limit = max_blocks;
    b->desc_dirty = true;
    return LBM_OK;
}

int lbm_batch_get_tile(lbm_batch_handle b, int32_t* tile) {
    if (!b || !tile) return fail(LBM_ERR_INVALID, "null argument");
    *tile = b->tile;
    return LBM_OK;
}

}  // extern "C"
