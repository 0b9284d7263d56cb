// closure.hip — learned closure for a 2x-coarsened step: a three-layer
// periodic 3x3x3 convolutional network (8 -> H -> H -> 5 channels, ReLU
// between layers) that maps the coarse state (Q 5, u 3) to a correction of
// the coarse step's Q update. Inference and training (forward, input and
// weight gradients, Adam) are hand-written kernels on the BF16 matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation; master weights and the
// optimizer state are FP32.
//
// Layouts
//   activations  BF16 [cell][C] (channels fastest), cell = (z*ny + y)*nx + x
//   weights      FP32 master, layer by layer: W[off][ci][co] then b[co], with
//                off = (dz+1)*9 + (dy+1)*3 + (dx+1) the stencil offset
//   packed       BF16 B-operand fragments [kchunk][ntile][lane][8]
// A 3x3x3 convolution is the GEMM out[cell][co] = sum_k A[cell][k] W[k][co]
// with k = off*Cin + ci and A[cell][k] = in[cell + d(off)][ci] gathered with
// periodic wrap. The input gradient is the same form with the flipped,
// transposed weights W'[off][co][ci] = W[26-off][ci][co]; the weight
// gradient is the GEMM dW[k][co] = sum_cell A[cell][k] dOut[cell][co].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbman3d.h"
#include "common.hpp"
#include "runtime.hpp"
#include "solver_state.hpp"

using namespace lbm;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int kOffsets = 27;
constexpr int kInC = 8;     // Q (5) + u (3)
constexpr int kOutC = 5;    // correction of Q

__device__ __forceinline__ __bf16 to_bf16(float v) { return (__bf16)v; }
__device__ __forceinline__ float from_bf16(__bf16 v) { return (float)v; }

__device__ __forceinline__ int wrap(int t, int n) { return t < 0 ? t + n : (t >= n ? t - n : t); }

// ---- forward convolution (also the input-gradient convolution) -------------
// A block of WZ wavefronts owns a 16 x 4 x WZ tile of output cells. It stages
// the periodic 18 x 6 x (WZ + 2) halo of the input (all CIN channels, BF16)
// in LDS once, then every stencil offset's A fragment is an LDS read; the
// 27-fold reuse of each input cell stays on chip. Wavefront w computes output
// plane z = w of the tile: 4 row tiles (y) of 16 cells (x) x (16 * NT) output
// channels, each weight fragment loaded once per K step for all 4 row tiles.
// Cin is 8 (four stencil offsets per 32-wide K step) or 32 (one offset per K
// step). Eight wavefronts per block (WZ = 8) amortise the halo staging over
// twice the cells of four and measured 25-30% faster on the 32-channel
// layers at 64^3.
//
// LDS records: one 16-byte vector per cell for Cin 8; four per cell for Cin
// 32, with the vector index XOR-swizzled by ((x >> 2) & 3) of the halo column
// so that the 16 lanes reading one row hit distinct banks.
//
// SRC: 0 = BF16 input [cell][CIN]; 1 = the closure input read straight from a
// simulation's FP32 fields (Q 5, u 3) times in_scale (Cin 8).
// EPI: 0 = store BF16 with bias + ReLU,
// 2 = store FP32 channels-first [co][cell] with bias (final layer),
// 3 = add alpha * (value + bias) / out_scale to the simulation's Q (final
// layer applied in place), 4 = store BF16 linear, zero where the stored
// activation `act` of the same cell and channel is not positive (the input
// gradient through a ReLU).
constexpr int kTX = 16, kTY = 4;

struct ConvIO {
    const __bf16* in;          // SRC 0
    const float* fin[8];       // SRC 1: Qxx, Qxy, Qxz, Qyy, Qyz, ux, uy, uz
    const float* in_scale;     // SRC 1
    void* out;                 // EPI 0-2, 4
    const __bf16* act;         // EPI 4: [cell][cout]
    float* q[5];               // EPI 3
    const float* out_scale;    // EPI 3
    float alpha;               // EPI 3
};

template <int CIN, int NT, int EPI, int WZ, int SRC>
__global__ void __launch_bounds__(64 * WZ)
conv_fwd_kernel(ConvIO io, const bf16x8* __restrict__ wpk, const float* __restrict__ bias, int cout, int nx, int ny,
                int nz) {
    static_assert(SRC == 0 || CIN == 8, "field input has 8 channels");
    constexpr int CV = CIN / 8;                  // 16-byte vectors per cell
    constexpr int HX = kTX + 2, HY = kTY + 2, HZ = WZ + 2, NH = HX * HY * HZ;
    __shared__ bf16x8 tile[NH * CV];
    const int ntx = nx / kTX, nty = ny / kTY;
    const int bx = blockIdx.x % ntx, by = (blockIdx.x / ntx) % nty, bz = blockIdx.x / (ntx * nty);
    const int x0 = bx * kTX - 1, y0 = by * kTY - 1, z0 = bz * WZ - 1;
    auto swz = [](int q, int hx) { return CV == 4 ? (q ^ ((hx >> 2) & 3)) : q; };
    for (int v = threadIdx.x; v < NH * CV; v += 64 * WZ) {
        const int hc = v / CV, q = v % CV;
        const int hx = hc % HX, hy = (hc / HX) % HY, hz = hc / (HX * HY);
        const long g = ((long)wrap(z0 + hz, nz) * ny + wrap(y0 + hy, ny)) * nx + wrap(x0 + hx, nx);
        bf16x8 r;
        if constexpr (SRC == 1) {
#pragma unroll
            for (int j = 0; j < 8; ++j) r[j] = to_bf16(io.fin[j][g] * io.in_scale[j]);
        } else {
            r = *reinterpret_cast<const bf16x8*>(io.in + g * CIN + q * 8);
        }
        tile[hc * CV + swz(q, hx)] = r;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, kb = lane >> 4;
    constexpr int KCH = CIN == 8 ? 7 : 27;   // K steps of 32 (8-channel input: 28 offsets, last zero)
    f32x4 acc[kTY][NT];
#pragma unroll
    for (int t = 0; t < kTY; ++t)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[t][n] = (f32x4){0.f, 0.f, 0.f, 0.f};
    // B fragments (packed weights, L2-resident) are loaded one K step ahead
    // so that their latency overlaps the MFMAs of the current step.
    bf16x8 bnext[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) bnext[n] = wpk[(long)n * 64 + lane];
    for (int kc = 0; kc < KCH; ++kc) {
        const int off = CIN == 8 ? kc * 4 + kb : kc;
        const int q = CIN == 8 ? 0 : kb;
        const int dz = off / 9, dy = (off / 3) % 3, dx = off % 3;   // halo coordinates shift by +1
        bf16x8 b[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n) b[n] = bnext[n];
        if (kc + 1 < KCH) {
#pragma unroll
            for (int n = 0; n < NT; ++n) bnext[n] = wpk[((long)(kc + 1) * NT + n) * 64 + lane];
        }
        const int hx = m + dx, qs = swz(q, hx);
#pragma unroll
        for (int t = 0; t < kTY; ++t) {
            bf16x8 a;
            if (off < kOffsets) {
                a = tile[(((w + dz) * HY + (t + dy)) * HX + hx) * CV + qs];
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) a[j] = to_bf16(0.f);
            }
#pragma unroll
            for (int n = 0; n < NT; ++n)
                acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[n], acc[t][n], 0, 0, 0);
        }
    }
    // D fragment: lane holds rows 4*kb + r (x within the row tile), column n*16 + m (channel).
    const long V = (long)nx * ny * nz;
    if constexpr (EPI == 2 || EPI == 3) {
        // FP32 channels-first output: in fragment order each store would
        // touch 4 cells 16 B apart per channel. The values go through LDS
        // (the input tile is no longer read) so that a thread finishes one
        // cell for every channel and a wavefront's accesses to one channel
        // array are 64 consecutive cells.
        static_assert(CIN == 32 && NT == 1, "final layer: 32 input channels, at most 16 outputs");
        constexpr int NC = WZ * kTY * kTX;
        static_assert(16 * NC * sizeof(float) <= sizeof(tile), "staging fits the input tile");
        float* so = reinterpret_cast<float*>(tile);
        __syncthreads();
        if (m < cout) {
            const float bco = bias[m];
#pragma unroll
            for (int t = 0; t < kTY; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) so[m * NC + (w * kTY + t) * kTX + 4 * kb + r] = acc[t][0][r] + bco;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < NC; i += 64 * WZ) {
            const int lx = i % kTX, ly = (i / kTX) % kTY, lz = i / (kTX * kTY);
            const long cell = ((long)(bz * WZ + lz) * ny + by * kTY + ly) * nx + bx * kTX + lx;
            for (int co = 0; co < cout; ++co) {
                const float v = so[co * NC + i];
                if (EPI == 3) io.q[co][cell] += io.alpha * v / io.out_scale[co];
                else reinterpret_cast<float*>(io.out)[(long)co * V + cell] = v;
            }
        }
        return;
    }
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int co = n * 16 + m;
        if (co >= cout) continue;
        const float bco = EPI == 4 ? 0.f : bias[co];
#pragma unroll
        for (int t = 0; t < kTY; ++t) {
            const long row = ((long)(bz * WZ + w) * ny + by * kTY + t) * nx + bx * kTX;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long cell = row + 4 * kb + r;
                float v = acc[t][n][r] + bco;
                if (EPI == 0) v = v > 0.f ? v : 0.f;
                if (EPI == 4 && !(from_bf16(io.act[cell * cout + co]) > 0.f)) v = 0.f;
                if (EPI == 3) io.q[co][cell] += io.alpha * v / io.out_scale[co];
                else if (EPI == 2) reinterpret_cast<float*>(io.out)[(long)co * V + cell] = v;
                else reinterpret_cast<__bf16*>(io.out)[cell * cout + co] = to_bf16(v);
            }
        }
    }
}

// ---- weight and bias gradient ---------------------------------------------
// dW[off][ci][co] = sum_cell in[cell + d(off)][ci] dOut[cell][co] and
// db[co] = sum_cell dOut[cell][co]: a GEMM whose K index runs over cells.
// A block of nine wavefronts takes 16 x 4 x 2-cell tiles in turn. For each
// tile it stages the periodic 20 x 6 x 4 input halo and the dOut tile in LDS
// channel-major (x fastest, one row per (z, y)), transposing the [cell][C]
// vectors of global memory on the way in. Wavefront w owns the three stencil
// offsets (dx = -1, 0, 1) of one (dy, dz) = (w % 3 - 1, w / 3 - 1): per
// 32-cell K step (two 16-cell rows) a lane reads the 8 cells of its row
// segment as one 16-byte LDS vector plus the dword on each side, and forms
// the fragments of the three x shifts with 16-bit funnel shifts
// (v_alignbit_b32), so every LDS read stays aligned. The dOut fragments are
// shared by the 3 x (CIN / 16) A fragments of the step. Wavefront 4
// (dy = dz = 0) also multiplies an all-ones A fragment into its dOut
// fragments: the bias gradient. A block accumulates over its tiles in
// registers and writes one partial [27 CIN + 1][cout] (weights, then the bias
// row); wgrad_reduce_kernel adds the partials of all blocks.
constexpr int kWgTX = 16, kWgTY = 4, kWgTZ = 2;
constexpr int kWgRL = 24;                                     // bf16 per LDS row (x0 - 2 .. x0 + 17, padded)
constexpr int kWgHY = kWgTY + 2, kWgHZ = kWgTZ + 2;
constexpr int kWgAStr = kWgHZ * kWgHY * kWgRL + 8;            // per input channel: 73 x 16 B (odd: no bank clash)
constexpr int kWgBStr = kWgTZ * kWgTY * kWgTX + 8;            // per output channel: 17 x 16 B
constexpr int kWgWaves = 9;

template <int CIN, int CS>
__global__ void __launch_bounds__(64 * kWgWaves)
conv_wgrad_kernel(const __bf16* __restrict__ in, const __bf16* __restrict__ dout, int cout,
                  float* __restrict__ part, int ntiles, int per_block, int nx, int ny, int nz) {
    constexpr int NTI = (CIN + 15) / 16, NTO = (CS + 15) / 16;
    constexpr int CVI = CIN / 8, CVO = CS / 8;
    __shared__ __attribute__((aligned(16))) __bf16 sA[CIN * kWgAStr + 8];
    __shared__ __attribute__((aligned(16))) __bf16 sB[CS * kWgBStr];
    uint32_t* sA32 = reinterpret_cast<uint32_t*>(sA);
    uint32_t* sB32 = reinterpret_cast<uint32_t*>(sB);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int m = lane & 15, kb = lane >> 4;
    const int dz = w / 3 - 1, dy = w % 3 - 1;
    const int ntx = nx / kWgTX;
    const int t0 = blockIdx.x * per_block, t1 = min(t0 + per_block, ntiles);

    f32x4 acc[3][NTI][NTO], accb[NTO];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int t = 0; t < NTI; ++t)
#pragma unroll
            for (int n = 0; n < NTO; ++n) acc[d][t][n] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int n = 0; n < NTO; ++n) accb[n] = (f32x4){0.f, 0.f, 0.f, 0.f};
    bf16x8 ones;
#pragma unroll
    for (int j = 0; j < 8; ++j) ones[j] = to_bf16(1.f);

    for (int tile = t0; tile < t1; ++tile) {
        // tiles z-fastest, so a block's consecutive tiles share halo planes in L2
        const int tz = tile % (nz / kWgTZ), txy = tile / (nz / kWgTZ);
        const int x0 = (txy % ntx) * kWgTX, y0 = (txy / ntx) * kWgTY, z0 = tz * kWgTZ;
        __syncthreads();   // the previous tile's LDS reads are done
        // input halo: 10 cell pairs per row x 24 rows x CVI vectors
        for (int it = tid; it < 10 * kWgHY * kWgHZ * CVI; it += 64 * kWgWaves) {
            const int cv = it % CVI, pr = it / CVI;
            const int p = pr % 10, row = pr / 10;
            const int hy = row % kWgHY, hz = row / kWgHY;
            const long rbase = ((long)wrap(z0 - 1 + hz, nz) * ny + wrap(y0 - 1 + hy, ny)) * nx;
            const int xa = wrap(x0 - 2 + 2 * p, nx), xb = wrap(x0 - 1 + 2 * p, nx);
            const u16x8 va = *reinterpret_cast<const u16x8*>(in + (rbase + xa) * CIN + cv * 8);
            const u16x8 vb = *reinterpret_cast<const u16x8*>(in + (rbase + xb) * CIN + cv * 8);
            const int e = 8 + row * kWgRL + 2 * p - 2;   // element index of x0 - 2 + 2p in channel 0
#pragma unroll
            for (int j = 0; j < 8; ++j) sA32[((cv * 8 + j) * kWgAStr + e) >> 1] = (uint32_t)va[j] | ((uint32_t)vb[j] << 16);
        }
        // dOut tile: 8 cell pairs per row x 8 rows x CVO vectors
        for (int it = tid; it < 8 * kWgTY * kWgTZ * CVO; it += 64 * kWgWaves) {
            const int cv = it % CVO, pr = it / CVO;
            const int p = pr % 8, row = pr / 8;
            const int yy = row % kWgTY, zz = row / kWgTY;
            const long c = ((long)(z0 + zz) * ny + y0 + yy) * nx + x0 + 2 * p;
            const u16x8 va = *reinterpret_cast<const u16x8*>(dout + c * CS + cv * 8);
            const u16x8 vb = *reinterpret_cast<const u16x8*>(dout + (c + 1) * CS + cv * 8);
            const int e = row * kWgTX + 2 * p;
#pragma unroll
            for (int j = 0; j < 8; ++j) sB32[((cv * 8 + j) * kWgBStr + e) >> 1] = (uint32_t)va[j] | ((uint32_t)vb[j] << 16);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kWgTY * kWgTZ / 2; ++s) {
            const int rr = 2 * s + (kb >> 1), xo = 8 * (kb & 1);
            const int yy = rr % kWgTY, zz = rr / kWgTY;
            bf16x8 b[NTO];
#pragma unroll
            for (int n = 0; n < NTO; ++n) {
                const int co = 16 * n + m;
                if (co < CS) {
                    b[n] = *reinterpret_cast<const bf16x8*>(sB + co * kWgBStr + rr * kWgTX + xo);
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j) b[n][j] = to_bf16(0.f);
                }
            }
#pragma unroll
            for (int t = 0; t < NTI; ++t) {
                const int ci = 16 * t + m;
                typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                u32x4 C = (u32x4){0u, 0u, 0u, 0u};
                uint32_t L = 0u, R = 0u;
                if (ci < CIN) {
                    const int e = ci * kWgAStr + 8 + ((zz + dz + 1) * kWgHY + (yy + dy + 1)) * kWgRL + xo;
                    C = *reinterpret_cast<const u32x4*>(sA + e);
                    L = sA32[(e - 2) >> 1];
                    R = sA32[(e + 8) >> 1];
                }
                const u32x4 Am = (u32x4){__builtin_amdgcn_alignbit(C[0], L, 16), __builtin_amdgcn_alignbit(C[1], C[0], 16),
                                         __builtin_amdgcn_alignbit(C[2], C[1], 16), __builtin_amdgcn_alignbit(C[3], C[2], 16)};
                const u32x4 Ap = (u32x4){__builtin_amdgcn_alignbit(C[1], C[0], 16), __builtin_amdgcn_alignbit(C[2], C[1], 16),
                                         __builtin_amdgcn_alignbit(C[3], C[2], 16), __builtin_amdgcn_alignbit(R, C[3], 16)};
                const bf16x8 a0 = __builtin_bit_cast(bf16x8, Am), a1 = __builtin_bit_cast(bf16x8, C),
                             a2 = __builtin_bit_cast(bf16x8, Ap);
#pragma unroll
                for (int n = 0; n < NTO; ++n) {
                    acc[0][t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b[n], acc[0][t][n], 0, 0, 0);
                    acc[1][t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b[n], acc[1][t][n], 0, 0, 0);
                    acc[2][t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b[n], acc[2][t][n], 0, 0, 0);
                }
            }
            if (w == 4) {
#pragma unroll
                for (int n = 0; n < NTO; ++n)
                    accb[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ones, b[n], accb[n], 0, 0, 0);
            }
        }
    }
    // D fragment: lane (m, kb) holds rows 4 kb + r (input channel within the
    // ci tile) of column m (output channel within the co tile).
    const int krows = 27 * CIN + 1;
    float* pb = part + (long)blockIdx.x * krows * cout;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int off = (dz + 1) * 9 + (dy + 1) * 3 + d;
#pragma unroll
        for (int t = 0; t < NTI; ++t)
#pragma unroll
            for (int n = 0; n < NTO; ++n)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int ci = 16 * t + 4 * kb + r, co = 16 * n + m;
                    if (ci < CIN && co < cout) pb[(off * CIN + ci) * cout + co] = acc[d][t][n][r];
                }
    }
    if (w == 4 && kb == 0) {
#pragma unroll
        for (int n = 0; n < NTO; ++n) {
            const int co = 16 * n + m;
            if (co < cout) pb[(krows - 1) * cout + co] = accb[n][0];
        }
    }
}

// Sum of the per-block partials [block][27 cin + 1][cout]: the weight rows go
// to gw ([off][ci][co]), the last row to gb. A block of 256 threads owns 64
// consecutive outputs; its four wavefronts add every fourth partial (eight
// loads in flight per lane), and LDS adds the four sums.
__global__ void __launch_bounds__(256)
wgrad_reduce_kernel(const float* __restrict__ part, int nblocks, int krows, int cout, float* __restrict__ gw,
                    float* __restrict__ gb) {
    __shared__ float red[4][64];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const long n = (long)krows * cout;
    const long i = (long)blockIdx.x * 64 + lane;
    float s = 0.f;
    if (i < n) {
        int b = g;
        for (; b + 28 < nblocks; b += 32) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = part[(long)(b + 4 * u) * n + i];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; b < nblocks; b += 4) s += part[(long)b * n + i];
    }
    red[g][lane] = s;
    __syncthreads();
    if (g == 0 && i < n) {
        const float t = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
        if (i < (long)(krows - 1) * cout) gw[i] = t;
        else gb[i - (long)(krows - 1) * cout] = t;
    }
}

// Mean-squared loss over V cells x 5 channels: dL/dpred = 2 (pred - y) / (5 V),
// written as BF16 [cell][8] (channels 5..7 zero) for the backward convolutions;
// per-block partial sums of the squared error.
__global__ void __launch_bounds__(256)
loss_grad_kernel(const float* __restrict__ pred, const float* __restrict__ target, long V, __bf16* __restrict__ dout,
                 float* __restrict__ partial) {
    __shared__ float red[256];
    float s = 0.f;
    const float scale = 2.0f / (kOutC * (float)V);
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < V; c += (long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int co = 0; co < 8; ++co) {
            float g = 0.f;
            if (co < kOutC) {
                const float d = pred[co * V + c] - target[co * V + c];
                s += d * d;
                g = scale * d;
            }
            dout[c * 8 + co] = to_bf16(g);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Pack FP32 master weights W[off][ci][co] into B fragments for a
// convolution kernel whose input has kcin channels per cell (K index
// off * kcin + c). transpose: pack the input-gradient weights
// W'[off][co][ci] = W[26 - off][ci][co]. One launch packs every job of the
// list (the forward fragments of the three layers and the input-gradient
// fragments of layers 2 and 3); job j covers the thread range
// [start[j], start[j + 1]).
struct PackJobs {
    const float* w[5];
    bf16x8* out[5];
    int cin[5], cout[5], transpose[5], kcin[5], ntiles[5];
    long start[6];
    int njobs;
};

__global__ void pack_kernel(PackJobs jobs) {
    const long gidx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gidx >= jobs.start[jobs.njobs]) return;
    int jb = 0;
    while (gidx >= jobs.start[jb + 1]) ++jb;
    const long idx = gidx - jobs.start[jb];   // (kc, t, lane)
    const int ntiles = jobs.ntiles[jb], kcin = jobs.kcin[jb], cin = jobs.cin[jb], cout = jobs.cout[jb];
    const int transpose = jobs.transpose[jb];
    const float* w = jobs.w[jb];
    const int lane = (int)(idx % 64), t = (int)((idx / 64) % ntiles), kc = (int)(idx / (64 * ntiles));
    const int ecin = transpose ? cout : cin, ecout = transpose ? cin : cout;
    const int n = t * 16 + (lane & 15);
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = kc * 32 + 8 * (lane >> 4) + j;
        const int off = k / kcin, c = k % kcin;
        float val = 0.f;
        if (off < kOffsets && c < ecin && n < ecout) {
            val = transpose ? w[((long)(26 - off) * cin + n) * cout + c] : w[((long)off * cin + c) * cout + n];
        }
        v[j] = to_bf16(val);
    }
    jobs.out[jb][idx] = v;
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// MFMA lane-layout check: D = A (16x32) * B (32x16) through the same
// fragment mapping the convolutions use.
__global__ void mfma_check_kernel(const float* a, const float* b, float* d) {
    const int lane = threadIdx.x;
    const int m = lane & 15, kb = lane >> 4;
    bf16x8 av, bv;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        av[j] = to_bf16(a[m * 32 + 8 * kb + j]);
        bv[j] = to_bf16(b[(8 * kb + j) * 16 + m]);
    }
    f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) d[(4 * kb + r) * 16 + m] = acc[r];
}

unsigned blocks_for(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct lbm_closure {
    int nx = 0, ny = 0, nz = 0, hidden = 32;
    long V = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    // parameters: offsets into the flat FP32 arrays
    long off_w[3], off_b[3], nparams = 0;
    int cin[3], cout[3];
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16x8* pk[3] = {nullptr, nullptr, nullptr};    // forward B fragments
    bf16x8* pkt[2] = {nullptr, nullptr};            // input-gradient B fragments of layers 2, 3
    int kch[3], nt[3], kcht[2], ntt[2];
    __bf16 *x = nullptr, *a1 = nullptr, *a2 = nullptr;   // activations [cell][C]
    __bf16 *d3 = nullptr, *d2 = nullptr, *d1 = nullptr;   // gradients [cell][C]
    float* pred = nullptr;                                // [5][cell]
    float* part = nullptr;                                // weight-gradient partials
    long part_elems = 0;
    float* loss_part = nullptr;
    int loss_blocks = 0;
    float* in_scale = nullptr;
    float* out_scale = nullptr;
    hipEvent_t ev_before = nullptr, ev_after = nullptr;   // order apply against the closure's own stream
    double h_in_scale[8] = {1, 1, 1, 1, 1, 1, 1, 1}, h_out_scale[5] = {1, 1, 1, 1, 1};
    std::vector<__bf16*> sx;   // stored samples: inputs
    std::vector<float*> sy;    // stored samples: targets [5][cell] normalised
    long step = 0;
    int wg_blocks = 256;   // weight-gradient blocks (partials) per layer
};

namespace {

#define CHK(expr)                                                                                \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(LBM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Runs a cleanup when the scope ends, so that the early returns of CHK
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

template <typename T>
int cl_alloc(lbm_closure* c, T** p, size_t elems) {
    void* q = nullptr;
    CHK(hipMalloc(&q, std::max<size_t>(elems, 1) * sizeof(T)));
    c->allocs.push_back(q);
    CHK(hipMemsetAsync(q, 0, std::max<size_t>(elems, 1) * sizeof(T), c->stream));
    *p = (T*)q;
    return LBM_OK;
}

int repack(lbm_closure* c) {
    PackJobs jobs{};
    long n = 0;
    auto add = [&](int l, int transpose, int kcin, int kch, int ntiles, bf16x8* out) {
        const int j = jobs.njobs++;
        jobs.w[j] = c->w + c->off_w[l];
        jobs.out[j] = out;
        jobs.cin[j] = c->cin[l];
        jobs.cout[j] = c->cout[l];
        jobs.transpose[j] = transpose;
        jobs.kcin[j] = kcin;
        jobs.ntiles[j] = ntiles;
        jobs.start[j] = n;
        n += (long)kch * ntiles * 64;
        jobs.start[j + 1] = n;
    };
    for (int l = 0; l < 3; ++l) add(l, 0, c->cin[l], c->kch[l], c->nt[l], c->pk[l]);
    // input of the transposed convolution: the layer's output gradient,
    // stored with 8 channels (padded) for the last layer, else cout
    for (int l = 1; l < 3; ++l)
        add(l, 1, c->cout[l] <= 8 ? 8 : c->cout[l], c->kcht[l - 1], c->ntt[l - 1], c->pkt[l - 1]);
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(n)), dim3(256), 0, c->stream, jobs);
    CHK(hipGetLastError());
    return LBM_OK;
}

// One convolution launch: blocks of 8 wavefronts (16 x 4 x 8 output cells)
// when nz is a multiple of 8, else of 4 (nz is a multiple of 4).
template <int CIN, int NT, int EPI, int SRC = 0>
void conv(const lbm_closure* c, const ConvIO& io, const bf16x8* wpk, const float* bias, int cout, hipStream_t s) {
    if (c->nz % 8 == 0)
        hipLaunchKernelGGL((conv_fwd_kernel<CIN, NT, EPI, 8, SRC>), dim3(c->V / (kTX * kTY * 8)), dim3(512), 0, s,
                           io, wpk, bias, cout, c->nx, c->ny, c->nz);
    else
        hipLaunchKernelGGL((conv_fwd_kernel<CIN, NT, EPI, 4, SRC>), dim3(c->V / (kTX * kTY * 4)), dim3(256), 0, s,
                           io, wpk, bias, cout, c->nx, c->ny, c->nz);
}

ConvIO io_bf16(const __bf16* in, void* out) {
    ConvIO io{};
    io.in = in;
    io.out = out;
    return io;
}

// Forward pass of the input x (BF16 [cell][8]) into c->pred, on stream s.
int forward(lbm_closure* c, const __bf16* x, hipStream_t s) {
    const int H = c->hidden;
    conv<8, 2, 0>(c, io_bf16(x, c->a1), c->pk[0], c->w + c->off_b[0], H, s);
    conv<32, 2, 0>(c, io_bf16(c->a1, c->a2), c->pk[1], c->w + c->off_b[1], H, s);
    conv<32, 1, 2>(c, io_bf16(c->a2, c->pred), c->pk[2], c->w + c->off_b[2], kOutC, s);
    CHK(hipGetLastError());
    return LBM_OK;
}

// Forward pass on a simulation's fields with the correction added to its Q
// in place: the first layer reads (Q, u) straight from the FP32 fields and
// the last adds alpha * prediction / out_scale to Q, on stream s.
int forward_apply(lbm_closure* c, const Fields<float>& f, double alpha, hipStream_t s) {
    const int H = c->hidden;
    ConvIO in{};
    for (int m = 0; m < 5; ++m) in.fin[m] = f.q[m];
    for (int m = 0; m < 3; ++m) in.fin[5 + m] = f.u[m];
    in.in_scale = c->in_scale;
    in.out = c->a1;
    conv<8, 2, 0, 1>(c, in, c->pk[0], c->w + c->off_b[0], H, s);
    conv<32, 2, 0>(c, io_bf16(c->a1, c->a2), c->pk[1], c->w + c->off_b[1], H, s);
    ConvIO out = io_bf16(c->a2, nullptr);
    for (int m = 0; m < 5; ++m) out.q[m] = f.q[m];
    out.out_scale = c->out_scale;
    out.alpha = (float)alpha;
    conv<32, 1, 3>(c, out, c->pk[2], c->w + c->off_b[2], kOutC, s);
    CHK(hipGetLastError());
    return LBM_OK;
}

// Weight and bias gradient of one layer (input `in` with cin channels per
// cell, output gradient `dout` with cs stored channels of which cout are
// real) into gw [off][ci][co] and gb [co].
int wgrad(lbm_closure* c, const __bf16* in, int cin, const __bf16* dout, int cs, int cout, float* gw, float* gb) {
    const int ntiles = (int)(c->V / (kWgTX * kWgTY * kWgTZ));
    const int nblk = std::min(ntiles, c->wg_blocks);
    const int per = (ntiles + nblk - 1) / nblk;
    const int blocks = (ntiles + per - 1) / per;
    const int krows = 27 * cin + 1;
    if ((long)blocks * krows * cout > c->part_elems) return fail(LBM_ERR_STATE, "wgrad scratch too small");
    const dim3 grid(blocks), block(64 * kWgWaves);
    if (cin == 8 && cs == 32)
        hipLaunchKernelGGL((conv_wgrad_kernel<8, 32>), grid, block, 0, c->stream, in, dout, cout, c->part, ntiles, per,
                           c->nx, c->ny, c->nz);
    else if (cin == 32 && cs == 32)
        hipLaunchKernelGGL((conv_wgrad_kernel<32, 32>), grid, block, 0, c->stream, in, dout, cout, c->part, ntiles,
                           per, c->nx, c->ny, c->nz);
    else if (cin == 32 && cs == 8)
        hipLaunchKernelGGL((conv_wgrad_kernel<32, 8>), grid, block, 0, c->stream, in, dout, cout, c->part, ntiles, per,
                           c->nx, c->ny, c->nz);
    else
        return fail(LBM_ERR_UNSUPPORTED, "closure layer shape");
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(blocks_for((long)krows * cout, 64)), dim3(256), 0, c->stream,
                       c->part, blocks, krows, cout, gw, gb);
    CHK(hipGetLastError());
    return LBM_OK;
}

// Forward + backward for one sample; gradient into c->g, loss (sum of
// squared errors / (5V)) returned through loss_out (synchronises).
int loss_and_grad(lbm_closure* c, const __bf16* x, const float* y, double* loss_out) {
    if (int r = forward(c, x, c->stream)) return r;
    const int H = c->hidden;
    const long V = c->V;
    hipLaunchKernelGGL(loss_grad_kernel, dim3(c->loss_blocks), dim3(256), 0, c->stream, c->pred, y, V, c->d3,
                       c->loss_part);
    // layer 3: weights, bias, input gradient (8-channel padded gradient in, H out)
    if (int r = wgrad(c, c->a2, H, c->d3, 8, kOutC, c->g + c->off_w[2], c->g + c->off_b[2])) return r;
    ConvIO g2 = io_bf16(c->d3, c->d2);
    g2.act = c->a2;
    conv<8, 2, 4>(c, g2, c->pkt[1], nullptr, H, c->stream);
    // layer 2
    if (int r = wgrad(c, c->a1, H, c->d2, H, H, c->g + c->off_w[1], c->g + c->off_b[1])) return r;
    ConvIO g1 = io_bf16(c->d2, c->d1);
    g1.act = c->a1;
    conv<32, 2, 4>(c, g1, c->pkt[0], nullptr, H, c->stream);
    // layer 1
    if (int r = wgrad(c, x, kInC, c->d1, H, H, c->g + c->off_w[0], c->g + c->off_b[0])) return r;
    CHK(hipGetLastError());
    if (loss_out) {
        std::vector<float> lp(c->loss_blocks);
        CHK(hipMemcpyAsync(lp.data(), c->loss_part, lp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        CHK(hipStreamSynchronize(c->stream));
        double s = 0.0;
        for (float f : lp) s += f;
        *loss_out = s / (kOutC * (double)V);
    }
    return LBM_OK;
}

// Closure correction of a simulation's Q, launched on the simulation's stream.
int apply_on(lbm_closure* c, lbm_solver* sim, double alpha) {
    return forward_apply(c, sim->d32, alpha, sim->stream);
}

int upload_x(lbm_closure* c, const double* x, __bf16* dst) {
    std::vector<__bf16> h(c->V * 8);
    for (long i = 0; i < c->V; ++i)
        for (int m = 0; m < 8; ++m) h[i * 8 + m] = (__bf16)(float)(x[m * c->V + i] * c->h_in_scale[m]);
    CHK(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(__bf16), hipMemcpyHostToDevice, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}
int upload_y(lbm_closure* c, const double* y, float* dst) {
    std::vector<float> h(c->V * kOutC);
    for (int m = 0; m < kOutC; ++m)
        for (long i = 0; i < c->V; ++i) h[m * c->V + i] = (float)(y[m * c->V + i] * c->h_out_scale[m]);
    CHK(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_closure_create(int32_t nx, int32_t ny, int32_t nz, int32_t hidden, uint64_t seed, lbm_closure_handle* out) {
    if (!out) return fail(LBM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (hidden != 32) return fail(LBM_ERR_UNSUPPORTED, "the closure kernels are built for 32 hidden channels");
    if (nx < 16 || ny < 4 || nz < 4 || nx % 16 != 0 || ny % 4 != 0 || nz % 4 != 0)
        return fail(LBM_ERR_INVALID, "closure grid must have nx a multiple of 16 and ny, nz multiples of 4");
    lbm_closure* c = new lbm_closure();
    auto bail = [&](int r) {
        lbm_closure_destroy(c);
        return r;
    };
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->hidden = hidden;
    c->V = (long)nx * ny * nz;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(LBM_ERR_HIP, "hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&c->ev_before, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_after, hipEventDisableTiming) != hipSuccess)
        return bail(fail(LBM_ERR_HIP, "hipEventCreate failed"));
    const int H = hidden;
    c->cin[0] = kInC, c->cout[0] = H;
    c->cin[1] = H, c->cout[1] = H;
    c->cin[2] = H, c->cout[2] = kOutC;
    long o = 0;
    for (int l = 0; l < 3; ++l) {
        c->off_w[l] = o;
        o += 27L * c->cin[l] * c->cout[l];
        c->off_b[l] = o;
        o += c->cout[l];
    }
    c->nparams = o;
    for (int l = 0; l < 3; ++l) {
        c->kch[l] = c->cin[l] == 8 ? 7 : 27;
        c->nt[l] = (c->cout[l] + 15) / 16;
    }
    // input-gradient convolutions: layer 3 (gradient has 8 padded channels in, H out), layer 2 (H in, H out)
    c->kcht[1] = 7;
    c->ntt[1] = (H + 15) / 16;
    c->kcht[0] = 27;
    c->ntt[0] = (H + 15) / 16;
    int r;
    if ((r = cl_alloc(c, &c->w, c->nparams)) || (r = cl_alloc(c, &c->g, c->nparams)) ||
        (r = cl_alloc(c, &c->m, c->nparams)) || (r = cl_alloc(c, &c->v, c->nparams)))
        return bail(r);
    for (int l = 0; l < 3; ++l)
        if ((r = cl_alloc(c, &c->pk[l], (size_t)c->kch[l] * c->nt[l] * 64))) return bail(r);
    for (int l = 0; l < 2; ++l)
        if ((r = cl_alloc(c, &c->pkt[l], (size_t)c->kcht[l] * c->ntt[l] * 64))) return bail(r);
    const size_t V = c->V;
    if ((r = cl_alloc(c, &c->x, V * 8)) || (r = cl_alloc(c, &c->a1, V * H)) || (r = cl_alloc(c, &c->a2, V * H)) ||
        (r = cl_alloc(c, &c->d3, V * 8)) || (r = cl_alloc(c, &c->d2, V * H)) || (r = cl_alloc(c, &c->d1, V * H)) ||
        (r = cl_alloc(c, &c->pred, V * kOutC)))
        return bail(r);
    c->part_elems = (long)c->wg_blocks * (27L * 32 + 1) * 32;
    if ((r = cl_alloc(c, &c->part, c->part_elems))) return bail(r);
    c->loss_blocks = 512;
    if ((r = cl_alloc(c, &c->loss_part, c->loss_blocks))) return bail(r);
    if ((r = cl_alloc(c, &c->in_scale, 8)) || (r = cl_alloc(c, &c->out_scale, 8))) return bail(r);
    // He-normal initial weights from the counter-based stream, zero biases.
    std::vector<float> w(c->nparams, 0.f);
    for (int l = 0; l < 3; ++l) {
        const double sd = std::sqrt(2.0 / (27.0 * c->cin[l])) * (l == 2 ? 0.1 : 1.0);
        const long n = 27L * c->cin[l] * c->cout[l];
        for (long i = 0; i < n; ++i) {
            const double u1 = std::max(1e-12, uniform01(seed, (uint64_t)(c->off_w[l] + i), 5));
            const double u2 = uniform01(seed, (uint64_t)(c->off_w[l] + i), 6);
            w[c->off_w[l] + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2));
        }
    }
    if (hipMemcpy(c->w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(LBM_ERR_HIP, "weight upload failed"));
    double ones[8] = {1, 1, 1, 1, 1, 1, 1, 1};
    if ((r = lbm_closure_set_normalization(c, ones, ones))) return bail(r);
    if ((r = repack(c))) return bail(r);
    *out = c;
    return LBM_OK;
}

int lbm_closure_destroy(lbm_closure_handle c) {
    if (!c) return LBM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : c->allocs) (void)hipFree(p);
    for (auto* p : c->sx) (void)hipFree(p);
    for (auto* p : c->sy) (void)hipFree(p);
    if (c->ev_before) (void)hipEventDestroy(c->ev_before);
    if (c->ev_after) (void)hipEventDestroy(c->ev_after);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LBM_OK;
}

int lbm_closure_param_count(lbm_closure_handle c, int64_t* n) {
    if (!c || !n) return fail(LBM_ERR_INVALID, "null argument");
    *n = c->nparams;
    return LBM_OK;
}

int lbm_closure_get_params(lbm_closure_handle c, double* w) {
    if (!c || !w) return fail(LBM_ERR_INVALID, "null argument");
    std::vector<float> h(c->nparams);
    CHK(hipMemcpyAsync(h.data(), c->w, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    for (long i = 0; i < c->nparams; ++i) w[i] = h[i];
    return LBM_OK;
}

int lbm_closure_set_params(lbm_closure_handle c, const double* w) {
    if (!c || !w) return fail(LBM_ERR_INVALID, "null argument");
    std::vector<float> h(c->nparams);
    for (long i = 0; i < c->nparams; ++i) h[i] = (float)w[i];
    CHK(hipMemcpyAsync(c->w, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    CHK(hipMemsetAsync(c->m, 0, c->nparams * sizeof(float), c->stream));
    CHK(hipMemsetAsync(c->v, 0, c->nparams * sizeof(float), c->stream));
    c->step = 0;
    if (int r = repack(c)) return r;
    CHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_closure_set_normalization(lbm_closure_handle c, const double* in_scale, const double* out_scale) {
    if (!c || !in_scale || !out_scale) return fail(LBM_ERR_INVALID, "null argument");
    float fi[8], fo[8] = {1, 1, 1, 1, 1, 1, 1, 1};
    for (int m = 0; m < 8; ++m) fi[m] = (float)(c->h_in_scale[m] = in_scale[m]);
    for (int m = 0; m < 5; ++m) fo[m] = (float)(c->h_out_scale[m] = out_scale[m]);
    CHK(hipMemcpyAsync(c->in_scale, fi, sizeof(fi), hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(c->out_scale, fo, sizeof(fo), hipMemcpyHostToDevice, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_closure_add_sample(lbm_closure_handle c, const double* x, const double* y, int32_t* index) {
    if (!c || !x || !y) return fail(LBM_ERR_INVALID, "null argument");
    __bf16* dx = nullptr;
    float* dy = nullptr;
    CHK(hipMalloc(&dx, c->V * 8 * sizeof(__bf16)));
    if (hipMalloc(&dy, c->V * kOutC * sizeof(float)) != hipSuccess) {
        (void)hipFree(dx);
        return fail(LBM_ERR_HIP, "hipMalloc failed");
    }
    c->sx.push_back(dx);
    c->sy.push_back(dy);
    if (int r = upload_x(c, x, dx)) return r;
    if (int r = upload_y(c, y, dy)) return r;
    if (index) *index = (int32_t)c->sx.size() - 1;
    return LBM_OK;
}

int lbm_closure_clear_samples(lbm_closure_handle c) {
    if (!c) return fail(LBM_ERR_INVALID, "null argument");
    CHK(hipStreamSynchronize(c->stream));
    for (auto* p : c->sx) (void)hipFree(p);
    for (auto* p : c->sy) (void)hipFree(p);
    c->sx.clear();
    c->sy.clear();
    return LBM_OK;
}

int lbm_closure_train(lbm_closure_handle c, int32_t iters, double lr, uint64_t seed, double* mean_loss) {
    if (!c || iters < 1) return fail(LBM_ERR_INVALID, "bad argument");
    if (c->sx.empty()) return fail(LBM_ERR_STATE, "no training samples");
    double acc = 0.0;
    const int ns = (int)c->sx.size();
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < iters; ++it) {
        const int s = (int)(uniform01(seed, (uint64_t)(c->step), 11) * ns) % ns;
        double loss = 0.0;
        const bool want = mean_loss != nullptr;
        if (int r = loss_and_grad(c, c->sx[s], c->sy[s], want ? &loss : nullptr)) return r;
        acc += loss;
        ++c->step;
        const float c1 = 1.f - std::pow(b1, (float)c->step), c2 = 1.f - std::pow(b2, (float)c->step);
        hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(c->nparams)), dim3(256), 0, c->stream, c->w, c->g, c->m, c->v,
                           c->nparams, (float)lr, b1, b2, eps, c1, c2);
        if (int r = repack(c)) return r;
    }
    CHK(hipStreamSynchronize(c->stream));
    if (mean_loss) *mean_loss = acc / iters;
    return LBM_OK;
}

int lbm_closure_loss_and_grad(lbm_closure_handle c, const double* x, const double* y, double* loss, double* grad) {
    if (!c || !x || !y || !loss) return fail(LBM_ERR_INVALID, "null argument");
    float* dy = nullptr;
    CHK(hipMalloc(&dy, c->V * kOutC * sizeof(float)));
    int r = upload_x(c, x, c->x);
    if (!r) r = upload_y(c, y, dy);
    if (!r) r = loss_and_grad(c, c->x, dy, loss);
    if (!r && grad) {
        std::vector<float> h(c->nparams);
        if (hipMemcpyAsync(h.data(), c->g, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            r = fail(LBM_ERR_HIP, "gradient download failed");
        for (long i = 0; i < c->nparams && !r; ++i) grad[i] = h[i];
    }
    (void)hipFree(dy);
    return r;
}

int lbm_closure_predict(lbm_closure_handle c, const double* x, double* y) {
    if (!c || !x || !y) return fail(LBM_ERR_INVALID, "null argument");
    if (int r = upload_x(c, x, c->x)) return r;
    if (int r = forward(c, c->x, c->stream)) return r;
    std::vector<float> h(c->V * kOutC);
    CHK(hipMemcpyAsync(h.data(), c->pred, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    for (int m = 0; m < kOutC; ++m)
        for (long i = 0; i < c->V; ++i) y[m * c->V + i] = h[m * c->V + i] / c->h_out_scale[m];
    return LBM_OK;
}

int lbm_closure_apply(lbm_closure_handle c, lbm_handle sim, double alpha) {
    if (!c || !sim) return fail(LBM_ERR_INVALID, "null argument");
    if (sim->fp64) return fail(LBM_ERR_UNSUPPORTED, "the closure corrects FP32 simulations");
    if (sim->cfg.nx != c->nx || sim->cfg.ny != c->ny || sim->cfg.nz != c->nz)
        return fail(LBM_ERR_INVALID, "closure and simulation grids differ");
    // Runs on the simulation's stream, after the closure's pending work and
    // before any later closure work; the host does not wait.
    CHK(hipEventRecord(c->ev_before, c->stream));
    CHK(hipStreamWaitEvent(sim->stream, c->ev_before, 0));
    if (int r = apply_on(c, sim, alpha)) return r;
    CHK(hipEventRecord(c->ev_after, sim->stream));
    CHK(hipStreamWaitEvent(c->stream, c->ev_after, 0));
    return LBM_OK;
}

int lbm_closure_run(lbm_closure_handle c, lbm_handle sim, int32_t n, double alpha, double* ms) {
    if (!c || !sim || n < 0) return fail(LBM_ERR_INVALID, "bad argument");
    if (sim->fp64) return fail(LBM_ERR_UNSUPPORTED, "the closure corrects FP32 simulations");
    if (sim->cfg.nx != c->nx || sim->cfg.ny != c->ny || sim->cfg.nz != c->nz)
        return fail(LBM_ERR_INVALID, "closure and simulation grids differ");
    if (n == 0) return LBM_OK;
    ensure_tuned(sim);
    hipEvent_t a = nullptr, b = nullptr;
    auto release = scope_exit([&] {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    });
    CHK(hipEventCreate(&a));
    CHK(hipEventCreate(&b));
    CHK(hipEventRecord(c->ev_before, c->stream));
    CHK(hipStreamWaitEvent(sim->stream, c->ev_before, 0));
    CHK(hipEventRecord(a, sim->stream));
    int r = LBM_OK;
    for (int t = 0; t < n && r == LBM_OK; ++t) {
        r = run_step(sim, t == n - 1, nullptr);
        if (r == LBM_OK) r = apply_on(c, sim, alpha);
    }
    CHK(hipEventRecord(b, sim->stream));
    CHK(hipEventRecord(c->ev_after, sim->stream));
    CHK(hipStreamWaitEvent(c->stream, c->ev_after, 0));
    CHK(hipEventSynchronize(b));
    float f = 0.f;
    CHK(hipEventElapsedTime(&f, a, b));
    if (ms) *ms = f;
    return r;
}

int lbm_closure_time_forward(lbm_closure_handle c, int32_t n, double* ms) {
    if (!c || !ms || n < 1) return fail(LBM_ERR_INVALID, "bad argument");
    hipEvent_t a = nullptr, b = nullptr;
    auto release = scope_exit([&] {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    });
    CHK(hipEventCreate(&a));
    CHK(hipEventCreate(&b));
    CHK(hipEventRecord(a, c->stream));
    for (int i = 0; i < n; ++i)
        if (int r = forward(c, c->x, c->stream)) return r;
    CHK(hipEventRecord(b, c->stream));
    CHK(hipEventSynchronize(b));
    float f = 0.f;
    CHK(hipEventElapsedTime(&f, a, b));
    *ms = f / n;
    return LBM_OK;
}

int lbm_debug_mfma_bf16(const float* a, const float* b, float* d) {
    if (!a || !b || !d) return fail(LBM_ERR_INVALID, "null argument");
    float *da = nullptr, *db = nullptr, *dd = nullptr;
    auto release = scope_exit([&] {
        (void)hipDeviceSynchronize();
        if (da) (void)hipFree(da);
        if (db) (void)hipFree(db);
        if (dd) (void)hipFree(dd);
    });
    CHK(hipMalloc(&da, 16 * 32 * sizeof(float)));
    CHK(hipMalloc(&db, 32 * 16 * sizeof(float)));
    CHK(hipMalloc(&dd, 16 * 16 * sizeof(float)));
    CHK(hipMemcpy(da, a, 16 * 32 * sizeof(float), hipMemcpyHostToDevice));
    CHK(hipMemcpy(db, b, 32 * 16 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mfma_check_kernel, dim3(1), dim3(64), 0, 0, da, db, dd);
    CHK(hipGetLastError());
    CHK(hipMemcpy(d, dd, 16 * 16 * sizeof(float), hipMemcpyDeviceToHost));
    return LBM_OK;
}

}  // extern "C"
