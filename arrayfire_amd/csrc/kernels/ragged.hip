// Ragged batch of independent FP32 2-D spatial convolutions in one persistent
// launch. Each problem is one af_convolve2(signal, filter, mode,
// AF_CONV_SPATIAL) call with AF_BATCH_NONE (src/api/c/convolve.cpp:228-241,
// src/backend/cpu/kernel/convolve.hpp:39-75): its own image size, filter size
// and DEFAULT/EXPAND mode.
//
// Work items are (problem, output tile) pairs listed by the host in
// descending cost order; workgroup g processes items g, g + G, g + 2G, ...
// (static round-robin over a cost-sorted list: balanced, no device counters,
// so the launch can be captured in a graph and replayed without any reset).
// Per item: the halo tile and the flipped filter are staged in LDS, then each
// lane computes a TX x TY block of outputs. FX (filter width) selects a
// specialised inner loop through a uniform switch; FY (filter height) is a
// runtime loop in which each LDS row segment is read once and applied to
// every output row it touches.
#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

struct RaggedProblem {
    const float *in;
    float *out;
    const float *filt;
    int S0, S1, O0, O1;
    int FX, FY, sx, sy;
    int nbx, nby;
};

constexpr int RG_MAXF = 15;  // largest filter side of the specialised loops
constexpr int RG_THREADS = 256;

// Tile variants (autotuned): output tile BX x BY, lane block TX x TY.
template <int V>
struct RVar;
template <>
struct RVar<0> {  // 128 x 64
    static constexpr int TX = 4, TY = 8, NTX = 32;
};
template <>
struct RVar<1> {  // 64 x 64
    static constexpr int TX = 4, TY = 4, NTX = 16;
};
template <>
struct RVar<2> {  // 64 x 32
    static constexpr int TX = 2, TY = 4, NTX = 32;
};

template <int V>
struct RGeo {
    static constexpr int TX = RVar<V>::TX, TY = RVar<V>::TY, NTX = RVar<V>::NTX;
    static constexpr int NTY = RG_THREADS / NTX;
    static constexpr int BX = TX * NTX, BY = TY * NTY;
    static constexpr int LW = BX + RG_MAXF - 1, LH = BY + RG_MAXF - 1;
    // row pitch: 16-byte aligned and 4 mod 8 floats, so that the rows a
    // 16-byte read pairs (lanes l and l + 32, 8 rows apart) sit 32 words
    // apart mod 64, on opposite halves of the LDS banks
    static constexpr int LS0 = (LW + 3) & ~3;
    static constexpr int LS = LS0 % 8 == 0 ? LS0 + 4 : LS0;
};

// acc[j][i] for output (tx*TX + i, ty*TY + j) of the tile; tile(lx+kx, ly+ky)
// * wf(ky, kx) with wf the flipped filter (row pitch RG_WP) in LDS. Row
// segments (TX a multiple of 4) and filter rows are read as whole 16-byte
// vectors: one broadcast read per 4 taps instead of one per tap.
constexpr int RG_WP = 16;  // filter row pitch in LDS (floats), >= RG_MAXF
template <int FX, int V>
__device__ inline void tile_fx(const float *tile, const float *wf, int FY, int LS,
                               float (&acc)[RGeo<V>::TY][RGeo<V>::TX]) {
    constexpr int TX = RGeo<V>::TX, TY = RGeo<V>::TY, NTX = RGeo<V>::NTX;
    constexpr int NV = (TX + FX - 1 + 3) / 4, NW = (FX + 3) / 4;
    const int tx = threadIdx.x % NTX, ty = threadIdx.x / NTX;
    const float *base = tile + (ty * TY) * LS + tx * TX;
    for (int r = 0; r < TY + FY - 1; ++r) {
        float v[4 * NV];
        if constexpr (TX % 4 == 0) {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const float4 f4 = lds_load16(base + r * LS + 4 * q);
                v[4 * q] = f4.x;
                v[4 * q + 1] = f4.y;
                v[4 * q + 2] = f4.z;
                v[4 * q + 3] = f4.w;
            }
        } else {
#pragma unroll
            for (int c = 0; c < TX + FX - 1; ++c) v[c] = base[r * LS + c];
        }
#pragma unroll
        for (int j = 0; j < TY; ++j) {
            const int ky = r - j;
            if (ky >= 0 && ky < FY) {
                float w[4 * NW];
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    const float4 f4 = lds_load16(wf + ky * RG_WP + 4 * q);
                    w[4 * q] = f4.x;
                    w[4 * q + 1] = f4.y;
                    w[4 * q + 2] = f4.z;
                    w[4 * q + 3] = f4.w;
                }
#pragma unroll
                for (int i = 0; i < TX; ++i)
#pragma unroll
                    for (int kx = 0; kx < FX; ++kx) acc[j][i] = __builtin_fmaf(v[i + kx], w[kx], acc[j][i]);
            }
        }
    }
}

// Any filter width up to RG_MAXF (even widths and widths without a
// specialised loop).
template <int V>
__device__ inline void tile_any(const float *tile, const float *wf, int FX, int FY, int LS,
                                float (&acc)[RGeo<V>::TY][RGeo<V>::TX]) {
    constexpr int TX = RGeo<V>::TX, TY = RGeo<V>::TY, NTX = RGeo<V>::NTX;
    const int tx = threadIdx.x % NTX, ty = threadIdx.x / NTX;
    const float *base = tile + (ty * TY) * LS + tx * TX;
    for (int j = 0; j < TY; ++j)
        for (int ky = 0; ky < FY; ++ky)
            for (int kx = 0; kx < FX; ++kx) {
                const float w = wf[ky * RG_WP + kx];
#pragma unroll
                for (int i = 0; i < TX; ++i)
                    acc[j][i] = __builtin_fmaf(base[(j + ky) * LS + i + kx], w, acc[j][i]);
            }
}

template <int V>
__global__ __launch_bounds__(RG_THREADS) void ragged_conv2(const RaggedProblem *__restrict__ probs,
                                                           const int2 *__restrict__ items,
                                                           long long nitems) {
    using G = RGeo<V>;
    __shared__ __attribute__((aligned(16))) float tile[G::LH * G::LS];
    __shared__ __attribute__((aligned(16))) float wf[RG_MAXF * RG_WP];
    for (long long it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 item = items[it];
        const RaggedProblem pr = probs[item.x];
        const int bx = item.y % pr.nbx, by = item.y / pr.nbx;
        const long long ox0 = (long long)bx * G::BX, oy0 = (long long)by * G::BY;
        const int LW = G::BX + pr.FX - 1, LH = G::BY + pr.FY - 1;
        __syncthreads();  // previous item's tile and filter no longer read
        for (int e = threadIdx.x; e < pr.FY * RG_WP; e += RG_THREADS) {
            const int ky = e / RG_WP, kx = e % RG_WP;  // padding taps (kx >= FX) are never read
            wf[e] = kx < pr.FX ? pr.filt[(pr.FY - 1 - ky) * pr.FX + (pr.FX - 1 - kx)] : 0.f;
        }
        load_halo_tile_f32<RG_THREADS, (G::LW + 6) / 4>(tile, pr.in, pr.S0, pr.S1,
                                                        ox0 + pr.sx - (pr.FX - 1),
                                                        oy0 + pr.sy - (pr.FY - 1), LH, LW, G::LS);
        __syncthreads();
        float acc[G::TY][G::TX];
#pragma unroll
        for (int j = 0; j < G::TY; ++j)
#pragma unroll
            for (int i = 0; i < G::TX; ++i) acc[j][i] = 0.f;
        switch (pr.FX) {
            case 3: tile_fx<3, V>(tile, wf, pr.FY, G::LS, acc); break;
            case 5: tile_fx<5, V>(tile, wf, pr.FY, G::LS, acc); break;
            case 7: tile_fx<7, V>(tile, wf, pr.FY, G::LS, acc); break;
            case 9: tile_fx<9, V>(tile, wf, pr.FY, G::LS, acc); break;
            case 11: tile_fx<11, V>(tile, wf, pr.FY, G::LS, acc); break;
            case 13: tile_fx<13, V>(tile, wf, pr.FY, G::LS, acc); break;
            case 15: tile_fx<15, V>(tile, wf, pr.FY, G::LS, acc); break;
            default: tile_any<V>(tile, wf, pr.FX, pr.FY, G::LS, acc); break;
        }
        const int tx = threadIdx.x % G::NTX, ty = threadIdx.x / G::NTX;
        const long long oxt = ox0 + tx * G::TX;
#pragma unroll
        for (int j = 0; j < G::TY; ++j) {
            const long long oy = oy0 + ty * G::TY + j;
            if (oy >= pr.O1) break;
            float *orow = pr.out + oy * pr.O0;
#pragma unroll
            for (int i = 0; i < G::TX; ++i)
                if (oxt + i < pr.O0) orow[oxt + i] = acc[j][i];
        }
    }
}

}  // namespace afhip

using namespace afhip;

// Output-tile edge of a ragged tile variant (for the host's work list).
extern "C" int afhip_k_ragged_tile(int variant, int *bx, int *by) {
    switch (variant) {
        case 0: *bx = RGeo<0>::BX; *by = RGeo<0>::BY; return K_OK;
        case 1: *bx = RGeo<1>::BX; *by = RGeo<1>::BY; return K_OK;
        case 2: *bx = RGeo<2>::BX; *by = RGeo<2>::BY; return K_OK;
        default: return K_ERR_ARG;
    }
}

extern "C" int afhip_k_convolve2_ragged(int variant, const void *problems, const void *items,
                                        long long nitems, int grid, void *stream) {
    if (nitems <= 0) return K_OK;
    if (grid < 1) return K_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const RaggedProblem *pp = (const RaggedProblem *)problems;
    const int2 *ip = (const int2 *)items;
    switch (variant) {
        case 0: ragged_conv2<0><<<grid, RG_THREADS, 0, s>>>(pp, ip, nitems); break;
        case 1: ragged_conv2<1><<<grid, RG_THREADS, 0, s>>>(pp, ip, nitems); break;
        case 2: ragged_conv2<2><<<grid, RG_THREADS, 0, s>>>(pp, ip, nitems); break;
        default: return K_ERR_ARG;
    }
    return launch_status();
}
