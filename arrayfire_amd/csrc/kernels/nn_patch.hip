// convolve2_nn forward (and the stride-1 data gradient, which is a forward
// convolution of dy with transposed, flipped weights) as an implicit GEMM
// that stages a channel-last input patch in LDS once per channel chunk and
// reuses it for every filter tap and every filter of the tile.
//
// Tile: 128 filters (M) x BN output pixels (N) of one image, BN = 256
// consecutive raster pixels; 8 waves as 2 (filters) x 4 (pixels), each wave a
// 64 x 64 block of MFMA 16x16 tiles with FP32 accumulators.
// Reduction: channel chunks of KC (32 for 16-bit operands, 16 for FP32) x
// filter taps. Per chunk the patch rows that the tile's pixels touch,
// [rows][cols][KC channels], are gathered from NCHW (coalesced along W) and
// stored transposed to LDS with the channels contiguous, padded to an 80-byte
// pixel pitch so that 16 lanes' 16-byte fragment reads hit distinct banks.
// Per (chunk, tap) step the 128 x KC weight tile comes from a pre-packed
// buffer (one 16-byte load per lane, prefetched a step ahead, double-buffered
// in LDS); B fragments are read straight from the patch at the tap's offset.
#include <cstdlib>
#include <mutex>
#include <type_traits>
#include <utility>

#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

typedef __bf16 pbf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 pf16x8 __attribute__((ext_vector_type(8)));
typedef float pf32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t pu16x8 __attribute__((ext_vector_type(8)));

enum PMode : int { PM_F32 = 0, PM_BF16 = 1, PM_F16 = 2, PM_SPLIT = 3 };

constexpr int P_BM = 128;       // filters per tile
constexpr int P_BN = 256;       // pixels per tile
constexpr int P_THREADS = 512;  // 8 waves

template <int MODE>
struct PCfg {
    static constexpr int KC = MODE == PM_F32 ? 16 : 32;      // channels per chunk
    static constexpr int ESZ = MODE == PM_F32 ? 4 : 2;        // operand bytes
    static constexpr int PLANES = MODE == PM_SPLIT ? 2 : 1;   // hi (+ lo)
    static constexpr int PIX_BYTES = 80;                      // LDS pitch per pixel
    static constexpr int A_TILE = P_BM * PIX_BYTES;           // bytes per A plane
};

struct PatchParams {
    int W, H, C, N;      // input [W,H,C,N]
    int OW, OH, K;       // output [OW,OH,K,N]
    int fw, fh, s0, s1, p0, p1, d0, d1;
    int RS, nchunks;     // taps, channel chunks
    int PW, PRmax;       // patch columns, max patch rows
    int tiles_per_img;   // ceil(OH*OW / BN)
    int ktiles;          // ceil(K / BM)
    long long ntiles;    // ktiles * tiles_per_img * N
    // pipelined kernel only
    int ngroups;         // ceil(RS / Q_TPS)
    int nseg;            // 16-pixel blocks per input row
    int items;           // 64 * nseg * PRmax 16-byte patch loads per stage
    int toffs[9];        // patch offset of tap t (static group schedules, RS <= 9)
};

__device__ inline uint16_t pf32_to_f16_bits(float v) {
    _Float16 h = (_Float16)v;
    return __builtin_bit_cast(uint16_t, h);
}

template <typename T>
__device__ inline float pld(const T *p, long long i);
template <>
__device__ inline float pld<float>(const float *p, long long i) {
    return p[i];
}
template <>
__device__ inline float pld<uint16_t>(const uint16_t *p, long long i) {  // bf16
    return bf16_bits_to_f32(p[i]);
}
template <>
__device__ inline float pld<_Float16>(const _Float16 *p, long long i) {
    return (float)p[i];
}

// ---------------------------------------------------------------- packing
// Packed weights: [ktile][chunk][tap][m (128)][KC] in the operand format
// (SPLIT: the hi plane, then the lo plane of the same shape). flip_fwd selects
// the forward weight w[fw-1-i, fh-1-j, ch, m]; otherwise the data-gradient
// weight w[i, j, m, ch] (transposed channels, flipped taps).
template <int MODE, typename TIn>
__global__ __launch_bounds__(256) void pack_weights(void *__restrict__ out,
                                                    const TIn *__restrict__ w, int M, int CH,
                                                    int fw, int fh, int wC, int nchunks,
                                                    int dgrad) {
    // one lane per packed element; blockIdx.x names the panel (ktile, chunk,
    // tap), whose P_BM x KC elements blockIdx.y covers 256 at a time with the
    // channel fastest, so a wave writes contiguous bytes
    constexpr int KC = PCfg<MODE>::KC;
    constexpr int PANEL = P_BM * KC;
    const int RS = fw * fh;
    const int pnl = blockIdx.x;
    const int t = pnl % RS, q = pnl / RS;
    const int cc = q % nchunks, kt = q / nchunks;
    const int i = t % fw, j = t / fw;
    const int idx = blockIdx.y * 256 + threadIdx.x;
    const int m = idx / KC, kc = idx % KC;
    const int mm = kt * P_BM + m, ch = cc * KC + kc;
    // fwd: w_af[fw-1-i, fh-1-j, ch, mm] (KCRS memory, channel stride RS);
    // dgrad: w_af[i, j, mm, ch] with mm the channel (filter stride wC*RS)
    const long long src = !dgrad ? ((long long)mm * wC + ch) * RS + (fh - 1 - j) * fw + (fw - 1 - i)
                                 : ((long long)ch * wC + mm) * RS + j * fw + i;
    const float v = (mm < M && ch < CH) ? pld(w, src) : 0.f;
    const long long e = (long long)pnl * PANEL + idx;
    if constexpr (MODE == PM_F32) {
        reinterpret_cast<float *>(out)[e] = v;
    } else {
        const uint16_t h = MODE == PM_F16 ? pf32_to_f16_bits(v) : f32_to_bf16_bits(v);
        reinterpret_cast<uint16_t *>(out)[e] = h;
        if constexpr (MODE == PM_SPLIT)
            reinterpret_cast<uint16_t *>(out)[(long long)gridDim.x * PANEL + e] =
                f32_to_bf16_bits(v - bf16_bits_to_f32(h));
    }
}

// ---------------------------------------------------------------- main kernel
// PASS (0 forward, 1 data gradient) changes nothing in the code; it gives the
// two uses distinct kernel names in profiles.
template <int MODE, typename TIn, typename TOut, int PASS>
__global__ __launch_bounds__(P_THREADS, 4) void conv_patch(TOut *__restrict__ y,
                                                        const TIn *__restrict__ x,
                                                        const char *__restrict__ wpk,
                                                        PatchParams p) {
    using Cfg = PCfg<MODE>;
    constexpr int KC = Cfg::KC;
    constexpr int PB = Cfg::PIX_BYTES;
    extern __shared__ __attribute__((aligned(16))) char smem[];
    // layout: A[2 buffers][PLANES][128 x 80 B], then patch[PLANES][rows*PW x 80 B]
    char *lA = smem;
    char *lP = smem + 2 * Cfg::PLANES * Cfg::A_TILE;
    const int patch_bytes = p.PRmax * p.PW * PB;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 2, wn = wave & 3;

    // tile -> (filter tile, image, pixel tile); pixel tiles fastest so that
    // neighbouring workgroups share input rows in L2
    long long t = blockIdx.x;
    const int ptile = (int)(t % p.tiles_per_img);
    t /= p.tiles_per_img;
    const int n = (int)(t % p.N);
    const int kt = (int)(t / p.N);
    const int P = p.OH * p.OW;
    const int q0 = ptile * P_BN;
    const int q1 = min(q0 + P_BN, P) - 1;
    const int oh_a = q0 / p.OW;
    const int ih0 = oh_a * p.s1 - p.p1;
    const int PR = (q1 / p.OW - oh_a) * p.s1 + (p.fh - 1) * p.d1 + 1;

    // B fragment bases: pixel j = wn*64 + ni*16 + lane%16 of the tile
    int bbase[4];
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
        int q = q0 + wn * 64 + ni * 16 + (lane & 15);
        if (q > q1) q = q1;  // dead pixels read a live one; never stored
        const int oh = q / p.OW, ow = q - (q / p.OW) * p.OW;
        bbase[ni] = (((oh - oh_a) * p.s1) * p.PW + ow * p.s0) * PB + (lane >> 4) * 16;
    }
    const int abase = (wm * 64 + (lane & 15)) * PB + (lane >> 4) * 16;

    pf32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = pf32x4{0.f, 0.f, 0.f, 0.f};

    // packed weight stream for this filter tile: steps (chunk, tap), each a
    // 128 x KC tile = 8 KB per plane; lane loads 16 B at offset tid*16
    const long long step_bytes = (long long)P_BM * KC * Cfg::ESZ;  // 8192
    const long long plane_bytes = (long long)p.ktiles * p.nchunks * p.RS * step_bytes;
    const char *wsrc = wpk + (long long)kt * p.nchunks * p.RS * step_bytes;
    const int nsteps = p.nchunks * p.RS;
    // A tile LDS position of this lane's 16 bytes: row tid/4, quarter tid%4
    const int a_st = (tid >> 2) * PB + (tid & 3) * 16;

    // weight tiles are prefetched two steps ahead: areg holds step s,
    // anext step s + 1 (its load was issued two steps earlier)
    uint4 areg[Cfg::PLANES], anext[Cfg::PLANES];
#pragma unroll
    for (int pl = 0; pl < Cfg::PLANES; ++pl) {
        areg[pl] = *reinterpret_cast<const uint4 *>(wsrc + pl * plane_bytes + tid * 16);
        anext[pl] = nsteps > 1 ? *reinterpret_cast<const uint4 *>(wsrc + pl * plane_bytes +
                                                                   step_bytes + tid * 16)
                               : areg[pl];
    }

    const long long HW = (long long)p.H * p.W;
    const TIn *xn = x + (long long)n * p.C * HW;
    
    for (int cc = 0; cc < p.nchunks; ++cc) {
        __syncthreads();  // previous chunk's patch no longer read
        // ---- gather the patch. Wave w owns channels w, w+8, ... of the chunk;
        // lane = column. The loads of two rows of every owned channel are
        // issued before any LDS store so that their latencies overlap.
        for (int col0 = 0; col0 < p.PW; col0 += 64) {
            const int col = col0 + lane;
            const int iw = col - p.p0;
            const bool inw = col < p.PW && iw >= 0 && iw < p.W;
            for (int r0 = 0; r0 < PR; r0 += 2) {
                float v[KC / 8][2];
#pragma unroll
                for (int cu = 0; cu < KC / 8; ++cu) {
                    const int ch = cc * KC + wave + 8 * cu;
#pragma unroll
                    for (int rr = 0; rr < 2; ++rr) {
                        const int ih = ih0 + r0 + rr;
                        v[cu][rr] = 0.f;
                        if (inw && ch < p.C && r0 + rr < PR && ih >= 0 && ih < p.H)
                            v[cu][rr] = pld(xn, (long long)ch * HW + (long long)ih * p.W + iw);
                    }
                }
                if (col < p.PW) {
#pragma unroll
                    for (int cu = 0; cu < KC / 8; ++cu)
#pragma unroll
                        for (int rr = 0; rr < 2; ++rr) {
                            if (r0 + rr >= PR) continue;
                            char *dst = lP + ((r0 + rr) * p.PW + col) * PB + (wave + 8 * cu) * Cfg::ESZ;
                            const float val = v[cu][rr];
                            if constexpr (MODE == PM_F32) {
                                *reinterpret_cast<float *>(dst) = val;
                            } else if constexpr (MODE == PM_BF16) {
                                *reinterpret_cast<uint16_t *>(dst) = f32_to_bf16_bits(val);
                            } else if constexpr (MODE == PM_F16) {
                                *reinterpret_cast<uint16_t *>(dst) = pf32_to_f16_bits(val);
                            } else {
                                const uint16_t hb = f32_to_bf16_bits(val);
                                *reinterpret_cast<uint16_t *>(dst) = hb;
                                *reinterpret_cast<uint16_t *>(dst + patch_bytes) =
                                    f32_to_bf16_bits(val - bf16_bits_to_f32(hb));
                            }
                        }
                }
            }
        }
        for (int tap = 0; tap < p.RS; ++tap) {
            const int s = cc * p.RS + tap;
            char *aw = lA + (s & 1) * Cfg::PLANES * Cfg::A_TILE;
#pragma unroll
            for (int pl = 0; pl < Cfg::PLANES; ++pl)
                *reinterpret_cast<uint4 *>(aw + pl * Cfg::A_TILE + a_st) = areg[pl];
            __syncthreads();
#pragma unroll
            for (int pl = 0; pl < Cfg::PLANES; ++pl) {
                areg[pl] = anext[pl];
                if (s + 2 < nsteps)
                    anext[pl] = *reinterpret_cast<const uint4 *>(wsrc + pl * plane_bytes +
                                                                 (s + 2) * step_bytes + tid * 16);
            }
            const int ti = tap % p.fw, tj = tap / p.fw;
            const int toff = (tj * p.d1 * p.PW + ti * p.d0) * PB;
            if constexpr (MODE == PM_F32) {
                // KC = 16: one float4 per lane covers the chunk, 4 MFMAs
                float4 af[4], bfr[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    af[i] = *reinterpret_cast<const float4 *>(aw + abase + i * 16 * PB);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    bfr[j] = *reinterpret_cast<const float4 *>(lP + bbase[j] + toff);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].x, bfr[j].x, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].y, bfr[j].y, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].z, bfr[j].z, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].w, bfr[j].w, acc[i][j], 0, 0, 0);
                    }
            } else {
                pu16x8 af[4], bfr[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    af[i] = *reinterpret_cast<const pu16x8 *>(aw + abase + i * 16 * PB);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    bfr[j] = *reinterpret_cast<const pu16x8 *>(lP + bbase[j] + toff);
                if constexpr (MODE == PM_SPLIT) {
                    pu16x8 al[4], bl[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        al[i] = *reinterpret_cast<const pu16x8 *>(aw + Cfg::A_TILE + abase + i * 16 * PB);
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        bl[j] = *reinterpret_cast<const pu16x8 *>(lP + patch_bytes + bbase[j] + toff);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                __builtin_bit_cast(pbf16x8, al[i]), __builtin_bit_cast(pbf16x8, bfr[j]),
                                acc[i][j], 0, 0, 0);
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                __builtin_bit_cast(pbf16x8, af[i]), __builtin_bit_cast(pbf16x8, bl[j]),
                                acc[i][j], 0, 0, 0);
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                __builtin_bit_cast(pbf16x8, af[i]), __builtin_bit_cast(pbf16x8, bfr[j]),
                                acc[i][j], 0, 0, 0);
                        }
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            if constexpr (MODE == PM_BF16)
                                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                    __builtin_bit_cast(pbf16x8, af[i]), __builtin_bit_cast(pbf16x8, bfr[j]),
                                    acc[i][j], 0, 0, 0);
                            else
                                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                                    __builtin_bit_cast(pf16x8, af[i]), __builtin_bit_cast(pf16x8, bfr[j]),
                                    acc[i][j], 0, 0, 0);
                        }
                }
            }
        }
    }

    // ---- epilogue: acc[i][j][v] = C[filter][pixel] with
    // filter = kt*128 + wm*64 + i*16 + 4*(lane/16) + v, pixel = q0 + wn*64 + j*16 + lane%16.
    // Each wave stages 16 filters x 64 pixels at a time in its own 4 KB of
    // LDS and writes them back as whole 64-pixel row segments (16 B per lane).
    __syncthreads();  // patch and A tiles are dead; reuse LDS
    float *stage = reinterpret_cast<float *>(smem) + wave * (16 * 64);
    const long long P64 = P;
    TOut *yn = y + (long long)n * p.K * P64;
    const int pix0 = q0 + wn * 64;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                stage[((lane >> 4) * 4 + v) * 64 + j * 16 + (lane & 15)] = acc[i][j][v];
        __syncthreads();
        // 16 rows x 64 pixels; lane reads 4 pixels: row = lane/16 + 4*rr
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int row = (lane >> 4) + 4 * rr;
            const int px = (lane & 15) * 4;
            const float4 val = *reinterpret_cast<const float4 *>(stage + row * 64 + px);
            const int k = kt * P_BM + wm * 64 + i * 16 + row;
            const int q = pix0 + px;
            if (k >= p.K) continue;
            TOut *dst = yn + (long long)k * P64 + q;
            const float vv[4] = {val.x, val.y, val.z, val.w};
            const bool full = q + 3 <= q1;
            if constexpr (sizeof(TOut) == 4) {
                if (full && ((P64 & 3) == 0)) {
                    *reinterpret_cast<float4 *>(dst) = val;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (q + e <= q1) dst[e] = vv[e];
                }
            } else {
                uint16_t h[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    h[e] = MODE == PM_F16 ? pf32_to_f16_bits(vv[e]) : f32_to_bf16_bits(vv[e]);
                uint16_t *d16 = reinterpret_cast<uint16_t *>(dst);
                if (full && ((P64 & 3) == 0)) {
                    *reinterpret_cast<uint2 *>(d16) =
                        make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (q + e <= q1) d16[e] = h[e];
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- pipelined kernel
// Same GEMM and packed weights as conv_patch, organised as a software
// pipeline for one 512-lane workgroup per CU:
//  * persistent: workgroup b walks tiles b', b' + G, ... (b' = XCD-contiguous
//    renumbering of b), so neighbouring tiles share halo rows in one L2;
//  * the patch of the next (tile, channel chunk) stage is read from HBM into
//    registers with 16-byte loads while the current stage runs its MFMAs, and
//    written to the other of two LDS patch buffers when the stage ends;
//  * weights are staged TPS taps at a time (one barrier per group instead of
//    one per tap), loaded two groups ahead;
//  * the MFMA operands are (patch, weights), so each lane's accumulators are
//    4 consecutive output pixels of one filter row: the epilogue stores them
//    directly (no LDS, no barrier) while the next tile's first patch and
//    weight groups are already in flight.
// LDS layout (patch and weights alike): 64 bytes per pixel (or filter row) =
// 4 groups of 16 bytes (8 bf16/f16 or 4 f32 channels); group g of pixel p sits
// at p*64 + ((g ^ ((p >> 1) & 3)) << 4). With that swizzle the MFMA fragment
// reads (16 consecutive pixels x 4 groups per ds_read_b128) run at the full
// LDS rate on gfx950 for any starting pixel; the patch row pitch is padded to
// OW mod 16 so that a fragment that wraps to the next output row still reads
// 16 positions congruent mod 16 to consecutive ones.
// Gather: each lane loads one 16-byte segment (SEG pixels of one channel) of
// NCHW; groups of SEG lanes hold SEG channels of the same segment and
// transpose it in registers (DPP / ds_swizzle exchanges), after which every
// lane writes one 16-byte channel group of one pixel.
// WM = 2: 128 filters x 256 pixels (8 waves as 2 x 4 of 64 x 64);
// WM = 1: 64 filters x 512 pixels (1 x 8), for GEMMs with at most 64 rows
// (the data gradient of a layer with 64 input channels).
// Requires W % SEG == 0 and a 16-byte aligned input, and operands stored in
// their input type (no conversion in the gather).
constexpr int Q_TPS = 3;    // max taps per weight-staging group (TPS = 3 when taps % 3 == 0, else 1)
constexpr int Q_THREADS = 512;
constexpr int Q_PB = 64;    // LDS bytes per pixel / filter row

template <int MODE, int WM>
struct QCfg {
    static constexpr int BM = 64 * WM;
    static constexpr int WN = 8 / WM;
    static constexpr int BN = 64 * WN;
    static constexpr int A_BUF = Q_TPS * BM * Q_PB;                               // bytes per weight buffer
    static constexpr int ASLOT = (Q_TPS * BM * 4 + Q_THREADS - 1) / Q_THREADS;  // 16-B loads per lane
    static constexpr int NSLOT = WM == 2 ? 4 : 8;  // 16-B patch loads per lane per stage (max)
};

__device__ inline int q_swz(int p, int g) { return (p << 6) + ((g ^ ((p >> 1) & 3)) << 4); }

// lane exchanges within groups of 2, 4 and 8 lanes
__device__ inline uint32_t q_x1(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
}
__device__ inline uint32_t q_x2(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
}
__device__ inline uint32_t q_x4(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x101F);  // lane ^ 4 (bit-mask mode)
}

// 8 x 8 transpose of 16-bit values over lanes l = 0..7 of each group: lane l
// holds row l (8 values, value e in half e%2 of dword e/2); afterwards lane l
// holds column l. Must run with all lanes active.
__device__ inline void q_tr8x8_b16(uint4 &v, int lane) {
    const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4;
    const uint32_t d0 = v.x, d1 = v.y, d2 = v.z, d3 = v.w;
    // columns 4-7 of rows 0-3 trade places with columns 0-3 of rows 4-7
    const uint32_t r0 = q_x4(b2 ? d0 : d2), r1 = q_x4(b2 ? d1 : d3);
    const uint32_t A0 = b2 ? r0 : d0, A1 = b2 ? r1 : d1, B0 = b2 ? d2 : r0, B1 = b2 ? d3 : r1;
    // then column pairs within each half
    const uint32_t u0 = q_x2(b1 ? A0 : A1), u1 = q_x2(b1 ? B0 : B1);
    const uint32_t V0 = b1 ? A1 : A0, V1 = b1 ? B1 : B0;
    // then single columns (16-bit halves)
    const uint32_t selS = b0 ? 0x05040100u : 0x07060302u, selK = b0 ? 0x07060302u : 0x05040100u;
    const uint32_t R0 = q_x1(__builtin_amdgcn_perm(V1, V0, selS)), R1 = q_x1(__builtin_amdgcn_perm(u1, u0, selS));
    const uint32_t K0 = __builtin_amdgcn_perm(V1, V0, selK), K1 = __builtin_amdgcn_perm(u1, u0, selK);
    const uint32_t P = b1 ? K1 : K0, Q = b1 ? R1 : R0, P2 = b1 ? K0 : K1, Q2 = b1 ? R0 : R1;
    const uint32_t x = b0 ? Q : P, y = b0 ? P : Q, x2 = b0 ? Q2 : P2, y2 = b0 ? P2 : Q2;
    v.x = __builtin_amdgcn_perm(y, x, 0x05040100u);
    v.y = __builtin_amdgcn_perm(y2, x2, 0x05040100u);
    v.z = __builtin_amdgcn_perm(y, x, 0x07060302u);
    v.w = __builtin_amdgcn_perm(y2, x2, 0x07060302u);
}

// 4 x 4 transpose of 32-bit values over lanes l = 0..3 of each group.
__device__ inline void q_tr4x4_b32(uint4 &v, int lane) {
    const bool b0 = lane & 1, b1 = lane & 2;
    const uint32_t d0 = v.x, d1 = v.y, d2 = v.z, d3 = v.w;
    const uint32_t r0 = q_x2(b1 ? d0 : d2), r1 = q_x2(b1 ? d1 : d3);
    const uint32_t a0 = b1 ? r0 : d0, a1 = b1 ? r1 : d1, c0 = b1 ? d2 : r0, c1 = b1 ? d3 : r1;
    const uint32_t u0 = q_x1(b0 ? a0 : a1), u1 = q_x1(b0 ? c0 : c1);
    const uint32_t k0 = b0 ? a1 : a0, k1 = b0 ? c1 : c0;
    v = b0 ? make_uint4(u0, k0, u1, k1) : make_uint4(k0, u0, k1, u1);
}

struct QTile {
    int id, n, kt, q0, q1, oh_a, ih0;
};

template <int I, int N, typename F>
__device__ __attribute__((always_inline)) inline void q_static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        q_static_for<I + 1, N>(f);
    }
}

template <int MODE, typename TIn, typename TOut, int WM, int TPS, int PASS, int NG>
__global__ __launch_bounds__(Q_THREADS, 2) void conv_pipe(TOut *__restrict__ y,
                                                         const TIn *__restrict__ x,
                                                         const char *__restrict__ wpk,
                                                         PatchParams p, int vec_out) {
    using Cfg = PCfg<MODE>;
    using Q = QCfg<MODE, WM>;
    constexpr int KC = Cfg::KC;
    constexpr int ESZ = Cfg::ESZ;
    static_assert(sizeof(TIn) == ESZ, "the gather stores operands in their input type");
    constexpr int BM = Q::BM, WN = Q::WN, BN = Q::BN;
    constexpr int SEG = 16 / ESZ;                 // pixels per 16-byte segment = channels per group
    constexpr int STEP = P_BM * KC * ESZ;         // packed bytes per (chunk, tap): 128 rows
    extern __shared__ __attribute__((aligned(16))) char smem[];
    char *lA = smem;                  // [2][Q_TPS][BM][64]
    char *lP = smem + 2 * Q::A_BUF;   // [2][PRmax][PW][64]   (PW = padded pitch)
    const int patch_bytes = p.PRmax * p.PW * Q_PB;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int g = lane >> 4;  // 16-byte channel group this lane reads / writes
    const int P = p.OH * p.OW;
    const long long HW = (long long)p.H * p.W;
    const int G = gridDim.x;
    const int lb = (int)xcd_remap(blockIdx.x, G);
    const int ntiles = (int)p.ntiles;    // < 2^31 (checked on the host)
    const int nrest = p.nseg * p.PRmax;  // (16-pixel block, row) pairs per stage; nseg = blocks per row

    // zero both patch buffers once: columns outside the gathered blocks are
    // never written again
    for (int i = tid; i < 2 * patch_bytes / 16; i += Q_THREADS)
        reinterpret_cast<uint4 *>(lP)[i] = make_uint4(0, 0, 0, 0);

    auto decode = [&](int id) __attribute__((always_inline)) {
        QTile t;
        t.id = id;
        const int pt = id % p.tiles_per_img, r = id / p.tiles_per_img;
        t.n = r % p.N;
        t.kt = r / p.N;
        t.q0 = pt * BN;
        t.q1 = min(t.q0 + BN, P) - 1;
        t.oh_a = t.q0 / p.OW;
        t.ih0 = t.oh_a * p.s1 - p.p1;
        return t;
    };

    // slot s of this lane: 16-pixel block / row pair rest = s*8 + wave
    // (wave-uniform), channel g*SEG + lane%SEG, segment (lane%16)/SEG of the block
    uint4 gx[Q::NSLOT];
    const int ch_l = g * SEG + (lane & (SEG - 1));
    const int px_l = (lane & 15) & ~(SEG - 1);
    // Loads go through raw buffer resources: an offset at or beyond the
    // resource's size reads zeros, so out-of-image segments and unused slots
    // need neither a branch nor a zeroing move (either would make the
    // compiler wait for every load in flight before reusing the registers).
    // The host keeps one image (C*H*W elements) and the packed weights below
    // 2^31 - 64 bytes.
    constexpr int kOOB = 0x7fffffc0;
    const int x_img_bytes = p.C * (int)HW * ESZ;
    auto issue_gather = [&](const QTile &t, int cc, bool live) __attribute__((always_inline)) {
        const auto xr = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<TIn *>(x + (long long)t.n * p.C * HW), (short)0, x_img_bytes, 0x00020000);
        const int c = cc * KC + ch_l;
#pragma unroll
        for (int s = 0; s < Q::NSLOT; ++s) {
            const int rest = s * 8 + wave;
            const int row = rest / p.nseg, blk = rest - (rest / p.nseg) * p.nseg;
            const int ih = t.ih0 + row, iw = blk * 16 + px_l;
            const bool ok = live && rest < nrest && ih >= 0 && ih < p.H && c < p.C && iw < p.W;
            const int off = ok ? (c * (int)HW + ih * p.W + iw) * ESZ : kOOB;
            gx[s] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, off, 0, 0));
        }
    };
    auto store_gather = [&](int buf) __attribute__((always_inline)) {
        char *dstb = lP + buf * patch_bytes;
#pragma unroll
        for (int s = 0; s < Q::NSLOT; ++s) {
            const int rest = s * 8 + wave;
            if (rest < nrest) {  // wave-uniform: the exchanges below see every lane
                uint4 v = gx[s];
                if constexpr (ESZ == 2)
                    q_tr8x8_b16(v, lane);
                else
                    q_tr4x4_b32(v, lane);
                const int row = rest / p.nseg, blk = rest - (rest / p.nseg) * p.nseg;
                const int col = p.p0 + blk * 16 + (lane & 15);
                if (col >= 0) *reinterpret_cast<uint4 *>(dstb + q_swz(row * p.PW + col, g)) = v;
            }
        }
    };

    // Weight groups are loaded two groups ahead into two register sets: set
    // k & 1 holds group k (the k-th group of this workgroup's sequence of
    // tiles lb, lb + G, ..., each tile nchunks * ngroups groups). A cursor
    // (tile step, chunk, group) walks two groups ahead of the compute.
    static_assert(Q::ASLOT <= 3, "weight slots");
    uint4 wa0 = make_uint4(0, 0, 0, 0), wa1 = wa0, wa2 = wa0, wb0 = wa0, wb1 = wa0, wb2 = wa0;
    int at = 0, acc_c = 0, ag = 0;        // cursor: tile step, chunk, group
    int akt = lb < ntiles ? (lb / p.tiles_per_img) / p.N : 0;
    auto advance = [&]() __attribute__((always_inline)) {
        if (++ag == p.ngroups) {
            ag = 0;
            if (++acc_c == p.nchunks) {
                acc_c = 0;
                ++at;
                const int id = lb + at * G;
                akt = id < ntiles ? (id / p.tiles_per_img) / p.N : 0;
            }
        }
    };
    // per-lane byte offset of weight slot s within a group (a = s*512 + tid;
    // lanes past the group's end read zeros and store nothing)
    const int w_total = p.ktiles * p.nchunks * p.RS * STEP;
    int w_lane[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int a = s * Q_THREADS + tid;
        const int tl = a / (BM * 4), r = (a >> 2) % BM, q = a & 3;
        w_lane[s] = a < TPS * BM * 4 ? tl * STEP + r * (KC * ESZ) + q * 16 : kOOB;
    }
    auto issue_A = [&](auto set) __attribute__((always_inline)) {  // loads the cursor's group, then advances it
        constexpr int S = decltype(set)::value;
        uint4 &r0 = S ? wb0 : wa0;
        uint4 &r1 = S ? wb1 : wa1;
        uint4 &r2 = S ? wb2 : wa2;
        // the group's uniform offset goes into the resource (scalar work);
        // a dead group gets an empty resource
        const bool live = lb + at * G < ntiles;
        const int src = ((akt * p.nchunks + acc_c) * p.RS + ag * TPS) * STEP;
        const auto wr = __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(wpk) + src, (short)0,
                                                          live ? w_total - src : 0, 0x00020000);
        r0 = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(wr, w_lane[0], 0, 0));
        if constexpr (Q::ASLOT > 1) r1 = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(wr, w_lane[1], 0, 0));
        if constexpr (Q::ASLOT > 2) r2 = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(wr, w_lane[2], 0, 0));
        advance();
    };
    auto store_A = [&](auto set, int buf) __attribute__((always_inline)) {
        constexpr int S = decltype(set)::value;
        constexpr int lim = TPS * BM * 4;
        auto st = [&](int s, const uint4 &v) __attribute__((always_inline)) {
            const int a = s * Q_THREADS + tid;
            const int tl = a / (BM * 4), r = (a >> 2) % BM, q = a & 3;
            if (a < lim) *reinterpret_cast<uint4 *>(lA + buf * Q::A_BUF + tl * BM * Q_PB + q_swz(r, q)) = v;
        };
        st(0, S ? wb0 : wa0);
        if constexpr (Q::ASLOT > 1) st(1, S ? wb1 : wa1);
        if constexpr (Q::ASLOT > 2) st(2, S ? wb2 : wa2);
    };
    using Set0 = std::integral_constant<int, 0>;
    using Set1 = std::integral_constant<int, 1>;
    using FragT = typename std::conditional<MODE == PM_F32, float4, pu16x8>::type;

    // Accumulators: the MFMA's A operand is the patch (16 pixels x 32
    // channels) and B the weights, so acc[i][j][v] = out[pixel][filter] with
    // pixel = q0 + wn*64 + j*16 + 4*(lane/16) + v and filter = kt*BM + wm*64 +
    // i*16 + lane%16: each lane owns 4 consecutive pixels of one filter row.
    pf32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = pf32x4{0.f, 0.f, 0.f, 0.f};

    // weight fragment rows wm*64 + i*16 + lane%16: the swizzle key depends on lane%16 only
    const int abase = q_swz(wm * 64 + (lane & 15), g);
    auto tile_bpos = [&](const QTile &t, int (&bp)[4]) __attribute__((always_inline)) {
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
            int q = t.q0 + wn * 64 + ni * 16 + (lane & 15);
            if (q > t.q1) q = t.q1;  // dead pixels read a live one; never stored
            const int oh = q / p.OW, ow = q - oh * p.OW;
            bp[ni] = ((oh - t.oh_a) * p.s1) * p.PW + ow * p.s0;
        }
    };
    // ---- epilogue of tile t: straight from the accumulators, 4 consecutive
    // pixels of one filter row per lane and fragment (no LDS, no barrier).
    // vec_out == 2 (4 | OH*OW, 16-byte aligned output, one output image under
    // 2^31 - 64 bytes): one buffer store per fragment, dead lanes dropped by
    // an offset past the resource's end; otherwise guarded stores.
    const int y_lane = ((wm * 64 + (lane & 15)) * P + wn * 64 + 4 * (lane >> 4)) * (int)sizeof(TOut);
    auto store_tile = [&](const QTile &t) __attribute__((always_inline)) {
        if (vec_out == 2) {
            const auto yr = __builtin_amdgcn_make_buffer_rsrc(y + (long long)t.n * p.K * (long long)P, (short)0,
                                                              p.K * P * (int)sizeof(TOut), 0x00020000);
            const int tofs = (t.kt * BM * P + t.q0) * (int)sizeof(TOut);
            const int krem = p.K - t.kt * BM - wm * 64 - (lane & 15);   // filter rows left for fragment row 0
            const int qrem = t.q1 - t.q0 - wn * 64 - 4 * (lane >> 4);   // pixels left for fragment column 0
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const bool ok = i * 16 < krem && j * 16 <= qrem;
                    const int off = ok ? y_lane + tofs + (i * 16 * P + j * 16) * (int)sizeof(TOut) : kOOB;
                    const pf32x4 v = acc[i][j];
                    if constexpr (sizeof(TOut) == 4) {
                        __builtin_amdgcn_raw_buffer_store_b128(
                            __builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned int, v), yr, off, 0, 0);
                    } else {
                        __attribute__((ext_vector_type(2))) unsigned int pk;
                        if constexpr (MODE == PM_F16) {
                            pk.x = (uint32_t)pf32_to_f16_bits(v[0]) | ((uint32_t)pf32_to_f16_bits(v[1]) << 16);
                            pk.y = (uint32_t)pf32_to_f16_bits(v[2]) | ((uint32_t)pf32_to_f16_bits(v[3]) << 16);
                        } else {
                            pk.x = (uint32_t)f32_to_bf16_bits(v[0]) | ((uint32_t)f32_to_bf16_bits(v[1]) << 16);
                            pk.y = (uint32_t)f32_to_bf16_bits(v[2]) | ((uint32_t)f32_to_bf16_bits(v[3]) << 16);
                        }
                        __builtin_amdgcn_raw_buffer_store_b64(pk, yr, off, 0, 0);
                    }
                    acc[i][j] = pf32x4{0.f, 0.f, 0.f, 0.f};
                }
            return;
        }
        TOut *yn = y + (long long)t.n * p.K * (long long)P;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int kf = t.kt * BM + wm * 64 + i * 16 + (lane & 15);
            if (kf >= p.K) continue;
            TOut *yrow = yn + (long long)kf * P;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int q = t.q0 + wn * 64 + j * 16 + 4 * (lane >> 4);
                if (q > t.q1) continue;
                const pf32x4 v = acc[i][j];
                if constexpr (sizeof(TOut) == 4) {
                    if (vec_out && q + 3 <= t.q1) {
                        *reinterpret_cast<float4 *>(yrow + q) = make_float4(v[0], v[1], v[2], v[3]);
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (q + e <= t.q1) yrow[q + e] = v[e];
                    }
                } else {
                    uint16_t hb[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) hb[e] = MODE == PM_F16 ? pf32_to_f16_bits(v[e]) : f32_to_bf16_bits(v[e]);
                    uint16_t *d16 = reinterpret_cast<uint16_t *>(yrow) + q;
                    if (vec_out && q + 3 <= t.q1) {
                        *reinterpret_cast<uint2 *>(d16) =
                            make_uint2((uint32_t)hb[0] | ((uint32_t)hb[1] << 16), (uint32_t)hb[2] | ((uint32_t)hb[3] << 16));
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (q + e <= t.q1) d16[e] = hb[e];
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = pf32x4{0.f, 0.f, 0.f, 0.f};
    };
    // MFMAs of one tap from fragments fa (weights) and fb (patch)
    auto mma_tap = [&](const FragT (&fa)[4], const FragT (&fb)[4]) __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if constexpr (MODE == PM_F32) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[j].x, fa[i].x, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[j].y, fa[i].y, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[j].z, fa[i].z, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[j].w, fa[i].w, acc[i][j], 0, 0, 0);
                } else if constexpr (MODE == PM_BF16) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(pbf16x8, fb[j]), __builtin_bit_cast(pbf16x8, fa[i]), acc[i][j], 0, 0, 0);
                } else {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                        __builtin_bit_cast(pf16x8, fb[j]), __builtin_bit_cast(pf16x8, fa[i]), acc[i][j], 0, 0, 0);
                }
            }
    };
    // fragments of one tap: weights from buffer aw (tap row block u), patch
    // from lp at pixel positions bp + to
    auto load_tap = [&](const char *aw, int u, const char *lp, const int (&bp)[4], int to, FragT (&fa)[4],
                        FragT (&fb)[4]) __attribute__((always_inline)) {
        const char *au = aw + u * BM * Q_PB + abase;
#pragma unroll
        for (int i = 0; i < 4; ++i) fa[i] = *reinterpret_cast<const FragT *>(au + i * 16 * Q_PB);
#pragma unroll
        for (int j = 0; j < 4; ++j) fb[j] = *reinterpret_cast<const FragT *>(lp + q_swz(bp[j] + to, g));
    };


    int sbuf = 0;         // patch buffer of the current stage
    int cc = 0, gq = 0;   // stage (channel chunk) and weight group within the current tile
    QTile cur = decode(lb);
    if (cur.id < ntiles) {
        issue_A(Set0{});
        issue_A(Set1{});
        issue_gather(cur, 0, true);
    }
    __syncthreads();  // zeroed patch buffers visible
    if (cur.id >= ntiles) return;  // uniform: lb is the same for the whole workgroup
    QTile nxt = decode(cur.id + G);

    // Patch fragment addresses. Runtime schedule: q_swz(bpos[j] + tap
    // offset) per tap. Static 3 x 3 schedule (fw = 3, d0 = 1, d1*PW a
    // multiple of 8, checked on the host): group gq holds filter row gq and
    // taps u = 0..2 of it; a row step of d1*PW pixels keeps the swizzle key
    // ((p >> 1) & 3), so the 3 column phases are swizzled once per tile and
    // a tap's address is pb[j][u] + gq * d1*PW*64.
    int bpos[4];      // patch position of this lane's pixel in fragment j (tap 0)
    int pb[4][3];     // static 3 x 3 schedule: swizzled address of fragment j, column phase u
    auto set_tile = [&]() __attribute__((always_inline)) {
        tile_bpos(cur, bpos);
        if constexpr (NG == 3) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int u = 0; u < 3; ++u) pb[j][u] = q_swz(bpos[j] + u, g);
        }
    };
    set_tile();
    const int rowb = p.d1 * p.PW * Q_PB;  // bytes per filter-row step in the patch

    // Start of a weight group with register set / LDS weight buffer S: at the
    // first group of a stage publish the stage's patch, then publish the
    // group's weights, barrier, refill the set with group k + 2 and (first
    // group) start the next stage's patch. Loads are issued unconditionally
    // (dead ones read zeros), so the number of loads between a load and its
    // use is the same on every path and the compiler waits only for it.
    auto begin_group = [&](auto set, auto first) __attribute__((always_inline)) {
        constexpr int S = decltype(set)::value;
        if (static_cast<bool>(first)) store_gather(sbuf);
        store_A(set, S);
        __syncthreads();
        issue_A(set);
        if (static_cast<bool>(first)) {
            const bool chunk = cc + 1 < p.nchunks;
            issue_gather(chunk ? cur : nxt, chunk ? cc + 1 : 0, chunk || nxt.id < ntiles);
        }
    };

    // TPS taps from weight buffer aw and patch lp, straight-line: fragments
    // of tap u + 1 are read from LDS while the MFMAs of tap u run.
    // paddr(u, j) is the patch byte offset of fragment j of the group's tap u.
    auto mma_group = [&](const char *aw, const char *lp, auto paddr) __attribute__((always_inline)) {
        FragT af[2][4], bfr[2][4];
        auto load_frags = [&](int u, FragT (&fa)[4], FragT (&fb)[4]) __attribute__((always_inline)) {
            const char *au = aw + u * BM * Q_PB + abase;
#pragma unroll
            for (int i = 0; i < 4; ++i) fa[i] = *reinterpret_cast<const FragT *>(au + i * 16 * Q_PB);
#pragma unroll
            for (int j = 0; j < 4; ++j) fb[j] = *reinterpret_cast<const FragT *>(lp + paddr(u, j));
        };
        load_frags(0, af[0], bfr[0]);
#pragma unroll
        for (int u = 0; u < TPS; ++u) {
            if (u + 1 < TPS) load_frags(u + 1, af[(u + 1) & 1], bfr[(u + 1) & 1]);
            mma_tap(af[u & 1], bfr[u & 1]);
        }
    };

    // End of a stage: flip the patch buffer; after the last stage of a tile
    // store it and move to the next. Returns true when no tile is left.
    auto end_stage = [&]() __attribute__((always_inline)) -> bool {
        sbuf ^= 1;
        if (++cc == p.nchunks) {
            cc = 0;
            store_tile(cur);
            cur = nxt;
            if (cur.id >= ntiles) return true;
            nxt = decode(cur.id + G);
            set_tile();
        }
        return false;
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    if constexpr (NG > 0) {
        // static schedule: NG groups per stage (p.ngroups == NG), unrolled, so
        // every register set, group index and tap offset is a constant; two
        // stages per iteration when NG is odd (the set parity alternates)
        auto stage = [&](auto par) __attribute__((always_inline)) -> bool {
            constexpr int P0 = decltype(par)::value;
            q_static_for<0, NG>([&](auto gi) __attribute__((always_inline)) {
                constexpr int GI = decltype(gi)::value;
                constexpr int S = (P0 + GI) & 1;
                begin_group(std::integral_constant<int, S>{}, std::integral_constant<bool, GI == 0>{});
                if constexpr (NG == 3) {
                    mma_group(lA + S * Q::A_BUF, lP + sbuf * patch_bytes + GI * rowb,
                              [&](int u, int j) __attribute__((always_inline)) { return pb[j][u]; });
                } else {
                    mma_group(lA + S * Q::A_BUF, lP + sbuf * patch_bytes,
                              [&](int u, int j) __attribute__((always_inline)) {
                                  return q_swz(bpos[j] + p.toffs[GI * TPS + u], g);
                              });
                }
            });
            return end_stage();
        };
        if constexpr (NG % 2 == 0) {
            while (!stage(I0{})) {
            }
        } else {
            while (!stage(I0{}) && !stage(I1{})) {
            }
        }
    } else {
        // runtime schedule: the loop alternates register sets S = 0, 1, so
        // each set stays a fixed group of registers (a copy between sets
        // would wait for the loads at once)
        auto group = [&](auto set) __attribute__((always_inline)) -> bool {
            constexpr int S = decltype(set)::value;
            begin_group(set, gq == 0);
            const int tap0 = gq * TPS;
            mma_group(lA + S * Q::A_BUF, lP + sbuf * patch_bytes, [&](int u, int j) __attribute__((always_inline)) {
                const int tap = tap0 + u;
                const int ti = tap % p.fw, tj = tap / p.fw;
                return q_swz(bpos[j] + tj * p.d1 * p.PW + ti * p.d0, g);
            });
            if (++gq == p.ngroups) {
                gq = 0;
                return end_stage();
            }
            return false;
        };
        while (!group(Set0{}) && !group(Set1{})) {
        }
    }
}

// ---------------------------------------------------------------- weights-resident kernel
// conv_wres: the forward pass, or the stride-1 data gradient (the same
// convolution of dy with the unflipped, transposed filters), for GEMMs whose
// reduction (NCH channel chunks of 32 x 9 taps) fits in registers, 16-bit
// operands.
// Unit of work: (tile of 16*NF*NQ filters, image, band of RB output rows);
// one persistent workgroup of 8 waves per CU walks units lb, lb + G, ... in
// filter-tile order. Wave w: wq = w % NQ picks 16*NF filters of the tile,
// pg = w / NQ one of PG = 8 / NQ pixel groups. Each wave keeps the B
// fragments of its filters for every k-step in VGPRs (NCH*9 x NF fragments
// of 4 VGPRs: NF = 2 for up to 64 channels, NF = 1 for up to 128), reloaded
// only when the unit's filter tile changes, so the MFMA loop reads nothing
// but the patch from LDS: per k-step 2 A fragments (16 pixels x 32 channels
// at the tap's offset, read a k-step ahead) feed 2*NF MFMAs, with no
// barrier. The band's patch (every channel chunk; channel-last
// 64-byte pixel rows swizzled as q_swz, the segment transposes of conv_pipe)
// arrives in two row pieces: the loads of both are issued at once, the first
// piece is stored, and the output rows it covers are computed while the
// second piece's loads land; then the second piece is stored and the
// remaining rows are computed. Wave w walks the 16-pixel blocks pg, pg + PG,
// ... in pairs, then singly, and stores each finished pair straight from the
// accumulators (4 consecutive output pixels of one filter per lane).
constexpr int WR_THREADS = 512;
constexpr int WR_WAVES = WR_THREADS / 64;

struct WresParams {
    int W, H, C, N, OW, OH, K;
    int s0, s1, p0, p1, d1;
    int PW, PR, RB;           // patch pitch (pixels, multiple of 8), patch rows, band rows
    int PRa, RBa;             // rows of the first patch piece, output rows it covers
    int nseg, nbands, ntw;    // 16-pixel input blocks per row, bands per image, filter tiles
    long long nunits;         // ntw * N * nbands
    int w_bytes;              // packed weight bytes (buffer resource size)
    int vec;                  // OW % 4 == 0 and one output image under 2^31 - 64 bytes: buffer stores
};

template <int MODE, typename TIn, typename TOut, int NCH, int NQ, int NF, int PASS>
__global__ __launch_bounds__(WR_THREADS, 1) void conv_wres(TOut *__restrict__ y, const TIn *__restrict__ x,
                                                           const char *__restrict__ wpk, WresParams p) {
    static_assert(MODE == PM_BF16 || MODE == PM_F16, "16-bit operands");
    static_assert(NCH * NF <= 4, "resident weights: at most 144 VGPRs");
    constexpr int RS = 9, NKS = NCH * RS, SEG = 8, KC = 32, PG = WR_WAVES / NQ, FPW = 16 * NF;
    constexpr int kOOB = 0x7fffffc0;
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform: slot arithmetic stays scalar
    const int wq = wave % NQ, pg = wave / NQ;
    const int g = lane >> 4;
    const int chunk_bytes = p.PR * p.PW * Q_PB;
    const long long HW = (long long)p.H * p.W;
    const int P = p.OH * p.OW;
    const int G = gridDim.x;
    const int lb = (int)xcd_remap(blockIdx.x, G);

    // zero the patch once: columns outside the gathered blocks stay zero
    for (int i = tid; i < NCH * chunk_bytes / 16; i += WR_THREADS)
        reinterpret_cast<uint4 *>(smem)[i] = make_uint4(0, 0, 0, 0);
    __syncthreads();   // before any wave stores gathered data over the zeros

    pu16x8 wb[NKS][NF];
    int cur_kt = -1;
    const auto wr = __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(wpk), (short)0, p.w_bytes, 0x00020000);
    auto load_weights = [&](int kt) __attribute__((always_inline)) {
        // packed [kt128][chunk][tap][128 rows][32 ch]; this wave's filters
        // f0 + fi * 16 + lane % 16 with f0 = kt * FPW * NQ + wq * FPW stay in
        // one 128-row panel
        const int f0 = kt * FPW * NQ + wq * FPW;
        const int k128 = f0 >> 7, row = (f0 & 127) + (lane & 15);
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
            for (int fi = 0; fi < NF; ++fi) {
                const int panel = (k128 * NCH + ks / RS) * RS + ks % RS;
                const int off = panel * (P_BM * KC * 2) + (row + fi * 16) * (KC * 2) + g * 16;
                wb[ks][fi] = __builtin_bit_cast(pu16x8, __builtin_amdgcn_raw_buffer_load_b128(wr, off, 0, 0));
            }
    };

    const int ch_l = g * SEG + (lane & (SEG - 1));
    const int px_l = (lane & 15) & ~(SEG - 1);
    const int x_img_bytes = p.C * (int)HW * 2;
    const int rowb = p.d1 * p.PW * Q_PB;   // bytes per filter-row step in the patch

    for (long long u = lb; u < p.nunits; u += G) {
        const int band = (int)(u % p.nbands);
        const long long r1 = u / p.nbands;
        const int n = (int)(r1 % p.N), kt = (int)(r1 / p.N);
        const int r0 = band * p.RB;
        const int rows = min(p.RB, p.OH - r0);
        const int Pu = rows * p.OW;            // output pixels of the unit
        const int ih0 = r0 * p.s1 - p.p1;
        // ---- gather: slot (chunk, patch row, 16-pixel block) of a piece of
        // patch rows [lo, lo + nr); WS slots per wave per round, slot s of a
        // round at rest = s0 + s * 8 + wave (decoded incrementally: uniform)
        const auto xr = __builtin_amdgcn_make_buffer_rsrc(const_cast<TIn *>(x + (long long)n * p.C * HW), (short)0,
                                                          x_img_bytes, 0x00020000);
        struct Slot {
            int cc, row, blk;
        };
        auto decode = [&](auto ws, Slot (&sl)[decltype(ws)::value], int s0, int lo, int nr)
                          __attribute__((always_inline)) {
            constexpr int WS = decltype(ws)::value;
            const int per_chunk = nr * p.nseg, rest = s0 + wave;
            int cc = rest / per_chunk, rem = rest - cc * per_chunk;
            int row = rem / p.nseg, blk = rem - row * p.nseg;
#pragma unroll
            for (int s = 0; s < WS; ++s) {
                sl[s] = Slot{cc, lo + row, blk};
                blk += WR_WAVES;
                while (blk >= p.nseg) {
                    blk -= p.nseg;
                    ++row;
                }
                while (row >= nr) {
                    row -= nr;
                    ++cc;
                }
            }
        };
        auto issue = [&](auto ws, uint4 (&gx)[decltype(ws)::value], const Slot (&sl)[decltype(ws)::value])
                         __attribute__((always_inline)) {
            constexpr int WS = decltype(ws)::value;
#pragma unroll
            for (int s = 0; s < WS; ++s) {
                const int c = sl[s].cc * KC + ch_l;
                const int ih = ih0 + sl[s].row, iw = sl[s].blk * 16 + px_l;
                const bool ok = sl[s].cc < NCH && ih >= 0 && ih < p.H && c < p.C && iw < p.W;
                const int off = ok ? (c * (int)HW + ih * p.W + iw) * 2 : kOOB;
                gx[s] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xr, off, 0, 0));
            }
        };
        auto store_slots = [&](auto ws, uint4 (&gx)[decltype(ws)::value], const Slot (&sl)[decltype(ws)::value])
                               __attribute__((always_inline)) {
            constexpr int WS = decltype(ws)::value;
#pragma unroll
            for (int s = 0; s < WS; ++s) {
                if (sl[s].cc < NCH) {  // wave-uniform: the transpose sees every lane
                    uint4 v = gx[s];
                    q_tr8x8_b16(v, lane);
                    const int col = p.p0 + sl[s].blk * 16 + (lane & 15);
                    if (col >= 0 && col < p.PW)
                        *reinterpret_cast<uint4 *>(smem + sl[s].cc * chunk_bytes + q_swz(sl[s].row * p.PW + col, g)) =
                            v;
                }
            }
        };
        using W8 = std::integral_constant<int, 8>;
        using W4 = std::integral_constant<int, 4>;
        const int PRa = min(p.PRa, p.PR), PRb = p.PR - PRa;
        const int slots_a = NCH * PRa * p.nseg, slots_b = NCH * PRb * p.nseg;
        const bool overlap_b = slots_b > 0 && slots_b <= 4 * WR_WAVES;   // one round of 4 slots
        // loads in the order they are needed: the first piece, the weights
        // (when the filter tile changes; the first MFMA waits for them), then
        // the second piece
        uint4 gb[4];
        {
            uint4 ga[8];
            Slot sa[8];
            decode(W8{}, sa, 0, 0, PRa);
            issue(W8{}, ga, sa);
            if (kt != cur_kt) {                // uniform across the workgroup
                load_weights(kt);
                cur_kt = kt;
            }
            if (overlap_b) {
                Slot sb[4];
                decode(W4{}, sb, 0, PRa, PRb);
                issue(W4{}, gb, sb);
            }
            store_slots(W8{}, ga, sa);
            for (int s0 = 8 * WR_WAVES; s0 < slots_a; s0 += 8 * WR_WAVES) {
                decode(W8{}, sa, s0, 0, PRa);
                issue(W8{}, ga, sa);
                store_slots(W8{}, ga, sa);
            }
        }
        __syncthreads();

        // ---- compute: 16-pixel blocks pg, pg + PG, ... of the unit; blocks
        // below nb1 read only the first piece
        const int nb = (Pu + 15) >> 4;
        const int nb1 = min(nb, (min(p.RBa, rows) * p.OW) >> 4);
        const int mine = nb > pg ? (nb - pg + PG - 1) / PG : 0;
        const int mine1 = nb1 > pg ? (nb1 - pg + PG - 1) / PG : 0;
        // the lane's pixel q = b * 16 + lane % 16 of its current block as (oh,
        // ow), walked PG * 16 pixels per block without a division
        int oh, ow;
        {
            const int q0 = pg * 16 + (lane & 15);
            oh = q0 / p.OW;
            ow = q0 - oh * p.OW;
        }
        auto next_block = [&](int (&pb)[3]) __attribute__((always_inline)) {
            int o = oh, w = ow;
            if (o >= rows) {            // dead pixels read the unit's last live one; never stored
                o = rows - 1;
                w = p.OW - 1;
            }
            const int base = o * p.s1 * p.PW + w * p.s0;
#pragma unroll
            for (int ph = 0; ph < 3; ++ph) pb[ph] = q_swz(base + ph, g);
            ow += PG * 16;
            while (ow >= p.OW) {
                ow -= p.OW;
                ++oh;
            }
        };
        // output offsets (elements, within image n) of this lane's filters;
        // -1 for filters past K
        int kfo[NF];
#pragma unroll
        for (int fi = 0; fi < NF; ++fi) {
            const int kf = kt * FPW * NQ + wq * FPW + fi * 16 + (lane & 15);
            kfo[fi] = kf < p.K ? kf * P + r0 * p.OW : -1;
        }
        TOut *yn = y + (long long)n * p.K * P;
        const auto yr = __builtin_amdgcn_make_buffer_rsrc(yn, (short)0, p.vec ? p.K * P * (int)sizeof(TOut) : 0,
                                                          0x00020000);
        auto store = [&](int b, int fi, const pf32x4 &v) __attribute__((always_inline)) {
            const int q = b * 16 + 4 * (lane >> 4);
            if (p.vec) {
                // OW % 4 == 0: the 4 pixels are all live or all dead, and
                // 16-byte (8-byte) aligned; dead lanes store past the end
                const int off = (kfo[fi] >= 0 && q < Pu) ? (kfo[fi] + q) * (int)sizeof(TOut) : kOOB;
                if constexpr (sizeof(TOut) == 4) {
                    __builtin_amdgcn_raw_buffer_store_b128(
                        __builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned int, v), yr, off, 0, 0);
                } else {
                    __attribute__((ext_vector_type(2))) unsigned int pk;
                    if constexpr (MODE == PM_F16) {
                        pk.x = (uint32_t)pf32_to_f16_bits(v[0]) | ((uint32_t)pf32_to_f16_bits(v[1]) << 16);
                        pk.y = (uint32_t)pf32_to_f16_bits(v[2]) | ((uint32_t)pf32_to_f16_bits(v[3]) << 16);
                    } else {
                        pk.x = (uint32_t)f32_to_bf16_bits(v[0]) | ((uint32_t)f32_to_bf16_bits(v[1]) << 16);
                        pk.y = (uint32_t)f32_to_bf16_bits(v[2]) | ((uint32_t)f32_to_bf16_bits(v[3]) << 16);
                    }
                    __builtin_amdgcn_raw_buffer_store_b64(pk, yr, off, 0, 0);
                }
                return;
            }
            if (kfo[fi] < 0 || q >= Pu) return;
            TOut *d = yn + kfo[fi] + q;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (q + k < Pu) {
                    if constexpr (sizeof(TOut) == 4)
                        d[k] = v[k];
                    else
                        reinterpret_cast<uint16_t *>(d)[k] =
                            MODE == PM_F16 ? pf32_to_f16_bits(v[k]) : f32_to_bf16_bits(v[k]);
                }
            }
        };
        auto mfma = [&](const pu16x8 &a, const pu16x8 &b, pf32x4 &acc) __attribute__((always_inline)) {
            if constexpr (MODE == PM_BF16)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(pbf16x8, a),
                                                              __builtin_bit_cast(pbf16x8, b), acc, 0, 0, 0);
            else
                acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(pf16x8, a),
                                                             __builtin_bit_cast(pf16x8, b), acc, 0, 0, 0);
        };
        // NB blocks at once: k-steps (chunk, filter row, column phase) fully
        // unrolled so every weight fragment is a fixed register; the A
        // fragments of k-step ks + 1 are read while the MFMAs of ks run
        auto run = [&](auto nbk, int i0) __attribute__((always_inline)) {
            constexpr int NB = decltype(nbk)::value;
            int pa[NB][3];
#pragma unroll
            for (int bi = 0; bi < NB; ++bi) next_block(pa[bi]);
            pf32x4 acc[NB][NF];
#pragma unroll
            for (int bi = 0; bi < NB; ++bi)
#pragma unroll
                for (int fi = 0; fi < NF; ++fi) acc[bi][fi] = pf32x4{0.f, 0.f, 0.f, 0.f};
            constexpr int PD = 2;   // k-steps of LDS reads in flight ahead of the MFMAs
            pu16x8 a[PD + 1][NB];
            auto fetch = [&](int ks, pu16x8 (&dst)[NB]) __attribute__((always_inline)) {
                const int cc = ks / RS, t = ks % RS;
                const char *base = smem + cc * chunk_bytes + (t / 3) * rowb;
#pragma unroll
                for (int bi = 0; bi < NB; ++bi) dst[bi] = *reinterpret_cast<const pu16x8 *>(base + pa[bi][t % 3]);
            };
#pragma unroll
            for (int ks = 0; ks < PD; ++ks) fetch(ks, a[ks]);
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks) {
                // the LDS reads of k-step ks + PD are issued before this
                // k-step's MFMAs (the barrier keeps the scheduler from sinking
                // them), so each read has PD k-steps of MFMAs to land
                if (ks + PD < NKS) fetch(ks + PD, a[(ks + PD) % (PD + 1)]);
                __builtin_amdgcn_sched_barrier(0x0006);   // only VALU / SALU work may cross
#pragma unroll
                for (int bi = 0; bi < NB; ++bi)
#pragma unroll
                    for (int fi = 0; fi < NF; ++fi) mfma(a[ks % (PD + 1)][bi], wb[ks][fi], acc[bi][fi]);
            }
#pragma unroll
            for (int bi = 0; bi < NB; ++bi)
#pragma unroll
                for (int fi = 0; fi < NF; ++fi) store(pg + PG * (i0 + bi), fi, acc[bi][fi]);
        };
        int i0 = 0;
        for (; i0 + 2 <= mine1; i0 += 2) run(std::integral_constant<int, 2>{}, i0);
        for (; i0 < mine1; ++i0) run(std::integral_constant<int, 1>{}, i0);
        if (PRb > 0) {
            // the second piece: its loads have been landing during the first
            // piece's MFMAs (or are issued now when they take several rounds)
            if (overlap_b) {
                Slot sb[4];
                decode(W4{}, sb, 0, PRa, PRb);
                store_slots(W4{}, gb, sb);
            } else {
                for (int s0 = 0; s0 < slots_b; s0 += 4 * WR_WAVES) {
                    Slot sb[4];
                    decode(W4{}, sb, s0, PRa, PRb);
                    issue(W4{}, gb, sb);
                    store_slots(W4{}, gb, sb);
                }
            }
            __syncthreads();
        }
        for (; i0 + 2 <= mine; i0 += 2) run(std::integral_constant<int, 2>{}, i0);
        for (; i0 < mine; ++i0) run(std::integral_constant<int, 1>{}, i0);
        __syncthreads();   // the next unit's gather overwrites the patch
    }
}

template <int MODE>
static size_t lds_bytes(const PatchParams &p) {
    using Cfg = PCfg<MODE>;
    const size_t main_loop =
        (size_t)2 * Cfg::PLANES * Cfg::A_TILE + (size_t)Cfg::PLANES * p.PRmax * p.PW * Cfg::PIX_BYTES;
    const size_t epilogue = (size_t)(P_THREADS / 64) * 16 * 64 * sizeof(float);
    return main_loop > epilogue ? main_loop : epilogue;
}

constexpr size_t kPatchLdsLimit = 160 * 1024;

// Geometry of the patch path for an output of (OW, OH) from an input of
// (W, H). Returns false when the patch does not fit in LDS.
template <int MODE>
static bool plan(PatchParams &p) {
    p.RS = p.fw * p.fh;
    p.nchunks = (p.C + PCfg<MODE>::KC - 1) / PCfg<MODE>::KC;
    p.PW = (p.OW - 1) * p.s0 + (p.fw - 1) * p.d0 + 1;
    const int P = p.OH * p.OW;
    const int span = P_BN >= P ? p.OH : (min(P_BN - 1, P - 1) / p.OW + 2);
    const int rows = min(span, p.OH);
    p.PRmax = (rows - 1) * p.s1 + (p.fh - 1) * p.d1 + 1;
    p.tiles_per_img = (P + P_BN - 1) / P_BN;
    p.ktiles = (p.K + P_BM - 1) / P_BM;
    p.ntiles = (long long)p.ktiles * p.tiles_per_img * p.N;
    return lds_bytes<MODE>(p) <= kPatchLdsLimit && p.PW > 0 && p.PRmax > 0;
}

// Geometry of the pipelined kernel (q, from the conv_patch plan p; q.PW is
// the padded patch row pitch). Returns false when the operands are converted
// on the way in, the input is not a whole number of aligned 16-byte segments
// per row, a stage's patch needs more than NSLOT loads per lane, or the two
// patch buffers and two weight buffers do not fit in LDS. Setting
// AFHIP_NN_PIPE=0 in the environment disables the pipelined kernel.
template <int MODE, typename TIn>
static bool plan_pipe(PatchParams &q, const PatchParams &p, const void *x, int &wm) {
    if constexpr (MODE == PM_SPLIT || sizeof(TIn) != PCfg<MODE>::ESZ) {
        return false;
    } else {
        const char *off = std::getenv("AFHIP_NN_PIPE");
        if (off && off[0] == '0') return false;
        constexpr int SEG = 16 / (int)sizeof(TIn);
        q = p;
        if (q.W % SEG != 0 || (reinterpret_cast<uintptr_t>(x) & 15) != 0) return false;
        wm = q.K <= 64 ? 1 : 2;
        const int BM = 64 * wm, BN = 64 * (8 / wm);
        const int P = q.OH * q.OW;
        const int span = BN >= P ? q.OH : (min(BN - 1, P - 1) / q.OW + 2);
        const int rows = min(span, q.OH);
        q.PRmax = (rows - 1) * q.s1 + (q.fh - 1) * q.d1 + 1;
        q.tiles_per_img = (P + BN - 1) / BN;
        q.ktiles = (q.K + BM - 1) / BM;
        q.ntiles = (long long)q.ktiles * q.tiles_per_img * q.N;
        q.ngroups = q.RS % Q_TPS == 0 ? q.RS / Q_TPS : q.RS;  // TPS = 3 or 1
        q.nseg = (q.W + 15) / 16;  // 16-pixel blocks per input row
        // padded row pitch: holds every gathered block; for stride 1 it is
        // congruent to OW mod 16 (fragments that wrap rows stay conflict-free)
        const int need = max(p.PW, q.p0 + q.nseg * 16);
        q.PW = q.s0 == 1 ? q.OW + 16 * ((need - q.OW + 15) / 16) : need;
        const int nslot = wm == 2 ? QCfg<MODE, 2>::NSLOT : QCfg<MODE, 1>::NSLOT;
        if ((long long)q.nseg * q.PRmax > 8LL * nslot || q.ntiles >= (1LL << 31) - (1LL << 20)) return false;
        q.items = 64 * q.nseg * q.PRmax;
        // buffer-resource offsets are 32-bit: one image and the packed
        // weights must stay below 2^31 - 64 bytes
        const long long img_bytes = (long long)q.C * q.H * q.W * (long long)sizeof(TIn);
        const long long w_bytes = (long long)q.ktiles * q.nchunks * q.RS * P_BM * PCfg<MODE>::KC * PCfg<MODE>::ESZ;
        if (img_bytes >= (1LL << 31) - 64 || w_bytes >= (1LL << 31) - 64) return false;
        for (int t = 0; t < 9; ++t) q.toffs[t] = t < q.RS ? (t / q.fw) * q.d1 * q.PW + (t % q.fw) * q.d0 : 0;
        const size_t lds = (size_t)2 * Q_TPS * BM * Q_PB + (size_t)2 * q.PRmax * q.PW * Q_PB;
        return lds <= kPatchLdsLimit;
    }
}

static int device_cu_count() {
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cus[dev] <= 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

template <int MODE, typename TIn, typename TOut>
static int launch_pipe(void *y, const void *x, const void *ws, const PatchParams &q, int wm, int dgrad,
                       hipStream_t s) {
    const int BM = 64 * wm;
    const size_t patch = (size_t)2 * q.PRmax * q.PW * Q_PB;
    // schedule: TPS = 3 taps per weight group when taps % 3 == 0, else 1;
    // a static (unrolled) group schedule for 3 x 3 taps (3 groups of 3, when
    // a filter-row step keeps the patch swizzle) and for 1 tap (1 group of
    // 1), a runtime one otherwise
    const bool t3 = q.RS % Q_TPS == 0;
    const bool s33 = q.RS == 9 && q.fw == 3 && q.d0 == 1 && (q.d1 * q.PW) % 8 == 0;
    const int sched = t3 ? (s33 ? 0 : 1) : (q.RS == 1 ? 2 : 3);
    const size_t lds = (size_t)2 * Q_TPS * BM * Q_PB + patch;
    const int P = q.OH * q.OW;
    using Fn = void (*)(TOut *, const TIn *, const char *, PatchParams, int);
    Fn fn;
#define AFHIP_PIPE_FN(WMv, DG)                                                                     \
    (sched == 0   ? conv_pipe<MODE, TIn, TOut, WMv, 3, DG, 3>                                     \
     : sched == 1 ? conv_pipe<MODE, TIn, TOut, WMv, 3, DG, 0>                                     \
     : sched == 2 ? conv_pipe<MODE, TIn, TOut, WMv, 1, DG, 1>                                     \
                  : conv_pipe<MODE, TIn, TOut, WMv, 1, DG, 0>)
    if (wm == 1)
        fn = dgrad ? AFHIP_PIPE_FN(1, 1) : AFHIP_PIPE_FN(1, 0);
    else
        fn = dgrad ? AFHIP_PIPE_FN(2, 1) : AFHIP_PIPE_FN(2, 0);
#undef AFHIP_PIPE_FN
    const int slot = ((wm - 1) * 2 + dgrad) * 4 + sched;
    static bool attr_set[16] = {false};
    if (!attr_set[slot]) {
        if (hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kPatchLdsLimit) != hipSuccess)
            return K_ERR_RUNTIME;
        attr_set[slot] = true;
    }
    // persistent grid: as many workgroups as can be resident at once
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, Q_THREADS, lds) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    long long grid = (long long)device_cu_count() * per_cu;
    if (grid > q.ntiles) grid = q.ntiles;
    // 1: whole 4-pixel groups and an aligned output; 2: also one output image
    // addressable with 32-bit buffer offsets
    int vec_out = (P % 4 == 0) && ((reinterpret_cast<uintptr_t>(y) & 15) == 0);
    if (vec_out && (long long)q.K * P * (long long)sizeof(TOut) < (1LL << 31) - 64) vec_out = 2;
    fn<<<dim3((unsigned)grid), Q_THREADS, lds, s>>>((TOut *)y, (const TIn *)x, (const char *)ws, q, vec_out);
    return launch_status();
}

// Geometry of the weights-resident forward kernel. Served: 16-bit operands
// stored as such, W a whole number of 8-pixel segments, a 16-byte aligned
// input, 3x3 filters with dilation 1 along W, C <= 64 (at most 2 chunks of 32
// channels: 18 k-steps of weights in registers), and a band of output rows
// whose patch (every chunk) fits in LDS. Filter tiles are 128 filters wide
// (4 x 32) when K > 64, else 64 (2 x 32, with 4 pixel groups). The band
// height RB minimises the units per CU x (blocks per unit + a per-unit
// overhead of about two blocks); the first patch piece covers the band's
// first ceil(RB / 2) output rows. AFHIP_NN_WRES=0 in the environment
// disables the kernel.
template <int MODE, typename TIn>
static bool plan_wres(WresParams &w, int &nch, int &nq, int &nf, const PatchParams &p, const void *x, int dgrad,
                      long long wsz) {
    if constexpr (MODE == PM_SPLIT || MODE == PM_F32 || sizeof(TIn) != 2) {
        return false;
    } else {
        const char *off = std::getenv("AFHIP_NN_WRES");
        if (off && off[0] == '0') return false;
        if (dgrad) {
            const char *offd = std::getenv("AFHIP_NN_WRES_DGRAD");
            if (offd && offd[0] == '0') return false;
        }
        // up to 64 channels with 32 filters per wave, up to 128 with 16
        if (p.fw != 3 || p.fh != 3 || p.d0 != 1 || p.C <= 0 || p.C > 128) return false;
        if (p.W % 8 != 0 || (reinterpret_cast<uintptr_t>(x) & 15) != 0) return false;
        if (p.p0 < 0 || p.p1 < 0) return false;
        nch = (p.C + 31) / 32;
        nf = nch <= 2 ? 2 : 1;
        nq = p.K > 32 * nf ? 4 : 2;
        w.W = p.W; w.H = p.H; w.C = p.C; w.N = p.N; w.OW = p.OW; w.OH = p.OH; w.K = p.K;
        w.s0 = p.s0; w.s1 = p.s1; w.p0 = p.p0; w.p1 = p.p1; w.d1 = p.d1;
        w.nseg = (p.W + 15) / 16;
        // pitch: every image column plus the widest tap reach, rounded up to a
        // multiple of 8 pixels so that a filter-row step keeps the swizzle;
        // gathered columns past the pitch (the dead end of the last 16-pixel
        // block) are not stored
        const int need = max((p.OW - 1) * p.s0 + 2 * p.d0 + 1, p.p0 + p.W);
        w.PW = (need + 7) & ~7;
        w.ntw = (p.K + 16 * nf * nq - 1) / (16 * nf * nq);
        const long long cus = device_cu_count();
        long long best = -1;
        for (int rb = 1; rb <= p.OH; ++rb) {
            const int pr = (rb - 1) * p.s1 + 2 * p.d1 + 1;
            if ((size_t)nch * pr * w.PW * Q_PB > kPatchLdsLimit) break;
            const long long nb = (p.OH + rb - 1) / rb;
            const long long units = (long long)w.ntw * p.N * nb;
            const long long per_cu = (units + cus - 1) / cus;
            const long long cost = per_cu * (((long long)rb * p.OW + 15) / 16 + 2);
            if (best < 0 || cost <= best) {
                best = cost;
                w.RB = rb;
            }
        }
        if (best < 0) return false;
        // AFHIP_NN_WRES_RB=<rows> in the environment forces the band height
        // (within the LDS budget)
        if (const char *rb_env = std::getenv("AFHIP_NN_WRES_RB")) {
            const int rb = std::atoi(rb_env);
            if (rb >= 1 && rb <= p.OH &&
                (size_t)nch * ((rb - 1) * p.s1 + 2 * p.d1 + 1) * w.PW * Q_PB <= kPatchLdsLimit)
                w.RB = rb;
        }
        w.PR = (w.RB - 1) * p.s1 + 2 * p.d1 + 1;
        w.RBa = (w.RB + 1) / 2;
        w.PRa = (w.RBa - 1) * p.s1 + 2 * p.d1 + 1;
        if (w.PRa >= w.PR) w.RBa = w.RB, w.PRa = w.PR;
        w.nbands = (p.OH + w.RB - 1) / w.RB;
        w.nunits = (long long)w.ntw * p.N * w.nbands;
        if (wsz >= (1LL << 31) - 64 || (long long)p.C * p.H * p.W * 2 >= (1LL << 31) - 64) return false;
        if ((long long)p.N * p.K * p.OH * p.OW >= (1LL << 40)) return false;
        w.w_bytes = (int)wsz;
        w.vec = 0;   // set by the launcher from the output
        return true;
    }
}

template <int MODE, typename TIn, typename TOut, int PASS>
static int launch_wres(void *y, const void *x, const void *ws, const WresParams &w, int nch, int nq, int nf,
                       hipStream_t s) {
    using Fn = void (*)(TOut *, const TIn *, const char *, WresParams);
    Fn fn;
    if (nf == 2)
        fn = nch == 1 ? (nq == 4 ? conv_wres<MODE, TIn, TOut, 1, 4, 2, PASS> : conv_wres<MODE, TIn, TOut, 1, 2, 2, PASS>)
                      : (nq == 4 ? conv_wres<MODE, TIn, TOut, 2, 4, 2, PASS> : conv_wres<MODE, TIn, TOut, 2, 2, 2, PASS>);
    else
        fn = nch == 3 ? (nq == 4 ? conv_wres<MODE, TIn, TOut, 3, 4, 1, PASS> : conv_wres<MODE, TIn, TOut, 3, 2, 1, PASS>)
                      : (nq == 4 ? conv_wres<MODE, TIn, TOut, 4, 4, 1, PASS> : conv_wres<MODE, TIn, TOut, 4, 2, 1, PASS>);
    const size_t lds = (size_t)nch * w.PR * w.PW * Q_PB;
    static bool attr_set[8] = {false, false, false, false, false, false, false, false};
    const int slot = (nch - 1) * 2 + (nq == 4);
    if (!attr_set[slot]) {
        if (hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kPatchLdsLimit) != hipSuccess)
            return K_ERR_RUNTIME;
        attr_set[slot] = true;
    }
    long long grid = device_cu_count();
    if (grid > w.nunits) grid = w.nunits;
    WresParams q = w;
    q.vec = (w.OW % 4 == 0) && (reinterpret_cast<uintptr_t>(y) & 15) == 0 &&
            (long long)w.K * w.OH * w.OW * (long long)sizeof(TOut) < (1LL << 31) - 64;
    fn<<<dim3((unsigned)grid), WR_THREADS, lds, s>>>((TOut *)y, (const TIn *)x, (const char *)ws, q);
    return launch_status();
}

// what: 1 = pack the weights into ws, 2 = run the convolution with the
// weights already packed in ws, 3 = both.
template <int MODE, typename TIn, typename TOut>
static int launch(void *y, const void *x, const void *w, PatchParams &p, int dgrad, void *ws,
                  int what, hipStream_t s) {
    using Cfg = PCfg<MODE>;
    if (what & 1) {
        // weights: K x C x fh x fw of the original filter; for dgrad the
        // GEMM's M is the original C and its channels are the original K
        const int wC = dgrad ? p.K : p.C;
        const long long panels = (long long)p.ktiles * p.nchunks * p.RS;
        if (panels * P_BM * Cfg::KC >= (1LL << 31)) return K_ERR_SIZE;
        const dim3 grid((unsigned)panels, P_BM * Cfg::KC / 256);
        pack_weights<MODE, TIn><<<grid, 256, 0, s>>>(ws, (const TIn *)w, p.K, p.C, p.fw, p.fh, wC,
                                                      p.nchunks, dgrad);
    }
    if (what & 2) {
        PatchParams q;
        int wm = 0;
        if constexpr (MODE != PM_SPLIT && sizeof(TIn) == PCfg<MODE>::ESZ) {
            if constexpr (MODE == PM_BF16 || MODE == PM_F16) {
                WresParams wr;
                int nch = 0, nq = 0, nf = 0;
                const long long wsz = (long long)p.ktiles * p.nchunks * p.RS * P_BM * PCfg<MODE>::KC * 2;
                if (plan_wres<MODE, TIn>(wr, nch, nq, nf, p, x, dgrad, wsz))
                    return dgrad ? launch_wres<MODE, TIn, TOut, 1>(y, x, ws, wr, nch, nq, nf, s)
                                 : launch_wres<MODE, TIn, TOut, 0>(y, x, ws, wr, nch, nq, nf, s);
            }
            if (plan_pipe<MODE, TIn>(q, p, x, wm)) return launch_pipe<MODE, TIn, TOut>(y, x, ws, q, wm, dgrad, s);
        }
        const size_t lds = lds_bytes<MODE>(p);
        auto fn = dgrad ? conv_patch<MODE, TIn, TOut, 1> : conv_patch<MODE, TIn, TOut, 0>;
        static bool attr_set[2] = {false, false};
        if (!attr_set[dgrad]) {
            if (hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kPatchLdsLimit) != hipSuccess)
                return K_ERR_RUNTIME;
            attr_set[dgrad] = true;
        }
        fn<<<dim3((unsigned)p.ntiles), P_THREADS, lds, s>>>((TOut *)y, (const TIn *)x,
                                                              (const char *)ws, p);
    }
    return launch_status();
}

static int mode_of(int in_type, int precision) {
    switch (precision) {
        case 0: return in_type == DT_F32 ? PM_F32 : in_type == DT_BF16 ? PM_BF16 : in_type == DT_F16 ? PM_F16 : -1;
        case 1: return PM_BF16;
        case 2: return PM_F16;
        case 3: return PM_SPLIT;
        case 4: return PM_F32;
        default: return -1;
    }
}

// Fills p for the GEMM "out[M=K', pixels] = conv(in, packed w)". For the
// forward pass the dims are the layer's; for the data gradient the input is
// dy [OW,OH,K,N], the output dx [W,H,C,N], the padding (f-1)d - p.
static void fill(PatchParams &p, int dgrad, const long long *xd, const long long *wd,
                 const long long *yd, const long long *st, const long long *pd, const long long *dl) {
    p.fw = (int)wd[0];
    p.fh = (int)wd[1];
    p.d0 = (int)dl[0];
    p.d1 = (int)dl[1];
    p.N = (int)xd[3];
    if (!dgrad) {
        p.W = (int)xd[0]; p.H = (int)xd[1]; p.C = (int)xd[2];
        p.OW = (int)yd[0]; p.OH = (int)yd[1]; p.K = (int)wd[3];
        p.s0 = (int)st[0]; p.s1 = (int)st[1];
        p.p0 = (int)pd[0]; p.p1 = (int)pd[1];
    } else {
        p.W = (int)yd[0]; p.H = (int)yd[1]; p.C = (int)wd[3];
        p.OW = (int)xd[0]; p.OH = (int)xd[1]; p.K = (int)wd[2];
        p.s0 = 1; p.s1 = 1;
        p.p0 = (p.fw - 1) * p.d0 - (int)pd[0];
        p.p1 = (p.fh - 1) * p.d1 - (int)pd[1];
    }
}

}  // namespace afhip

using namespace afhip;

// Returns the workspace bytes the patch path needs (packed weights), or 0 when
// the geometry is not served by it (the caller then uses the gather path).
extern "C" long long afhip_k_conv2_nn_patch_workspace(int pass, int in_type, int precision,
                                                      const long long *xd, const long long *wd,
                                                      const long long *yd, const long long *st,
                                                      const long long *pd, const long long *dl) {
    if (pass == 1 && (st[0] != 1 || st[1] != 1)) return 0;
    const int mode = mode_of(in_type, precision);
    if (mode < 0) return 0;
    PatchParams p;
    fill(p, pass, xd, wd, yd, st, pd, dl);
    if (p.OW <= 0 || p.OH <= 0 || p.C <= 0 || p.K <= 0 || p.N <= 0) return 0;
    bool ok;
    long long kc;
    switch (mode) {
        case PM_F32: ok = plan<PM_F32>(p); kc = PCfg<PM_F32>::KC * 4; break;
        case PM_BF16: ok = plan<PM_BF16>(p); kc = 64; break;
        case PM_F16: ok = plan<PM_F16>(p); kc = 64; break;
        default: ok = plan<PM_SPLIT>(p); kc = 128; break;
    }
    if (!ok) return 0;
    return (long long)p.ktiles * p.nchunks * p.RS * P_BM * kc;
}

template <typename T>
struct PTag {
    using type = T;
};

extern "C" long long afhip_k_conv2_nn_patch_tiles(int pass, int in_type, int precision, const long long *xd,
                                                  const long long *wd, const long long *yd, const long long *st,
                                                  const long long *pd, const long long *dl) {
    if (afhip_k_conv2_nn_patch_workspace(pass, in_type, precision, xd, wd, yd, st, pd, dl) <= 0) return 0;
    const int mode = mode_of(in_type, precision);
    PatchParams p;
    fill(p, pass, xd, wd, yd, st, pd, dl);
    auto tiles = [&](auto md, auto tin) -> long long {
        constexpr int MD = decltype(md)::value;
        using TIn = typename decltype(tin)::type;
        plan<MD>(p);
        if constexpr (MD != PM_SPLIT && sizeof(TIn) == PCfg<MD>::ESZ) {
            if constexpr (MD == PM_BF16 || MD == PM_F16) {
                WresParams wr;
                int nch = 0, nq = 0, nf = 0;
                const long long wsz = (long long)p.ktiles * p.nchunks * p.RS * P_BM * PCfg<MD>::KC * 2;
                if (plan_wres<MD, TIn>(wr, nch, nq, nf, p, nullptr, pass, wsz)) return wr.nunits;
            }
            PatchParams q;
            int wm = 0;
            if (plan_pipe<MD, TIn>(q, p, nullptr, wm)) return q.ntiles;
        }
        return p.ntiles;
    };
    using I = std::integral_constant<int, PM_F32>;
    switch (mode) {
        case PM_F32: return tiles(I{}, PTag<float>{});
        case PM_BF16:
            return in_type == DT_BF16 ? tiles(std::integral_constant<int, PM_BF16>{}, PTag<uint16_t>{})
                                      : tiles(std::integral_constant<int, PM_BF16>{}, PTag<float>{});
        case PM_F16:
            return in_type == DT_F16 ? tiles(std::integral_constant<int, PM_F16>{}, PTag<_Float16>{})
                                     : tiles(std::integral_constant<int, PM_F16>{}, PTag<float>{});
        default: return tiles(std::integral_constant<int, PM_SPLIT>{}, PTag<float>{});
    }
}

static int patch_call(int what, int pass, int in_type, int out_type, int precision, void *out,
                      const void *in, const void *w, const long long *xd, const long long *wd,
                      const long long *yd, const long long *st, const long long *pd,
                      const long long *dl, void *workspace, long long workspace_bytes, void *stream) {
    const long long need = afhip_k_conv2_nn_patch_workspace(pass, in_type, precision, xd, wd, yd, st, pd, dl);
    if (need == 0) return K_ERR_NOT_SUPPORTED;
    if (workspace == nullptr || workspace_bytes < need) return K_ERR_ARG;
    const int mode = mode_of(in_type, precision);
    PatchParams p;
    fill(p, pass, xd, wd, yd, st, pd, dl);
    hipStream_t s = (hipStream_t)stream;
    switch (mode) {
        case PM_F32: plan<PM_F32>(p); break;
        case PM_BF16: plan<PM_BF16>(p); break;
        case PM_F16: plan<PM_F16>(p); break;
        default: plan<PM_SPLIT>(p); break;
    }
#define GO(MD, TI, TO) return launch<MD, TI, TO>(out, in, w, p, pass, workspace, what, s)
    if (in_type == DT_F32 && out_type == DT_F32) {
        switch (mode) {
            case PM_F32: GO(PM_F32, float, float);
            case PM_BF16: GO(PM_BF16, float, float);
            case PM_F16: GO(PM_F16, float, float);
            default: GO(PM_SPLIT, float, float);
        }
    }
    if (in_type == DT_BF16 && mode == PM_BF16) {
        if (out_type == DT_F32) GO(PM_BF16, uint16_t, float);
        if (out_type == DT_BF16) GO(PM_BF16, uint16_t, uint16_t);
    }
    if (in_type == DT_F16 && mode == PM_F16) {
        if (out_type == DT_F32) GO(PM_F16, _Float16, float);
        if (out_type == DT_F16) GO(PM_F16, _Float16, uint16_t);
    }
#undef GO
    return K_ERR_TYPE;
}

extern "C" int afhip_k_conv2_nn_patch(int pass, int in_type, int out_type, int precision, void *out,
                                      const void *in, const void *w, const long long *xd,
                                      const long long *wd, const long long *yd, const long long *st,
                                      const long long *pd, const long long *dl, void *workspace,
                                      long long workspace_bytes, void *stream) {
    return patch_call(3, pass, in_type, out_type, precision, out, in, w, xd, wd, yd, st, pd, dl,
                      workspace, workspace_bytes, stream);
}

extern "C" int afhip_k_conv2_nn_patch_pack(int pass, int in_type, int precision, const void *w,
                                           const long long *xd, const long long *wd,
                                           const long long *yd, const long long *st,
                                           const long long *pd, const long long *dl, void *workspace,
                                           long long workspace_bytes, void *stream) {
    // the output type does not matter for packing; pass the input's
    const int ot = in_type == DT_F32 ? DT_F32 : in_type;
    return patch_call(1, pass, in_type, ot, precision, nullptr, nullptr, w, xd, wd, yd, st, pd, dl,
                      workspace, workspace_bytes, stream);
}

extern "C" int afhip_k_conv2_nn_patch_run(int pass, int in_type, int out_type, int precision, void *out,
                                          const void *in, const long long *xd, const long long *wd,
                                          const long long *yd, const long long *st, const long long *pd,
                                          const long long *dl, const void *workspace,
                                          long long workspace_bytes, void *stream) {
    return patch_call(2, pass, in_type, out_type, precision, out, in, nullptr, xd, wd, yd, st, pd, dl,
                      const_cast<void *>(workspace), workspace_bytes, stream);
}

// ---------------------------------------------------------------- no-alloc variant
// afhip_k_conv2_nn_patch_noalloc: the same launches with the packed-weight
// workspace in a file-scope device array (one copy per GPU, held until the
// process ends) instead of a caller or library allocation, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops). Its
// device address is looked up once per device with that device current, and
// only that address is passed to kernels. Calls that overlap on one GPU
// (different streams, or host threads) share the array and race; calls
// queued in order on one stream do not. Nothing in it needs to start at
// zero: the pack kernel writes every byte the convolution reads.
namespace afhip {
constexpr long long kNoAllocScratchBytes = 16LL << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64] = {};

static int noalloc_scratch(void **ptr) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return K_ERR_RUNTIME;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return K_ERR_RUNTIME;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return K_OK;
}
}  // namespace afhip

extern "C" long long afhip_k_conv2_nn_patch_noalloc_limit(void) { return kNoAllocScratchBytes; }

extern "C" int afhip_k_conv2_nn_patch_noalloc(int pass, int in_type, int out_type, int precision, void *out,
                                              const void *in, const void *w, const long long *xd,
                                              const long long *wd, const long long *yd, const long long *st,
                                              const long long *pd, const long long *dl, void *stream) {
    const long long need = afhip_k_conv2_nn_patch_workspace(pass, in_type, precision, xd, wd, yd, st, pd, dl);
    if (need == 0) return K_ERR_NOT_SUPPORTED;
    if (need > kNoAllocScratchBytes) return K_ERR_SIZE;
    void *ws = nullptr;
    const int rc = noalloc_scratch(&ws);
    if (rc != K_OK) return rc;
    return patch_call(3, pass, in_type, out_type, precision, out, in, w, xd, wd, yd, st, pd, dl, ws,
                      kNoAllocScratchBytes, stream);
}
