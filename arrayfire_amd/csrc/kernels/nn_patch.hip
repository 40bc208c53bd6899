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
                                                    int ktiles, int dgrad) {
    constexpr int KC = PCfg<MODE>::KC;
    const int RS = fw * fh;
    const long long plane = (long long)ktiles * nchunks * RS * P_BM * KC;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < plane;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e;
        const int c = (int)(r % KC);
        r /= KC;
        const int m = (int)(r % P_BM);
        r /= P_BM;
        const int t = (int)(r % RS);
        r /= RS;
        const int cc = (int)(r % nchunks);
        const int kt = (int)(r / nchunks);
        const int mm = kt * P_BM + m, ch = cc * KC + c;
        const int i = t % fw, j = t / fw;
        float v = 0.f;
        if (mm < M && ch < CH) {
            long long idx;
            if (!dgrad)  // w_af[fw-1-i, fh-1-j, ch, mm], KCRS memory (mm*C + ch)
                idx = (((long long)mm * wC + ch) * fh + (fh - 1 - j)) * fw + (fw - 1 - i);
            else         // w_af[i, j, mm, ch] with mm the channel, ch the filter
                idx = (((long long)ch * wC + mm) * fh + j) * fw + i;
            v = pld(w, idx);
        }
        if constexpr (MODE == PM_F32) {
            reinterpret_cast<float *>(out)[e] = v;
        } else if constexpr (MODE == PM_BF16) {
            reinterpret_cast<uint16_t *>(out)[e] = f32_to_bf16_bits(v);
        } else if constexpr (MODE == PM_F16) {
            reinterpret_cast<uint16_t *>(out)[e] = pf32_to_f16_bits(v);
        } else {
            const uint16_t hb = f32_to_bf16_bits(v);
            reinterpret_cast<uint16_t *>(out)[e] = hb;
            reinterpret_cast<uint16_t *>(out)[plane + e] = f32_to_bf16_bits(v - bf16_bits_to_f32(hb));
        }
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

// what: 1 = pack the weights into ws, 2 = run the convolution with the
// weights already packed in ws, 3 = both.
template <int MODE, typename TIn, typename TOut>
static int launch(void *y, const void *x, const void *w, PatchParams &p, int dgrad, void *ws,
                  int what, hipStream_t s) {
    using Cfg = PCfg<MODE>;
    if (what & 1) {
        const long long plane = (long long)p.ktiles * p.nchunks * p.RS * P_BM * Cfg::KC;
        // weights: K x C x fh x fw of the original filter; for dgrad the
        // GEMM's M is the original C and its channels are the original K
        const int wC = dgrad ? p.K : p.C;
        long long blocks = (plane + 255) / 256;
        if (blocks > kMaxGridBlocks) blocks = kMaxGridBlocks;
        pack_weights<MODE, TIn><<<(unsigned)blocks, 256, 0, s>>>(ws, (const TIn *)w, p.K, p.C, p.fw,
                                                                   p.fh, wC, p.nchunks, p.ktiles, dgrad);
    }
    if (what & 2) {
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
