// convolve2_nn filter gradient on the gfx950 matrix cores, 16-bit operands
// (BF16 or FP16) with FP32 accumulation:
//   dW[fw-1-i, fh-1-j, c, k] = sum_{n, oh, ow} x[n][c][oh*s1-p1+j*d1][ow*s0-p0+i] * dy[n][k][oh][ow]
//
// GEMM view per tap (i, j): D[k][c] += A[k][pixels] * B[pixels][c], with the
// reduction over output pixels. Both operands stay in their natural NCHW row
// layout: A is a block of dy rows, B the rows of x the block touches. A lane's
// 8 reduction values are 8 consecutive output columns of one row (columns are
// padded to a multiple of 8 with dy = 0), so its A fragment is one aligned
// 16-byte LDS read, and its B fragment is 8 x values at column stride S0
// starting at an offset that depends only on the tap. B fragments for all FW
// taps of a filter row come from one window of dword reads, realigned with
// v_alignbyte (S0 = 1) or v_perm (S0 = 2): the window is read once and serves
// FW taps x 4 filter tiles of MFMAs.
//
// Workgroup: 4 waves; tile = 64 filters x 64 channels x all taps; wave w owns
// channels 16w..16w+15 and all 64 filters (4 MFMA row tiles), so its
// accumulators are 4 x FW*fh tiles of 16x16. The reduction over N*OH*OW is cut
// into blocks of R_oh output rows of one image, staged in LDS; the blocks are
// split across workgroups, whose FP32 partial tiles a second kernel sums in a
// fixed order (deterministic).
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

typedef __bf16 wbf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 wf16x8 __attribute__((ext_vector_type(8)));
typedef float wf32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t wu32x4 __attribute__((ext_vector_type(4)));

constexpr int WG_BK = 64, WG_BC = 64, WG_THREADS = 256;
constexpr size_t WG_LDS_LIMIT = 150 * 1024;
constexpr size_t WG_PC_LDS = 160 * 1024;  // two block buffers of the producer/consumer kernel

struct WGParams {
    int W, H, C, N, OW, OH, K;
    int fh, s1, p0, p1, d1;
    int OWp, segrow;     // padded output width, 8-pixel segments per row
    int R_oh, nseg, steps, dyp;  // rows per block, segments, MFMA steps, dy LDS pitch
    int PR, PADL, PWpad; // patch rows, left pad (multiple of 8), patch pitch
    int units_per_img, total_units, nsplit, upsplit;
    int ktiles, ctiles, RS;
    int is_f16;
    int pc;              // 1: producer/consumer kernel with two block buffers
    int bufb;            // bytes per block buffer (16-byte multiple)
    int dy_lpr_log2;     // producer copy: log2 of lanes per dy row (>= segrow)
    int x_rpr;           // producer copy: x rows per round (256 / (PWpad / 8))
    int esz;             // operand bytes: 2 (BF16 / FP16) or 4 (FP32, wgrad_f32)
};

__device__ inline uint4 pack8(const uint16_t *h) {
    return make_uint4(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16),
                      h[4] | ((uint32_t)h[5] << 16), h[6] | ((uint32_t)h[7] << 16));
}

// Stages one block (dy rows and the x rows they touch) in LDS. With ALIGNED
// (W and OW multiples of 8) every 8-value chunk is one aligned 16-byte load or
// lies wholly outside its row; otherwise chunks are gathered element-wise.
// NB chunks per thread are loaded before any is stored so that their
// latencies overlap; the other workgroup on the CU computes meanwhile.
template <bool ALIGNED>
__device__ inline uint4 dy_chunk(const uint16_t *__restrict__ dy, const WGParams &p, int e, int n,
                                 int oh0, int kt, int segs, long long OHW) {
    const int m = e / segs, seg = e - (e / segs) * segs;
    const int k = kt * WG_BK + m;
    if (m >= WG_BK || seg >= p.nseg || k >= p.K) return make_uint4(0, 0, 0, 0);
    const int oh = oh0 + seg / p.segrow;
    const int ow0 = (seg % p.segrow) * 8;
    if (oh >= p.OH) return make_uint4(0, 0, 0, 0);
    const uint16_t *src = dy + ((long long)n * p.K + k) * OHW + (long long)oh * p.OW + ow0;
    if constexpr (ALIGNED) {
        return *reinterpret_cast<const uint4 *>(src);
    } else {
        uint16_t h[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) h[t] = (ow0 + t < p.OW) ? src[t] : (uint16_t)0;
        return pack8(h);
    }
}

template <bool ALIGNED>
__device__ inline uint4 x_chunk(const uint16_t *__restrict__ x, const WGParams &p, int e, int n,
                                int oh0, int ct, int qrow, int xchunks, long long HW) {
    if (e >= xchunks) return make_uint4(0, 0, 0, 0);
    const int cr = e / qrow, qq = e - (e / qrow) * qrow;
    const int c = cr / p.PR, r = cr - (cr / p.PR) * p.PR;
    const int ch = ct * WG_BC + c;
    const int ih = oh0 * p.s1 - p.p1 + r;
    const int iw0 = qq * 8 - p.PADL;
    if (ch >= p.C || ih < 0 || ih >= p.H || iw0 + 8 <= 0 || iw0 >= p.W) return make_uint4(0, 0, 0, 0);
    const uint16_t *src = x + ((long long)n * p.C + ch) * HW + (long long)ih * p.W;
    if constexpr (ALIGNED) {
        return *reinterpret_cast<const uint4 *>(src + iw0);
    } else {
        uint16_t h[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int iw = iw0 + t;
            h[t] = (iw >= 0 && iw < p.W) ? src[iw] : (uint16_t)0;
        }
        return pack8(h);
    }
}

template <bool ALIGNED, int NT = WG_THREADS, int NBA = 8>
__device__ inline void load_direct(const uint16_t *__restrict__ x, const uint16_t *__restrict__ dy,
                                   uint16_t *lds_dy, uint16_t *lds_x, const WGParams &p, int n, int oh0,
                                   int kt, int ct, int tid = (int)threadIdx.x) {
    constexpr int NB = ALIGNED ? NBA : 4;
    constexpr int WG_THREADS = NT;  // threads taking part in the copy
    const long long HW = (long long)p.H * p.W, OHW = (long long)p.OH * p.OW;
    const int segs = p.steps * 4;
    const int dchunks = WG_BK * segs;
    for (int e0 = 0; e0 < dchunks; e0 += NB * WG_THREADS) {
        uint4 v[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) v[q] = dy_chunk<ALIGNED>(dy, p, e0 + q * WG_THREADS + tid, n, oh0, kt, segs, OHW);
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int e = e0 + q * WG_THREADS + tid;
            if (e < dchunks) {
                const int m = e / segs, seg = e - (e / segs) * segs;
                *reinterpret_cast<uint4 *>(lds_dy + m * p.dyp + seg * 8) = v[q];
            }
        }
    }
    const int qrow = p.PWpad / 8;
    const int xchunks = WG_BC * p.PR * qrow;
    for (int e0 = 0; e0 < xchunks; e0 += NB * WG_THREADS) {
        uint4 v[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) v[q] = x_chunk<ALIGNED>(x, p, e0 + q * WG_THREADS + tid, n, oh0, ct, qrow, xchunks, HW);
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int e = e0 + q * WG_THREADS + tid;
            if (e < xchunks) {
                const int cr = e / qrow, qq = e - (e / qrow) * qrow;
                *reinterpret_cast<uint4 *>(lds_x + cr * p.PWpad + qq * 8) = v[q];
            }
        }
    }
}

// B fragments of the FW taps of one filter row from a window of dword reads.
// rowp: the x row in LDS; e0: element index of the first value for tap i=0
// (its parity `odd` is the same for every lane).
template <int FW, int S0>
__device__ inline void b_frags(const uint16_t *rowp, int e0, wu32x4 (&bf)[FW]) {
    constexpr int ND = S0 == 1 ? (8 + FW + 1) / 2 + 1 : (16 + FW + 1) / 2 + 1;
    const uint32_t *dw = reinterpret_cast<const uint32_t *>(rowp) + (e0 >> 1);
    uint32_t d[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) d[q] = dw[q];
    // `odd` (the parity of e0) is the same for every lane: branch on it so
    // that every register index below is a compile-time constant
    if ((e0 & 1) == 0) {
#pragma unroll
        for (int i = 0; i < FW; ++i) {
            if constexpr (S0 == 1) {
                // out dword m = elements (i + 2m, i + 2m + 1)
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    bf[i][m] = (i & 1) ? __builtin_amdgcn_alignbyte(d[(i + 1) / 2 + m], d[(i - 1) / 2 + m], 2)
                                       : d[i / 2 + m];
            } else {
                // out dword m = elements (i + 4m, i + 4m + 2): one half of
                // dwords i/2 + 2m and i/2 + 2m + 1
                const uint32_t sel = (i & 1) ? 0x07060302u : 0x05040100u;
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    bf[i][m] = __builtin_amdgcn_perm(d[i / 2 + 2 * m + 1], d[i / 2 + 2 * m], sel);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < FW; ++i) {
            // element offset o = i + 1
            if constexpr (S0 == 1) {
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    bf[i][m] = (i & 1) ? d[(i + 1) / 2 + m]
                                       : __builtin_amdgcn_alignbyte(d[i / 2 + 1 + m], d[i / 2 + m], 2);
            } else {
                const uint32_t sel = (i & 1) ? 0x05040100u : 0x07060302u;
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    bf[i][m] = __builtin_amdgcn_perm(d[(i + 1) / 2 + 2 * m + 1], d[(i + 1) / 2 + 2 * m], sel);
            }
        }
    }
}

template <int FW, int FH, int S0, int F16, bool ALIGNED>
__global__ __launch_bounds__(WG_THREADS, 2) void wgrad_rows(float *__restrict__ out,
                                                            const uint16_t *__restrict__ x,
                                                            const uint16_t *__restrict__ dy,
                                                            WGParams p) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    uint16_t *lds_dy = reinterpret_cast<uint16_t *>(smem);
    uint16_t *lds_x = lds_dy + WG_BK * p.dyp;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    long long bid = blockIdx.x;
    const int split = (int)(bid % p.nsplit);
    bid /= p.nsplit;
    const int ct = (int)(bid % p.ctiles);
    const int kt = (int)(bid / p.ctiles);

    constexpr int RS = FW * FH;
    wf32x4 acc[4][RS];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int t = 0; t < RS; ++t) acc[mi][t] = wf32x4{0.f, 0.f, 0.f, 0.f};

    const int g = lane >> 4, r16 = lane & 15;
    const int cl = wave * 16 + r16;  // lane's channel within the tile
    const int u0 = split * p.upsplit;
    const int u1 = min(u0 + p.upsplit, p.total_units);
    for (int u = u0; u < u1; ++u) {
        const int n = u / p.units_per_img;
        __syncthreads();  // the previous block is no longer read
        load_direct<ALIGNED>(x, dy, lds_dy, lds_x, p, n, (u - n * p.units_per_img) * p.R_oh, kt, ct);
        __syncthreads();
        for (int s = 0; s < p.steps; ++s) {
            int seg = s * 4 + g;
            // A fragments: dy rows mi*16 + r16, pixels seg*8 .. +7
            wu32x4 af[4];
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
                af[mi] = *reinterpret_cast<const wu32x4 *>(lds_dy + (mi * 16 + r16) * p.dyp + seg * 8);
            if (seg >= p.nseg) seg = p.nseg - 1;  // dummy segment: A is zero there
            const int ro = seg / p.segrow;
            const int ow0 = (seg - ro * p.segrow) * 8;
            const int e0 = ow0 * S0 + (p.PADL - p.p0);
#pragma unroll
            for (int j = 0; j < FH; ++j) {
                const uint16_t *rowp = lds_x + (cl * p.PR + ro * p.s1 + j * p.d1) * p.PWpad;
                wu32x4 bf[FW];
                b_frags<FW, S0>(rowp, e0, bf);
#pragma unroll
                for (int i = 0; i < FW; ++i)
#pragma unroll
                    for (int mi = 0; mi < 4; ++mi) {
                        if constexpr (F16)
                            acc[mi][j * FW + i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                                __builtin_bit_cast(wf16x8, af[mi]), __builtin_bit_cast(wf16x8, bf[i]),
                                acc[mi][j * FW + i], 0, 0, 0);
                        else
                            acc[mi][j * FW + i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                __builtin_bit_cast(wbf16x8, af[mi]), __builtin_bit_cast(wbf16x8, bf[i]),
                                acc[mi][j * FW + i], 0, 0, 0);
                    }
            }
        }
    }

    // acc[mi][t][v] = D[filter][channel] for tap t = j*FW + i:
    // filter = kt*64 + mi*16 + 4*(lane/16) + v, channel = ct*64 + wave*16 + lane%16.
    // Stored in the lane-contiguous partial layout (one 256-byte store per
    // (mi, t, v)); wgrad_sum maps it to dW.
    float *dst = out + ((long long)split * p.ktiles * p.ctiles + kt * p.ctiles + ct) * (4096LL * RS) +
                 (long long)wave * (1024 * RS) + lane;
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int t = 0; t < RS; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) dst[((mi * RS + t) * 4 + v) * 64] = acc[mi][t][v];
}

// Producer copy of one block for ALIGNED geometries (W and OW multiples of
// 8): 256 lanes laid out as rows x 16-byte chunks. A lane's row advances by a
// constant stride each round (no division), addresses are 32-bit offsets into
// one image's buffer resource, and out-of-image chunks read zeros (offset past
// the resource's end), so each chunk costs a few VALU operations. NB rounds of
// loads are in flight before their LDS stores. The dy chunks of dummy
// segments (nseg..steps*4) are never written: they are zeroed once.
template <int NB>
__device__ inline void pc_copy(const uint16_t *__restrict__ x, const uint16_t *__restrict__ dy, uint16_t *ldy,
                               uint16_t *lx, const WGParams &p, int n, int oh0, int kt, int ct, int t) {
    constexpr int kOOB = 0x7fffffc0;
    // ---- dy: WG_BK * R_oh rows (filter m, output row r) of segrow chunks
    {
        const auto rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint16_t *>(dy + (long long)n * p.K * p.OH * p.OW), (short)0, p.K * p.OH * p.OW * 2, 0x00020000);
        const int rpr = WG_THREADS >> p.dy_lpr_log2;  // rows per round
        const int j = t & ((1 << p.dy_lpr_log2) - 1);
        const int nrow = WG_BK * p.R_oh;
        const int row0 = t >> p.dy_lpr_log2;
        int m = row0 / p.R_oh, r = row0 - (row0 / p.R_oh) * p.R_oh;
        const int dm = rpr / p.R_oh, dr = rpr - (rpr / p.R_oh) * p.R_oh;
        const bool jok = j < p.segrow;
        for (int base = 0; base < nrow; base += NB * rpr) {
            uint4 v[NB];
            int dst[NB];
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const int k = kt * WG_BK + m, oh = oh0 + r;
                const bool in = base + q * rpr + row0 < nrow && jok;
                const bool ok = in && k < p.K && oh < p.OH;
                v[q] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(
                                                     rs, ok ? ((k * p.OH + oh) * p.OW + j * 8) * 2 : kOOB, 0, 0));
                dst[q] = in ? m * p.dyp + (r * p.segrow + j) * 8 : -1;
                m += dm;
                r += dr;
                if (r >= p.R_oh) {
                    r -= p.R_oh;
                    ++m;
                }
            }
#pragma unroll
            for (int q = 0; q < NB; ++q)
                if (dst[q] >= 0) *reinterpret_cast<uint4 *>(ldy + dst[q]) = v[q];
        }
    }
    // ---- x: WG_BC * PR rows (channel c, patch row r) of qrow chunks, qrow
    // lanes per row, x_rpr rows per round (lanes past x_rpr * qrow idle)
    {
        const auto rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint16_t *>(x + (long long)n * p.C * p.H * p.W), (short)0, p.C * p.H * p.W * 2, 0x00020000);
        const int qrow = p.PWpad / 8, rpr = p.x_rpr;
        const int qq = t % qrow;
        const int nrow = WG_BC * p.PR;
        const int row0 = t / qrow;
        const bool lane_ok = row0 < rpr;
        int c = row0 / p.PR, r = row0 - (row0 / p.PR) * p.PR;
        const int dc = rpr / p.PR, dr = rpr - (rpr / p.PR) * p.PR;
        const int iw0 = qq * 8 - p.PADL;
        const bool wok = iw0 >= 0 && iw0 < p.W;  // a chunk never straddles a row end (W % 8 == 0)
        const int ih00 = oh0 * p.s1 - p.p1;
        for (int base = 0; base < nrow; base += NB * rpr) {
            uint4 v[NB];
            int dst[NB];
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const int row = base + q * rpr + row0;
                const int ch = ct * WG_BC + c, ih = ih00 + r;
                const bool in = lane_ok && row < nrow;
                const bool ok = in && wok && ch < p.C && ih >= 0 && ih < p.H;
                v[q] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(
                                                     rs, ok ? ((ch * p.H + ih) * p.W + iw0) * 2 : kOOB, 0, 0));
                dst[q] = in ? row * p.PWpad + qq * 8 : -1;
                c += dc;
                r += dr;
                if (r >= p.PR) {
                    r -= p.PR;
                    ++c;
                }
            }
#pragma unroll
            for (int q = 0; q < NB; ++q)
                if (dst[q] >= 0) *reinterpret_cast<uint4 *>(lx + dst[q]) = v[q];
        }
    }
}

// Producer/consumer form of wgrad_rows: 8 waves, the first 4 compute exactly
// as in wgrad_rows, the other 4 copy the next block (dy rows and the x rows
// they touch) into the second of two LDS block buffers meanwhile, with 16
// chunks per lane in flight. One barrier per block hands the buffers over,
// so HBM reads overlap the MFMAs instead of alternating with them.
template <int FW, int FH, int S0, int F16, bool ALIGNED>
__global__ __launch_bounds__(2 * WG_THREADS, 2) void wgrad_pc(float *__restrict__ out,
                                                              const uint16_t *__restrict__ x,
                                                              const uint16_t *__restrict__ dy,
                                                              WGParams p) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool producer = wave >= 4;
    long long bid = blockIdx.x;
    const int split = (int)(bid % p.nsplit);
    bid /= p.nsplit;
    const int ct = (int)(bid % p.ctiles);
    const int kt = (int)(bid / p.ctiles);
    const int u0 = split * p.upsplit;
    const int u1 = min(u0 + p.upsplit, p.total_units);
    uint16_t *const lds_dy0 = reinterpret_cast<uint16_t *>(smem);
    uint16_t *const lds_dy1 = reinterpret_cast<uint16_t *>(smem + p.bufb);
    const int xoff = WG_BK * p.dyp;  // x block after the dy block, in elements
    // Producers and consumers run separate loops with the same barriers (one
    // before the first block, one after each block), so the accumulators are
    // live only on the consumer side.
    if (producer) {
        const int t = tid - WG_THREADS;
        if constexpr (ALIGNED) {
            // dummy dy segments (nseg .. steps*4) of both buffers: zero once
            const int dseg = p.steps * 4 - p.nseg;
            for (int e = t; e < 2 * WG_BK * dseg; e += WG_THREADS) {
                const int bsel = e / (WG_BK * dseg), rem = e - bsel * (WG_BK * dseg);
                const int m = rem / dseg, sg = p.nseg + (rem - (rem / dseg) * dseg);
                *reinterpret_cast<uint4 *>((bsel ? lds_dy1 : lds_dy0) + m * p.dyp + sg * 8) = make_uint4(0, 0, 0, 0);
            }
        }
        for (int i = -1; u0 + i < u1; ++i) {
            const int u = u0 + i + 1;
            if (u < u1) {
                uint16_t *ldy = (i & 1) ? lds_dy0 : lds_dy1;  // block i + 1 goes to buffer (i + 1) & 1
                const int n = u / p.units_per_img;
                if constexpr (ALIGNED)
                    pc_copy<8>(x, dy, ldy, ldy + xoff, p, n, (u - n * p.units_per_img) * p.R_oh, kt, ct, t);
                else
                    load_direct<ALIGNED, WG_THREADS, 16>(x, dy, ldy, ldy + xoff, p, n,
                                                         (u - n * p.units_per_img) * p.R_oh, kt, ct, t);
            }
            __syncthreads();  // block i consumed (i >= 0), block i + 1 published
        }
        return;
    }

    constexpr int RS = FW * FH;
    wf32x4 acc[4][RS];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int t = 0; t < RS; ++t) acc[mi][t] = wf32x4{0.f, 0.f, 0.f, 0.f};

    const int g = lane >> 4, r16 = lane & 15;
    const int cl = wave * 16 + r16;  // lane's channel within the tile
    __syncthreads();
    for (int i = 0; u0 + i < u1; ++i) {
        const uint16_t *ldy = (i & 1) ? lds_dy1 : lds_dy0;
        const uint16_t *lx = ldy + xoff;
        for (int s = 0; s < p.steps; ++s) {
            int seg = s * 4 + g;
            wu32x4 af[4];
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
                af[mi] = *reinterpret_cast<const wu32x4 *>(ldy + (mi * 16 + r16) * p.dyp + seg * 8);
            if (seg >= p.nseg) seg = p.nseg - 1;  // dummy segment: A is zero there
            const int ro = seg / p.segrow;
            const int ow0 = (seg - ro * p.segrow) * 8;
            const int e0 = ow0 * S0 + (p.PADL - p.p0);
#pragma unroll
            for (int j = 0; j < FH; ++j) {
                const uint16_t *rowp = lx + (cl * p.PR + ro * p.s1 + j * p.d1) * p.PWpad;
                wu32x4 bf[FW];
                b_frags<FW, S0>(rowp, e0, bf);
#pragma unroll
                for (int ii = 0; ii < FW; ++ii)
#pragma unroll
                    for (int mi = 0; mi < 4; ++mi) {
                        if constexpr (F16)
                            acc[mi][j * FW + ii] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                                __builtin_bit_cast(wf16x8, af[mi]), __builtin_bit_cast(wf16x8, bf[ii]),
                                acc[mi][j * FW + ii], 0, 0, 0);
                        else
                            acc[mi][j * FW + ii] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                __builtin_bit_cast(wbf16x8, af[mi]), __builtin_bit_cast(wbf16x8, bf[ii]),
                                acc[mi][j * FW + ii], 0, 0, 0);
                    }
            }
        }
        __syncthreads();  // block i consumed, block i + 1 published
    }

    // acc[mi][t][v] = D[filter][channel] for tap t = j*FW + i:
    // filter = kt*64 + mi*16 + 4*(lane/16) + v, channel = ct*64 + wave*16 + lane%16.
    // Stored in the lane-contiguous partial layout (one 256-byte store per
    // (mi, t, v)); wgrad_sum maps it to dW.
    float *dst = out + ((long long)split * p.ktiles * p.ctiles + kt * p.ctiles + ct) * (4096LL * RS) +
                 (long long)wave * (1024 * RS) + lane;
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int t = 0; t < RS; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) dst[((mi * RS + t) * 4 + v) * 64] = acc[mi][t][v];
}

// ---------------------------------------------------------------- FP32
// The FP32 filter gradient on v_mfma_f32_16x16x4f32 with the same blocking
// (64 filters x 64 channels x all taps per workgroup, row blocks of output
// pixels staged in LDS, FP32 split partials in the lane-contiguous layout
// that wgrad_sum adds). One reduction value per lane per MFMA: in step s,
// lane group g = lane / 16 owns the 8-pixel segment 4s + g and MFMA q of a
// tap takes its pixel q. A lane reads its 8 dy values per filter row tile
// as two 16-byte LDS reads and, per filter row, one window of 7*S0 + FW x
// values that serves all FW taps. The FP32 MFMA (32 cycles per SIMD) bounds
// the kernel; one workgroup per CU stages a block while the next waits.
template <bool ALIGNED>
__device__ inline void load_f32(const float *__restrict__ x, const float *__restrict__ dy, float *lds_dy,
                                float *lds_x, const WGParams &p, int n, int oh0, int kt, int ct) {
    constexpr int NB = 8;
    const int tid = threadIdx.x;
    const long long HW = (long long)p.H * p.W, OHW = (long long)p.OH * p.OW;
    // dy: WG_BK rows of steps*4 segments of 8 pixels, as 4-value chunks
    const int cpr = p.steps * 8;
    const int dchunks = WG_BK * cpr;
    for (int e0 = 0; e0 < dchunks; e0 += NB * WG_THREADS) {
        float4 v[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int e = e0 + q * WG_THREADS + tid;
            const int m = e / cpr, cj = e - (e / cpr) * cpr;
            const int seg = cj >> 1;
            const int k = kt * WG_BK + m;
            const int oh = oh0 + seg / p.segrow;
            const int ow0 = (seg % p.segrow) * 8 + (cj & 1) * 4;
            const bool ok = e < dchunks && seg < p.nseg && k < p.K && oh < p.OH;
            const float *src = dy + ((long long)n * p.K + k) * OHW + (long long)oh * p.OW + ow0;
            if constexpr (ALIGNED) {
                v[q] = (ok && ow0 < p.OW) ? *reinterpret_cast<const float4 *>(src) : make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                float h[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) h[t] = (ok && ow0 + t < p.OW) ? src[t] : 0.f;
                v[q] = make_float4(h[0], h[1], h[2], h[3]);
            }
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int e = e0 + q * WG_THREADS + tid;
            if (e < dchunks) {
                const int m = e / cpr, cj = e - (e / cpr) * cpr;
                *reinterpret_cast<float4 *>(lds_dy + m * p.dyp + cj * 4) = v[q];
            }
        }
    }
    // x: WG_BC channels x PR rows x PWpad/4 chunks; LDS column q holds image
    // column q - PADL
    const int qrow = p.PWpad / 4;
    const int xchunks = WG_BC * p.PR * qrow;
    for (int e0 = 0; e0 < xchunks; e0 += NB * WG_THREADS) {
        float4 v[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int e = e0 + q * WG_THREADS + tid;
            const int cr = e / qrow, qq = e - (e / qrow) * qrow;
            const int c = cr / p.PR, r = cr - (cr / p.PR) * p.PR;
            const int ch = ct * WG_BC + c;
            const int ih = oh0 * p.s1 - p.p1 + r;
            const int iw0 = qq * 4 - p.PADL;
            const bool ok = e < xchunks && ch < p.C && ih >= 0 && ih < p.H;
            const float *src = x + ((long long)n * p.C + ch) * HW + (long long)ih * p.W;
            if constexpr (ALIGNED) {
                v[q] = (ok && iw0 >= 0 && iw0 < p.W) ? *reinterpret_cast<const float4 *>(src + iw0)
                                                     : make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                float h[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int iw = iw0 + t;
                    h[t] = (ok && iw >= 0 && iw < p.W) ? src[iw] : 0.f;
                }
                v[q] = make_float4(h[0], h[1], h[2], h[3]);
            }
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int e = e0 + q * WG_THREADS + tid;
            if (e < xchunks) *reinterpret_cast<float4 *>(lds_x + e * 4) = v[q];
        }
    }
}

template <int FW, int FH, int S0, bool ALIGNED>
__global__ __launch_bounds__(WG_THREADS, 2) void wgrad_f32(float *__restrict__ out, const float *__restrict__ x,
                                                           const float *__restrict__ dy, WGParams p) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    float *lds_dy = reinterpret_cast<float *>(smem);
    float *lds_x = lds_dy + WG_BK * p.dyp;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    long long bid = blockIdx.x;
    const int split = (int)(bid % p.nsplit);
    bid /= p.nsplit;
    const int ct = (int)(bid % p.ctiles);
    const int kt = (int)(bid / p.ctiles);

    constexpr int RS = FW * FH;
    constexpr int NWIN = 7 * S0 + FW;
    wf32x4 acc[4][RS];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int t = 0; t < RS; ++t) acc[mi][t] = wf32x4{0.f, 0.f, 0.f, 0.f};

    const int g = lane >> 4, r16 = lane & 15;
    const int cl = wave * 16 + r16;  // lane's channel within the tile
    const int u0 = split * p.upsplit;
    const int u1 = min(u0 + p.upsplit, p.total_units);
    for (int u = u0; u < u1; ++u) {
        const int n = u / p.units_per_img;
        __syncthreads();  // the previous block is no longer read
        load_f32<ALIGNED>(x, dy, lds_dy, lds_x, p, n, (u - n * p.units_per_img) * p.R_oh, kt, ct);
        __syncthreads();
        for (int s = 0; s < p.steps; ++s) {
            int seg = s * 4 + g;
            float a[4][8];
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) {
                const float4 *ap = reinterpret_cast<const float4 *>(lds_dy + (mi * 16 + r16) * p.dyp + seg * 8);
                const float4 lo = ap[0], hi = ap[1];
                a[mi][0] = lo.x; a[mi][1] = lo.y; a[mi][2] = lo.z; a[mi][3] = lo.w;
                a[mi][4] = hi.x; a[mi][5] = hi.y; a[mi][6] = hi.z; a[mi][7] = hi.w;
            }
            if (seg >= p.nseg) seg = p.nseg - 1;  // dummy segment: A is zero there
            const int ro = seg / p.segrow;
            const int ow0 = (seg - ro * p.segrow) * 8;
            const int e0 = ow0 * S0 + (p.PADL - p.p0);
#pragma unroll
            for (int j = 0; j < FH; ++j) {
                const float *rowp = lds_x + (cl * p.PR + ro * p.s1 + j * p.d1) * p.PWpad + e0;
                float b[NWIN];
#pragma unroll
                for (int q = 0; q < NWIN; ++q) b[q] = rowp[q];
#pragma unroll
                for (int q = 0; q < 8; ++q)
#pragma unroll
                    for (int i = 0; i < FW; ++i)
#pragma unroll
                        for (int mi = 0; mi < 4; ++mi)
                            acc[mi][j * FW + i] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                a[mi][q], b[q * S0 + i], acc[mi][j * FW + i], 0, 0, 0);
            }
        }
    }

    // same partial layout as wgrad_rows
    float *dst = out + ((long long)split * p.ktiles * p.ctiles + kt * p.ctiles + ct) * (4096LL * RS) +
                 (long long)wave * (1024 * RS) + lane;
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int t = 0; t < RS; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) dst[((mi * RS + t) * 4 + v) * 64] = acc[mi][t][v];
}

// Sums nsplit partial dW buffers and converts to the output type. A split's
// partials are tiles (kt, ct) of [wave][mi][tap][v][lane] floats (4096 * RS
// per tile); every partial element is read coalesced and written to the dW
// element it maps to.
// dW index of partial element e, or -1 for the padding of a partial tile
__device__ inline long long partial_to_dw(long long e, const WGParams &p) {
    const int lane = (int)(e & 63), v = (int)((e >> 6) & 3);
    long long rest = e >> 8;
    const int t = (int)(rest % p.RS);
    rest /= p.RS;
    const int mi = (int)(rest & 3), wave = (int)((rest >> 2) & 3);
    const long long tile = rest >> 4;
    const int kt = (int)(tile / p.ctiles), ct = (int)(tile % p.ctiles);
    const int k = kt * WG_BK + mi * 16 + (lane >> 4) * 4 + v;
    const int c = ct * WG_BC + wave * 16 + (lane & 15);
    if (k >= p.K || c >= p.C) return -1;
    return ((long long)k * p.C + c) * p.RS + (p.RS - 1 - t);
}

// 1024 threads = 64 columns of 4 consecutive partial elements (16-byte loads)
// x 16 split groups; group q sums splits [q*nsplit/16, (q+1)*nsplit/16) in
// order with all of its loads in flight (up to 8 per batch), then the 16
// group sums are added in a fixed pairwise tree: the order of the additions
// never depends on scheduling.
constexpr int WS_GROUPS = 16;
template <typename TOut>
__global__ __launch_bounds__(64 * WS_GROUPS) void wgrad_sum(TOut *__restrict__ out, const float *__restrict__ part,
                                                            WGParams p) {
    __shared__ float4 red[WS_GROUPS][64];
    const long long total = (long long)p.ktiles * p.ctiles * 4096 * p.RS;  // a multiple of 256
    const int col = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int q0 = (int)((long long)grp * p.nsplit / WS_GROUPS);
    const int q1 = (int)((long long)(grp + 1) * p.nsplit / WS_GROUPS);
    for (long long e0 = blockIdx.x * 256LL; e0 < total; e0 += (long long)gridDim.x * 256) {
        const long long e = e0 + col * 4;
        float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int q = q0; q < q1; q += 8) {
            float4 v[8];
#pragma unroll
            for (int t = 0; t < 8; ++t)
                v[t] = q + t < q1 ? *reinterpret_cast<const float4 *>(part + (long long)(q + t) * total + e)
                                  : make_float4(0.f, 0.f, 0.f, 0.f);
            s.x += ((v[0].x + v[1].x) + (v[2].x + v[3].x)) + ((v[4].x + v[5].x) + (v[6].x + v[7].x));
            s.y += ((v[0].y + v[1].y) + (v[2].y + v[3].y)) + ((v[4].y + v[5].y) + (v[6].y + v[7].y));
            s.z += ((v[0].z + v[1].z) + (v[2].z + v[3].z)) + ((v[4].z + v[5].z) + (v[6].z + v[7].z));
            s.w += ((v[0].w + v[1].w) + (v[2].w + v[3].w)) + ((v[4].w + v[5].w) + (v[6].w + v[7].w));
        }
        red[grp][col] = s;
        __syncthreads();
#pragma unroll
        for (int h = WS_GROUPS / 2; h >= 1; h >>= 1) {
            if (grp < h) {
                const float4 a = red[grp][col], b = red[grp + h][col];
                red[grp][col] = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
            }
            __syncthreads();
        }
        if (grp == 0) {
            const float4 t4 = red[0][col];
            const float r[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long long o = partial_to_dw(e + i, p);
                if (o < 0) continue;
                if constexpr (sizeof(TOut) == 4)
                    out[o] = r[i];
                else if constexpr (std::is_same<TOut, _Float16>::value)
                    out[o] = (_Float16)r[i];
                else
                    out[o] = f32_to_bf16_bits(r[i]);
            }
        }
        __syncthreads();
    }
}

static size_t wg_lds(const WGParams &p) {
    return ((size_t)WG_BK * p.dyp + (size_t)WG_BC * p.PR * p.PWpad) * (size_t)p.esz;
}

// Plans the blocking; returns false when the geometry is outside this kernel.
static bool wg_plan(WGParams &p, int fw, int s0, int d0) {
    if (d0 != 1 || (s0 != 1 && s0 != 2) || (fw != 1 && fw != 3) || (p.fh != 1 && p.fh != 3))
        return false;
    p.RS = fw * p.fh;
    if (p.OW <= 0 || p.OH <= 0 || p.N <= 0 || p.C <= 0 || p.K <= 0) return false;
    p.OWp = (p.OW + 7) & ~7;
    p.segrow = p.OWp / 8;
    p.PADL = (p.p0 + 7) & ~7;
    // widest column read: last segment start + 8 values at stride s0 + taps,
    // plus the dword window slack
    const int maxe = (p.OWp - 8) * s0 + (p.PADL - p.p0) + 8 * s0 + fw + 4;
    p.PWpad = ((maxe + 7) & ~7) + 8;
    if (p.PADL - p.p0 < 0) return false;
    // rows per block: up to about 512 padded pixels, within the LDS budget
    // (one workgroup per CU: taller blocks re-read fewer halo rows of x).
    // First choice: the producer/consumer kernel, two block buffers in 160 KB;
    // otherwise one buffer in WG_LDS_LIMIT.
    int r0 = 512 / p.OWp;
    if (r0 < 1) r0 = 1;
    if (r0 > p.OH) r0 = p.OH;
    // producer copy layout (pc_copy): lanes per dy row a power of two >= segrow,
    // x rows per round; one image of x and of dy within 32-bit byte offsets
    p.dy_lpr_log2 = 0;
    while ((1 << p.dy_lpr_log2) < p.segrow) ++p.dy_lpr_log2;
    p.x_rpr = WG_THREADS / (p.PWpad / 8);
    // AFHIP_NN_WGRAD_PC=0 in the environment selects the single-buffer kernel
    const char *pc_env = std::getenv("AFHIP_NN_WGRAD_PC");
    const bool pc_ok = !(pc_env && pc_env[0] == '0') && p.esz == 2 && p.dy_lpr_log2 <= 8 && p.x_rpr >= 1 &&
                       (long long)p.C * p.H * p.W * 2 < (1LL << 31) - 64 &&
                       (long long)p.K * p.OH * p.OW * 2 < (1LL << 31) - 64;
    // AFHIP_NN_WGRAD_F32_LDS=<bytes> in the environment sets another LDS
    // budget for the FP32 kernel (up to WG_LDS_LIMIT; 81920 or less lets two
    // workgroups share a CU)
    size_t lds_limit = WG_LDS_LIMIT;
    if (p.esz == 4) {
        const char *le = std::getenv("AFHIP_NN_WGRAD_F32_LDS");
        if (le && le[0] >= '1' && le[0] <= '9') lds_limit = std::min<size_t>(WG_LDS_LIMIT, std::atol(le));
    }
    int r = 0;
    for (int pc = pc_ok ? 1 : 0; pc >= 0 && r < 1; --pc) {
        for (r = r0; r >= 1; --r) {
            p.R_oh = r;
            p.nseg = r * p.segrow;
            p.steps = (p.nseg + 3) / 4;
            p.dyp = p.steps * 32 + 8;
            p.PR = (r - 1) * p.s1 + (p.fh - 1) * p.d1 + 1;
            p.pc = pc;
            p.bufb = (int)((wg_lds(p) + 15) & ~(size_t)15);
            if (pc ? 2 * (size_t)p.bufb <= WG_PC_LDS : wg_lds(p) <= lds_limit) break;
        }
    }
    if (r < 1) return false;
    p.units_per_img = (p.OH + p.R_oh - 1) / p.R_oh;
    p.total_units = p.units_per_img * p.N;
    p.ktiles = (p.K + WG_BK - 1) / WG_BK;
    p.ctiles = (p.C + WG_BC - 1) / WG_BC;
    const int tiles = p.ktiles * p.ctiles;
    // Split count: one workgroup per CU (measured fastest on the benchmark
    // layers, whose units cost far more than their partial tiles).
    // AFHIP_NN_WGRAD_SPLIT=<n> in the environment forces n splits.
    const int cus = cu_count_current();
    int ns = (cus + tiles - 1) / tiles;
    const char *env = std::getenv("AFHIP_NN_WGRAD_SPLIT");
    if (env && env[0] >= '1' && env[0] <= '9') ns = std::atoi(env);
    if (ns > p.total_units) ns = p.total_units;
    if (ns < 1) ns = 1;
    p.upsplit = (p.total_units + ns - 1) / ns;
    p.nsplit = (p.total_units + p.upsplit - 1) / p.upsplit;
    return true;
}

static void wg_fill(WGParams &p, const long long *xd, const long long *wd, const long long *yd,
                    const long long *st, const long long *pd, const long long *dl) {
    p.W = (int)xd[0];
    p.H = (int)xd[1];
    p.C = (int)xd[2];
    p.N = (int)xd[3];
    p.OW = (int)yd[0];
    p.OH = (int)yd[1];
    p.K = (int)wd[3];
    p.fh = (int)wd[1];
    p.s1 = (int)st[1];
    p.p0 = (int)pd[0];
    p.p1 = (int)pd[1];
    p.d1 = (int)dl[1];
    p.esz = 2;
}

template <int FW, int FH, int S0, int F16>
static int wg_launch(float *part, const void *x, const void *dy, const WGParams &p, hipStream_t s) {
    const bool aligned = (p.W & 7) == 0 && (p.OW & 7) == 0;
    if (p.pc) {
        auto fp = aligned ? wgrad_pc<FW, FH, S0, F16, true> : wgrad_pc<FW, FH, S0, F16, false>;
        static bool attr_pc[2] = {false, false};
        if (!attr_pc[aligned]) {
            if (hipFuncSetAttribute((const void *)fp, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)WG_PC_LDS) != hipSuccess)
                return K_ERR_RUNTIME;
            attr_pc[aligned] = true;
        }
        const long long grid = (long long)p.ktiles * p.ctiles * p.nsplit;
        fp<<<dim3((unsigned)grid), 2 * WG_THREADS, 2 * (size_t)p.bufb, s>>>(part, (const uint16_t *)x,
                                                                          (const uint16_t *)dy, p);
        return launch_status();
    }
    auto fn = aligned ? wgrad_rows<FW, FH, S0, F16, true> : wgrad_rows<FW, FH, S0, F16, false>;
    static bool attr[2] = {false, false};
    if (!attr[aligned]) {
        if (hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)WG_LDS_LIMIT) != hipSuccess)
            return K_ERR_RUNTIME;
        attr[aligned] = true;
    }
    const long long grid = (long long)p.ktiles * p.ctiles * p.nsplit;
    fn<<<dim3((unsigned)grid), WG_THREADS, wg_lds(p), s>>>(part, (const uint16_t *)x,
                                                          (const uint16_t *)dy, p);
    return launch_status();
}

template <int FW, int FH, int S0>
static int wf_launch(float *part, const void *x, const void *dy, const WGParams &p, hipStream_t s) {
    const bool aligned = (p.W & 3) == 0 && (p.OW & 3) == 0;
    auto fn = aligned ? wgrad_f32<FW, FH, S0, true> : wgrad_f32<FW, FH, S0, false>;
    static bool attr[2] = {false, false};
    if (!attr[aligned]) {
        if (hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)WG_LDS_LIMIT) != hipSuccess)
            return K_ERR_RUNTIME;
        attr[aligned] = true;
    }
    const long long grid = (long long)p.ktiles * p.ctiles * p.nsplit;
    fn<<<dim3((unsigned)grid), WG_THREADS, wg_lds(p), s>>>(part, (const float *)x, (const float *)dy, p);
    return launch_status();
}

}  // namespace afhip

using namespace afhip;

extern "C" long long afhip_k_conv2_nn_wgrad_rows_workspace(int in_type, int precision,
                                                          const long long *xd, const long long *wd,
                                                          const long long *yd, const long long *st,
                                                          const long long *pd, const long long *dl) {
    const bool bf = (in_type == DT_BF16 && (precision == 0 || precision == 1));
    const bool hf = (in_type == DT_F16 && (precision == 0 || precision == 2));
    const bool ff = (in_type == DT_F32 && (precision == 0 || precision == 4));
    if (!bf && !hf && !ff) return 0;
    // AFHIP_NN_WGRAD_F32=0 in the environment keeps FP32 on the gather kernel
    const char *env = std::getenv("AFHIP_NN_WGRAD_F32");
    if (ff && env && env[0] == '0') return 0;
    WGParams p;
    wg_fill(p, xd, wd, yd, st, pd, dl);
    p.esz = ff ? 4 : 2;
    if (!wg_plan(p, (int)wd[0], (int)st[0], (int)dl[0])) return 0;
    return (long long)p.nsplit * p.ktiles * p.ctiles * 4096 * p.RS * 4;
}

// what: 1 = partial tiles into the workspace, 2 = sum them into dw, 3 = both
static int rows_call(int what, int in_type, int out_type, int precision, void *dw, const void *x,
                     const void *dy, const long long *xd, const long long *wd, const long long *yd,
                     const long long *st, const long long *pd, const long long *dl, void *workspace,
                     long long workspace_bytes, void *stream) {
    const long long need = afhip_k_conv2_nn_wgrad_rows_workspace(in_type, precision, xd, wd, yd, st, pd, dl);
    if (need == 0) return K_ERR_NOT_SUPPORTED;
    if (workspace == nullptr || workspace_bytes < need) return K_ERR_ARG;
    WGParams p;
    wg_fill(p, xd, wd, yd, st, pd, dl);
    const int fw = (int)wd[0], s0 = (int)st[0];
    p.esz = in_type == DT_F32 ? 4 : 2;
    wg_plan(p, fw, s0, (int)dl[0]);
    p.is_f16 = in_type == DT_F16;
    hipStream_t s = (hipStream_t)stream;
    float *part = (float *)workspace;
    if ((what & 1) && p.esz == 4) {
        int rc;
#define WFS(FWV, FHV) rc = s0 == 1 ? wf_launch<FWV, FHV, 1>(part, x, dy, p, s) : wf_launch<FWV, FHV, 2>(part, x, dy, p, s)
        if (fw == 3 && p.fh == 3) { WFS(3, 3); }
        else if (fw == 3) { WFS(3, 1); }
        else if (p.fh == 3) { WFS(1, 3); }
        else { WFS(1, 1); }
#undef WFS
        if (rc) return rc;
    } else if (what & 1) {
        int rc;
#define WGO(FWV, FHV, S0V)                                                                    \
    rc = p.is_f16 ? wg_launch<FWV, FHV, S0V, 1>(part, x, dy, p, s)                           \
                  : wg_launch<FWV, FHV, S0V, 0>(part, x, dy, p, s)
#define WGS(FWV, FHV)            \
    if (s0 == 1) WGO(FWV, FHV, 1); \
    else WGO(FWV, FHV, 2);
        if (fw == 3 && p.fh == 3) { WGS(3, 3) }
        else if (fw == 3) { WGS(3, 1) }
        else if (p.fh == 3) { WGS(1, 3) }
        else { WGS(1, 1) }
#undef WGS
#undef WGO
        if (rc) return rc;
    }
    if (what & 2) {
        const long long total = (long long)p.ktiles * p.ctiles * 4096 * p.RS;
        long long blocks = total / 256;
        if (blocks > kMaxGridBlocks) blocks = kMaxGridBlocks;
        if (in_type == DT_F32 && out_type != DT_F32)
            return K_ERR_TYPE;
        if (out_type == DT_F32)
            wgrad_sum<float><<<(unsigned)blocks, 64 * WS_GROUPS, 0, s>>>((float *)dw, part, p);
        else if (out_type == DT_BF16 && in_type == DT_BF16)
            wgrad_sum<uint16_t><<<(unsigned)blocks, 64 * WS_GROUPS, 0, s>>>((uint16_t *)dw, part, p);
        else if (out_type == DT_F16 && in_type == DT_F16)
            wgrad_sum<_Float16><<<(unsigned)blocks, 64 * WS_GROUPS, 0, s>>>((_Float16 *)dw, part, p);
        else
            return K_ERR_TYPE;
    }
    return launch_status();
}

extern "C" int afhip_k_conv2_nn_wgrad_rows(int in_type, int out_type, int precision, void *dw,
                                           const void *x, const void *dy, const long long *xd,
                                           const long long *wd, const long long *yd,
                                           const long long *st, const long long *pd,
                                           const long long *dl, void *workspace,
                                           long long workspace_bytes, void *stream) {
    return rows_call(3, in_type, out_type, precision, dw, x, dy, xd, wd, yd, st, pd, dl, workspace,
                     workspace_bytes, stream);
}

extern "C" int afhip_k_conv2_nn_wgrad_rows_partial(int in_type, int precision, const void *x,
                                                   const void *dy, const long long *xd,
                                                   const long long *wd, const long long *yd,
                                                   const long long *st, const long long *pd,
                                                   const long long *dl, void *workspace,
                                                   long long workspace_bytes, void *stream) {
    return rows_call(1, in_type, DT_F32, precision, nullptr, x, dy, xd, wd, yd, st, pd, dl, workspace,
                     workspace_bytes, stream);
}

extern "C" int afhip_k_conv2_nn_wgrad_rows_reduce(int in_type, int out_type, int precision, void *dw,
                                                  const long long *xd, const long long *wd,
                                                  const long long *yd, const long long *st,
                                                  const long long *pd, const long long *dl,
                                                  const void *workspace, long long workspace_bytes,
                                                  void *stream) {
    return rows_call(2, in_type, out_type, precision, dw, nullptr, nullptr, xd, wd, yd, st, pd, dl,
                     const_cast<void *>(workspace), workspace_bytes, stream);
}
