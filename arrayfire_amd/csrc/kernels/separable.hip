// Separable 2-D convolution (af_convolve2_sep).
//
// Reference semantics (src/backend/cpu/kernel/convolve.hpp:213-289): a column
// pass along dim0 into a T-typed temp, then a row pass along dim1; the tap is
// converted to T before the product; batch over dims 2 and 3.
//
//  * sep_pass<InT,AccT,DIM>: one pass, one lane per output (the two-launch
//    form, used for non-f32 types and when the caller asks for the temp).
//  * sep_fused_f32<F>: both passes in one kernel, for filters of at most F
//    taps (F = 3, 5, ..., 17, 33). A workgroup owns a 64 (dim0) x 64 (dim1)
//    output tile. It stages the
//    (64+cf-1) x (64+rf-1) input halo in LDS, runs the column pass into a
//    second LDS tile of 64 x (64+F-1) (rounded to float, the reference's temp
//    precision), then the row pass to HBM: one read and one write of every
//    element.
//  * sep_strip_f32<F>: the same two passes walked down 64-column strips with
//    a temp-row ring in LDS, for rows of whole 16-byte vectors (see the
//    kernel).
#include <stdlib.h>

#include <type_traits>

#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

struct SepParams {
    dim_t sd[4], td[4], od[4];
    dim_t flen;
    int expand;
};

template <typename InT, typename AccT, int DIM>
__global__ __launch_bounds__(256) void sep_pass(InT *__restrict__ out,
                                                const InT *__restrict__ in,
                                                const AccT *__restrict__ filt,
                                                SepParams p) {
    // Products are rounded before they are added, as on the reference CPU.
#pragma clang fp contract(off)
    // Pass 0 reads the signal (dims sd) and writes the temp (dims td);
    // pass 1 reads the temp and writes the output (dims od).
    const dim_t *id = DIM == 0 ? p.sd : p.td;
    const dim_t *od = DIM == 0 ? p.td : p.od;
    const dim_t n = od[0] * od[1] * od[2] * od[3];
    const dim_t half = p.expand ? 0 : (p.flen >> 1);
    for (dim_t lin = blockIdx.x * (dim_t)blockDim.x + threadIdx.x; lin < n;
         lin += (dim_t)gridDim.x * blockDim.x) {
        const dim_t i = lin % od[0];
        dim_t r = lin / od[0];
        const dim_t j = r % od[1];
        const dim_t b = r / od[1];  // b2 + b3 * d2, same for all three arrays
        const InT *ip = in + b * id[0] * id[1];
        const dim_t ci = i + (DIM == 0 ? half : 0);
        const dim_t cj = j + (DIM == 1 ? half : 0);
        AccT acc = AccT(0);
        for (dim_t f = 0; f < p.flen; ++f) {
            const InT fv = (InT)filt[f];
            InT s = InT(0);
            if (DIM == 0) {
                const dim_t x = ci - f;
                if (x >= 0 && x < id[0] && cj < id[1]) s = ip[cj * id[0] + x];
            } else {
                const dim_t y = cj - f;
                if (ci < id[0] && y >= 0 && y < id[1]) s = ip[y * id[0] + ci];
            }
            acc += AccT(s * fv);
        }
        out[lin] = (InT)acc;
    }
}

constexpr int SF_B = 64;        // output tile edge
constexpr int SF_MAXF = AFHIP_SEP_FUSED_MAX_TAPS;  // longest fused filter

struct SepFusedParams {
    dim_t S0, S1, O0, O1;
    dim_t nbx, nby, nblk;
    int cf, rf, sx, sy;
};

// Both passes with the tap count bucketed to F >= max(cf, rf): the loops are
// unrolled to F and taps k >= cf (k >= rf) are skipped by uniform branches
// (never multiplied by 0: the image may hold Inf or NaN); EXACT (cf == rf ==
// F) has no branches. Taps are uniform: they live in scalar registers.
//  * column pass: a lane owns 8 consecutive dim0 outputs of one halo row and
//    slides over a register window of 8 + F - 1 inputs read as float4s;
//  * row pass: a lane owns one dim0 column and 16 consecutive dim1 outputs;
//    the 16 + F - 1 temp values it reads are loaded once into registers.
template <int F, bool EXACT>
__global__ __launch_bounds__(256) void sep_fused_f32(float *__restrict__ out,
                                                     const float *__restrict__ in,
                                                     const float *__restrict__ cfilt,
                                                     const float *__restrict__ rfilt,
                                                     SepFusedParams p) {
    constexpr int B = SF_B;
    constexpr int L = B + F - 1;                     // halo rows / columns
    constexpr int TX = 8, TY = 16;
    constexpr int NW = (TX + F - 1 + 3) / 4;         // float4s per column-pass window
    constexpr int LP = ((B - TX + 4 * NW) > L ? (B - TX + 4 * NW) : L + 3) & ~3;  // xin pitch
    __shared__ __attribute__((aligned(16))) float xin[L * LP];
    __shared__ __attribute__((aligned(16))) float tmid[L * B];

    const dim_t blk = xcd_remap(blockIdx.x, p.nblk);
    const dim_t bx = blk % p.nbx;
    const dim_t t = blk / p.nbx;
    const dim_t by = t % p.nby;
    const dim_t b = t / p.nby;
    in += b * p.S0 * p.S1;
    out += b * p.O0 * p.O1;

    const int cf = p.cf, rf = p.rf;
    float wc[F], wr[F];  // flipped taps, zero beyond the filter length
#pragma unroll
    for (int k = 0; k < F; ++k) {
        wc[k] = k < cf ? cfilt[cf - 1 - k] : 0.f;
        wr[k] = k < rf ? rfilt[rf - 1 - k] : 0.f;
    }

    const dim_t ox0 = bx * B, oy0 = by * B;
    const dim_t ix0 = ox0 + p.sx - (cf - 1), iy0 = oy0 + p.sy - (rf - 1);
    const int LWr = B + cf - 1, LHr = B + rf - 1;  // halo the filters read
    load_halo_tile_f32<256, (L + 6) / 4>(xin, in, p.S0, p.S1, ix0, iy0, LHr, LWr, LP);
    // zero the cells outside the halo (read into registers, never used)
    for (int i = threadIdx.x; i < L * LP; i += 256) {
        const int r = i / LP, c = i - (i / LP) * LP;
        if (r >= LHr || c >= LWr) xin[i] = 0.f;
    }
    __syncthreads();

    // Column pass: tmid(x, r) = sum_k xin(x + k, r) * wc[k], r < L.
    {
        const int x0 = (threadIdx.x & 7) * TX;
        for (int r = threadIdx.x >> 3; r < L; r += 32) {
            float win[4 * NW];
            const float4 *src = reinterpret_cast<const float4 *>(xin + r * LP + x0);
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const float4 v = src[q];
                win[4 * q] = v.x;
                win[4 * q + 1] = v.y;
                win[4 * q + 2] = v.z;
                win[4 * q + 3] = v.w;
            }
            float acc[TX];
#pragma unroll
            for (int e = 0; e < TX; ++e) acc[e] = 0.f;
#pragma unroll
            for (int k = 0; k < F; ++k)
                if (EXACT || k < cf)  // uniform: taps past the filter are skipped, not multiplied by 0
#pragma unroll
                    for (int e = 0; e < TX; ++e) acc[e] = __builtin_fmaf(win[e + k], wc[k], acc[e]);
            float4 *dst = reinterpret_cast<float4 *>(tmid + r * B + x0);
            dst[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
            dst[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
        }
    }
    __syncthreads();

    // Row pass: out(x, y) = sum_k tmid(x, y + k) * wr[k].
    {
        const int x = threadIdx.x & 63;
        const int y0 = (threadIdx.x >> 6) * TY;
        float v[TY + F - 1];  // the column's temp values this lane's outputs read
#pragma unroll
        for (int q = 0; q < TY + F - 1; ++q) v[q] = tmid[(y0 + q) * B + x];
        float acc[TY];
#pragma unroll
        for (int q = 0; q < TY; ++q) acc[q] = 0.f;
#pragma unroll
        for (int k = 0; k < F; ++k)
            if (EXACT || k < rf)  // uniform, as in the column pass
#pragma unroll
                for (int q = 0; q < TY; ++q) acc[q] = __builtin_fmaf(v[q + k], wr[k], acc[q]);
        const dim_t gx = ox0 + x;
        if (gx < p.O0) {
#pragma unroll
            for (int q = 0; q < TY; ++q) {
                const dim_t gy = oy0 + y0 + q;
                if (gy < p.O1) out[gy * p.O0 + gx] = acc[q];
            }
        }
    }
}

// Strip form of the fused FP32 kernel, for inputs whose rows are whole
// 16-byte vectors (S0 % 4 == 0, 16-byte aligned base, images below 2 GiB).
// A workgroup owns a 64-column output strip over p.seg output rows of one
// image and walks down it in steps of 16 rows:
//  * the column pass of step u turns the 16 input rows iy0 + 16u .. +15 (in
//    LDS, xin) into 16 temp rows of a ring in LDS, so the rf - 1 halo rows
//    are computed once per strip segment, not once per output tile; each temp
//    row is written twice, RING rows apart, so that a row-pass read sequence
//    never wraps;
//  * the row pass of output rows 16(u - P) .. +15 reads the ring once the
//    rows it needs are there (P = ceil((F - 1) / 16) steps behind);
//  * the input rows of step u + 2 are loaded into registers while step u
//    runs (buffer loads: rows and 16-byte column groups outside the image
//    read zeros) and written to xin between the two passes of step u + 1.
// Sums run in the same order as sep_fused_f32 (flipped taps ascending); taps
// past the filter length are skipped by uniform branches (EXACT: cf == rf ==
// F, no branches), never multiplied by 0. The xin origin is the strip's
// first input column floored to a multiple of 4; the column pass reads its
// window with aligned 16-byte loads (lanes 16 bytes apart, so no LDS bank
// conflicts) and applies the taps at the remaining offset, a compile-time
// constant in each of the four copies of the step loop.
struct SepStripParams {
    dim_t S0, S1, O0, O1;
    dim_t nbx, nseg, nblk;
    int cf, rf, sx, sy;
    int seg;  // output rows per workgroup (a multiple of 16)
};

template <int F, bool EXACT>
__global__ __launch_bounds__(256) void sep_strip_f32(float *__restrict__ out,
                                                     const float *__restrict__ in,
                                                     const float *__restrict__ cfilt,
                                                     const float *__restrict__ rfilt,
                                                     SepStripParams p) {
    constexpr int H = F - 1;
    constexpr int P = (H + 15) / 16;                    // column-only steps before the first row pass
    constexpr int RING = 16 * (P + 1) <= 32 ? 32 : 64;  // temp rows kept (power of two)
    constexpr int XQ = (64 + H + 3 + 3) / 4;            // float4s loaded per input row
    // xin row pitch (floats): at least the row, and 16 mod 32 so that the two
    // rows a 16-byte window read pairs (lanes l and l + 32, rows 2 apart) sit
    // on opposite halves of the 64 banks
    constexpr int XP = (4 * XQ + 15) / 32 * 32 + 16;
    constexpr int NSLOT = (16 * XQ + 255) / 256;        // 16-byte loads per lane per step
    constexpr int kOOB = 0x7fffffc0;
    __shared__ __attribute__((aligned(16))) float xin[16 * XP];
    __shared__ __attribute__((aligned(16))) float ring[2 * RING * 64];

    const dim_t blk = xcd_remap(blockIdx.x, p.nblk);
    const int bx = (int)(blk % p.nbx);
    const dim_t t = blk / p.nbx;
    const int sg = (int)(t % p.nseg);
    const dim_t b = t / p.nseg;
    const float *img = in + b * p.S0 * p.S1;
    out += b * p.O0 * p.O1;

    const int S0 = (int)p.S0, S1 = (int)p.S1, O0 = (int)p.O0, O1 = (int)p.O1;
    const int cf = p.cf, rf = p.rf;
    const int ox0 = bx * 64, oy0 = sg * p.seg;
    const int cx = ox0 + p.sx - (cf - 1);             // first input column the strip reads
    const int a0 = cx - (((cx % 4) + 4) % 4);         // floored to a multiple of 4
    const int off = cx - a0;                          // 0..3, the same for every strip
    const int iy0 = oy0 + p.sy - (rf - 1);
    const int nsteps = ((O1 - oy0 < p.seg ? O1 - oy0 : p.seg) + 15) / 16;
    const int U = nsteps + P;                         // column-pass steps

    float wc[F], wr[F];  // flipped taps (only k < cf, k < rf are read)
#pragma unroll
    for (int k = 0; k < F; ++k) {
        wc[k] = cfilt[k < cf ? cf - 1 - k : 0];
        wr[k] = rfilt[k < rf ? rf - 1 - k : 0];
    }

    // this lane's load slots: row s_row, byte offset s_col in the image row
    const auto rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(img), (short)0, S0 * S1 * 4,
                                                        0x00020000);
    int s_row[NSLOT], s_col[NSLOT], s_lds[NSLOT], s_off[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        const int slot = s * 256 + (int)threadIdx.x;
        const int r = slot / XQ, q = slot - (slot / XQ) * XQ;
        const int gx = a0 + 4 * q;
        const bool ok = slot < 16 * XQ && gx >= 0 && gx < S0;  // a float4 is wholly in or out
        s_row[s] = ok ? r : -(1 << 28);                      // a dead slot's row is never in range
        s_col[s] = ok ? gx * 4 : 0;
        s_off[s] = ok ? r * (S0 * 4) + gx * 4 : kOOB;        // from the block's first row
        s_lds[s] = slot < 16 * XQ ? r * XP + 4 * q : -1;
    }
    float4 pre[NSLOT];
    auto issue = [&](int u) __attribute__((always_inline)) {
        const int row0 = iy0 + 16 * u;
        if (row0 >= 0 && row0 + 16 <= S1) {
            // the block's 16 rows are inside the image: the row goes in the
            // uniform offset, the lane offsets are fixed
#pragma unroll
            for (int s = 0; s < NSLOT; ++s)
                pre[s] = __builtin_bit_cast(
                    float4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, s_off[s], row0 * (S0 * 4), 0));
        } else {
#pragma unroll
            for (int s = 0; s < NSLOT; ++s) {
                const int gy = row0 + s_row[s];
                const int o = (unsigned)gy < (unsigned)S1 ? gy * (S0 * 4) + s_col[s] : kOOB;
                pre[s] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, o, 0, 0));
            }
        }
    };
    auto commit = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int s = 0; s < NSLOT; ++s)
            if (s_lds[s] >= 0) lds_store16(xin + s_lds[s], pre[s]);  // s_lds is a multiple of 4
    };

    issue(0);
    commit();
    if (U > 1) issue(1);
    __syncthreads();

    const int cr = threadIdx.x >> 4, cx0 = 4 * (threadIdx.x & 15);  // column pass: row, 4 outputs
    const int rx = threadIdx.x & 63;                                  // row pass: column, 4 outputs
    const int ry0 = 4 * __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);  // wave-uniform
    // results: buffer stores at a fixed lane offset (past the end for columns
    // beyond the output) and a uniform row offset
    const auto orsrc = __builtin_amdgcn_make_buffer_rsrc(out, (short)0, O0 * O1 * 4, 0x00020000);
    const int o_lane = ox0 + rx < O0 ? (ox0 + rx) * 4 : kOOB;
    const float *wsrc = xin + cr * XP + cx0;  // 16-byte aligned; the window starts OFF floats in
    // One copy of the step loop per column offset, so that the window is read
    // with aligned 16-byte loads and indexed with compile-time offsets.
    auto steps = [&](auto offc) __attribute__((always_inline)) {
        constexpr int OFF = decltype(offc)::value;
        constexpr int NWQ = (OFF + 4 + H + 3) / 4;  // float4s per column-pass window
        static_assert(60 + 4 * NWQ <= XP, "window stays in its xin row");
        for (int u = 0; u < U; ++u) {
            {   // column pass: temp row 16u + cr, outputs cx0 .. cx0 + 3
                float win[4 * NWQ];
#pragma unroll
                for (int q = 0; q < NWQ; ++q) {
                    const float4 f = lds_load16(wsrc + 4 * q);
                    win[4 * q] = f.x;
                    win[4 * q + 1] = f.y;
                    win[4 * q + 2] = f.z;
                    win[4 * q + 3] = f.w;
                }
                float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < F; ++k)
                    if (EXACT || k < cf) {  // uniform
                        if constexpr (!EXACT) asm volatile("");  // keeps the branch (no select of products)
#pragma unroll
                        for (int e = 0; e < 4; ++e) acc[e] = __builtin_fmaf(win[OFF + e + k], wc[k], acc[e]);
                    }
                const float4 v = make_float4(acc[0], acc[1], acc[2], acc[3]);
                float *dst = ring + ((16 * u + cr) & (RING - 1)) * 64 + cx0;
                *reinterpret_cast<float4 *>(dst) = v;
                *reinterpret_cast<float4 *>(dst + RING * 64) = v;
            }
            __syncthreads();
            if (u + 1 < U) commit();
            if (u + 2 < U) issue(u + 2);
            if (u >= P) {  // row pass: output rows 16(u - P) + ry0 .. + 3 of column rx
                const int y0 = 16 * (u - P) + ry0;
                const float *src = ring + (((16 * (u - P)) & (RING - 1)) + ry0) * 64 + rx;
                float v[4 + H];
#pragma unroll
                for (int q = 0; q < 4 + H; ++q) v[q] = src[q * 64];
                float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < F; ++k)
                    if (EXACT || k < rf) {  // uniform
                        if constexpr (!EXACT) asm volatile("");
#pragma unroll
                        for (int i = 0; i < 4; ++i) acc[i] = __builtin_fmaf(v[i + k], wr[k], acc[i]);
                    }
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (oy0 + y0 + i < O1)  // uniform
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, acc[i]), orsrc, o_lane,
                                                              (oy0 + y0 + i) * (O0 * 4), 0);
            }
            __syncthreads();
        }
    };
    switch (off) {
        case 0: steps(std::integral_constant<int, 0>{}); break;
        case 1: steps(std::integral_constant<int, 1>{}); break;
        case 2: steps(std::integral_constant<int, 2>{}); break;
        default: steps(std::integral_constant<int, 3>{}); break;
    }
}

// The strip kernel takes images whose rows are whole 16-byte vectors from a
// 16-byte aligned base, input and output images below 2 GiB each (its buffer
// offsets are 32-bit).
// AFHIP_SEP_STRIP=0 in the environment selects the tile kernel instead.
static bool sep_strip_ok(const long long *sdims, const long long *odims, const void *in) {
    static const bool enabled = [] {
        const char *e = getenv("AFHIP_SEP_STRIP");
        return !(e && e[0] == '0');
    }();
    return enabled && sdims[0] % 4 == 0 && ((uintptr_t)in & 15) == 0 &&
           sdims[0] * sdims[1] * 4 < 0x7fffffc0LL && odims[0] * odims[1] * 4 < 0x7fffffc0LL;
}

static unsigned grid_for(dim_t n) {
    dim_t b = (n + 255) / 256;
    if (b > kMaxGridBlocks) b = kMaxGridBlocks;
    if (b < 1) b = 1;
    return (unsigned)b;
}

template <typename InT, typename AccT>
static int launch_two_pass(void *out, const void *in, const void *cf, dim_t cflen,
                           const void *rf, dim_t rflen, SepParams p, void *tmp,
                           hipStream_t s) {
    const dim_t nt = p.td[0] * p.td[1] * p.td[2] * p.td[3];
    const dim_t no = p.od[0] * p.od[1] * p.od[2] * p.od[3];
    if (no == 0) return K_OK;
    if (!tmp) return K_ERR_ARG;
    p.flen = cflen;
    sep_pass<InT, AccT, 0><<<grid_for(nt), 256, 0, s>>>((InT *)tmp, (const InT *)in,
                                                          (const AccT *)cf, p);
    p.flen = rflen;
    sep_pass<InT, AccT, 1><<<grid_for(no), 256, 0, s>>>((InT *)out, (const InT *)tmp,
                                                          (const AccT *)rf, p);
    return launch_status();
}

}  // namespace afhip

using namespace afhip;

extern "C" int afhip_k_convolve2_sep(int dtype, void *out, const long long *odims,
                                     const void *in, const long long *sdims,
                                     const void *cfilt, long long cflen,
                                     const void *rfilt, long long rflen, int expand,
                                     void *tmp, void *stream) {
    if (cflen < 1 || rflen < 1) return K_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    SepParams p;
    for (int d = 0; d < 4; ++d) {
        p.sd[d] = sdims[d];
        p.od[d] = odims[d];
        p.td[d] = sdims[d];
    }
    if (expand) p.td[0] += cflen - 1;
    p.expand = expand;
    p.flen = 0;
    if (dtype == DT_F32 && tmp == nullptr && cflen <= SF_MAXF && rflen <= SF_MAXF &&
        sep_strip_ok(sdims, odims, in)) {
        SepStripParams f;
        f.S0 = sdims[0];
        f.S1 = sdims[1];
        f.O0 = odims[0];
        f.O1 = odims[1];
        f.nbx = (f.O0 + 63) / 64;
        f.seg = strip_rows(f.O1, f.nbx * odims[2] * odims[3]);
        f.nseg = (f.O1 + f.seg - 1) / f.seg;
        f.nblk = f.nbx * f.nseg * odims[2] * odims[3];
        f.cf = (int)cflen;
        f.rf = (int)rflen;
        f.sx = expand ? 0 : (int)(cflen >> 1);
        f.sy = expand ? 0 : (int)(rflen >> 1);
        if (f.nblk == 0) return K_OK;
        if (f.nblk > 0x7fffffffLL) return K_ERR_SIZE;
        const int fm = (int)(cflen > rflen ? cflen : rflen);
#define SEPS(FB)                                                                                           \
    do {                                                                                                   \
        if (cflen == FB && rflen == FB)                                                                    \
            sep_strip_f32<FB, true><<<dim3((unsigned)f.nblk), 256, 0, s>>>(                                \
                (float *)out, (const float *)in, (const float *)cfilt, (const float *)rfilt, f);           \
        else                                                                                               \
            sep_strip_f32<FB, false><<<dim3((unsigned)f.nblk), 256, 0, s>>>(                               \
                (float *)out, (const float *)in, (const float *)cfilt, (const float *)rfilt, f);           \
    } while (0)
        if (fm <= 3) SEPS(3);
        else if (fm <= 5) SEPS(5);
        else if (fm <= 7) SEPS(7);
        else if (fm <= 9) SEPS(9);
        else if (fm <= 11) SEPS(11);
        else if (fm <= 13) SEPS(13);
        else if (fm <= 15) SEPS(15);
        else if (fm <= 17) SEPS(17);
        else SEPS(SF_MAXF);
#undef SEPS
        return launch_status();
    }
    if (dtype == DT_F32 && tmp == nullptr && cflen <= SF_MAXF && rflen <= SF_MAXF) {
        SepFusedParams f;
        f.S0 = sdims[0];
        f.S1 = sdims[1];
        f.O0 = odims[0];
        f.O1 = odims[1];
        f.nbx = (f.O0 + SF_B - 1) / SF_B;
        f.nby = (f.O1 + SF_B - 1) / SF_B;
        f.nblk = f.nbx * f.nby * odims[2] * odims[3];
        f.cf = (int)cflen;
        f.rf = (int)rflen;
        f.sx = expand ? 0 : (int)(cflen >> 1);
        f.sy = expand ? 0 : (int)(rflen >> 1);
        if (f.nblk == 0) return K_OK;
        if (f.nblk > 0x7fffffffLL) return K_ERR_SIZE;
        const int fm = (int)(cflen > rflen ? cflen : rflen);
#define SEPF(FB)                                                                                           \
    do {                                                                                                   \
        if (cflen == FB && rflen == FB)                                                                    \
            sep_fused_f32<FB, true><<<dim3((unsigned)f.nblk), 256, 0, s>>>(                                \
                (float *)out, (const float *)in, (const float *)cfilt, (const float *)rfilt, f);           \
        else                                                                                               \
            sep_fused_f32<FB, false><<<dim3((unsigned)f.nblk), 256, 0, s>>>(                               \
                (float *)out, (const float *)in, (const float *)cfilt, (const float *)rfilt, f);           \
    } while (0)
        if (fm <= 3) SEPF(3);
        else if (fm <= 5) SEPF(5);
        else if (fm <= 7) SEPF(7);
        else if (fm <= 9) SEPF(9);
        else if (fm <= 11) SEPF(11);
        else if (fm <= 13) SEPF(13);
        else if (fm <= 15) SEPF(15);
        else if (fm <= 17) SEPF(17);
        else SEPF(SF_MAXF);
#undef SEPF
        return launch_status();
    }
    switch (dtype) {
        case DT_F32: return launch_two_pass<float, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_F64: return launch_two_pass<double, double>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_C32: return launch_two_pass<cfloat, cfloat>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_C64: return launch_two_pass<cdouble, cdouble>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_S32: return launch_two_pass<int32_t, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_U32: return launch_two_pass<uint32_t, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_S16: return launch_two_pass<int16_t, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_U16: return launch_two_pass<uint16_t, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_S64: return launch_two_pass<long long, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_U64: return launch_two_pass<unsigned long long, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_U8: return launch_two_pass<uint8_t, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_S8: return launch_two_pass<int8_t, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        case DT_B8: return launch_two_pass<int8_t, float>(out, in, cfilt, cflen, rfilt, rflen, p, tmp, s);
        default: return K_ERR_TYPE;
    }
}
