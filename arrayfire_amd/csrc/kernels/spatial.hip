// Spatial-domain convolution kernels (convolve1/2/3 with every batch kind).
//
// Semantics (reference CPU backend, src/backend/cpu/kernel/convolve.hpp):
//   y[o] = sum_f x[o + s - f] * w[f],  s = EXPAND ? 0 : floor(F/2),
// zero outside x, accumulated in AccT, stored as InT(acc); batch stepping as
// convolve_nd (:132-211). The host passes per-kind batch extents and steps.
//
// Implementations:
//  * conv2_strip<FX,FY,OFF>: FP32 rank-2 kernel for odd filter sides 3..9 on
//    rows of whole 16-byte vectors; walks 128-column strips down the image
//    with an input-row ring in LDS and register prefetch (see the kernel).
//  * conv3_tile<F,F,F>: FP32 rank-3 kernel for 3x3x3 and 5x5x5 filters,
//    32x8x8 output blocks.
//  * conv1_tile: FP32 rank-1 kernel for filters of up to 2048 taps, every
//    batch kind: 1024 outputs per 256-lane workgroup from a window staged in
//    LDS, 4 outputs per lane, taps read as LDS broadcasts.
//  * conv2_tile<FX,FY>: FP32 rank-2 kernel. A 128x64 output tile per
//    256-thread workgroup; the (128+FX-1)x(64+FY-1) input halo tile is staged
//    in LDS with zero borders so the tap loops are branch-free; each lane keeps
//    a 4x8 block of outputs in registers, reads each LDS row segment once
//    (ds_read_b128) and applies all the taps that touch it; the flipped filter
//    lives in SGPRs (uniform scalar loads).
//  * conv_generic: any rank, dtype, filter size and batch kind; one lane per
//    output, grid-stride loop. Used where no tiled variant applies.
#include <stdlib.h>

#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

struct SpatialParams {
    dim_t od[4], sd[4], fd[4];
    dim_t os[4], ss[4], fs[4];
    dim_t out_step[4], in_step[4], filt_step[4], batch[4];
    int rank, expand;
};

// The product is rounded before it is added (no fused multiply-add), as the
// reference's accum += AccT(s_val * f) is evaluated on the CPU: truncation to
// an integer output then gives the reference's value bit for bit.
template <typename InT, typename AccT>
__device__ inline AccT mul_acc(InT s, AccT f) {
#pragma clang fp contract(off)
    return AccT(s * f);
}
template <>
__device__ inline float mul_acc<int8_t, float>(int8_t s, float f) {
#pragma clang fp contract(off)
    return float(s) * f;
}

template <typename InT, typename AccT>
__device__ inline InT to_out(AccT a) {
    return InT(a);
}

template <typename InT, typename AccT, int RANK>
__global__ __launch_bounds__(256) void conv_generic(InT *__restrict__ out,
                                                    const InT *__restrict__ in,
                                                    const AccT *__restrict__ filt,
                                                    SpatialParams p) {
#pragma clang fp contract(off)
    dim_t ext[3] = {p.od[0], RANK > 1 ? p.od[1] : 1, RANK > 2 ? p.od[2] : 1};
    const dim_t nsp = ext[0] * ext[1] * ext[2];
    const dim_t total = nsp * p.batch[1] * p.batch[2] * p.batch[3];
    dim_t st[3];
    for (int d = 0; d < 3; ++d) st[d] = p.expand ? 0 : p.fd[d] / 2;
    for (dim_t lin = blockIdx.x * (dim_t)blockDim.x + threadIdx.x; lin < total;
         lin += (dim_t)gridDim.x * blockDim.x) {
        dim_t r = lin;
        const dim_t o0 = r % ext[0];
        r /= ext[0];
        const dim_t o1 = r % ext[1];
        r /= ext[1];
        const dim_t o2 = r % ext[2];
        r /= ext[2];
        const dim_t b1 = r % p.batch[1];
        r /= p.batch[1];
        const dim_t b2 = r % p.batch[2];
        const dim_t b3 = r / p.batch[2];
        const InT *ip = in + b1 * p.in_step[1] + b2 * p.in_step[2] + b3 * p.in_step[3];
        const AccT *fp = filt + b1 * p.filt_step[1] + b2 * p.filt_step[2] + b3 * p.filt_step[3];
        InT *op = out + b1 * p.out_step[1] + b2 * p.out_step[2] + b3 * p.out_step[3];
        const dim_t i0 = o0 + st[0], i1 = o1 + st[1], i2 = o2 + st[2];
        AccT acc = AccT(0);
        const dim_t F1 = RANK > 1 ? p.fd[1] : 1, F2 = RANK > 2 ? p.fd[2] : 1;
        for (dim_t w2 = 0; w2 < F2; ++w2) {
            const dim_t x2 = i2 - w2;
            const bool v2 = RANK < 3 || (x2 >= 0 && x2 < p.sd[2]);
            for (dim_t w1 = 0; w1 < F1; ++w1) {
                const dim_t x1 = i1 - w1;
                const bool v1 = v2 && (RANK < 2 || (x1 >= 0 && x1 < p.sd[1]));
                const dim_t soff = (RANK > 2 ? x2 * p.ss[2] : 0) + (RANK > 1 ? x1 * p.ss[1] : 0);
                const dim_t foff = (RANK > 2 ? w2 * p.fs[2] : 0) + (RANK > 1 ? w1 * p.fs[1] : 0);
                for (dim_t w0 = 0; w0 < p.fd[0]; ++w0) {
                    const dim_t x0 = i0 - w0;
                    InT s = InT(0);
                    if (v1 && x0 >= 0 && x0 < p.sd[0]) s = ip[soff + x0];
                    acc += mul_acc<InT, AccT>(s, fp[foff + w0 * p.fs[0]]);
                }
            }
        }
        const dim_t ooff = o0 + (RANK > 1 ? o1 * p.os[1] : 0) + (RANK > 2 ? o2 * p.os[2] : 0);
        op[ooff] = to_out<InT, AccT>(acc);
    }
}

// ---------------------------------------------------------------- FP32 2-D
struct Tile2Params {
    dim_t S0, S1, O0, O1;   // signal and output extents of one slice
    dim_t in_step2, in_step3, out_step2, out_step3, filt_step2, filt_step3;
    dim_t batch2;           // extent of batch dim 2 (dim 3 is the remainder)
    dim_t nbx, nby, nblk;   // tiles along dim0, dim1, total workgroups
    int sx, sy;             // centring offsets (0 for EXPAND)
    int seg;                // conv2_strip: output rows per workgroup (a multiple of 16)
};

constexpr int T2_TX = 4, T2_TY = 8, T2_NTX = 32, T2_NTY = 8;
constexpr int T2_BX = T2_TX * T2_NTX, T2_BY = T2_TY * T2_NTY;

template <int FX, int FY>
__global__ __launch_bounds__(256) void conv2_tile(float *__restrict__ out,
                                                  const float *__restrict__ in,
                                                  const float *__restrict__ filt,
                                                  Tile2Params p) {
    constexpr int TX = T2_TX, TY = T2_TY, NTX = T2_NTX;
    constexpr int BX = T2_BX, BY = T2_BY;
    constexpr int LW = BX + FX - 1, LH = BY + FY - 1;
    constexpr int LS = (LW + 3) & ~3;  // row pitch, 16-byte aligned
    constexpr int NV = (TX + FX - 1 + 3) / 4;  // float4 reads per row segment
    __shared__ __attribute__((aligned(16))) float tile[LH * LS];

    const dim_t blk = xcd_remap(blockIdx.x, p.nblk);
    const dim_t bx = blk % p.nbx;
    const dim_t t = blk / p.nbx;
    const dim_t by = t % p.nby;
    const dim_t b = t / p.nby;
    const dim_t b2 = b % p.batch2, b3 = b / p.batch2;
    in += b2 * p.in_step2 + b3 * p.in_step3;
    out += b2 * p.out_step2 + b3 * p.out_step3;
    filt += b2 * p.filt_step2 + b3 * p.filt_step3;

    // Flipped taps: out(lx,ly) = sum tile(lx+kx, ly+ky) * w(FX-1-kx, FY-1-ky).
    float wf[FY][FX];
#pragma unroll
    for (int ky = 0; ky < FY; ++ky)
#pragma unroll
        for (int kx = 0; kx < FX; ++kx) wf[ky][kx] = filt[(FY - 1 - ky) * FX + (FX - 1 - kx)];

    const dim_t ox0 = bx * BX, oy0 = by * BY;
    const dim_t ix0 = ox0 + p.sx - (FX - 1), iy0 = oy0 + p.sy - (FY - 1);
    load_halo_tile_f32<256, (LW + 6) / 4>(tile, in, p.S0, p.S1, ix0, iy0, LH, LW, LS);
    __syncthreads();

    const int tx = threadIdx.x % NTX, ty = threadIdx.x / NTX;
    float acc[TY][TX];
#pragma unroll
    for (int j = 0; j < TY; ++j)
#pragma unroll
        for (int i = 0; i < TX; ++i) acc[j][i] = 0.f;

    const float *base = tile + (ty * TY) * LS + tx * TX;
#pragma unroll
    for (int r = 0; r < TY + FY - 1; ++r) {
        float v[NV * 4];
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const float4 f4 = *reinterpret_cast<const float4 *>(base + r * LS + 4 * q);
            v[4 * q + 0] = f4.x;
            v[4 * q + 1] = f4.y;
            v[4 * q + 2] = f4.z;
            v[4 * q + 3] = f4.w;
        }
#pragma unroll
        for (int j = 0; j < TY; ++j) {
            const int ky = r - j;
            if (ky >= 0 && ky < FY) {
#pragma unroll
                for (int i = 0; i < TX; ++i)
#pragma unroll
                    for (int kx = 0; kx < FX; ++kx)
                        acc[j][i] = __builtin_fmaf(v[i + kx], wf[ky][kx], acc[j][i]);
            }
        }
    }

    const dim_t oxt = ox0 + tx * TX;
#pragma unroll
    for (int j = 0; j < TY; ++j) {
        const dim_t oy = oy0 + ty * TY + j;
        if (oy >= p.O1) break;
        float *orow = out + oy * p.O0;
        if ((p.O0 & 3) == 0 && oxt + TX <= p.O0) {
            *reinterpret_cast<float4 *>(orow + oxt) =
                make_float4(acc[j][0], acc[j][1], acc[j][2], acc[j][3]);
        } else {
#pragma unroll
            for (int i = 0; i < TX; ++i)
                if (oxt + i < p.O0) orow[oxt + i] = acc[j][i];
        }
    }
}

// Strip form of conv2_tile for inputs whose rows are whole 16-byte vectors
// (S0 % 4 == 0, 16-byte aligned slices, slices below 2 GiB). A workgroup owns
// a 128-column output strip over p.seg output rows of one slice and walks
// down it in steps of 16 output rows. Input rows live in a ring of 4 blocks
// of 16 rows in LDS; output step s reads blocks s and s + 1 (FY <= 17):
//  * at the start of step s, block s + 2 (loaded into registers during step
//    s - 1) is written to its ring slot, whose previous block s - 2 no step
//    after s - 1 reads, and the loads of block s + 3 are issued;
//  * the taps of step s then run from LDS; one barrier ends the step.
// Rows and 16-byte column groups outside the slice read zeros (buffer loads).
// A ring row holds the input columns from ix0 floored to a multiple of 4, so
// loaded float4s are stored and read whole and aligned (lanes 16 bytes apart:
// no LDS bank conflicts); OFF = ix0 mod 4, a compile-time constant, shifts
// the taps within each lane's window. Each lane computes 4 columns x 2 rows
// with the flipped filter in scalar registers, summing in the same order as
// conv2_tile.
template <int FX, int FY, int OFF>
__global__ __launch_bounds__(256) void conv2_strip(float *__restrict__ out,
                                                   const float *__restrict__ in,
                                                   const float *__restrict__ filt,
                                                   Tile2Params p) {
    constexpr int BX = 128, TX = 4, TY = 2;
    constexpr int LW = BX + FX - 1;                    // input columns a strip reads
    constexpr int XQ = (LW + 3 + 3) / 4;               // float4s loaded per input row
    // ring row pitch (floats): at least the row, and 16 mod 32 so that the
    // two rows a 16-byte read pairs (lanes l and l + 32, rows 2 apart) sit on
    // opposite halves of the 64 banks
    constexpr int LS = (4 * XQ + 15) / 32 * 32 + 16;
    constexpr int NSLOT = (16 * XQ + 255) / 256;       // 16-byte loads per lane per block
    constexpr int NV = (OFF + TX + FX - 1 + 3) / 4;    // float4 reads per row segment
    constexpr int kOOB = 0x7fffffc0;
    static_assert(FY <= 17, "two ring blocks per output step");
    static_assert(4 * 31 + 4 * NV <= LS, "row segment reads stay in the row");
    __shared__ __attribute__((aligned(16))) float ring[64 * LS];

    const dim_t blk = xcd_remap(blockIdx.x, p.nblk);
    const int bx = (int)(blk % p.nbx);
    const dim_t t = blk / p.nbx;
    const int sg = (int)(t % p.nby);
    const dim_t b = t / p.nby;
    const dim_t b2 = b % p.batch2, b3 = b / p.batch2;
    const float *img = in + b2 * p.in_step2 + b3 * p.in_step3;
    out += b2 * p.out_step2 + b3 * p.out_step3;
    filt += b2 * p.filt_step2 + b3 * p.filt_step3;

    float wf[FY][FX];
#pragma unroll
    for (int ky = 0; ky < FY; ++ky)
#pragma unroll
        for (int kx = 0; kx < FX; ++kx) wf[ky][kx] = filt[(FY - 1 - ky) * FX + (FX - 1 - kx)];

    const int S0 = (int)p.S0, S1 = (int)p.S1, O0 = (int)p.O0, O1 = (int)p.O1;
    const int ox0 = bx * BX, oy0 = sg * p.seg;
    const int ix0 = ox0 + p.sx - (FX - 1), iy0 = oy0 + p.sy - (FY - 1);
    const int a0 = ix0 - (((ix0 % 4) + 4) % 4);        // ix0 floored to a multiple of 4
    const int nsteps = ((O1 - oy0 < p.seg ? O1 - oy0 : p.seg) + 15) / 16;
    const int nblocks = nsteps + 1;                    // input blocks the segment reads

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
        s_lds[s] = slot < 16 * XQ ? r * LS + 4 * q : -1;
    }
    float4 pre[NSLOT], pre1[NSLOT];
    auto issue = [&](float4 (&v)[NSLOT], int bl) __attribute__((always_inline)) {
        const int row0 = iy0 + 16 * bl;
        if (row0 >= 0 && row0 + 16 <= S1) {
            // the block's 16 rows are inside the image: the row goes in the
            // uniform offset, the lane offsets are fixed
#pragma unroll
            for (int s = 0; s < NSLOT; ++s)
                v[s] = __builtin_bit_cast(
                    float4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, s_off[s], row0 * (S0 * 4), 0));
        } else {
#pragma unroll
            for (int s = 0; s < NSLOT; ++s) {
                const int gy = row0 + s_row[s];
                const int o = (unsigned)gy < (unsigned)S1 ? gy * (S0 * 4) + s_col[s] : kOOB;
                v[s] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, o, 0, 0));
            }
        }
    };
    auto commit = [&](const float4 (&v)[NSLOT], int bl) __attribute__((always_inline)) {
        float *dst = ring + (bl & 3) * (16 * LS);
#pragma unroll
        for (int s = 0; s < NSLOT; ++s)
            if (s_lds[s] >= 0) lds_store16(dst + s_lds[s], v[s]);  // s_lds is a multiple of 4
    };

    issue(pre, 0);
    issue(pre1, 1);
    commit(pre, 0);
    commit(pre1, 1);
    if (nblocks > 2) issue(pre, 2);
    __syncthreads();

    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int oxt = ox0 + tx * TX;
    const bool vec_out = (O0 & 3) == 0 && oxt + TX <= O0;
    for (int st = 0; st < nsteps; ++st) {
        if (st + 2 < nblocks) commit(pre, st + 2);
        if (st + 3 < nblocks) issue(pre, st + 3);
        float acc[TY][TX];
#pragma unroll
        for (int j = 0; j < TY; ++j)
#pragma unroll
            for (int i = 0; i < TX; ++i) acc[j][i] = 0.f;
        const int rb = 16 * st + ty * TY;              // first input row (segment-relative)
#pragma unroll
        for (int r = 0; r < TY + FY - 1; ++r) {
            const float *row = ring + ((rb + r) & 63) * LS + tx * TX;
            float v[NV * 4];
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const float4 f4 = lds_load16(row + 4 * q);
                v[4 * q + 0] = f4.x;
                v[4 * q + 1] = f4.y;
                v[4 * q + 2] = f4.z;
                v[4 * q + 3] = f4.w;
            }
#pragma unroll
            for (int j = 0; j < TY; ++j) {
                const int ky = r - j;
                if (ky >= 0 && ky < FY) {
#pragma unroll
                    for (int i = 0; i < TX; ++i)
#pragma unroll
                        for (int kx = 0; kx < FX; ++kx)
                            acc[j][i] = __builtin_fmaf(v[OFF + i + kx], wf[ky][kx], acc[j][i]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < TY; ++j) {
            const int oy = oy0 + rb + j;
            if (oy < O1) {
                float *orow = out + (dim_t)oy * O0;
                if (vec_out) {
                    *reinterpret_cast<float4 *>(orow + oxt) =
                        make_float4(acc[j][0], acc[j][1], acc[j][2], acc[j][3]);
                } else {
#pragma unroll
                    for (int i = 0; i < TX; ++i)
                        if (oxt + i < O0) orow[oxt + i] = acc[j][i];
                }
            }
        }
        __syncthreads();
    }
}

// LDS bytes of conv2_tile<FX,FY>; must stay within 64 KiB per workgroup.
constexpr int tile2_lds_bytes(int fx, int fy) {
    return (T2_BY + fy - 1) * (((T2_BX + fx - 1) + 3) & ~3) * 4;
}

typedef void (*Tile2Fn)(float *, const float *, const float *, Tile2Params);

template <int FX, int FY>
constexpr Tile2Fn tile2_entry() {
    return conv2_tile<FX, FY>;
}

// Filter sizes with a register-tiled FP32 2-D variant: every combination up
// to 5x5, and the odd squares up to 9x9 (larger filters go generic for now).
static Tile2Fn tile2_lookup(dim_t fx, dim_t fy) {
#define T2CASE(a, b) \
    if (fx == a && fy == b) return tile2_entry<a, b>();
#define T2ROW(a) T2CASE(a, 1) T2CASE(a, 2) T2CASE(a, 3) T2CASE(a, 4) T2CASE(a, 5)
    T2ROW(1) T2ROW(2) T2ROW(3) T2ROW(4) T2ROW(5)
    T2CASE(7, 7) T2CASE(9, 9)
#undef T2ROW
#undef T2CASE
    return nullptr;
}

static void fill_params(SpatialParams &p, const dim_t *od, const dim_t *sd,
                        const dim_t *fd, int kind, int rank, int expand) {
    for (int d = 0; d < 4; ++d) {
        p.od[d] = od[d];
        p.sd[d] = sd[d];
        p.fd[d] = fd[d];
    }
    p.os[0] = p.ss[0] = p.fs[0] = 1;
    for (int d = 1; d < 4; ++d) {
        p.os[d] = p.os[d - 1] * od[d - 1];
        p.ss[d] = p.ss[d - 1] * sd[d - 1];
        p.fs[d] = p.fs[d - 1] * fd[d - 1];
    }
    // Batch stepping per kind: src/backend/cpu/kernel/convolve.hpp:160-180.
    for (int i = 0; i < 4; ++i) {
        p.out_step[i] = p.in_step[i] = p.filt_step[i] = 0;
        p.batch[i] = 1;
    }
    for (int i = 1; i < 4; ++i) {
        switch (kind) {
            case 2:  // AF_BATCH_LHS
                p.out_step[i] = p.os[i];
                p.in_step[i] = p.ss[i];
                if (i >= rank) p.batch[i] = sd[i];
                break;
            case 4:  // AF_BATCH_SAME
                p.out_step[i] = p.os[i];
                p.in_step[i] = p.ss[i];
                p.filt_step[i] = p.fs[i];
                if (i >= rank) p.batch[i] = sd[i];
                break;
            case 3:  // AF_BATCH_RHS
                p.out_step[i] = p.os[i];
                p.filt_step[i] = p.fs[i];
                if (i >= rank) p.batch[i] = fd[i];
                break;
            default: break;
        }
    }
    p.rank = rank;
    p.expand = expand;
}

template <typename InT, typename AccT>
static int launch_generic(void *out, const void *in, const void *filt,
                          const SpatialParams &p, hipStream_t s) {
    dim_t nsp = 1;
    for (int d = 0; d < p.rank; ++d) nsp *= p.od[d];
    const dim_t total = nsp * p.batch[1] * p.batch[2] * p.batch[3];
    if (total == 0) return K_OK;
    dim_t blocks = (total + 255) / 256;
    if (blocks > kMaxGridBlocks) blocks = kMaxGridBlocks;
    switch (p.rank) {
        case 1:
            conv_generic<InT, AccT, 1><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
                (InT *)out, (const InT *)in, (const AccT *)filt, p);
            break;
        case 2:
            conv_generic<InT, AccT, 2><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
                (InT *)out, (const InT *)in, (const AccT *)filt, p);
            break;
        case 3:
            conv_generic<InT, AccT, 3><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
                (InT *)out, (const InT *)in, (const AccT *)filt, p);
            break;
        default: return K_ERR_ARG;
    }
    return launch_status();
}

// Odd filter sides 3..9 have a strip variant (conv2_strip). Its column
// offset OFF = (sx - (FX - 1)) mod 4 takes one value for DEFAULT (sx = FX / 2)
// and one for EXPAND (sx = 0); only those two are built.
template <int FX, int FY>
static Tile2Fn strip2_of(int off) {
    constexpr int DEF = ((-(FX - 1) / 2) % 4 + 4) % 4, EXP = ((-(FX - 1)) % 4 + 4) % 4;
    return off == DEF ? conv2_strip<FX, FY, DEF> : off == EXP ? conv2_strip<FX, FY, EXP> : nullptr;
}
template <int FX>
static Tile2Fn strip2_fy(dim_t fy, int off) {
    switch (fy) {
        case 3: return strip2_of<FX, 3>(off);
        case 5: return strip2_of<FX, 5>(off);
        case 7: return strip2_of<FX, 7>(off);
        case 9: return strip2_of<FX, 9>(off);
        default: return nullptr;
    }
}
static Tile2Fn strip2_lookup(dim_t fx, dim_t fy, int sx) {
    const int off = (int)((((sx - (fx - 1)) % 4) + 4) % 4);
    switch (fx) {
        case 3: return strip2_fy<3>(fy, off);
        case 5: return strip2_fy<5>(fy, off);
        case 7: return strip2_fy<7>(fy, off);
        case 9: return strip2_fy<9>(fy, off);
        default: return nullptr;
    }
}

// conv2_strip takes slices whose rows are whole 16-byte vectors from 16-byte
// aligned addresses, below 2 GiB each (its buffer offsets are 32-bit).
// AFHIP_CONV2_STRIP=0 in the environment selects conv2_tile instead.
static bool strip2_ok(const SpatialParams &p, const void *in) {
    static const bool enabled = [] {
        const char *e = getenv("AFHIP_CONV2_STRIP");
        return !(e && e[0] == '0');
    }();
    return enabled && p.sd[0] % 4 == 0 && p.in_step[2] % 4 == 0 && p.in_step[3] % 4 == 0 &&
           ((uintptr_t)in & 15) == 0 && p.sd[0] * p.sd[1] * 4 < 0x7fffffc0LL;
}

// fn: the conv2_tile variant (may be null when strip is not); strip: the
// conv2_strip variant, used when not null.
static int launch_tile2(float *out, const float *in, const float *filt,
                        const SpatialParams &p, Tile2Fn fn, Tile2Fn strip, hipStream_t s) {
    Tile2Params t;
    t.S0 = p.sd[0];
    t.S1 = p.sd[1];
    t.O0 = p.od[0];
    t.O1 = p.od[1];
    t.in_step2 = p.in_step[2];
    t.in_step3 = p.in_step[3];
    t.out_step2 = p.out_step[2];
    t.out_step3 = p.out_step[3];
    t.filt_step2 = p.filt_step[2];
    t.filt_step3 = p.filt_step[3];
    t.batch2 = p.batch[2];
    if (strip) fn = strip;
    t.nbx = (t.O0 + T2_BX - 1) / T2_BX;  // both kernels are 128 columns wide
    t.seg = strip ? strip_rows(t.O1, t.nbx * p.batch[2] * p.batch[3]) : 0;
    t.nby = strip ? (t.O1 + t.seg - 1) / t.seg : (t.O1 + T2_BY - 1) / T2_BY;
    const dim_t nb = p.batch[2] * p.batch[3];
    t.nblk = t.nbx * t.nby * nb;
    t.sx = p.expand ? 0 : (int)(p.fd[0] / 2);
    t.sy = p.expand ? 0 : (int)(p.fd[1] / 2);
    if (t.nblk == 0) return K_OK;
    if (t.nblk > 0x7fffffffLL) return K_ERR_SIZE;
    fn<<<dim3((unsigned)t.nblk), dim3(256), 0, s>>>(out, in, filt, t);
    return launch_status();
}

// ---------------------------------------------------------------- FP32 3-D
struct Tile3Params {
    dim_t S0, S1, S2, O0, O1, O2;  // signal and output extents of one volume
    dim_t in_step3, out_step3, filt_step3;
    dim_t nbx, nby, nbz, nblk;     // blocks along dim0, dim1, dim2, total workgroups
    int sx, sy, sz;                // centring offsets (0 for EXPAND)
};

// FP32 rank-3 kernel for small filters: a 32 x 8 x 8 output block per
// 256-lane workgroup. The (32+FX-1) x (8+FY-1) x (8+FZ-1) input halo is
// loaded in one batch of 16-byte loads (every slot's load in flight before
// any LDS store; rows read as aligned float4s, zeros outside the volume) and
// scattered into LDS so that tile column c is input column ix0 + c. Each lane
// computes 4 consecutive x outputs on 2 planes, reading each LDS row segment
// once (ds_read_b128) for all the taps that touch it; the flipped filter
// lives in SGPRs.
constexpr int T3_BX = 32, T3_BY = 8, T3_BZ = 8;

template <int FX, int FY, int FZ>
__global__ __launch_bounds__(256) void conv3_tile(float *__restrict__ out,
                                                  const float *__restrict__ in,
                                                  const float *__restrict__ filt,
                                                  Tile3Params p) {
    constexpr int LX = T3_BX + FX - 1, LY = T3_BY + FY - 1, LZ = T3_BZ + FZ - 1;
    // row pitch: 16-byte aligned, and rows 4 apart (lanes l and l + 32 of a
    // 16-byte read) 32 words apart mod 64, on opposite halves of the banks
    constexpr int LS = (((LX + 3) & ~3) * 4) % 64 == 32 ? ((LX + 3) & ~3)
                     : ((((LX + 3) & ~3) + 4) * 4) % 64 == 32 ? ((LX + 3) & ~3) + 4
                     : ((((LX + 3) & ~3) + 8) * 4) % 64 == 32 ? ((LX + 3) & ~3) + 8 : ((LX + 3) & ~3) + 12;
    constexpr int NQ = (LX + 6) / 4;              // float4 slots per row (origin floored to 4)
    constexpr int SLOTS = LZ * LY * NQ;
    constexpr int NB = (SLOTS + 255) / 256;       // loads per lane, all in flight at once
    constexpr int NV = (4 + FX - 1 + 3) / 4;      // float4 reads per row segment
    static_assert(4 * 7 + 4 * NV <= LS, "row segment reads stay in the row");
    __shared__ __attribute__((aligned(16))) float tile[LZ * LY * LS];

    const dim_t blk = xcd_remap(blockIdx.x, p.nblk);
    const dim_t bx = blk % p.nbx;
    dim_t t = blk / p.nbx;
    const dim_t by = t % p.nby;
    t /= p.nby;
    const dim_t bz = t % p.nbz;
    const dim_t b3 = t / p.nbz;
    in += b3 * p.in_step3;
    out += b3 * p.out_step3;
    filt += b3 * p.filt_step3;

    float wf[FZ][FY][FX];
#pragma unroll
    for (int kz = 0; kz < FZ; ++kz)
#pragma unroll
        for (int ky = 0; ky < FY; ++ky)
#pragma unroll
            for (int kx = 0; kx < FX; ++kx)
                wf[kz][ky][kx] = filt[((FZ - 1 - kz) * FY + (FY - 1 - ky)) * FX + (FX - 1 - kx)];

    const dim_t ox0 = bx * T3_BX, oy0 = by * T3_BY, oz0 = bz * T3_BZ;
    const dim_t ix0 = ox0 + p.sx - (FX - 1), iy0 = oy0 + p.sy - (FY - 1), iz0 = oz0 + p.sz - (FZ - 1);
    const dim_t a0 = ix0 - (((ix0 % 4) + 4) % 4);
    const int off = (int)(ix0 - a0);
    const bool vec = (p.S0 & 3) == 0;
    {
        float4 v[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int slot = u * 256 + (int)threadIdx.x;
            const int q = slot % NQ, r = (slot / NQ) % LY, z = slot / (NQ * LY);
            v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            const dim_t gx = a0 + 4 * q, gy = iy0 + r, gz = iz0 + z;
            if (slot < SLOTS && gy >= 0 && gy < p.S1 && gz >= 0 && gz < p.S2) {
                const float *row = in + (gz * p.S1 + gy) * p.S0;
                if (vec && gx >= 0 && gx + 3 < p.S0) {
                    v[u] = *reinterpret_cast<const float4 *>(row + gx);
                } else {
                    float e4[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) e4[e] = (gx + e >= 0 && gx + e < p.S0) ? row[gx + e] : 0.f;
                    v[u] = make_float4(e4[0], e4[1], e4[2], e4[3]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int slot = u * 256 + (int)threadIdx.x;
            const int q = slot % NQ, rz = slot / NQ;
            const int c0 = 4 * q - off;
            if (slot < SLOTS) {
                const float e4[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (c0 + e >= 0 && c0 + e < LS) tile[rz * LS + c0 + e] = c0 + e < LX ? e4[e] : 0.f;
            }
        }
    }
    __syncthreads();

    const int tx = threadIdx.x & 7, ty = (threadIdx.x >> 3) & 7, tz = threadIdx.x >> 6;  // 8 x 8 x 4
    float acc[2][4];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[k][i] = 0.f;
#pragma unroll
    for (int zz = 0; zz < 2 + FZ - 1; ++zz)
#pragma unroll
        for (int ky = 0; ky < FY; ++ky) {
            const float *row = tile + ((2 * tz + zz) * LY + ty + ky) * LS + 4 * tx;
            float w4[4 * NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const float4 f4 = lds_load16(row + 4 * q);
                w4[4 * q] = f4.x;
                w4[4 * q + 1] = f4.y;
                w4[4 * q + 2] = f4.z;
                w4[4 * q + 3] = f4.w;
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int kz = zz - k;
                if (kz >= 0 && kz < FZ) {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int kx = 0; kx < FX; ++kx) acc[k][i] = __builtin_fmaf(w4[i + kx], wf[kz][ky][kx], acc[k][i]);
                }
            }
        }

    const dim_t ox = ox0 + 4 * tx, oy = oy0 + ty;
    if (oy >= p.O1) return;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const dim_t oz = oz0 + 2 * tz + k;
        if (oz >= p.O2) continue;
        float *orow = out + (oz * p.O1 + oy) * p.O0;
        if ((p.O0 & 3) == 0 && ox + 4 <= p.O0) {
            *reinterpret_cast<float4 *>(orow + ox) = make_float4(acc[k][0], acc[k][1], acc[k][2], acc[k][3]);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (ox + i < p.O0) orow[ox + i] = acc[k][i];
        }
    }
}

static int launch_tile3(float *out, const float *in, const float *filt, const SpatialParams &p, hipStream_t s) {
    Tile3Params t;
    t.S0 = p.sd[0];
    t.S1 = p.sd[1];
    t.S2 = p.sd[2];
    t.O0 = p.od[0];
    t.O1 = p.od[1];
    t.O2 = p.od[2];
    t.in_step3 = p.in_step[3];
    t.out_step3 = p.out_step[3];
    t.filt_step3 = p.filt_step[3];
    t.nbx = (t.O0 + T3_BX - 1) / T3_BX;
    t.nby = (t.O1 + T3_BY - 1) / T3_BY;
    t.nbz = (t.O2 + T3_BZ - 1) / T3_BZ;
    t.nblk = t.nbx * t.nby * t.nbz * p.batch[3];
    t.sx = p.expand ? 0 : (int)(p.fd[0] / 2);
    t.sy = p.expand ? 0 : (int)(p.fd[1] / 2);
    t.sz = p.expand ? 0 : (int)(p.fd[2] / 2);
    if (t.nblk == 0) return K_OK;
    if (t.nblk > 0x7fffffffLL) return K_ERR_SIZE;
    if (p.fd[0] == 5)
        conv3_tile<5, 5, 5><<<dim3((unsigned)t.nblk), dim3(256), 0, s>>>(out, in, filt, t);
    else
        conv3_tile<3, 3, 3><<<dim3((unsigned)t.nblk), dim3(256), 0, s>>>(out, in, filt, t);
    return launch_status();
}

// ---------------------------------------------------------------- FP32 1-D
// y[o] = sum_w x[o + st - w] * f[w] for one slice: the workgroup's 1024
// outputs need x[o0 + st - (F-1) .. o0 + st + 1023], staged in LDS with zeros
// outside the signal; lane l computes outputs 4l..4l+3 with w ascending,
// sliding a 4-value register window one LDS read per tap. Each product is
// rounded before it is added, in the reference's order, so the result is the
// reference's bit for bit (af_iir feeds it to a recurrence that amplifies any
// rounding difference).
constexpr int T1_OUT = 1024;
constexpr int T1_MAXF = 2048;

struct Tile1Params {
    dim_t S0, O0, F;                    // signal, output and filter lengths of one slice
    dim_t in_step[4], out_step[4], filt_step[4], batch[4];
    dim_t ntiles, nblk;                 // tiles per slice, workgroups
    int st;                             // centring offset (0 for EXPAND)
};

__global__ __launch_bounds__(256) void conv1_tile(float *__restrict__ out, const float *__restrict__ in,
                                                  const float *__restrict__ filt, Tile1Params p) {
    __shared__ float xs[T1_OUT + T1_MAXF];
    __shared__ float wt[T1_MAXF];
    const dim_t blk = blockIdx.x;
    const dim_t tile = blk % p.ntiles;
    dim_t sl = blk / p.ntiles;
    const dim_t b1 = sl % p.batch[1];
    sl /= p.batch[1];
    const dim_t b2 = sl % p.batch[2], b3 = sl / p.batch[2];
    in += b1 * p.in_step[1] + b2 * p.in_step[2] + b3 * p.in_step[3];
    out += b1 * p.out_step[1] + b2 * p.out_step[2] + b3 * p.out_step[3];
    filt += b1 * p.filt_step[1] + b2 * p.filt_step[2] + b3 * p.filt_step[3];
    const int F = (int)p.F;
    const dim_t o0 = tile * T1_OUT;
    const dim_t base = o0 + p.st - (F - 1);
    const int nx = T1_OUT + F - 1;
    for (int k = threadIdx.x; k < nx; k += 256) {
        const dim_t g = base + k;
        xs[k] = (g >= 0 && g < p.S0) ? in[g] : 0.f;
    }
    for (int k = threadIdx.x; k < F; k += 256) wt[k] = filt[k];
    __syncthreads();
    const int l4 = 4 * (int)threadIdx.x;
    // window v[k] = xs[l4 + k + F - 1 - w] for the current tap w
    float v0 = xs[l4 + F - 1], v1 = xs[l4 + F], v2 = xs[l4 + F + 1], v3 = xs[l4 + F + 2];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    for (int w = 0; w < F; ++w) {
#pragma clang fp contract(off)
        const float f = wt[w];
        a0 = a0 + v0 * f;
        a1 = a1 + v1 * f;
        a2 = a2 + v2 * f;
        a3 = a3 + v3 * f;
        v3 = v2;
        v2 = v1;
        v1 = v0;
        if (w + 1 < F) v0 = xs[l4 + F - 2 - w];
    }
    const dim_t o = o0 + l4;
    if (o + 3 < p.O0 && ((reinterpret_cast<uintptr_t>(out + o) & 15) == 0)) {
        *reinterpret_cast<float4 *>(out + o) = make_float4(a0, a1, a2, a3);
    } else {
        const float a[4] = {a0, a1, a2, a3};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (o + k < p.O0) out[o + k] = a[k];
    }
}

static int launch_tile1(float *out, const float *in, const float *filt, const SpatialParams &p, hipStream_t s) {
    Tile1Params t;
    t.S0 = p.sd[0];
    t.O0 = p.od[0];
    t.F = p.fd[0];
    for (int d = 0; d < 4; ++d) {
        t.in_step[d] = p.in_step[d];
        t.out_step[d] = p.out_step[d];
        t.filt_step[d] = p.filt_step[d];
        t.batch[d] = p.batch[d];
    }
    t.st = p.expand ? 0 : (int)(p.fd[0] / 2);
    t.ntiles = (t.O0 + T1_OUT - 1) / T1_OUT;
    t.nblk = t.ntiles * p.batch[1] * p.batch[2] * p.batch[3];
    if (t.nblk == 0) return K_OK;
    if (t.nblk > 0x7fffffffLL) return K_ERR_SIZE;
    conv1_tile<<<dim3((unsigned)t.nblk), dim3(256), 0, s>>>(out, in, filt, t);
    return launch_status();
}

}  // namespace afhip

using namespace afhip;

extern "C" int afhip_k_convolve(int dtype, void *out, const long long *odims,
                                const void *in, const long long *sdims,
                                const void *filt, const long long *fdims,
                                int kind, int rank, int expand, void *stream) {
    if (rank < 1 || rank > 3 || kind < 1 || kind > 4) return K_ERR_ARG;
    SpatialParams p;
    fill_params(p, odims, sdims, fdims, kind, rank, expand);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == DT_F32 && rank == 2 && p.batch[1] == 1) {
        const int sx = expand ? 0 : (int)(fdims[0] / 2);
        Tile2Fn strip = strip2_ok(p, in) ? strip2_lookup(fdims[0], fdims[1], sx) : nullptr;
        Tile2Fn fn = tile2_lookup(fdims[0], fdims[1]);
        if (fn || strip) return launch_tile2((float *)out, (const float *)in, (const float *)filt, p, fn, strip, s);
    }
    if (dtype == DT_F32 && rank == 3 && p.batch[1] == 1 && p.batch[2] == 1 && fdims[0] == fdims[1] &&
        fdims[1] == fdims[2] && (fdims[0] == 3 || fdims[0] == 5))
        return launch_tile3((float *)out, (const float *)in, (const float *)filt, p, s);
    // AFHIP_CONV1_TILE=0 in the environment keeps FP32 1-D on conv_generic
    static const bool tile1_off = [] {
        const char *e = getenv("AFHIP_CONV1_TILE");
        return e && e[0] == '0';
    }();
    if (dtype == DT_F32 && rank == 1 && fdims[0] >= 1 && fdims[0] <= T1_MAXF && !tile1_off)
        return launch_tile1((float *)out, (const float *)in, (const float *)filt, p, s);
    switch (dtype) {
        case DT_F32: return launch_generic<float, float>(out, in, filt, p, s);
        case DT_F64: return launch_generic<double, double>(out, in, filt, p, s);
        case DT_C32: return launch_generic<cfloat, cfloat>(out, in, filt, p, s);
        case DT_C64: return launch_generic<cdouble, cdouble>(out, in, filt, p, s);
        case DT_S32: return launch_generic<int32_t, float>(out, in, filt, p, s);
        case DT_U32: return launch_generic<uint32_t, float>(out, in, filt, p, s);
        case DT_S16: return launch_generic<int16_t, float>(out, in, filt, p, s);
        case DT_U16: return launch_generic<uint16_t, float>(out, in, filt, p, s);
        case DT_S64: return launch_generic<long long, float>(out, in, filt, p, s);
        case DT_U64: return launch_generic<unsigned long long, float>(out, in, filt, p, s);
        case DT_U8: return launch_generic<uint8_t, float>(out, in, filt, p, s);
        case DT_S8: return launch_generic<int8_t, float>(out, in, filt, p, s);
        case DT_B8: return launch_generic<int8_t, float>(out, in, filt, p, s);
        default: return K_ERR_TYPE;
    }
}
