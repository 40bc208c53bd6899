// Spatial-domain convolution kernels (convolve1/2/3 with every batch kind).
//
// Semantics (reference CPU backend, src/backend/cpu/kernel/convolve.hpp):
//   y[o] = sum_f x[o + s - f] * w[f],  s = EXPAND ? 0 : floor(F/2),
// zero outside x, accumulated in AccT, stored as InT(acc); batch stepping as
// convolve_nd (:132-211). The host passes per-kind batch extents and steps.
//
// Two implementations:
//  * conv2_tile<FX,FY>: FP32 rank-2 kernel. A 128x64 output tile per
//    256-thread workgroup; the (128+FX-1)x(64+FY-1) input halo tile is staged
//    in LDS with zero borders so the tap loops are branch-free; each lane keeps
//    a 4x8 block of outputs in registers, reads each LDS row segment once
//    (ds_read_b128) and applies all the taps that touch it; the flipped filter
//    lives in SGPRs (uniform scalar loads).
//  * conv_generic: any rank, dtype, filter size and batch kind; one lane per
//    output, grid-stride loop. Used where no tiled variant applies.
#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

struct SpatialParams {
    dim_t od[4], sd[4], fd[4];
    dim_t os[4], ss[4], fs[4];
    dim_t out_step[4], in_step[4], filt_step[4], batch[4];
    int rank, expand;
};

template <typename InT, typename AccT>
__device__ inline AccT mul_acc(InT s, AccT f) {
    return AccT(s * f);
}
template <>
__device__ inline float mul_acc<int8_t, float>(int8_t s, float f) {
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

static int launch_tile2(float *out, const float *in, const float *filt,
                        const SpatialParams &p, Tile2Fn fn, hipStream_t s) {
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
    t.nbx = (t.O0 + T2_BX - 1) / T2_BX;
    t.nby = (t.O1 + T2_BY - 1) / T2_BY;
    const dim_t nb = p.batch[2] * p.batch[3];
    t.nblk = t.nbx * t.nby * nb;
    t.sx = p.expand ? 0 : (int)(p.fd[0] / 2);
    t.sy = p.expand ? 0 : (int)(p.fd[1] / 2);
    if (t.nblk == 0) return K_OK;
    if (t.nblk > 0x7fffffffLL) return K_ERR_SIZE;
    fn<<<dim3((unsigned)t.nblk), dim3(256), 0, s>>>(out, in, filt, t);
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
        if (Tile2Fn fn = tile2_lookup(fdims[0], fdims[1]))
            return launch_tile2((float *)out, (const float *)in, (const float *)filt, p, fn, s);
    }
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
