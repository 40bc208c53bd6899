// convolve2_nn forward, data gradient and filter gradient as implicit GEMMs
// on the gfx950 matrix cores.
//
// Reference semantics (src/backend/cpu/convolve.cpp:123-243; closed forms):
//   fwd   y[ow,oh,k,n]  = sum_{c,i,j} x[ow*s0-p0+i*d0, oh*s1-p1+j*d1, c, n]
//                                     * w[fw-1-i, fh-1-j, c, k]
//   dgrad dx[x,y,c,n]   = sum over (k,ow,oh,i,j) hitting (x,y) of dy*w(flipped)
//   wgrad dw[fw-1-i, fh-1-j, c, k] = sum_{n,ow,oh} x[...] * dy[ow,oh,k,n]
// ArrayFire dims [W,H,C,N] are row-major NCHW; [fw,fh,C,K] is KCRS.
//
// GEMM views (C[M x Ncol] += A[M x R] * B[R x Ncol]):
//   fwd   M = K,  Ncol = N*OH*OW, R = (tap, C)   A = w,  B = im2col(x)
//   dgrad M = C,  Ncol = N*H*W,   R = (tap, K)   A = w,  B = col(dy) gathered
//                                                through the stride phases
//   wgrad M = K,  Ncol = C*fh*fw, R = N*OH*OW    A = dy, B = im2col(x)^T.
// Layers with too few output tiles to fill the chip split R across
// workgroups (FP32 partial tiles, then a second pass that sums them in split
// order: deterministic). A strided data gradient orders its columns by stride
// phase (x mod s0, y mod s1), each phase padded to whole tiles, and visits
// only the taps that reach that phase: about 1/(s0*s1) of the slices.
// Nothing is materialised: each 32-deep slice of R is gathered from HBM into
// registers, converted to the operand format and stored to LDS with R
// contiguous per row, where the MFMA fragments read it with ds_read_b128.
//
// Workgroup: 256 lanes (4 waves as 2x2), output tile 128 x 128, each wave a
// 64 x 64 block of 4x4 MFMA 16x16 tiles. Operand formats (MODE):
//   F32   v_mfma_f32_16x16x4_f32, FP32 operands
//   BF16  v_mfma_f32_16x16x32_bf16, operands rounded to bf16 (RNE)
//   F16   v_mfma_f32_16x16x32_f16
//   SPLIT hi = bf16(v), lo = bf16(v - hi); hi*hi' + hi*lo' + lo*hi' on BF16
// Accumulation is FP32 in every mode. One LDS buffer; the next slice's global
// loads are in flight in registers while the current slice is multiplied.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8_t __attribute__((ext_vector_type(8)));

enum Mode : int { M_F32 = 0, M_BF16 = 1, M_F16 = 2, M_SPLIT = 3 };
enum Pass : int { P_FWD = 0, P_DGRAD = 1, P_WGRAD = 2 };

// Strided data gradients order their columns by stride phase when there are
// at most this many phases (s0 * s1).
constexpr int kMaxPhases = 16;

struct NNParams {
    // x [W,H,C,N], w [fw,fh,C,K], y [OW,OH,K,N]
    int W, H, C, N, fw, fh, K, OW, OH;
    int s0, s1, p0, p1, d0, d1;
    int RS;              // fw*fh
    long long M, Ncol;   // GEMM extents
    int R;               // reduction extent (wgrad) -- informational
    int kchunks;         // 32-deep slices per tap (fwd: ceil(C/32), dgrad: ceil(K/32))
    long long ksteps;    // total slices (phased data gradient: the most of any phase)
    long long ksteps_per_split;
    int nsplit;          // > 1: FP32 partial tiles per split, summed by nn_split_reduce / wgrad_reduce
    long long tilesM, tilesN;
    // Phased data gradient (stride > 1): the input pixels (x, y) with
    // x = s0*mx + rx, y = s1*my + ry form phase ph = ry*s0 + rx; each phase's
    // N*My*Mx columns start at column tile ph_tile0[ph] and only the taps
    // (i, j) with (rx + p0 - i*d0) and (ry + p1 - j*d1) divisible by the
    // strides meet them, so only those taps are visited.
    int phased;
    int nph;
    int ph_tile0[kMaxPhases + 1];
    // Channel-last operands (forward and data gradient, operands stored in
    // the MMA type): packed weights Wp [RS][M][Rcp] and the activation (x,
    // or dy for the data gradient) as [N][rows][cols][Rcp], Rcp = the
    // reduction channels rounded up to the 32-deep slice, zero past them.
    int cl;
    int Rcp;
    int wp_bytes, act_bytes;  // sizes of the channel-last copies (each below 2^31 - 64)
    int nsplit_req;           // > 0: the caller's split count (forward / data gradient)
};

// Phase extents: pixels of residue r along an axis of length L and stride s,
// and the taps of a filter (length f, dilation d, padding pd) that reach them.
__host__ __device__ inline int phase_len(int L, int s, int r) { return r < L ? (L - r + s - 1) / s : 0; }
__host__ __device__ inline bool phase_tap(int r, int pd, int i, int d, int s) {
    const int v = (r + pd - i * d) % s;
    return v == 0;
}
__host__ __device__ inline int phase_taps(int f, int r, int pd, int d, int s) {
    int n = 0;
    for (int i = 0; i < f; ++i) n += phase_tap(r, pd, i, d, s) ? 1 : 0;
    return n;
}
// The u-th tap (in increasing order) that reaches residue r.
__host__ __device__ inline int phase_tap_at(int f, int r, int pd, int d, int s, int u) {
    for (int i = 0; i < f; ++i)
        if (phase_tap(r, pd, i, d, s)) {
            if (u == 0) return i;
            --u;
        }
    return 0;
}

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int NT = 256;

// ------------------------------------------------------------- element IO
struct bf16s {
    uint16_t b;
};
template <typename T>
__device__ inline float ld_f(const T *p, long long i);
template <>
__device__ inline float ld_f<float>(const float *p, long long i) {
    return p[i];
}
template <>
__device__ inline float ld_f<bf16s>(const bf16s *p, long long i) {
    return bf16_bits_to_f32(p[i].b);
}
template <>
__device__ inline float ld_f<_Float16>(const _Float16 *p, long long i) {
    return (float)p[i];
}

template <typename T>
__device__ inline void st_f(T *p, long long i, float v);
template <>
__device__ inline void st_f<float>(float *p, long long i, float v) {
    p[i] = v;
}
template <>
__device__ inline void st_f<bf16s>(bf16s *p, long long i, float v) {
    p[i].b = f32_to_bf16_bits(v);
}
template <>
__device__ inline void st_f<_Float16>(_Float16 *p, long long i, float v) {
    p[i] = (_Float16)v;
}

__device__ inline uint16_t f32_to_f16_bits(float v) {
    _Float16 h = (_Float16)v;
    return __builtin_bit_cast(uint16_t, h);
}

// ------------------------------------------------------------- LDS layout
// Each operand tile is 128 rows x 32 R-values. 16-bit formats use a row pitch
// of 40 elements (80 B), FP32 a pitch of 36 (144 B); the pad staggers rows
// across LDS banks for the 16-byte fragment reads.
template <int MODE>
struct Lds {
    static constexpr int PITCH = MODE == M_F32 ? 36 : 40;
    static constexpr int ESZ = MODE == M_F32 ? 4 : 2;
    static constexpr int PARTS = MODE == M_SPLIT ? 2 : 1;  // hi, lo
    static constexpr int TILE_BYTES = 128 * PITCH * ESZ;
    static constexpr int BYTES = 2 * PARTS * TILE_BYTES;    // A and B
};

// Writes 16 values (one row, R offset kk0) of an operand tile.
template <int MODE>
__device__ inline void store_row16(char *tile, int row, int kk0, const float (&v)[16]) {
    constexpr int P = Lds<MODE>::PITCH;
    if constexpr (MODE == M_F32) {
        float *t = reinterpret_cast<float *>(tile) + row * P + kk0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            *reinterpret_cast<float4 *>(t + 4 * q) =
                make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else if constexpr (MODE == M_BF16 || MODE == M_F16) {
        uint16_t *t = reinterpret_cast<uint16_t *>(tile) + row * P + kk0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            u16x8_t h;
#pragma unroll
            for (int e = 0; e < 8; ++e)
                h[e] = MODE == M_BF16 ? f32_to_bf16_bits(v[8 * q + e]) : f32_to_f16_bits(v[8 * q + e]);
            *reinterpret_cast<u16x8_t *>(t + 8 * q) = h;
        }
    } else {  // SPLIT: hi tile then lo tile
        uint16_t *th = reinterpret_cast<uint16_t *>(tile) + row * P + kk0;
        uint16_t *tl = reinterpret_cast<uint16_t *>(tile + Lds<MODE>::TILE_BYTES) + row * P + kk0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            u16x8_t hh, ll;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float x = v[8 * q + e];
                const uint16_t hb = f32_to_bf16_bits(x);
                hh[e] = hb;
                ll[e] = f32_to_bf16_bits(x - bf16_bits_to_f32(hb));
            }
            *reinterpret_cast<u16x8_t *>(th + 8 * q) = hh;
            *reinterpret_cast<u16x8_t *>(tl + 8 * q) = ll;
        }
    }
}

// One 32-deep slice: acc[4][4] += A(wave rows) * B(wave cols).
template <int MODE>
__device__ inline void mma_slice(const char *tA, const char *tB, int wm, int wn, int lane,
                                 f32x4_t (&acc)[4][4]) {
    constexpr int P = Lds<MODE>::PITCH;
    const int r16 = lane & 15, g = lane >> 4;
    if constexpr (MODE == M_F32) {
        const float *a = reinterpret_cast<const float *>(tA);
        const float *b = reinterpret_cast<const float *>(tB);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float4 af[4], bfr[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = *reinterpret_cast<const float4 *>(a + (wm * 64 + i * 16 + r16) * P + h * 16 + 4 * g);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                bfr[j] = *reinterpret_cast<const float4 *>(b + (wn * 64 + j * 16 + r16) * P + h * 16 + 4 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].x, bfr[j].x, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].y, bfr[j].y, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].z, bfr[j].z, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i].w, bfr[j].w, acc[i][j], 0, 0, 0);
                }
        }
    } else {
        const uint16_t *a = reinterpret_cast<const uint16_t *>(tA);
        const uint16_t *b = reinterpret_cast<const uint16_t *>(tB);
        u16x8_t af[4], bfr[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            af[i] = *reinterpret_cast<const u16x8_t *>(a + (wm * 64 + i * 16 + r16) * P + 8 * g);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            bfr[j] = *reinterpret_cast<const u16x8_t *>(b + (wn * 64 + j * 16 + r16) * P + 8 * g);
        if constexpr (MODE == M_SPLIT) {
            constexpr int TB = Lds<MODE>::TILE_BYTES / 2;  // elements per tile
            u16x8_t al[4], bl[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                al[i] = *reinterpret_cast<const u16x8_t *>(a + TB + (wm * 64 + i * 16 + r16) * P + 8 * g);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                bl[j] = *reinterpret_cast<const u16x8_t *>(b + TB + (wn * 64 + j * 16 + r16) * P + 8 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const bf16x8_t ah = __builtin_bit_cast(bf16x8_t, af[i]);
                    const bf16x8_t bh = __builtin_bit_cast(bf16x8_t, bfr[j]);
                    const bf16x8_t alo = __builtin_bit_cast(bf16x8_t, al[i]);
                    const bf16x8_t blo = __builtin_bit_cast(bf16x8_t, bl[j]);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bh, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, blo, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[i][j], 0, 0, 0);
                }
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if constexpr (MODE == M_BF16) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16x8_t, af[i]), __builtin_bit_cast(bf16x8_t, bfr[j]),
                            acc[i][j], 0, 0, 0);
                    } else {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                            __builtin_bit_cast(f16x8_t, af[i]), __builtin_bit_cast(f16x8_t, bfr[j]),
                            acc[i][j], 0, 0, 0);
                    }
                }
        }
    }
}

// ------------------------------------------------------------- gathers
// Each lane owns one row of the A tile and one row of the B tile and fills
// 16 consecutive R positions of each (kk0 = 0 or 16).

// fwd / wgrad helper: output pixel q -> (n, oh, ow)
struct Pix {
    int n, a, b;  // image, row, column
};
__device__ inline Pix decode(long long q, int rows, int cols) {
    Pix p;
    const long long per = (long long)rows * cols;
    p.n = (int)(q / per);
    const int r = (int)(q - (long long)p.n * per);
    p.a = r / cols;
    p.b = r - p.a * cols;
    return p;
}

template <typename TIn, int PASS>
struct Gather;

// ---- forward: A = w (row k), B = im2col(x) (row = output pixel)
template <typename TIn>
struct Gather<TIn, P_FWD> {
    const TIn *x, *w;
    const NNParams &p;
    long long arow, brow;  // global row indices
    bool bvalid;
    Pix bp;
    __device__ Gather(const TIn *x_, const TIn *w_, const NNParams &p_, long long ar, long long br, int)
        : x(x_), w(w_), p(p_), arow(ar), brow(br) {
        bvalid = br < p.Ncol;
        bp = decode(bvalid ? br : 0, p.OH, p.OW);
    }
    // slice -> (tap, channel block)
    __device__ void load(long long ks, int kk0, float (&va)[16], float (&vb)[16]) const {
        const int t = (int)(ks / p.kchunks);
        const int c0 = (int)(ks - (long long)t * p.kchunks) * BK + kk0;
        const int i = t % p.fw, j = t / p.fw;
        // A: w[k][c][fh-1-j][fw-1-i]
        const bool ak = arow < p.K;
        const long long abase = (arow * p.C) * p.RS + (p.RS - 1 - t);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int c = c0 + e;
            va[e] = (ak && c < p.C) ? ld_f(w, abase + (long long)c * p.RS) : 0.f;
        }
        // B: x[n][c][oh*s1-p1+j*d1][ow*s0-p0+i*d0]
        const int ih = bp.a * p.s1 - p.p1 + j * p.d1;
        const int iw = bp.b * p.s0 - p.p0 + i * p.d0;
        const bool ok = bvalid && ih >= 0 && ih < p.H && iw >= 0 && iw < p.W;
        const long long hw = (long long)p.H * p.W;
        const long long bbase = ((long long)bp.n * p.C) * hw + (long long)ih * p.W + iw;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int c = c0 + e;
            vb[e] = (ok && c < p.C) ? ld_f(x, bbase + (long long)c * hw) : 0.f;
        }
    }
};

// ---- data gradient: A = w (row c), B = dy gathered (row = input pixel)
template <typename TIn>
struct Gather<TIn, P_DGRAD> {
    const TIn *dy, *w;
    const NNParams &p;
    long long arow;
    bool bvalid;
    int bn, by, bx;     // this lane's B row: image and input pixel
    int rx, ry, Ti;     // phased: the tile's phase and its tap count along W
    __device__ Gather(const TIn *dy_, const TIn *w_, const NNParams &p_, long long ar, long long br, int ph)
        : dy(dy_), w(w_), p(p_), arow(ar) {
        if (p.phased) {
            rx = ph % p.s0;
            ry = ph / p.s0;
            const int Mx = phase_len(p.W, p.s0, rx), My = phase_len(p.H, p.s1, ry);
            bvalid = br < (long long)p.N * My * Mx;
            const Pix q = decode(bvalid ? br : 0, My, Mx);
            bn = q.n;
            by = p.s1 * q.a + ry;
            bx = p.s0 * q.b + rx;
            Ti = phase_taps(p.fw, rx, p.p0, p.d0, p.s0);
        } else {
            rx = ry = 0;
            Ti = p.fw;
            bvalid = br < p.Ncol;
            const Pix q = decode(bvalid ? br : 0, p.H, p.W);
            bn = q.n;
            by = q.a;
            bx = q.b;
        }
    }
    __device__ void load(long long ks, int kk0, float (&va)[16], float (&vb)[16]) const {
        const int u = (int)(ks / p.kchunks);
        const int k0 = (int)(ks - (long long)u * p.kchunks) * BK + kk0;
        int i, j;
        if (p.phased) {
            const int ui = u % Ti;
            i = phase_tap_at(p.fw, rx, p.p0, p.d0, p.s0, ui);
            j = phase_tap_at(p.fh, ry, p.p1, p.d1, p.s1, u / Ti);
        } else {
            i = u % p.fw;
            j = u / p.fw;
        }
        const int t = j * p.fw + i;
        // A: w[k][c][fh-1-j][fw-1-i], row = c, R = k
        const bool ac = arow < p.C;
        const long long crs = (long long)p.C * p.RS;
        const long long abase = arow * p.RS + (p.RS - 1 - t);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int k = k0 + e;
            va[e] = (ac && k < p.K) ? ld_f(w, abase + (long long)k * crs) : 0.f;
        }
        // B: dy[n][k][oh][ow] where oh*s1 - p1 + j*d1 = y, ow*s0 - p0 + i*d0 = x
        const int nh = by + p.p1 - j * p.d1;
        const int nw = bx + p.p0 - i * p.d0;
        bool ok = bvalid && nh >= 0 && nw >= 0 && (nh % p.s1) == 0 && (nw % p.s0) == 0;
        const int oh = ok ? nh / p.s1 : 0, ow = ok ? nw / p.s0 : 0;
        ok = ok && oh < p.OH && ow < p.OW;
        const long long ohw = (long long)p.OH * p.OW;
        const long long bbase = ((long long)bn * p.K) * ohw + (long long)oh * p.OW + ow;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int k = k0 + e;
            vb[e] = (ok && k < p.K) ? ld_f(dy, bbase + (long long)k * ohw) : 0.f;
        }
    }
};

// ---- filter gradient: A = dy (row k, R = output pixel), B = x (row = (c,tap))
template <typename TIn>
struct Gather<TIn, P_WGRAD> {
    const TIn *x, *dy;
    const NNParams &p;
    long long arow, brow;
    int bc, bi, bj;  // B row -> channel and tap (i along W, j along H)
    bool bvalid;
    __device__ Gather(const TIn *x_, const TIn *dy_, const NNParams &p_, long long ar, long long br, int)
        : x(x_), dy(dy_), p(p_), arow(ar), brow(br) {
        bvalid = br < p.Ncol;
        const int r = bvalid ? (int)br : 0;
        bc = r / p.RS;
        const int t = r - bc * p.RS;
        bi = t % p.fw;
        bj = t / p.fw;
    }
    __device__ void load(long long ks, int kk0, float (&va)[16], float (&vb)[16]) const {
        const long long q0 = ks * BK + kk0;
        const long long Rtot = (long long)p.N * p.OH * p.OW;
        Pix pp = decode(q0 < Rtot ? q0 : 0, p.OH, p.OW);
        const long long ohw = (long long)p.OH * p.OW, hw = (long long)p.H * p.W;
        const bool ak = arow < p.K;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const bool inr = q0 + e < Rtot;
            // A: dy[n][k][oh][ow]
            va[e] = (ak && inr) ? ld_f(dy, ((long long)pp.n * p.K + arow) * ohw + (long long)pp.a * p.OW + pp.b)
                                : 0.f;
            // B: x[n][c][oh*s1-p1+j*d1][ow*s0-p0+i*d0]
            const int ih = pp.a * p.s1 - p.p1 + bj * p.d1;
            const int iw = pp.b * p.s0 - p.p0 + bi * p.d0;
            const bool ok = bvalid && inr && ih >= 0 && ih < p.H && iw >= 0 && iw < p.W;
            vb[e] = ok ? ld_f(x, ((long long)pp.n * p.C + bc) * hw + (long long)ih * p.W + iw) : 0.f;
            // next pixel
            if (++pp.b == p.OW) {
                pp.b = 0;
                if (++pp.a == p.OH) {
                    pp.a = 0;
                    ++pp.n;
                }
            }
        }
    }
};

// Slices of the channel-last path in flight per lane (register sets).
constexpr int CL_DEPTH = 4;

// ---- channel-last gathers: each lane's 16 R values are one contiguous run
// of 16 channels (2 or 4 aligned 16-byte loads) for A and for B.
template <typename T, int PASS>
struct GatherCL {
    static constexpr int Q = (int)sizeof(T);  // 16-byte vectors per 16 values
    const T *act, *wp;
    const NNParams &p;
    long long arow;
    bool bvalid;
    int bn, ba, bb;   // image; output row, column (forward) or input y, x (data gradient)
    int rx, ry, Ti;   // phased data gradient: the tile's phase, its tap count along W
    __device__ GatherCL(const T *act_, const T *wp_, const NNParams &p_, long long ar, long long br, int ph)
        : act(act_), wp(wp_), p(p_), arow(ar) {
        rx = ry = 0;
        Ti = p.fw;
        if (PASS == P_DGRAD && p.phased) {
            rx = ph % p.s0;
            ry = ph / p.s0;
            const int Mx = phase_len(p.W, p.s0, rx), My = phase_len(p.H, p.s1, ry);
            bvalid = br < (long long)p.N * My * Mx;
            const Pix q = decode(bvalid ? br : 0, My, Mx);
            bn = q.n;
            ba = p.s1 * q.a + ry;
            bb = p.s0 * q.b + rx;
            Ti = phase_taps(p.fw, rx, p.p0, p.d0, p.s0);
        } else {
            bvalid = br < p.Ncol;
            const Pix q = PASS == P_FWD ? decode(bvalid ? br : 0, p.OH, p.OW) : decode(bvalid ? br : 0, p.H, p.W);
            bn = q.n;
            ba = q.a;
            bb = q.b;
        }
    }
    // Buffer loads: an operand outside the image (or a row past M) reads
    // zeros from an offset past the resource's end, so every lane issues the
    // same loads and the prefetch ring's waits count them exactly.
    __device__ void load(long long ks, int kk0, uint4 (&ra)[Q], uint4 (&rb)[Q]) const {
        constexpr int kOOB = 0x7fffffc0;
        const int u = (int)(ks / p.kchunks);
        const int c0 = (int)(ks - (long long)u * p.kchunks) * BK + kk0;
        int i, j;
        if (PASS == P_DGRAD && p.phased) {
            i = phase_tap_at(p.fw, rx, p.p0, p.d0, p.s0, u % Ti);
            j = phase_tap_at(p.fh, ry, p.p1, p.d1, p.s1, u / Ti);
        } else {
            i = u % p.fw;
            j = u / p.fw;
        }
        const int t = j * p.fw + i;
        const auto rw = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(wp), (short)0, p.wp_bytes, 0x00020000);
        const auto ra_ = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(act), (short)0, p.act_bytes, 0x00020000);
        const int offa = arow < p.M ? (int)((((long long)t * p.M + arow) * p.Rcp + c0) * (long long)sizeof(T)) : kOOB;
        bool ok;
        long long off;
        if constexpr (PASS == P_FWD) {
            const int ih = ba * p.s1 - p.p1 + j * p.d1;
            const int iw = bb * p.s0 - p.p0 + i * p.d0;
            ok = bvalid && ih >= 0 && ih < p.H && iw >= 0 && iw < p.W;
            off = (((long long)bn * p.H + ih) * p.W + iw) * p.Rcp + c0;
        } else {
            const int nh = ba + p.p1 - j * p.d1;
            const int nw = bb + p.p0 - i * p.d0;
            ok = bvalid && nh >= 0 && nw >= 0 && (nh % p.s1) == 0 && (nw % p.s0) == 0;
            const int oh = ok ? nh / p.s1 : 0, ow = ok ? nw / p.s0 : 0;
            ok = ok && oh < p.OH && ow < p.OW;
            off = (((long long)bn * p.OH + oh) * p.OW + ow) * p.Rcp + c0;
        }
        const int offb = ok ? (int)(off * (long long)sizeof(T)) : kOOB;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            ra[q] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rw, offa + 16 * q, 0, 0));
            rb[q] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(ra_, offb + 16 * q, 0, 0));
        }
    }
};

// Writes one lane's 16 raw values (row, R offset kk0) of an operand tile.
template <int MODE>
__device__ inline void store_raw16(char *tile, int row, int kk0, const uint4 *v) {
    constexpr int P = Lds<MODE>::PITCH, E = Lds<MODE>::ESZ;
    uint4 *t = reinterpret_cast<uint4 *>(tile + (row * P + kk0) * E);
#pragma unroll
    for (int q = 0; q < E; ++q) t[q] = v[q];
}

// Wp [RS][M][Rcp]: forward M = K, R = C, Wp[t][k][c] = w[k][c][fh-1-j][fw-1-i];
// data gradient M = C, R = K, Wp[t][c][k] = the same element (t = j*fw + i).
// Both channel-last copies in one launch: the first nbw workgroups pack the
// weights into Wp (above), the others transpose the activation:
// dst [N][rows][cols][Cp] <- src [N][Cd][rows][cols] (NCHW to channel-last),
// zero for the channels past Cd, one lane writing 16 bytes of channels.
template <typename T>
__global__ __launch_bounds__(256) void nn_cl_copies(T *__restrict__ wp, const T *__restrict__ w, NNParams p,
                                                    int dgrad, int nbw, T *__restrict__ dst,
                                                    const T *__restrict__ src, int cols, int rows, int Cd) {
    if ((int)blockIdx.x < nbw) {
        const int Rc = dgrad ? p.K : p.C;
        const long long total = (long long)p.RS * p.M * p.Rcp;
        for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
             e += (long long)nbw * blockDim.x) {
            const int r = (int)(e % p.Rcp);
            const long long mt = e / p.Rcp;
            const int m = (int)(mt % p.M);
            const int t = (int)(mt / p.M);
            const int i = t % p.fw, j = t / p.fw;
            T v{};
            if (r < Rc) {
                const int k = dgrad ? r : m, c = dgrad ? m : r;
                v = w[(((long long)k * p.C + c) * p.fh + (p.fh - 1 - j)) * p.fw + (p.fw - 1 - i)];
            }
            wp[e] = v;
        }
        return;
    }
    constexpr int G = 16 / (int)sizeof(T);
    const int Cp = p.Rcp, N = p.N;
    const int cg = Cp / G;
    const long long hw = (long long)rows * cols;
    const long long total = (long long)N * hw * cg;
    const long long nbt = (long long)gridDim.x - nbw;
    for (long long e = (blockIdx.x - nbw) * (long long)blockDim.x + threadIdx.x; e < total;
         e += nbt * blockDim.x) {
        const int g = (int)(e % cg);
        const long long pix = e / cg;  // n * hw + y * cols + x
        const long long n = pix / hw, yx = pix - n * hw;
        T v[G];
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const int c = g * G + q;
            v[q] = c < Cd ? src[(n * Cd + c) * hw + yx] : T{};
        }
        uint4 o;
        __builtin_memcpy(&o, v, 16);
        *reinterpret_cast<uint4 *>(dst + pix * Cp + g * G) = o;
    }
}

// ------------------------------------------------------------- kernel
template <int MODE, typename TIn, typename TOut, int PASS, bool CL = false>
__global__ __launch_bounds__(NT) void nn_gemm(TOut *__restrict__ out, const TIn *__restrict__ in0,
                                              const TIn *__restrict__ in1, float *__restrict__ partial,
                                              NNParams p) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    char *tA = smem;
    char *tB = smem + Lds<MODE>::BYTES / 2;

    // Tile order: along Ncol fastest so concurrent workgroups share A rows.
    const long long tiles = p.tilesM * p.tilesN;
    const long long bid = blockIdx.x;
    const long long split = bid / tiles;
    const long long tl = bid - split * tiles;
    const long long tm = tl / p.tilesN, tn = tl - tm * p.tilesN;
    const long long m0 = tm * BM;
    // phased data gradient: the tile's phase and its column offset in it
    int ph = 0;
    long long n0 = tn * BN;
    long long ksteps = p.ksteps, kps = p.ksteps_per_split;
    if (PASS == P_DGRAD && p.phased) {
        while (ph + 1 < p.nph && tn >= p.ph_tile0[ph + 1]) ++ph;
        n0 = (tn - p.ph_tile0[ph]) * BN;
        const int rx = ph % p.s0, ry = ph / p.s0;
        ksteps = (long long)phase_taps(p.fw, rx, p.p0, p.d0, p.s0) * phase_taps(p.fh, ry, p.p1, p.d1, p.s1) *
                 p.kchunks;
        kps = (ksteps + p.nsplit - 1) / p.nsplit;
    }

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lrow = tid & 127, kk0 = (tid >> 7) * 16;

    f32x4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};

    const long long ks0 = split * kps;
    long long ks1 = ks0 + kps;
    if (ks1 > ksteps) ks1 = ksteps;

    if constexpr (CL) {
        // in0 = the channel-last activation, in1 = the packed weights. A ring
        // of CL_DEPTH register sets keeps the loads of the next CL_DEPTH - 1
        // slices in flight while a slice is stored and multiplied.
        // Every lane issues the same loads in every slot (slices past the
        // range reload the last one), so the waits before each store count
        // exactly the CL_DEPTH - 1 slices still in flight.
        GatherCL<TIn, PASS> g(in0, in1, p, m0 + lrow, n0 + lrow, ph);
        constexpr int Q = GatherCL<TIn, PASS>::Q;
        if (ks0 < ks1) {
            uint4 ra[CL_DEPTH][Q], rb[CL_DEPTH][Q];
            const long long last = ks1 - 1;
#pragma unroll
            for (int d = 0; d < CL_DEPTH; ++d) g.load(min(ks0 + d, last), kk0, ra[d], rb[d]);
            for (long long ks = ks0; ks < ks1; ks += CL_DEPTH) {
#pragma unroll
                for (int d = 0; d < CL_DEPTH; ++d) {
                    const bool live = ks + d < ks1;
                    __syncthreads();
                    if (live) {
                        store_raw16<MODE>(tA, lrow, kk0, ra[d]);
                        store_raw16<MODE>(tB, lrow, kk0, rb[d]);
                    }
                    __syncthreads();
                    g.load(min(ks + d + CL_DEPTH, last), kk0, ra[d], rb[d]);
                    if (live) mma_slice<MODE>(tA, tB, wm, wn, lane, acc);
                }
            }
        }
    } else {
        Gather<TIn, PASS> g(in0, in1, p, m0 + lrow, n0 + lrow, ph);
        float va[16], vb[16];
        if (ks0 < ks1) g.load(ks0, kk0, va, vb);
        for (long long ks = ks0; ks < ks1; ++ks) {
            __syncthreads();
            store_row16<MODE>(tA, lrow, kk0, va);
            store_row16<MODE>(tB, lrow, kk0, vb);
            __syncthreads();
            if (ks + 1 < ks1) g.load(ks + 1, kk0, va, vb);
            mma_slice<MODE>(tA, tB, wm, wn, lane, acc);
        }
    }

    // Epilogue. acc[i][j][v] is C[row][col] with
    // row = m0 + wm*64 + i*16 + 4*(lane/16) + v, col = n0 + wn*64 + j*16 + lane%16.
    // Split reductions store FP32 partial tiles (row-major over the padded
    // columns; nn_split_reduce / wgrad_reduce sum them in split order).
    const int g4 = (lane >> 4) * 4, c16 = lane & 15;
    if (p.nsplit > 1) {
        const long long pc0 = (PASS == P_DGRAD && p.phased) ? (long long)p.ph_tile0[ph] * BN : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long col = n0 + wn * 64 + j * 16 + c16;
            if (pc0 + col >= p.Ncol) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const long long row = m0 + wm * 64 + i * 16 + g4 + v;
                    if (row < p.M) partial[(split * p.M + row) * p.Ncol + pc0 + col] = acc[i][j][v];
                }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long col = n0 + wn * 64 + j * 16 + c16;
        long long cbase;
        long long rstride;
        if constexpr (PASS == P_FWD) {
            if (col >= p.Ncol) continue;
            const Pix q = decode(col, p.OH, p.OW);
            const long long ohw = (long long)p.OH * p.OW;
            cbase = (long long)q.n * p.K * ohw + (long long)q.a * p.OW + q.b;
            rstride = ohw;
        } else if constexpr (PASS == P_DGRAD) {
            const long long hw = (long long)p.H * p.W;
            if (p.phased) {
                const int rx = ph % p.s0, ry = ph / p.s0;
                const int Mx = phase_len(p.W, p.s0, rx), My = phase_len(p.H, p.s1, ry);
                if (col >= (long long)p.N * My * Mx) continue;
                const Pix q = decode(col, My, Mx);
                cbase = (long long)q.n * p.C * hw + (long long)(p.s1 * q.a + ry) * p.W + (p.s0 * q.b + rx);
            } else {
                if (col >= p.Ncol) continue;
                const Pix q = decode(col, p.H, p.W);
                cbase = (long long)q.n * p.C * hw + (long long)q.a * p.W + q.b;
            }
            rstride = hw;
        } else {
            if (col >= p.Ncol) continue;
            // dW[fw-1-i, fh-1-j, c, k]: element (k*C + c)*RS + RS-1-t
            const int c = (int)(col / p.RS);
            const int t = (int)(col - (long long)c * p.RS);
            cbase = (long long)c * p.RS + (p.RS - 1 - t);
            rstride = (long long)p.C * p.RS;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long row = m0 + wm * 64 + i * 16 + g4 + v;
                if (row >= p.M) continue;
                st_f(out, cbase + row * rstride, acc[i][j][v]);
            }
    }
}

// Second stage of a split forward or data gradient: sums the nsplit partial
// tiles of each output element in split order (deterministic) and stores it
// in the output layout. Padding columns of the phased layout are skipped.
template <typename TOut, int PASS>
__global__ __launch_bounds__(256) void nn_split_reduce(TOut *__restrict__ out, const float *__restrict__ partial,
                                                       NNParams p) {
    const long long total = p.M * p.Ncol;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long row = e / p.Ncol, col = e - row * p.Ncol;
        long long dst;
        if constexpr (PASS == P_FWD) {
            const Pix q = decode(col, p.OH, p.OW);
            const long long ohw = (long long)p.OH * p.OW;
            dst = ((long long)q.n * p.K + row) * ohw + (long long)q.a * p.OW + q.b;
        } else {
            const long long hw = (long long)p.H * p.W;
            if (p.phased) {
                const long long tn = col / BN;
                int ph = 0;
                while (ph + 1 < p.nph && tn >= p.ph_tile0[ph + 1]) ++ph;
                const long long lc = col - (long long)p.ph_tile0[ph] * BN;
                const int rx = ph % p.s0, ry = ph / p.s0;
                const int Mx = phase_len(p.W, p.s0, rx), My = phase_len(p.H, p.s1, ry);
                if (lc >= (long long)p.N * My * Mx) continue;
                const Pix q = decode(lc, My, Mx);
                dst = ((long long)q.n * p.C + row) * hw + (long long)(p.s1 * q.a + ry) * p.W + (p.s0 * q.b + rx);
            } else {
                const Pix q = decode(col, p.H, p.W);
                dst = ((long long)q.n * p.C + row) * hw + (long long)q.a * p.W + q.b;
            }
        }
        float s = 0.f;
        for (int sp = 0; sp < p.nsplit; ++sp) s += partial[sp * total + e];
        st_f(out, dst, s);
    }
}

// Deterministic second stage of the split filter gradient: sums the nsplit
// partial tiles in split order and scatters into dW's layout.
template <typename TOut>
__global__ __launch_bounds__(256) void wgrad_reduce(TOut *__restrict__ out,
                                                    const float *__restrict__ partial,
                                                    NNParams p) {
    const long long total = p.M * p.Ncol;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int sp = 0; sp < p.nsplit; ++sp) s += partial[sp * total + e];
        const long long row = e / p.Ncol, col = e - row * p.Ncol;
        const int c = (int)(col / p.RS);
        const int t = (int)(col - (long long)c * p.RS);
        st_f(out, row * ((long long)p.C * p.RS) + (long long)c * p.RS + (p.RS - 1 - t), s);
    }
}

// ------------------------------------------------------------- host side
static int setup(NNParams &p, int pass, const long long *xd, const long long *wd,
                 const long long *yd, const long long *st, const long long *pd,
                 const long long *dl) {
    p.W = (int)xd[0];
    p.H = (int)xd[1];
    p.C = (int)xd[2];
    p.N = (int)xd[3];
    p.fw = (int)wd[0];
    p.fh = (int)wd[1];
    p.K = (int)wd[3];
    p.OW = (int)yd[0];
    p.OH = (int)yd[1];
    p.s0 = (int)st[0];
    p.s1 = (int)st[1];
    p.p0 = (int)pd[0];
    p.p1 = (int)pd[1];
    p.d0 = (int)dl[0];
    p.d1 = (int)dl[1];
    if (wd[2] != xd[2] || yd[2] != wd[3] || yd[3] != xd[3]) return K_ERR_SIZE;
    if (p.s0 < 1 || p.s1 < 1 || p.d0 < 1 || p.d1 < 1) return K_ERR_ARG;
    p.RS = p.fw * p.fh;
    p.nsplit = 1;
    p.phased = 0;
    p.nph = 1;
    p.cl = 0;
    p.Rcp = 0;
    p.wp_bytes = p.act_bytes = 0;
    p.nsplit_req = 0;
    for (int i = 0; i <= kMaxPhases; ++i) p.ph_tile0[i] = 0;
    if (pass == P_FWD) {
        p.M = p.K;
        p.Ncol = (long long)p.N * p.OH * p.OW;
        p.kchunks = (p.C + BK - 1) / BK;
        p.ksteps = (long long)p.RS * p.kchunks;
    } else if (pass == P_DGRAD) {
        p.M = p.C;
        p.Ncol = (long long)p.N * p.H * p.W;
        p.kchunks = (p.K + BK - 1) / BK;
        p.ksteps = (long long)p.RS * p.kchunks;
        // AFHIP_NN_PHASED=0 in the environment keeps the unphased order
        const char *env = std::getenv("AFHIP_NN_PHASED");
        const bool off = env && env[0] == '0';
        if (!off && (p.s0 > 1 || p.s1 > 1) && p.s0 * p.s1 <= kMaxPhases) {
            p.phased = 1;
            p.nph = p.s0 * p.s1;
            long long t0 = 0, kmax = 0;
            for (int ph = 0; ph < p.nph; ++ph) {
                const int rx = ph % p.s0, ry = ph / p.s0;
                const long long cols = (long long)p.N * phase_len(p.H, p.s1, ry) * phase_len(p.W, p.s0, rx);
                p.ph_tile0[ph] = (int)t0;
                t0 += (cols + BN - 1) / BN;
                const long long ks = (long long)phase_taps(p.fw, rx, p.p0, p.d0, p.s0) *
                                     phase_taps(p.fh, ry, p.p1, p.d1, p.s1) * p.kchunks;
                if (ks > kmax) kmax = ks;
            }
            if (t0 >= (1LL << 31)) return K_ERR_SIZE;
            p.ph_tile0[p.nph] = (int)t0;
            p.Ncol = t0 * BN;
            p.ksteps = kmax;
        }
    } else {
        p.M = p.K;
        p.Ncol = (long long)p.C * p.RS;
        p.kchunks = 0;
        const long long R = (long long)p.N * p.OH * p.OW;
        p.ksteps = (R + BK - 1) / BK;
    }
    p.ksteps_per_split = p.ksteps;
    p.tilesM = (p.M + BM - 1) / BM;
    p.tilesN = (p.Ncol + BN - 1) / BN;
    return K_OK;
}

// Split count for the filter gradient: enough workgroups to cover the chip
// about twice, each with at least 16 slices of R.
static int wgrad_splits(const NNParams &p) {
    const long long tiles = p.tilesM * p.tilesN;
    long long want = (512 + tiles - 1) / tiles;
    long long maxs = p.ksteps / 16;
    if (want > maxs) want = maxs;
    if (want < 1) want = 1;
    return (int)want;
}

// The channel-last path runs when the operands are stored in the MMA's
// operand type (FP32, BF16 or FP16 arrays at their own precision).
template <int MODE, typename TIn>
constexpr bool cl_eligible() {
    return (MODE == M_F32 && std::is_same<TIn, float>::value) ||
           (MODE == M_BF16 && std::is_same<TIn, bf16s>::value) ||
           (MODE == M_F16 && std::is_same<TIn, _Float16>::value);
}

// Workspace of the forward / data-gradient GEMM: packed weights and the
// channel-last activation (esz > 0: element bytes of the channel-last path),
// then the split partials; every part 256-byte aligned.
// AFHIP_NN_CL=0 in the environment disables the channel-last path.
struct GemmLayout {
    int cl = 0, Rcp = 0, wp_bytes = 0, act_bytes = 0;
    long long off_w = 0, off_act = 0, off_part = 0, total = 0;
};
static GemmLayout gemm_layout(const NNParams &p, int pass, int esz, int nsplit) {
    GemmLayout L;
    auto al = [](long long b) { return (b + 255) & ~255LL; };
    const char *env = std::getenv("AFHIP_NN_CL");
    const long long Rcp = (long long)p.kchunks * BK;
    const long long rows = pass == P_FWD ? p.H : p.OH, cols = pass == P_FWD ? p.W : p.OW;
    const long long wb = (long long)p.RS * p.M * Rcp * esz, ab = (long long)p.N * rows * cols * Rcp * esz;
    // the copies are read through buffer resources (32-bit byte offsets)
    if (esz > 0 && !(env && env[0] == '0') && wb < (1LL << 31) - 64 && ab < (1LL << 31) - 64) {
        L.cl = 1;
        L.Rcp = (int)Rcp;
        L.wp_bytes = (int)wb;
        L.act_bytes = (int)ab;
        L.off_w = 0;
        L.off_act = al(wb);
        L.off_part = L.off_act + al(ab);
    }
    L.total = L.off_part + (nsplit > 1 ? (long long)nsplit * p.M * p.Ncol * 4 : 0);
    return L;
}

// Split count for the forward and data gradient: layers whose output tiles
// leave most CUs idle split the reduction (FP32 partial tiles, a second
// summing pass), about two workgroups per CU, each with at least 4 slices,
// the partials within 256 MiB. AFHIP_NN_SPLIT=<n> in the environment forces
// n (1 = unsplit).
static int gemm_splits(const NNParams &p) {
    const long long tiles = p.tilesM * p.tilesN;
    long long want;
    const char *env = std::getenv("AFHIP_NN_SPLIT");
    if (env && env[0] >= '1' && env[0] <= '9') {
        want = std::atoll(env);
    } else if (p.nsplit_req > 0) {
        want = p.nsplit_req;
    } else {
        const long long cus = cu_count_current();
        if (tiles * 2 >= cus) return 1;
        want = (2 * cus + tiles - 1) / tiles;
        const long long maxs = p.ksteps / 4;
        if (want > maxs) want = maxs;
    }
    while (want > 1 && want * p.M * p.Ncol * 4 > (256LL << 20)) --want;
    if (want > p.ksteps) want = p.ksteps;
    if (want < 1) want = 1;
    return (int)want;
}

template <int MODE, typename TIn, typename TOut>
static int run(int pass, void *out, const void *a0, const void *a1, NNParams &p, void *ws,
               long long ws_bytes, hipStream_t s) {
    if (p.M == 0 || p.Ncol == 0) return K_OK;
    const int smem = Lds<MODE>::BYTES;
    const long long tiles = p.tilesM * p.tilesN;
    if (pass == P_FWD || pass == P_DGRAD) {
        p.nsplit = gemm_splits(p);
        constexpr bool kCL = cl_eligible<MODE, TIn>();
        GemmLayout L = gemm_layout(p, pass, kCL ? (int)sizeof(TIn) : 0, p.nsplit);
        if (ws == nullptr || ws_bytes < L.total) {
            // a short workspace: drop the split, then the channel-last copies
            p.nsplit = 1;
            L = gemm_layout(p, pass, kCL ? (int)sizeof(TIn) : 0, 1);
            if (ws == nullptr || ws_bytes < L.total) L = gemm_layout(p, pass, 0, 1);
        }
        p.ksteps_per_split = (p.ksteps + p.nsplit - 1) / p.nsplit;
        char *wsb = (char *)ws;
        float *part = p.nsplit > 1 ? (float *)(wsb + L.off_part) : nullptr;
        const unsigned grid = (unsigned)(tiles * p.nsplit);
        if constexpr (kCL) {
            if (L.cl) {
                p.cl = 1;
                p.Rcp = L.Rcp;
                p.wp_bytes = L.wp_bytes;
                p.act_bytes = L.act_bytes;
                TIn *wp = (TIn *)(wsb + L.off_w), *act = (TIn *)(wsb + L.off_act);
                const long long nw = (long long)p.RS * p.M * p.Rcp;
                const long long b1 = std::min<long long>((nw + 255) / 256, kMaxGridBlocks / 2);
                const int rows = pass == P_FWD ? p.H : p.OH, cols = pass == P_FWD ? p.W : p.OW;
                const int Cd = pass == P_FWD ? p.C : p.K;
                const long long na = (long long)p.N * rows * cols * (p.Rcp / (16 / (int)sizeof(TIn)));
                const long long b2 = std::min<long long>((na + 255) / 256, kMaxGridBlocks / 2);
                nn_cl_copies<TIn><<<(unsigned)(b1 + b2), 256, 0, s>>>(wp, (const TIn *)a1, p, pass == P_DGRAD,
                                                                      (int)b1, act, (const TIn *)a0, cols, rows,
                                                                      Cd);
                auto fn = pass == P_FWD ? nn_gemm<MODE, TIn, TOut, P_FWD, true> : nn_gemm<MODE, TIn, TOut, P_DGRAD, true>;
                fn<<<grid, NT, smem, s>>>((TOut *)out, act, wp, part, p);
            }
        }
        if (!p.cl) {
            auto fn = pass == P_FWD ? nn_gemm<MODE, TIn, TOut, P_FWD> : nn_gemm<MODE, TIn, TOut, P_DGRAD>;
            fn<<<grid, NT, smem, s>>>((TOut *)out, (const TIn *)a0, (const TIn *)a1, part, p);
        }
        if (p.nsplit > 1) {
            long long blocks = (p.M * p.Ncol + 255) / 256;
            if (blocks > kMaxGridBlocks) blocks = kMaxGridBlocks;
            auto red = pass == P_FWD ? nn_split_reduce<TOut, P_FWD> : nn_split_reduce<TOut, P_DGRAD>;
            red<<<dim3((unsigned)blocks), 256, 0, s>>>((TOut *)out, part, p);
        }
    } else {
        if (p.ksteps == 0) {
            // empty reduction: dW = 0
            return afhip_k_fill(sizeof(TOut) == 4 ? DT_F32 : (std::is_same<TOut, bf16s>::value ? DT_BF16 : DT_F16),
                                out, 0.0, p.M * p.Ncol, s);
        }
        p.nsplit = wgrad_splits(p);
        const long long need = (long long)p.nsplit * p.M * p.Ncol * 4;
        if (p.nsplit > 1 && (ws == nullptr || ws_bytes < need)) p.nsplit = 1;
        p.ksteps_per_split = (p.ksteps + p.nsplit - 1) / p.nsplit;
        p.nsplit = (int)((p.ksteps + p.ksteps_per_split - 1) / p.ksteps_per_split);
        nn_gemm<MODE, TIn, TOut, P_WGRAD><<<dim3((unsigned)(tiles * p.nsplit)), NT, smem, s>>>(
            (TOut *)out, (const TIn *)a0, (const TIn *)a1, (float *)ws, p);
        if (p.nsplit > 1) {
            long long blocks = (p.M * p.Ncol + 255) / 256;
            if (blocks > kMaxGridBlocks) blocks = kMaxGridBlocks;
            wgrad_reduce<TOut><<<dim3((unsigned)blocks), 256, 0, s>>>((TOut *)out, (const float *)ws, p);
        }
    }
    return launch_status();
}

static int dispatch(int pass, int in_type, int out_type, int precision, void *out,
                    const void *a0, const void *a1, NNParams &p, void *ws, long long wsb,
                    hipStream_t s) {
    // precision: 0 default, 1 bf16, 2 f16, 3 split, 4 f32 (afhip_precision)
    int mode;
    switch (precision) {
        case 0:
            mode = in_type == DT_F32 ? M_F32 : in_type == DT_BF16 ? M_BF16 : in_type == DT_F16 ? M_F16 : -1;
            break;
        case 1: mode = M_BF16; break;
        case 2: mode = M_F16; break;
        case 3: mode = M_SPLIT; break;
        case 4: mode = M_F32; break;
        default: return K_ERR_ARG;
    }
    if (mode < 0) return K_ERR_TYPE;
#define GO(MD, TI, TO) return run<MD, TI, TO>(pass, out, a0, a1, p, ws, wsb, s)
    if (in_type == DT_F32) {
        if (out_type != DT_F32) return K_ERR_TYPE;
        switch (mode) {
            case M_F32: GO(M_F32, float, float);
            case M_BF16: GO(M_BF16, float, float);
            case M_F16: GO(M_F16, float, float);
            case M_SPLIT: GO(M_SPLIT, float, float);
        }
    } else if (in_type == DT_BF16) {
        if (mode != M_BF16) return K_ERR_TYPE;
        if (out_type == DT_F32) GO(M_BF16, bf16s, float);
        if (out_type == DT_BF16) GO(M_BF16, bf16s, bf16s);
    } else if (in_type == DT_F16) {
        if (mode != M_F16) return K_ERR_TYPE;
        if (out_type == DT_F32) GO(M_F16, _Float16, float);
        if (out_type == DT_F16) GO(M_F16, _Float16, _Float16);
    }
#undef GO
    return K_ERR_TYPE;
}

}  // namespace afhip

using namespace afhip;

extern "C" long long afhip_k_conv2_nn_gemm_split_workspace(int pass, int in_type, int precision,
                                                           const long long *xdims, const long long *wdims,
                                                           const long long *ydims, const long long *stride,
                                                           const long long *pad, const long long *dil,
                                                           int nsplit) {
    if (pass != P_FWD && pass != P_DGRAD) return 0;
    NNParams p;
    if (setup(p, pass, xdims, wdims, ydims, stride, pad, dil)) return 0;
    if (p.M == 0 || p.Ncol == 0) return 0;
    p.nsplit_req = nsplit > 0 ? nsplit : 0;
    // channel-last path: operands stored at their own precision
    const bool cl = (in_type == DT_F32 && (precision == 0 || precision == 4)) ||
                    (in_type == DT_BF16 && (precision == 0 || precision == 1)) ||
                    (in_type == DT_F16 && (precision == 0 || precision == 2));
    const int esz = cl ? (in_type == DT_F32 ? 4 : 2) : 0;
    return gemm_layout(p, pass, esz, gemm_splits(p)).total;
}

extern "C" int afhip_k_conv2_nn_gemm_split(int pass, int in_type, int out_type, int precision, void *out,
                                           const void *in, const void *w, const long long *xdims,
                                           const long long *wdims, const long long *ydims,
                                           const long long *stride, const long long *pad, const long long *dil,
                                           int nsplit, void *workspace, long long workspace_bytes, void *stream) {
    if (pass != P_FWD && pass != P_DGRAD) return K_ERR_ARG;
    NNParams p;
    int rc = setup(p, pass, xdims, wdims, ydims, stride, pad, dil);
    if (rc) return rc;
    p.nsplit_req = nsplit > 0 ? nsplit : 0;
    return dispatch(pass, in_type, out_type, precision, out, in, w, p, workspace, workspace_bytes,
                    (hipStream_t)stream);
}

extern "C" long long afhip_k_conv2_nn_gemm_workspace(int pass, int in_type, int precision, const long long *xdims,
                                                     const long long *wdims, const long long *ydims,
                                                     const long long *stride, const long long *pad,
                                                     const long long *dil) {
    return afhip_k_conv2_nn_gemm_split_workspace(pass, in_type, precision, xdims, wdims, ydims, stride, pad, dil,
                                                 0);
}

extern "C" int afhip_k_conv2_nn_gemm(int pass, int in_type, int out_type, int precision, void *out,
                                     const void *in, const void *w, const long long *xdims,
                                     const long long *wdims, const long long *ydims,
                                     const long long *stride, const long long *pad, const long long *dil,
                                     void *workspace, long long workspace_bytes, void *stream) {
    return afhip_k_conv2_nn_gemm_split(pass, in_type, out_type, precision, out, in, w, xdims, wdims, ydims, stride,
                                       pad, dil, 0, workspace, workspace_bytes, stream);
}

extern "C" int afhip_k_conv2_nn_fwd(int in_type, int out_type, int precision, void *y,
                                    const void *x, const void *w, const long long *xdims,
                                    const long long *wdims, const long long *ydims,
                                    const long long *stride, const long long *pad,
                                    const long long *dil, void *stream) {
    return afhip_k_conv2_nn_gemm(P_FWD, in_type, out_type, precision, y, x, w, xdims, wdims, ydims, stride, pad,
                                 dil, nullptr, 0, stream);
}

extern "C" int afhip_k_conv2_nn_dgrad(int in_type, int out_type, int precision, void *dx,
                                      const void *dy, const void *w, const long long *xdims,
                                      const long long *wdims, const long long *ydims,
                                      const long long *stride, const long long *pad,
                                      const long long *dil, void *stream) {
    return afhip_k_conv2_nn_gemm(P_DGRAD, in_type, out_type, precision, dx, dy, w, xdims, wdims, ydims, stride,
                                 pad, dil, nullptr, 0, stream);
}

extern "C" long long afhip_k_conv2_nn_wgrad_workspace(const long long *xdims, const long long *wdims,
                                                      const long long *ydims, const long long *stride,
                                                      const long long *pad, const long long *dil) {
    NNParams p;
    if (setup(p, P_WGRAD, xdims, wdims, ydims, stride, pad, dil)) return 0;
    const int ns = wgrad_splits(p);
    return ns > 1 ? (long long)ns * p.M * p.Ncol * 4 : 0;
}

extern "C" int afhip_k_conv2_nn_wgrad(int in_type, int out_type, int precision, void *dw,
                                      const void *x, const void *dy, const long long *xdims,
                                      const long long *wdims, const long long *ydims,
                                      const long long *stride, const long long *pad,
                                      const long long *dil, void *workspace,
                                      long long workspace_bytes, void *stream) {
    NNParams p;
    int rc = setup(p, P_WGRAD, xdims, wdims, ydims, stride, pad, dil);
    if (rc) return rc;
    return dispatch(P_WGRAD, in_type, out_type, precision, dw, x, dy, p, workspace, workspace_bytes,
                    (hipStream_t)stream);
}
