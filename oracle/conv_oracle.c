/*
 * CPU oracle for ArrayFire's convolution path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke entry point and bench.py's CPU-baseline leg may load
 * this library, and only as the checker (or as the timed CPU baseline). The
 * product library never links it.
 *
 * This is a restatement, in plain C, of the reference CPU backend's loop nests
 * with the same index rules, accumulation types and rounding points:
 *
 *   one2one_1d / one2one_2d / one2one_3d ... src/backend/cpu/kernel/convolve.hpp:19-130
 *   convolve_nd (batch stepping) ........... src/backend/cpu/kernel/convolve.hpp:132-211
 *   convolve2_separable / convolve2 ........ src/backend/cpu/kernel/convolve.hpp:213-289
 *   unwrap_dim (im2col, d = 0) ............. src/backend/cpu/kernel/unwrap.hpp:20-80
 *   wrap_dim_dilated (col2im, d = 0) ....... src/backend/cpu/kernel/wrap.hpp:83-145
 *
 * Every function is instantiated for the reference's (InT, AccT) pairs
 * through a macro: every af_dtype the spatial path accepts.
 * The double instantiation of the spatial loops is also the FP64 brute-force
 * definition of the spatial convolution (it is the textbook sum with no
 * reordering).
 *
 * Arrays are ArrayFire-style column-major 4-D with contiguous strides
 * [1, d0, d0*d1, d0*d1*d2]. All dims are int64[4].
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off so that a*b+c is
 * evaluated as two rounded operations, as the reference source spells it).
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t dim_t;

enum { BATCH_NONE = 1, BATCH_LHS, BATCH_RHS, BATCH_SAME, BATCH_DIFF,
       BATCH_UNSUPPORTED };

static void strides_of(const dim_t *d, dim_t *s) {
    s[0] = 1;
    s[1] = d[0];
    s[2] = d[0] * d[1];
    s[3] = d[0] * d[1] * d[2];
}

/* Instantiates the spatial loops for one (InT, AccT) pair.
 * CONV_CAST_IN converts an accumulator back to InT (the reference's InT(accum)). */
#define DEFINE_SPATIAL(SUF, InT, AccT)                                          \
static void one2one_1d_##SUF(InT *optr, const InT *iptr, const AccT *fptr,      \
                             const dim_t *oD, const dim_t *sD, const dim_t *fD, \
                             const dim_t *sS, int expand) {                     \
    dim_t start = expand ? 0 : fD[0] / 2;                                       \
    dim_t end = expand ? oD[0] : start + sD[0];                                 \
    for (dim_t i = start; i < end; ++i) {                                       \
        AccT accum = 0;                                                         \
        for (dim_t f = 0; f < fD[0]; ++f) {                                     \
            dim_t iIdx = i - f;                                                 \
            InT s_val = (iIdx >= 0 && iIdx < sD[0]) ? iptr[iIdx * sS[0]]        \
                                                    : (InT)0;                   \
            accum += (AccT)(s_val * fptr[f]);                                   \
        }                                                                       \
        optr[i - start] = (InT)accum;                                           \
    }                                                                           \
}                                                                               \
static void one2one_2d_##SUF(InT *optr, const InT *iptr, const AccT *fptr,      \
                             const dim_t *oD, const dim_t *sD, const dim_t *fD, \
                             const dim_t *oS, const dim_t *sS, const dim_t *fS, \
                             int expand) {                                      \
    dim_t jStart = expand ? 0 : fD[1] / 2;                                      \
    dim_t jEnd = expand ? oD[1] : jStart + sD[1];                               \
    dim_t iStart = expand ? 0 : fD[0] / 2;                                      \
    dim_t iEnd = expand ? oD[0] : iStart + sD[0];                               \
    for (dim_t j = jStart; j < jEnd; ++j) {                                     \
        dim_t joff = (j - jStart) * oS[1];                                      \
        for (dim_t i = iStart; i < iEnd; ++i) {                                 \
            AccT accum = 0;                                                     \
            for (dim_t wj = 0; wj < fD[1]; ++wj) {                              \
                dim_t jIdx = j - wj;                                            \
                dim_t w_joff = wj * fS[1];                                      \
                dim_t s_joff = jIdx * sS[1];                                    \
                int isJValid = (jIdx >= 0 && jIdx < sD[1]);                     \
                for (dim_t wi = 0; wi < fD[0]; ++wi) {                          \
                    dim_t iIdx = i - wi;                                        \
                    InT s_val = 0;                                              \
                    if (isJValid && iIdx >= 0 && iIdx < sD[0])                  \
                        s_val = iptr[s_joff + iIdx * sS[0]];                    \
                    accum += (AccT)(s_val * fptr[w_joff + wi * fS[0]]);         \
                }                                                               \
            }                                                                   \
            optr[joff + i - iStart] = (InT)accum;                               \
        }                                                                       \
    }                                                                           \
}                                                                               \
static void one2one_3d_##SUF(InT *optr, const InT *iptr, const AccT *fptr,      \
                             const dim_t *oD, const dim_t *sD, const dim_t *fD, \
                             const dim_t *oS, const dim_t *sS, const dim_t *fS, \
                             int expand) {                                      \
    dim_t kStart = expand ? 0 : fD[2] / 2;                                      \
    dim_t kEnd = expand ? oD[2] : kStart + sD[2];                               \
    dim_t jStart = expand ? 0 : fD[1] / 2;                                      \
    dim_t jEnd = expand ? oD[1] : jStart + sD[1];                               \
    dim_t iStart = expand ? 0 : fD[0] / 2;                                      \
    dim_t iEnd = expand ? oD[0] : iStart + sD[0];                               \
    for (dim_t k = kStart; k < kEnd; ++k) {                                     \
        dim_t koff = (k - kStart) * oS[2];                                      \
        for (dim_t j = jStart; j < jEnd; ++j) {                                 \
            dim_t joff = (j - jStart) * oS[1];                                  \
            for (dim_t i = iStart; i < iEnd; ++i) {                             \
                AccT accum = 0;                                                 \
                for (dim_t wk = 0; wk < fD[2]; ++wk) {                          \
                    dim_t kIdx = k - wk;                                        \
                    dim_t w_koff = wk * fS[2];                                  \
                    dim_t s_koff = kIdx * sS[2];                                \
                    int isKValid = (kIdx >= 0 && kIdx < sD[2]);                 \
                    for (dim_t wj = 0; wj < fD[1]; ++wj) {                      \
                        dim_t jIdx = j - wj;                                    \
                        dim_t w_joff = wj * fS[1];                              \
                        dim_t s_joff = jIdx * sS[1];                            \
                        int isJValid = (jIdx >= 0 && jIdx < sD[1]);             \
                        for (dim_t wi = 0; wi < fD[0]; ++wi) {                  \
                            dim_t iIdx = i - wi;                                \
                            InT s_val = 0;                                      \
                            if (isKValid && isJValid && iIdx >= 0 &&            \
                                iIdx < sD[0])                                   \
                                s_val = iptr[s_koff + s_joff + iIdx * sS[0]];   \
                            accum += (AccT)(s_val *                             \
                                            fptr[w_koff + w_joff + wi * fS[0]]);\
                        }                                                       \
                    }                                                           \
                }                                                               \
                optr[koff + joff + i - iStart] = (InT)accum;                    \
            }                                                                   \
        }                                                                       \
    }                                                                           \
}                                                                               \
/* convolve_nd: batch stepping per kind, kernel/convolve.hpp:132-211 */         \
void orc_convolve_nd_##SUF(InT *optr, const dim_t *oD, const InT *iptr,         \
                           const dim_t *sD, const AccT *fptr, const dim_t *fD,  \
                           int kind, int rank, int expand) {                    \
    dim_t oS[4], sS[4], fS[4];                                                  \
    strides_of(oD, oS); strides_of(sD, sS); strides_of(fD, fS);                 \
    dim_t out_step[4] = {0, 0, 0, 0}, in_step[4] = {0, 0, 0, 0};                \
    dim_t filt_step[4] = {0, 0, 0, 0}, batch[4] = {0, 1, 1, 1};                 \
    for (int i = 1; i < 4; ++i) {                                               \
        switch (kind) {                                                         \
        case BATCH_LHS:                                                         \
            out_step[i] = oS[i]; in_step[i] = sS[i];                            \
            if (i >= rank) batch[i] = sD[i];                                    \
            break;                                                              \
        case BATCH_SAME:                                                        \
            out_step[i] = oS[i]; in_step[i] = sS[i]; filt_step[i] = fS[i];      \
            if (i >= rank) batch[i] = sD[i];                                    \
            break;                                                              \
        case BATCH_RHS:                                                         \
            out_step[i] = oS[i]; filt_step[i] = fS[i];                          \
            if (i >= rank) batch[i] = fD[i];                                    \
            break;                                                              \
        default: break;                                                         \
        }                                                                       \
    }                                                                           \
    for (dim_t b3 = 0; b3 < batch[3]; ++b3)                                     \
        for (dim_t b2 = 0; b2 < batch[2]; ++b2)                                 \
            for (dim_t b1 = 0; b1 < batch[1]; ++b1) {                           \
                InT *o = optr + b1 * out_step[1] + b2 * out_step[2] +           \
                         b3 * out_step[3];                                      \
                const InT *in = iptr + b1 * in_step[1] + b2 * in_step[2] +      \
                                b3 * in_step[3];                                \
                const AccT *f = fptr + b1 * filt_step[1] + b2 * filt_step[2] +  \
                                b3 * filt_step[3];                              \
                switch (rank) {                                                 \
                case 1: one2one_1d_##SUF(o, in, f, oD, sD, fD, sS, expand);     \
                    break;                                                      \
                case 2: one2one_2d_##SUF(o, in, f, oD, sD, fD, oS, sS, fS,      \
                                         expand);                               \
                    break;                                                      \
                case 3: one2one_3d_##SUF(o, in, f, oD, sD, fD, oS, sS, fS,      \
                                         expand);                               \
                    break;                                                      \
                }                                                               \
            }                                                                   \
}                                                                               \
/* One pass of the separable convolution, kernel/convolve.hpp:213-255.         \
 * The tap is converted to InT before the product, as the reference does. */   \
static void sep_pass_##SUF(InT *optr, const InT *iptr, const AccT *fptr,        \
                           const dim_t *oD, const dim_t *sD, dim_t fDim,        \
                           const dim_t *oS, int convDim, int expand) {          \
    for (dim_t j = 0; j < oD[1]; ++j) {                                         \
        dim_t jOff = j * oS[1];                                                 \
        dim_t cj = j + (convDim == 1) * (expand ? 0 : fDim >> 1);               \
        for (dim_t i = 0; i < oD[0]; ++i) {                                     \
            dim_t iOff = i * oS[0];                                             \
            dim_t ci = i + (convDim == 0) * (expand ? 0 : fDim >> 1);           \
            AccT accum = 0;                                                     \
            for (dim_t f = 0; f < fDim; ++f) {                                  \
                InT f_val = (InT)fptr[f];                                       \
                InT s_val;                                                      \
                if (convDim == 0) {                                             \
                    dim_t offi = ci - f;                                        \
                    int ok = offi >= 0 && offi < sD[0] && cj >= 0 &&            \
                             cj < sD[1];                                        \
                    s_val = ok ? iptr[cj * sD[0] + offi] : (InT)0;              \
                } else {                                                        \
                    dim_t offj = cj - f;                                        \
                    int ok = ci >= 0 && ci < sD[0] && offj >= 0 &&              \
                             offj < sD[1];                                      \
                    s_val = ok ? iptr[offj * sD[0] + ci] : (InT)0;              \
                }                                                               \
                accum += (AccT)(s_val * f_val);                                 \
            }                                                                   \
            optr[iOff + jOff] = (InT)accum;                                     \
        }                                                                       \
    }                                                                           \
}                                                                               \
/* cpu::convolve2 (separable) host + kernel::convolve2,                        \
 * src/backend/cpu/convolve.cpp:68-99, kernel/convolve.hpp:257-289.            \
 * The temp array is InT-typed (a rounding point). */                          \
int orc_convolve2_sep_##SUF(InT *optr, const dim_t *oD, const InT *iptr,        \
                            const dim_t *sD, const AccT *cf, dim_t cflen,       \
                            const AccT *rf, dim_t rflen, int expand) {          \
    dim_t tD[4] = {sD[0], sD[1], sD[2], sD[3]};                                 \
    if (expand) tD[0] += cflen - 1;                                             \
    dim_t oS[4], sS[4], tS[4];                                                  \
    strides_of(oD, oS); strides_of(sD, sS); strides_of(tD, tS);                 \
    InT *temp = (InT *)malloc(sizeof(InT) * (size_t)tD[0] * tD[1]);             \
    if (!temp) return -1;                                                       \
    for (dim_t b3 = 0; b3 < oD[3]; ++b3)                                        \
        for (dim_t b2 = 0; b2 < oD[2]; ++b2) {                                  \
            const InT *ip = iptr + b2 * sS[2] + b3 * sS[3];                     \
            InT *op = optr + b2 * oS[2] + b3 * oS[3];                           \
            sep_pass_##SUF(temp, ip, cf, tD, sD, cflen, tS, 0, expand);         \
            sep_pass_##SUF(op, temp, rf, oD, tD, rflen, oS, 1, expand);         \
        }                                                                       \
    free(temp);                                                                 \
    return 0;                                                                   \
}

/* The (InT, AccT) pairs of src/api/c/convolve.cpp:155-209 (c32 -> c32,
 * c64 -> c64, f32 -> f32, f64 -> f64, every integer type and b8 -> float).
 * b8 is the reference's `char`; the complex types are C99's, whose product
 * (ac - bd, ad + bc) with -ffp-contract=off rounds as std::complex does. */
typedef float _Complex cf32_t;
typedef double _Complex cf64_t;
DEFINE_SPATIAL(f32, float, float)
DEFINE_SPATIAL(f64, double, double)
DEFINE_SPATIAL(c32, cf32_t, cf32_t)
DEFINE_SPATIAL(c64, cf64_t, cf64_t)
DEFINE_SPATIAL(s32, int32_t, float)
DEFINE_SPATIAL(u32, uint32_t, float)
DEFINE_SPATIAL(s16, int16_t, float)
DEFINE_SPATIAL(u16, uint16_t, float)
DEFINE_SPATIAL(s64, int64_t, float)
DEFINE_SPATIAL(u64, uint64_t, float)
DEFINE_SPATIAL(u8, uint8_t, float)
DEFINE_SPATIAL(s8, int8_t, float)
DEFINE_SPATIAL(b8, char, float)

/* im2col with d = 0 (retCols = false), src/backend/cpu/kernel/unwrap.hpp:20-80
 * and the output shape of src/backend/cpu/unwrap.cpp:23-42:
 *   out dims = [nx*ny, wx*wy, C, N], out[col, y*wx + x, c, n]. */
#define DEFINE_UNWRAP(SUF, T)                                                   \
void orc_unwrap_##SUF(T *outPtr, const T *inPtr, const dim_t *idims,            \
                      dim_t wx, dim_t wy, dim_t sx, dim_t sy, dim_t px,         \
                      dim_t py, dim_t dx, dim_t dy) {                           \
    dim_t nx = 1 + (idims[0] + 2 * px - (((wx - 1) * dx) + 1)) / sx;            \
    dim_t ny = 1 + (idims[1] + 2 * py - (((wy - 1) * dy) + 1)) / sy;            \
    dim_t odims[4] = {nx * ny, wx * wy, idims[2], idims[3]};                    \
    dim_t is[4], os[4];                                                         \
    strides_of(idims, is); strides_of(odims, os);                               \
    for (dim_t w = 0; w < odims[3]; w++)                                        \
        for (dim_t z = 0; z < odims[2]; z++) {                                  \
            const T *iptr = inPtr + w * is[3] + z * is[2];                      \
            T *optr_ = outPtr + w * os[3] + z * os[2];                          \
            for (dim_t col = 0; col < odims[0]; col++) {                        \
                T *optr = optr_ + col * os[0];                                  \
                dim_t winy = col / nx, winx = col % nx;                         \
                dim_t spx = winx * sx - px, spy = winy * sy - py;               \
                for (dim_t y = 0; y < wy; y++) {                                \
                    dim_t ypad = spy + y * dy;                                  \
                    for (dim_t x = 0; x < wx; x++) {                            \
                        dim_t xpad = spx + x * dx;                              \
                        dim_t oloc = (y * wx + x) * os[1];                      \
                        if (xpad >= 0 && xpad < idims[0] && ypad >= 0 &&        \
                            ypad < idims[1])                                    \
                            optr[oloc] = iptr[ypad * is[1] + xpad * is[0]];     \
                        else                                                    \
                            optr[oloc] = (T)0;                                  \
                    }                                                           \
                }                                                               \
            }                                                                   \
        }                                                                       \
}                                                                               \
/* col2im with d = 0, src/backend/cpu/kernel/wrap.hpp:83-145; out is zeroed   \
 * first as src/backend/cpu/wrap.cpp:67 does. The serial += scatter order is   \
 * kept: (n, c, col, y, x). in dims = [nx*ny, wx*wy, C, N]. */                 \
void orc_wrap_dilated_##SUF(T *outPtr, const dim_t *odims, const T *inPtr,      \
                            const dim_t *idims, dim_t wx, dim_t wy, dim_t sx,   \
                            dim_t sy, dim_t px, dim_t py, dim_t dx, dim_t dy) { \
    dim_t is[4], os[4];                                                         \
    strides_of(idims, is); strides_of(odims, os);                               \
    memset(outPtr, 0, sizeof(T) * (size_t)(odims[0] * odims[1] * odims[2] *     \
                                           odims[3]));                          \
    dim_t nx = 1 + (odims[0] + 2 * px - (((wx - 1) * dx) + 1)) / sx;            \
    for (dim_t w = 0; w < idims[3]; w++)                                        \
        for (dim_t z = 0; z < idims[2]; z++) {                                  \
            const T *iptr_ = inPtr + w * is[3] + z * is[2];                     \
            T *optr = outPtr + w * os[3] + z * os[2];                           \
            for (dim_t col = 0; col < idims[0]; col++) {                        \
                const T *iptr = iptr_ + col * is[0];                            \
                dim_t winy = col / nx, winx = col % nx;                         \
                dim_t spx = winx * sx - px, spy = winy * sy - py;               \
                for (dim_t y = 0; y < wy; y++) {                                \
                    dim_t ypad = spy + y * dy;                                  \
                    for (dim_t x = 0; x < wx; x++) {                            \
                        dim_t xpad = spx + x * dx;                              \
                        dim_t iloc = (y * wx + x) * is[1];                      \
                        if (xpad >= 0 && xpad < odims[0] && ypad >= 0 &&        \
                            ypad < odims[1])                                    \
                            optr[ypad * os[1] + xpad * os[0]] += iptr[iloc];    \
                    }                                                           \
                }                                                               \
            }                                                                   \
        }                                                                       \
}

DEFINE_UNWRAP(f32, float)
DEFINE_UNWRAP(f64, double)
