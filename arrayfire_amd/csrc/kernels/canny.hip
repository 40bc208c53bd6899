// Canny edge detector: af_canny (src/api/c/canny.cpp:198-285), a caller of
// the separable convolution (its 5-tap smoothing, canny.cpp:203-209), with
// the CPU backend's kernels as the semantics (src/backend/cpu/kernel/
// sobel.hpp:20-70, canny.hpp:18-186, histogram.hpp:19-47):
//  * canny_grad: 3x3 Sobel derivatives of the smoothed slice (reflect101
//    borders; dx along dim0, dy along dim1) and the L1 or L2 magnitude;
//  * canny_nms: interior pixels keep their magnitude when it exceeds both
//    magnitudes interpolated along the gradient direction, others are 0;
//  * canny_minmax_part + canny_minmax_final: the suppressed map's min and
//    max (MANUAL normalisation, AUTO_OTSU bin count);
//  * canny_hist + canny_otsu: per slice, the histogram over [0, max] and
//    Otsu's threshold (bin index of the first largest between-class
//    variance), its scans and sums in float in bin order on one lane;
//  * canny_candidates: strong / weak pixels against the thresholds;
//  * canny_seed + canny_sweep: edge tracking by hysteresis. The reference
//    starts a trace at every strong pixel its scan visits (linear indices
//    d0+1 .. d0+1+(d0-1)*(d1-1)-1 of a slice) and marks weak pixels reached
//    through the 8 linear-offset neighbours; the marked set is the closure of
//    the seeds over weak pixels, computed here by repeated sweeps that mark a
//    weak pixel next to a marked one, until a sweep changes nothing.
// Arithmetic is uncontracted and in the reference's order, so every
// comparison sees the reference's values.
#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

static unsigned cgrid(long long n) {
    long long b = (n + 255) / 256;
    if (b > kMaxGridBlocks) b = kMaxGridBlocks;
    return (unsigned)(b < 1 ? 1 : b);
}

__device__ inline int reflect101(int idx, int end) { return abs(end - abs(end - idx)); }

__global__ __launch_bounds__(256) void canny_grad(float *__restrict__ gx, float *__restrict__ gy,
                                                  float *__restrict__ mag, const float *__restrict__ x, int d0,
                                                  int d1, long long ns, int is_fast) {
#pragma clang fp contract(off)
    const long long per = (long long)d0 * d1, n = per * ns;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const long long sl = e / per;
        const int r = (int)(e - sl * per);
        const int i = r % d0, j = r / d0;
        const float *p = x + sl * per;
        const int im = reflect101(i - 1, d0 - 1), ip = reflect101(i + 1, d0 - 1);
        const int jm = reflect101(j - 1, d1 - 1), jp = reflect101(j + 1, d1 - 1);
        const float NW = p[jm * d0 + im], SW = p[jm * d0 + ip], NE = p[jp * d0 + im], SE = p[jp * d0 + ip];
        const float N = p[j * d0 + im], S = p[j * d0 + ip], W = p[jm * d0 + i], E = p[jp * d0 + i];
        const float dx = ((SW + SE) - (NW + NE)) + 2.0f * (S - N);
        const float dy = ((NE + SE) - (NW + SW)) + 2.0f * (E - W);
        gx[e] = dx;
        gy[e] = dy;
        mag[e] = is_fast ? fabsf(dx) + fabsf(dy) : sqrtf(dx * dx + dy * dy);
    }
}

__global__ __launch_bounds__(256) void canny_nms(float *__restrict__ out, const float *__restrict__ mag,
                                                 const float *__restrict__ gxp, const float *__restrict__ gyp,
                                                 int d0, int d1, long long ns) {
#pragma clang fp contract(off)
    const long long per = (long long)d0 * d1, n = per * ns;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const long long sl = e / per;
        const int r = (int)(e - sl * per);
        const int i = r % d0, j = r / d0;
        float res = 0.f;
        if (i >= 1 && i < d0 - 1 && j >= 1 && j < d1 - 1) {
            const float *m = mag + e;
            const float curr = m[0];
            if (curr != 0.f) {
                const float se = m[d0 + 1], nw = m[-d0 - 1], ea = m[1], we = m[-1];
                const float ne = m[-d0 + 1], sw = m[d0 - 1], no = m[-d0], so = m[d0];
                const float dx = gxp[e], dy = gyp[e];
                float a1, a2, b1, b2, alpha;
                if (dx >= 0) {
                    if (dy >= 0) {
                        const bool g = (dx - dy) >= 0;
                        a1 = g ? ea : so;
                        a2 = g ? we : no;
                        b1 = se;
                        b2 = nw;
                        alpha = g ? dy / dx : dx / dy;
                    } else {
                        const bool g = (dx + dy) >= 0;
                        a1 = g ? ea : no;
                        a2 = g ? we : so;
                        b1 = ne;
                        b2 = sw;
                        alpha = g ? -dy / dx : dx / -dy;
                    }
                } else {
                    if (dy >= 0) {
                        const bool g = (dx + dy) >= 0;
                        a1 = g ? so : we;
                        a2 = g ? no : ea;
                        b1 = sw;
                        b2 = ne;
                        alpha = g ? -dx / dy : dy / -dx;
                    } else {
                        const bool g = (-dx + dy) >= 0;
                        a1 = g ? we : no;
                        a2 = g ? ea : so;
                        b1 = nw;
                        b2 = se;
                        alpha = g ? dy / dx : dx / dy;
                    }
                }
                const float mag1 = (1.0f - alpha) * a1 + alpha * b1;
                const float mag2 = (1.0f - alpha) * a2 + alpha * b2;
                if (curr > mag1 && curr > mag2) res = curr;
            }
        }
        out[e] = res;
    }
}

// mm[0] = min, mm[1] = max over n values in two launches: each of up to
// kMinMaxBlocks workgroups reduces a grid-stride share into part[2b],
// part[2b+1], then one workgroup reduces those (min and max are exact, so
// the grouping does not change the values).
constexpr int kMinMaxBlocks = 1024;

__device__ inline void block_minmax(float &lo, float &hi) {
    __shared__ float smin[256], smax[256];
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + h]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + h]);
        }
        __syncthreads();
    }
    lo = smin[0];
    hi = smax[0];
}

__global__ __launch_bounds__(256) void canny_minmax_part(float *__restrict__ part, const float *__restrict__ v,
                                                         long long n) {
    float lo = INFINITY, hi = -INFINITY;
    for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
        lo = fminf(lo, v[e]);
        hi = fmaxf(hi, v[e]);
    }
    block_minmax(lo, hi);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = lo;
        part[2 * blockIdx.x + 1] = hi;
    }
}

__global__ __launch_bounds__(256) void canny_minmax_final(float *__restrict__ mm, const float *__restrict__ part,
                                                          int nb) {
    float lo = INFINITY, hi = -INFINITY;
    for (int b = threadIdx.x; b < nb; b += 256) {
        lo = fminf(lo, part[2 * b]);
        hi = fmaxf(hi, part[2 * b + 1]);
    }
    block_minmax(lo, hi);
    if (threadIdx.x == 0) {
        mm[0] = lo;
        mm[1] = hi;
    }
}

// hist [ns][nbins]: bin = (int)((v - 0) / step) clamped to [0, nbins-1]
__global__ __launch_bounds__(256) void canny_hist(unsigned *__restrict__ hist, const float *__restrict__ v,
                                                  long long per, long long ns, int nbins, float step) {
#pragma clang fp contract(off)
    const long long n = per * ns;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const long long sl = e / per;
        int b = (int)((v[e] - 0.0f) / step);
        b = max(b, 0);
        b = min(b, nbins - 1);
        atomicAdd(hist + sl * nbins + b, 1u);
    }
}

// One lane per slice: Otsu's threshold (src/api/c/canny.cpp:82-147) as the
// bin index of the largest between-class variance (MinMaxOp<af_max_t>,
// src/backend/cpu/kernel/ireduce.hpp:47-63: element 0 first, a NaN there
// read as the lowest float; larger magnitudes win, NaN never does).
__global__ void canny_otsu(float *__restrict__ thr, const unsigned *__restrict__ hist, int nbins, long long ns,
                           float total) {
#pragma clang fp contract(off)
    const long long sl = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (sl >= ns) return;
    const unsigned *h = hist + sl * nbins;
    // weightedSum: the sum of w_b * f_b in bin order
    float ws = 0.f;
    for (int b = 0; b < nbins; ++b) ws = ((float)b * ((float)h[b] / total)) + ws;
    float cum = 0.f, cprod = 0.f;
    float best = 0.f;
    int best_i = 0;
    for (int b = 0; b < nbins - 1; ++b) {
        const float f = (float)h[b] / total;
        cum = cum + f;
        cprod = cprod + (float)b * f;
        const float qL = cum;
        const float qH = 1.0f - qL;
        const float muL = cprod / qL;
        const float muH = (ws - cprod) / qH;
        const float diff = muL - muH;
        const float sig = (diff * diff) * (qL * qH);
        if (b == 0) {
            best = sig != sig ? -3.402823466e38f : sig;
            best_i = 0;
        } else if (fabsf(sig) > fabsf(best)) {
            best = sig;
            best_i = b;
        }
    }
    thr[sl] = (float)best_i;
}

// strong / weak (uint8). mode 0 (MANUAL): g = (v - lo) / (hi - lo) against
// t1, t2; mode 1 (AUTO_OTSU): v against T1 = T2 * t1, T2 = thr[slice].
__global__ __launch_bounds__(256) void canny_candidates(unsigned char *__restrict__ strong,
                                                        unsigned char *__restrict__ weak,
                                                        const float *__restrict__ v, const float *__restrict__ mm,
                                                        const float *__restrict__ thr, long long per, long long ns,
                                                        float t1, float t2, int mode) {
#pragma clang fp contract(off)
    const long long n = per * ns;
    const float lo = mm[0], hi = mm[1];
    const float den = hi - lo;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        float g, T1, T2;
        if (mode == 0) {
            g = (v[e] - lo) / den;
            T1 = t1;
            T2 = t2;
        } else {
            g = v[e];
            T2 = thr[e / per];
            T1 = T2 * t1;
        }
        strong[e] = g >= T2;
        weak[e] = (g >= T1) && (g < T2);
    }
}

// out = strong pixels at the reference's scan positions of each slice
__global__ __launch_bounds__(256) void canny_seed(unsigned char *__restrict__ out,
                                                  const unsigned char *__restrict__ strong, int d0, int d1,
                                                  long long ns) {
    const long long per = (long long)d0 * d1, n = per * ns;
    const long long t0 = d0 + 1, t1 = t0 + (long long)(d0 - 1) * (d1 - 1);
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const long long r = e % per;
        out[e] = (r >= t0 && r < t1 && strong[e]) ? 1 : 0;
    }
}

// One sweep: a weak, unmarked pixel with a marked linear neighbour (offsets
// +-1, +-d0, +-d0+-1 inside the slice) is marked; *changed records it.
__global__ __launch_bounds__(256) void canny_sweep(unsigned char *__restrict__ out,
                                                   const unsigned char *__restrict__ weak, int d0, int d1,
                                                   long long ns, unsigned *__restrict__ changed) {
    const long long per = (long long)d0 * d1, n = per * ns;
    unsigned ch = 0;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        if (!weak[e] || out[e]) continue;
        const long long sl = e / per, r = e - sl * per;
        const unsigned char *o = out + sl * per;
        const long long nb[8] = {r - d0 - 1, r - d0, r - d0 + 1, r - 1, r + 1, r + d0 - 1, r + d0, r + d0 + 1};
        bool hit = false;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (nb[k] >= 0 && nb[k] < per && o[nb[k]]) hit = true;
        if (hit) {
            out[e] = 1;
            ch = 1;
        }
    }
    if (ch) changed[0] = 1;
}

}  // namespace afhip

using namespace afhip;

extern "C" int afhip_k_canny_grad(float *gx, float *gy, float *mag, const float *smooth, long long d0, long long d1,
                                  long long ns, int is_fast, void *stream) {
    if (d0 < 2 || d1 < 2 || d0 * d1 >= (1LL << 31)) return K_ERR_SIZE;
    canny_grad<<<cgrid(d0 * d1 * ns), 256, 0, (hipStream_t)stream>>>(gx, gy, mag, smooth, (int)d0, (int)d1, ns,
                                                                     is_fast);
    return launch_status();
}

extern "C" int afhip_k_canny_nms(float *out, const float *mag, const float *gx, const float *gy, long long d0,
                                 long long d1, long long ns, void *stream) {
    if (d0 < 2 || d1 < 2 || d0 * d1 >= (1LL << 31)) return K_ERR_SIZE;
    canny_nms<<<cgrid(d0 * d1 * ns), 256, 0, (hipStream_t)stream>>>(out, mag, gx, gy, (int)d0, (int)d1, ns);
    return launch_status();
}

extern "C" int afhip_k_canny_minmax(float *mm, const float *v, long long n, float *part, void *stream) {
    if (n < 1 || part == nullptr) return K_ERR_ARG;
    long long nb = (n + 256LL * 16 - 1) / (256LL * 16);  // about 16 values per lane
    if (nb > kMinMaxBlocks) nb = kMinMaxBlocks;
    canny_minmax_part<<<(unsigned)nb, 256, 0, (hipStream_t)stream>>>(part, v, n);
    canny_minmax_final<<<1, 256, 0, (hipStream_t)stream>>>(mm, part, (int)nb);
    return launch_status();
}

extern "C" int afhip_k_canny_otsu(float *thr, unsigned *hist, const float *v, long long per, long long ns,
                                  int nbins, float maxval, void *stream) {
    if (nbins < 1) return K_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    // hist must be zeroed by the caller (afhip_k_fill)
    const float step = (float)(((double)maxval - 0.0) / (double)(float)nbins);
    canny_hist<<<cgrid(per * ns), 256, 0, s>>>(hist, v, per, ns, nbins, step);
    canny_otsu<<<(unsigned)((ns + 63) / 64), 64, 0, s>>>(thr, hist, nbins, ns, (float)per);
    return launch_status();
}

extern "C" int afhip_k_canny_candidates(unsigned char *strong, unsigned char *weak, const float *v,
                                        const float *mm, const float *thr, long long per, long long ns, float t1,
                                        float t2, int mode, void *stream) {
    canny_candidates<<<cgrid(per * ns), 256, 0, (hipStream_t)stream>>>(strong, weak, v, mm, thr, per, ns, t1, t2,
                                                                       mode);
    return launch_status();
}

extern "C" int afhip_k_canny_seed(unsigned char *out, const unsigned char *strong, long long d0, long long d1,
                                  long long ns, void *stream) {
    canny_seed<<<cgrid(d0 * d1 * ns), 256, 0, (hipStream_t)stream>>>(out, strong, (int)d0, (int)d1, ns);
    return launch_status();
}

extern "C" int afhip_k_canny_sweep(unsigned char *out, const unsigned char *weak, long long d0, long long d1,
                                   long long ns, unsigned *changed, void *stream) {
    canny_sweep<<<cgrid(d0 * d1 * ns), 256, 0, (hipStream_t)stream>>>(out, weak, (int)d0, (int)d1, ns, changed);
    return launch_status();
}
