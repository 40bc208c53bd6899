// FP64 convolve2_nn and its two adjoints, evaluated directly from the closed
// forms (one lane per output element). The reference accepts f64 in
// af_convolve2_nn / af_convolve2_gradient_nn (src/api/c/convolve.cpp:378-381,
// 459-463); this path keeps that dtype working. The FP32 and 16-bit paths use
// the MFMA implicit-GEMM kernels in nn.hip.
#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

struct NN64 {
    int W, H, C, N, fw, fh, K, OW, OH, s0, s1, p0, p1, d0, d1;
};

__global__ __launch_bounds__(256) void nn64_fwd(double *__restrict__ y, const double *__restrict__ x,
                                                const double *__restrict__ w, NN64 p) {
    const long long total = (long long)p.OW * p.OH * p.K * p.N;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e;
        const int ow = (int)(r % p.OW);
        r /= p.OW;
        const int oh = (int)(r % p.OH);
        r /= p.OH;
        const int k = (int)(r % p.K);
        const int n = (int)(r / p.K);
        double acc = 0.0;
        for (int c = 0; c < p.C; ++c)
            for (int j = 0; j < p.fh; ++j) {
                const int ih = oh * p.s1 - p.p1 + j * p.d1;
                if (ih < 0 || ih >= p.H) continue;
                for (int i = 0; i < p.fw; ++i) {
                    const int iw = ow * p.s0 - p.p0 + i * p.d0;
                    if (iw < 0 || iw >= p.W) continue;
                    const double xv = x[(((long long)n * p.C + c) * p.H + ih) * p.W + iw];
                    const double wv =
                        w[(((long long)k * p.C + c) * p.fh + (p.fh - 1 - j)) * p.fw + (p.fw - 1 - i)];
                    acc += xv * wv;
                }
            }
        y[e] = acc;
    }
}

__global__ __launch_bounds__(256) void nn64_dgrad(double *__restrict__ dx, const double *__restrict__ dy,
                                                  const double *__restrict__ w, NN64 p) {
    const long long total = (long long)p.W * p.H * p.C * p.N;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e;
        const int xw = (int)(r % p.W);
        r /= p.W;
        const int xh = (int)(r % p.H);
        r /= p.H;
        const int c = (int)(r % p.C);
        const int n = (int)(r / p.C);
        double acc = 0.0;
        for (int j = 0; j < p.fh; ++j) {
            const int nh = xh + p.p1 - j * p.d1;
            if (nh < 0 || nh % p.s1) continue;
            const int oh = nh / p.s1;
            if (oh >= p.OH) continue;
            for (int i = 0; i < p.fw; ++i) {
                const int nw = xw + p.p0 - i * p.d0;
                if (nw < 0 || nw % p.s0) continue;
                const int ow = nw / p.s0;
                if (ow >= p.OW) continue;
                for (int k = 0; k < p.K; ++k)
                    acc += dy[(((long long)n * p.K + k) * p.OH + oh) * p.OW + ow] *
                           w[(((long long)k * p.C + c) * p.fh + (p.fh - 1 - j)) * p.fw + (p.fw - 1 - i)];
            }
        }
        dx[e] = acc;
    }
}

__global__ __launch_bounds__(256) void nn64_wgrad(double *__restrict__ dw, const double *__restrict__ x,
                                                  const double *__restrict__ dy, NN64 p) {
    const long long total = (long long)p.fw * p.fh * p.C * p.K;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e;
        const int fi = (int)(r % p.fw);  // = fw-1-i
        r /= p.fw;
        const int fj = (int)(r % p.fh);  // = fh-1-j
        r /= p.fh;
        const int c = (int)(r % p.C);
        const int k = (int)(r / p.C);
        const int i = p.fw - 1 - fi, j = p.fh - 1 - fj;
        double acc = 0.0;
        for (int n = 0; n < p.N; ++n)
            for (int oh = 0; oh < p.OH; ++oh) {
                const int ih = oh * p.s1 - p.p1 + j * p.d1;
                if (ih < 0 || ih >= p.H) continue;
                for (int ow = 0; ow < p.OW; ++ow) {
                    const int iw = ow * p.s0 - p.p0 + i * p.d0;
                    if (iw < 0 || iw >= p.W) continue;
                    acc += x[(((long long)n * p.C + c) * p.H + ih) * p.W + iw] *
                           dy[(((long long)n * p.K + k) * p.OH + oh) * p.OW + ow];
                }
            }
        dw[e] = acc;
    }
}

}  // namespace afhip

using namespace afhip;

extern "C" int afhip_k_conv2_nn_f64(int pass, void *out, const void *a0, const void *a1,
                                    const long long *xd, const long long *wd, const long long *yd,
                                    const long long *st, const long long *pd, const long long *dl,
                                    void *stream) {
    NN64 p;
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
    if (p.s0 < 1 || p.s1 < 1 || p.d0 < 1 || p.d1 < 1) return K_ERR_ARG;
    long long total;
    if (pass == 0)
        total = (long long)p.OW * p.OH * p.K * p.N;
    else if (pass == 1)
        total = (long long)p.W * p.H * p.C * p.N;
    else
        total = (long long)p.fw * p.fh * p.C * p.K;
    if (total == 0) return K_OK;
    long long blocks = (total + 255) / 256;
    if (blocks > kMaxGridBlocks) blocks = kMaxGridBlocks;
    hipStream_t s = (hipStream_t)stream;
    if (pass == 0)
        nn64_fwd<<<(unsigned)blocks, 256, 0, s>>>((double *)out, (const double *)a0, (const double *)a1, p);
    else if (pass == 1)
        nn64_dgrad<<<(unsigned)blocks, 256, 0, s>>>((double *)out, (const double *)a0, (const double *)a1, p);
    else if (pass == 2)
        nn64_wgrad<<<(unsigned)blocks, 256, 0, s>>>((double *)out, (const double *)a0, (const double *)a1, p);
    else
        return K_ERR_ARG;
    return launch_status();
}
