// Strided data gradient of convolve2_nn by polyphase decomposition
// (conv2DataGradient, src/backend/cpu/convolve.cpp:179-211, for stride > 1
// and dilation 1).
//
// Per axis, with filter length f, stride s and padding p, the forward pass is
// y[o] = sum_i x[s*o - p + i] * w[f-1-i]. Its adjoint at an input position
// u = s*m + r (residue r < s) only meets the taps i = i0 + s*t with
// i0 = (r + p) mod s, at o = m + c0 - t with c0 = (r + p - i0) / s:
//     dx[s*m + r] = sum_t dy[m + c0 - t] * w[f-1-i0-s*t],   t < T_r.
// That is a stride-1 convolution of dy with the T_r-tap filter
// g[t] = w[f-1-i0-s*t] padded by T_r - 1 ("full" output), read at m + c0.
// In 2-D each of the s0*s1 residue pairs is one such convolution with a
// Tx x Ty sub-filter whose channel dims are transposed (dy's K channels in,
// dx's C channels out): G[tx, ty, k, c] = w[fw-1-i0x-s0*tx, fh-1-i0y-s1*ty, c, k].
// The phases run on the patch-staged forward kernels (afhip_k_conv2_nn_patch,
// pass 0) one after the other, each followed by a scatter of its cropped
// output into the interleaved dx. Every dx element belongs to exactly one
// phase and is that phase's complete sum, rounded once, as in the reference.
#include <algorithm>
#include <cstdlib>

#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

struct PolyAxis {
    int T, i0, c0;  // taps, first tap, dy offset
};

static PolyAxis poly_axis(long long f, long long s, long long p, long long r) {
    PolyAxis a;
    a.i0 = (int)((r + p) % s);
    a.c0 = (int)((r + p - a.i0) / s);
    a.T = a.i0 < f ? (int)((f - 1 - a.i0) / s + 1) : 0;
    return a;
}

template <typename E>
__global__ __launch_bounds__(256) void poly_weights(E *__restrict__ G, const E *__restrict__ w, int fw, int fh,
                                                    int C, int K, int Tx, int Ty, int i0x, int i0y, int sx, int sy) {
    // G [Tx, Ty, K, C] <- w [fw, fh, C, K] at (fw-1-i0x-sx*tx, fh-1-i0y-sy*ty, c, k)
    const long long total = (long long)Tx * Ty * K * C;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e;
        const int tx = (int)(r % Tx);
        r /= Tx;
        const int ty = (int)(r % Ty);
        r /= Ty;
        const int k = (int)(r % K);
        const int c = (int)(r / K);
        const int ix = fw - 1 - i0x - sx * tx, iy = fh - 1 - i0y - sy * ty;
        G[e] = w[(((long long)k * C + c) * fh + iy) * fw + ix];
    }
}

template <typename E>
__global__ __launch_bounds__(256) void poly_scatter(E *__restrict__ dx, const E *__restrict__ yp, int W, int H,
                                                    long long CN, int Mx, int My, int PWp, int PHp, int sx, int sy,
                                                    int rx, int ry, int c0x, int c0y) {
    // dx [W, H, C*N] at (sx*m + rx, sy*n + ry) <- yp [PWp, PHp, C*N] at (m + c0x, n + c0y)
    const long long total = (long long)Mx * My * CN;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e;
        const int m = (int)(r % Mx);
        r /= Mx;
        const int n = (int)(r % My);
        const long long cn = r / My;
        // positions past the full convolution's output meet no output
        // element of the forward pass: their gradient is zero
        const int mm = m + c0x, nn = n + c0y;
        dx[(cn * H + (sy * n + ry)) * W + (sx * m + rx)] =
            (mm < PWp && nn < PHp) ? yp[(cn * PHp + nn) * PWp + mm] : E{};
    }
}

template <typename E>
__global__ __launch_bounds__(256) void poly_zero(E *__restrict__ dx, int W, int H, long long CN, int Mx, int My,
                                                 int sx, int sy, int rx, int ry) {
    const long long total = (long long)Mx * My * CN;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e;
        const int m = (int)(r % Mx);
        r /= Mx;
        const int n = (int)(r % My);
        const long long cn = r / My;
        dx[(cn * H + (sy * n + ry)) * W + (sx * m + rx)] = E{};
    }
}

static unsigned grid_of(long long n) {
    long long b = (n + 255) / 256;
    if (b > kMaxGridBlocks) b = kMaxGridBlocks;
    return (unsigned)(b < 1 ? 1 : b);
}

static int esize(int t) {
    switch (t) {
        case DT_F32: return 4;
        case DT_BF16: case DT_F16: return 2;
        default: return 0;
    }
}

static long long align256(long long b) { return (b + 255) & ~255LL; }

// Byte sizes of the three workspace parts (largest over the phases): the
// sub-filter, the forward kernel's packed weights, the phase output.
struct PolyWs {
    long long g, pack, yp, total;
};

// The phases run one after the other, each on workgroups of 128 output
// channels x 256 pixels of one image: a layer whose phases give fewer of
// those than there are CUs leaves the chip mostly idle and is served faster
// by the gather kernel, which the caller then uses.
static int cu_count() {
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

static bool poly_plan(int in_type, int out_type, int precision, const long long *xd, const long long *wd,
                      const long long *yd, const long long *st, const long long *pd, const long long *dl,
                      PolyWs &ws) {
    if (dl[0] != 1 || dl[1] != 1 || (st[0] == 1 && st[1] == 1) || st[0] < 1 || st[1] < 1) return false;
    if (pd[0] < 0 || pd[1] < 0) return false;
    const long long tiles = ((xd[2] + 127) / 128) * ((yd[0] * yd[1] + 255) / 256) * yd[3];
    const char *force = std::getenv("AFHIP_NN_POLY");
    if (!(force && force[0] == '1') && tiles < cu_count()) return false;
    const int ei = esize(in_type), eo = esize(out_type);
    if (!ei || !eo) return false;
    ws = PolyWs{0, 0, 0, 0};
    for (long long ry = 0; ry < st[1]; ++ry)
        for (long long rx = 0; rx < st[0]; ++rx) {
            const PolyAxis ax = poly_axis(wd[0], st[0], pd[0], rx), ay = poly_axis(wd[1], st[1], pd[1], ry);
            if (ax.T == 0 || ay.T == 0) continue;
            // phase forward: x' = dy [OW, OH, K, N], w' = G [Tx, Ty, K, C], pad (T-1), stride 1
            const long long xp[4] = {yd[0], yd[1], yd[2], yd[3]};
            const long long gp[4] = {ax.T, ay.T, wd[3], wd[2]};
            const long long ypd[4] = {yd[0] + ax.T - 1, yd[1] + ay.T - 1, wd[2], yd[3]};
            const long long one[2] = {1, 1}, pad[2] = {ax.T - 1, ay.T - 1};
            const long long pk = afhip_k_conv2_nn_patch_workspace(0, in_type, precision, xp, gp, ypd, one, pad, one);
            if (pk <= 0) return false;
            ws.g = std::max(ws.g, align256(gp[0] * gp[1] * gp[2] * gp[3] * ei));
            ws.pack = std::max(ws.pack, align256(pk));
            ws.yp = std::max(ws.yp, align256(ypd[0] * ypd[1] * ypd[2] * ypd[3] * eo));
        }
    ws.total = ws.g + ws.pack + ws.yp;
    return ws.total > 0;
}

template <typename E>
static int run_poly(int in_type, int out_type, int precision, void *dx, const void *dy, const void *w,
                    const long long *xd, const long long *wd, const long long *yd, const long long *st,
                    const long long *pd, char *ws, const PolyWs &sz, hipStream_t s) {
    char *G = ws, *pack = ws + sz.g;
    void *yp = ws + sz.g + sz.pack;
    const long long CN = xd[2] * xd[3];
    for (long long ry = 0; ry < st[1]; ++ry)
        for (long long rx = 0; rx < st[0]; ++rx) {
            const PolyAxis ax = poly_axis(wd[0], st[0], pd[0], rx), ay = poly_axis(wd[1], st[1], pd[1], ry);
            const int Mx = (int)((xd[0] - rx + st[0] - 1) / st[0]), My = (int)((xd[1] - ry + st[1] - 1) / st[1]);
            if (Mx <= 0 || My <= 0) continue;
            const long long nout = (long long)Mx * My * CN;
            if (ax.T == 0 || ay.T == 0) {   // no tap reaches these positions
                if (out_type == DT_F32)
                    poly_zero<float><<<grid_of(nout), 256, 0, s>>>((float *)dx, (int)xd[0], (int)xd[1], CN, Mx, My,
                                                                    (int)st[0], (int)st[1], (int)rx, (int)ry);
                else
                    poly_zero<uint16_t><<<grid_of(nout), 256, 0, s>>>((uint16_t *)dx, (int)xd[0], (int)xd[1], CN,
                                                                       Mx, My, (int)st[0], (int)st[1], (int)rx,
                                                                       (int)ry);
                continue;
            }
            const long long ng = (long long)ax.T * ay.T * wd[3] * wd[2];
            poly_weights<E><<<grid_of(ng), 256, 0, s>>>((E *)G, (const E *)w, (int)wd[0], (int)wd[1], (int)wd[2],
                                                        (int)wd[3], ax.T, ay.T, ax.i0, ay.i0, (int)st[0], (int)st[1]);
            int rc = launch_status();
            if (rc) return rc;
            const long long xp[4] = {yd[0], yd[1], yd[2], yd[3]};
            const long long gp[4] = {ax.T, ay.T, wd[3], wd[2]};
            const long long ypd[4] = {yd[0] + ax.T - 1, yd[1] + ay.T - 1, wd[2], yd[3]};
            const long long one[2] = {1, 1}, pad[2] = {ax.T - 1, ay.T - 1};
            rc = afhip_k_conv2_nn_patch(0, in_type, out_type, precision, yp, dy, G, xp, gp, ypd, one, pad, one, pack,
                                        sz.pack, s);
            if (rc) return rc;
            if (out_type == DT_F32)
                poly_scatter<float><<<grid_of(nout), 256, 0, s>>>((float *)dx, (const float *)yp, (int)xd[0],
                                                                   (int)xd[1], CN, Mx, My, (int)ypd[0], (int)ypd[1],
                                                                   (int)st[0], (int)st[1], (int)rx, (int)ry, ax.c0,
                                                                   ay.c0);
            else
                poly_scatter<uint16_t><<<grid_of(nout), 256, 0, s>>>(
                    (uint16_t *)dx, (const uint16_t *)yp, (int)xd[0], (int)xd[1], CN, Mx, My, (int)ypd[0],
                    (int)ypd[1], (int)st[0], (int)st[1], (int)rx, (int)ry, ax.c0, ay.c0);
            rc = launch_status();
            if (rc) return rc;
        }
    return K_OK;
}

}  // namespace afhip

using namespace afhip;

extern "C" long long afhip_k_conv2_nn_dgrad_poly_workspace(int in_type, int out_type, int precision,
                                                          const long long *xd, const long long *wd,
                                                          const long long *yd, const long long *st,
                                                          const long long *pd, const long long *dl) {
    PolyWs ws;
    if (!poly_plan(in_type, out_type, precision, xd, wd, yd, st, pd, dl, ws)) return 0;
    return ws.total;
}

extern "C" int afhip_k_conv2_nn_dgrad_poly(int in_type, int out_type, int precision, void *dx, const void *dy,
                                           const void *w, const long long *xd, const long long *wd,
                                           const long long *yd, const long long *st, const long long *pd,
                                           const long long *dl, void *workspace, long long workspace_bytes,
                                           void *stream) {
    PolyWs ws;
    if (!poly_plan(in_type, out_type, precision, xd, wd, yd, st, pd, dl, ws)) return K_ERR_NOT_SUPPORTED;
    if (workspace == nullptr || workspace_bytes < ws.total) return K_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (esize(in_type) == 4)
        return run_poly<float>(in_type, out_type, precision, dx, dy, w, xd, wd, yd, st, pd, (char *)workspace, ws, s);
    return run_poly<uint16_t>(in_type, out_type, precision, dx, dy, w, xd, wd, yd, st, pd, (char *)workspace, ws, s);
}
