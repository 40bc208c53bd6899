/*
 * Kernel-level C ABI of libafhip.so: raw device pointers, int64 dims in
 * ArrayFire's column-major order (dim0 fastest, contiguous strides), an
 * explicit HIP stream (hipStream_t passed as void*), and an af_err return
 * code (0 on success; 202 bad argument, 203 bad size, 204 unsupported dtype,
 * 103 HIP launch failure). Nothing here allocates except where a workspace
 * is documented; no exception crosses it.
 *
 * These are the entry points the af_* dispatch (the headers under include/af) calls after
 * it has validated its arguments and computed output dims. They are also
 * what a host that manages its own device memory (the Python test wrapper,
 * a torch.autograd.Function) binds directly. Each replaces one backend
 * template of the reference, cited below.
 */
#pragma once

#ifdef __cplusplus
#define AFHIP_KAPI extern "C" __attribute__((visibility("default")))
#else
#define AFHIP_KAPI __attribute__((visibility("default")))
#endif

/* detail::convolve<T, accT> (src/backend/cpu/convolve.hpp:16-18;
 * src/backend/cpu/convolve.cpp:38-66 and kernel/convolve.hpp:132-211).
 * dtype is the signal's af_dtype; filt holds accT elements (float for f32 and
 * the integer types, double for f64, complex float/double for c32/c64).
 * kind is the AF_BATCH_KIND value (NONE=1, LHS=2, RHS=3, SAME=4); odims must
 * be the output dims the reference computes for (sdims, fdims, kind, rank,
 * expand). */
AFHIP_KAPI int afhip_k_convolve(int dtype, void *out, const long long *odims,
                                const void *in, const long long *sdims,
                                const void *filt, const long long *fdims,
                                int kind, int rank, int expand, void *stream);

/* Longest column or row filter of the fused single-kernel FP32 separable
 * path; longer filters (and every other dtype) need the tmp buffer. */
#define AFHIP_SEP_FUSED_MAX_TAPS 33

/* detail::convolve2<T, accT> separable (src/backend/cpu/convolve.hpp:20-22;
 * src/backend/cpu/convolve.cpp:68-99, kernel/convolve.hpp:213-289).
 * col/row filters hold cflen/rflen accT elements; odims = sdims, plus
 * cflen-1 on dim0 and rflen-1 on dim1 when expand. tmp, when not NULL, is a
 * device buffer of sdims[0]+(expand?cflen-1:0) x sdims[1] x sdims[2] x
 * sdims[3] elements that receives the column pass (the reference's T-typed
 * temp); when NULL the two passes are fused in one kernel and the
 * intermediate is rounded to T in registers without being stored. */
AFHIP_KAPI int afhip_k_convolve2_sep(int dtype, void *out, const long long *odims,
                                     const void *in, const long long *sdims,
                                     const void *cfilt, long long cflen,
                                     const void *rfilt, long long rflen,
                                     int expand, void *tmp, void *stream);

/* detail::fftconvolve<T> (src/backend/cpu/fftconvolve.hpp:16-19,
 * src/backend/cpu/fftconvolve.cpp:42-203) and the DIFF / complex fallback of
 * src/api/c/fftconvolve.cpp:52-121: frequency-domain convolution of rank 1-3.
 * dtype is the signal's af_dtype; ftype the filter's (the batched path takes
 * the filter already converted to the signal's type, as castArray<T> does;
 * the fallback any type). kind is the AF_BATCH_KIND (NONE=1 .. DIFF=5); odims
 * the reference's output dims (for DIFF, max(s, f) past the rank). The
 * workspace holds the padded spectra: query its size first (-1: a padded
 * axis above 2^24). Integer outputs are rounded to nearest (batched kinds) or
 * truncated (DIFF), as the reference does. */
AFHIP_KAPI long long afhip_k_fft_convolve_workspace(int dtype, const long long *sdims,
                                                    const long long *fdims,
                                                    const long long *odims, int kind,
                                                    int rank);
AFHIP_KAPI int afhip_k_fft_convolve(int dtype, void *out, const long long *odims,
                                    const void *in, const long long *sdims, int ftype,
                                    const void *filt, const long long *fdims, int kind,
                                    int rank, int expand, void *workspace,
                                    long long workspace_bytes, void *stream);

/* Element type conversion between af_dtype values (and 64 = bfloat16) over n
 * contiguous elements. */
AFHIP_KAPI int afhip_k_cast(int out_type, void *out, int in_type,
                            const void *in, long long n, void *stream);

/* Fills n elements of dtype with value. */
AFHIP_KAPI int afhip_k_fill(int dtype, void *out, double value, long long n,
                            void *stream);

/* y = T(accT(x) * scale) elementwise: the scalar/scalar case of the separable
 * convolution (src/api/c/convolve.cpp:57-68). */
AFHIP_KAPI int afhip_k_scale(int dtype, void *out, const void *in,
                             double scale, long long n, void *stream);

/* y = x * (re + i im) elementwise for complex x (c32 or c64, the product
 * formed in the signal's precision with each real product rounded before the
 * sum): the scalar/scalar separable case for complex signals
 * (src/api/c/convolve.cpp:57-68 with accT = cfloat/cdouble). */
AFHIP_KAPI int afhip_k_scale_complex(int dtype, void *out, const void *in,
                                     double re, double im, long long n,
                                     void *stream);

/* Precision codes below are afhip_precision values (include/af/hip.h):
 * 0 default (by in_type), 1 BF16 MFMA, 2 FP16 MFMA, 3 split-BF16, 4 FP32.
 * in_type is f32 (0), f16 (12) or bfloat16 (64); out_type f32, or the 16-bit
 * input type. Dims: x [W,H,C,N], w [fw,fh,C,K], y [OW,OH,K,N]; stride, pad and
 * dil hold two entries (dim0 = W, dim1 = H). */

/* detail::convolve2<T> (src/backend/cpu/convolve.hpp:24-26;
 * src/backend/cpu/convolve.cpp:123-167): y = conv2_nn(x, w). */
AFHIP_KAPI int afhip_k_conv2_nn_fwd(int in_type, int out_type, int precision,
                                    void *y, const void *x, const void *w,
                                    const long long *xdims,
                                    const long long *wdims,
                                    const long long *ydims,
                                    const long long *stride,
                                    const long long *pad, const long long *dil,
                                    void *stream);

/* detail::conv2DataGradient<T> (src/backend/cpu/convolve.hpp:28-33;
 * src/backend/cpu/convolve.cpp:179-211): dx [W,H,C,N] from dy [OW,OH,K,N]. */
AFHIP_KAPI int afhip_k_conv2_nn_dgrad(int in_type, int out_type, int precision,
                                      void *dx, const void *dy, const void *w,
                                      const long long *xdims,
                                      const long long *wdims,
                                      const long long *ydims,
                                      const long long *stride,
                                      const long long *pad,
                                      const long long *dil, void *stream);

/* The gather GEMM behind afhip_k_conv2_nn_fwd (pass 0: in = x, out = y) and
 * afhip_k_conv2_nn_dgrad (pass 1: in = dy, out = dx), for any geometry, with
 * a workspace: operands stored at their own precision are first copied
 * channel-last (weights packed per tap, the activation NCHW -> NHWC) so that
 * every gather is a run of 16-byte loads; layers whose output tiles leave
 * most CUs idle split the reduction into FP32 partial tiles that a second
 * pass sums in a fixed order. Strided data gradients visit only the taps of
 * each stride phase. The query returns the workspace bytes (0: none needed);
 * a NULL or short workspace runs unsplit and without the channel-last
 * copies. Same dims and codes as afhip_k_conv2_nn_fwd/_dgrad
 * (src/backend/cpu/convolve.cpp:123-211). */
AFHIP_KAPI long long afhip_k_conv2_nn_gemm_workspace(
    int pass, int in_type, int precision, const long long *xdims,
    const long long *wdims, const long long *ydims, const long long *stride,
    const long long *pad, const long long *dil);
AFHIP_KAPI int afhip_k_conv2_nn_gemm(int pass, int in_type, int out_type,
                                     int precision, void *out, const void *in,
                                     const void *w, const long long *xdims,
                                     const long long *wdims,
                                     const long long *ydims,
                                     const long long *stride,
                                     const long long *pad,
                                     const long long *dil, void *workspace,
                                     long long workspace_bytes, void *stream);

/* afhip_k_conv2_nn_gemm with the reduction split nsplit ways (nsplit <= 0:
 * the built-in rule), for the autotuner; the query sizes that workspace. */
AFHIP_KAPI long long afhip_k_conv2_nn_gemm_split_workspace(
    int pass, int in_type, int precision, const long long *xdims,
    const long long *wdims, const long long *ydims, const long long *stride,
    const long long *pad, const long long *dil, int nsplit);
AFHIP_KAPI int afhip_k_conv2_nn_gemm_split(int pass, int in_type, int out_type,
                                           int precision, void *out,
                                           const void *in, const void *w,
                                           const long long *xdims,
                                           const long long *wdims,
                                           const long long *ydims,
                                           const long long *stride,
                                           const long long *pad,
                                           const long long *dil, int nsplit,
                                           void *workspace,
                                           long long workspace_bytes,
                                           void *stream);

/* Workspace bytes afhip_k_conv2_nn_wgrad uses for its split reduction over
 * N*OH*OW (0 when it runs unsplit). */
AFHIP_KAPI long long afhip_k_conv2_nn_wgrad_workspace(
    const long long *xdims, const long long *wdims, const long long *ydims,
    const long long *stride, const long long *pad, const long long *dil);

/* detail::conv2FilterGradient<T> (src/backend/cpu/convolve.hpp:35-40;
 * src/backend/cpu/convolve.cpp:213-243): dw [fw,fh,C,K] from x and dy.
 * workspace may be NULL (or smaller than the query): the reduction then runs
 * unsplit, which is correct but slower. */
AFHIP_KAPI int afhip_k_conv2_nn_wgrad(int in_type, int out_type, int precision,
                                      void *dw, const void *x, const void *dy,
                                      const long long *xdims,
                                      const long long *wdims,
                                      const long long *ydims,
                                      const long long *stride,
                                      const long long *pad,
                                      const long long *dil, void *workspace,
                                      long long workspace_bytes, void *stream);

/* FP64 forward (pass 0), data gradient (pass 1) and filter gradient (pass 2)
 * of convolve2_nn, evaluated directly from the closed forms; the f64 cases of
 * the three templates above. Operands: pass 0 (x, w) -> y; pass 1 (dy, w) ->
 * dx; pass 2 (x, dy) -> dw. */
AFHIP_KAPI int afhip_k_conv2_nn_f64(int pass, void *out, const void *a0,
                                    const void *a1, const long long *xdims,
                                    const long long *wdims,
                                    const long long *ydims,
                                    const long long *stride,
                                    const long long *pad, const long long *dil,
                                    void *stream);

/* Patch-staged implicit GEMM for the forward pass (pass 0) and the stride-1
 * data gradient (pass 1), the fast path behind afhip_k_conv2_nn_fwd/_dgrad.
 * The workspace query returns the bytes of packed weights the call needs, or
 * 0 when the geometry is not served (stride != 1 for pass 1, or an input
 * patch that does not fit in LDS); the call then returns 301 and the caller
 * uses afhip_k_conv2_nn_fwd/_dgrad. For pass 0 `in` is x and `out` is y; for
 * pass 1 `in` is dy and `out` is dx. */
AFHIP_KAPI long long afhip_k_conv2_nn_patch_workspace(
    int pass, int in_type, int precision, const long long *xdims,
    const long long *wdims, const long long *ydims, const long long *stride,
    const long long *pad, const long long *dil);
/* Workgroups the patch path would launch for this geometry (its persistent
 * kernels: at most the tile count), or 0 when it does not serve it. The
 * dispatch sends layers with too few tiles to fill the chip to the split
 * gather GEMM (afhip_k_conv2_nn_gemm). */
AFHIP_KAPI long long afhip_k_conv2_nn_patch_tiles(
    int pass, int in_type, int precision, const long long *xdims,
    const long long *wdims, const long long *ydims, const long long *stride,
    const long long *pad, const long long *dil);
AFHIP_KAPI int afhip_k_conv2_nn_patch(int pass, int in_type, int out_type,
                                      int precision, void *out, const void *in,
                                      const void *w, const long long *xdims,
                                      const long long *wdims,
                                      const long long *ydims,
                                      const long long *stride,
                                      const long long *pad,
                                      const long long *dil, void *workspace,
                                      long long workspace_bytes, void *stream);

/* No counterpart in the reference: afhip_k_conv2_nn_patch without the
 * workspace arguments and without any allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The packed weights go to a per-GPU static device
 * array of afhip_k_conv2_nn_patch_noalloc_limit() bytes (16 MiB); a layer
 * whose workspace query exceeds it returns 203. Calls that overlap on one
 * GPU share that array and race; calls in order on one stream do not. */
AFHIP_KAPI long long afhip_k_conv2_nn_patch_noalloc_limit(void);
AFHIP_KAPI int afhip_k_conv2_nn_patch_noalloc(int pass, int in_type, int out_type,
                                              int precision, void *out,
                                              const void *in, const void *w,
                                              const long long *xdims,
                                              const long long *wdims,
                                              const long long *ydims,
                                              const long long *stride,
                                              const long long *pad,
                                              const long long *dil, void *stream);

/* Row-staged filter gradient for 16-bit operands (bfloat16 or binary16
 * storage, FP32 accumulation): the fast path behind afhip_k_conv2_nn_wgrad
 * for filters of width 1 or 3, dilation 1 along W, stride 1 or 2 along W and
 * at most 9 taps. The workspace holds the FP32 partial tiles of the split
 * reduction; the query returns 0 when the geometry or dtype is not served
 * (the call then returns 301). */
AFHIP_KAPI long long afhip_k_conv2_nn_wgrad_rows_workspace(
    int in_type, int precision, const long long *xdims, const long long *wdims,
    const long long *ydims, const long long *stride, const long long *pad,
    const long long *dil);
AFHIP_KAPI int afhip_k_conv2_nn_wgrad_rows(int in_type, int out_type,
                                           int precision, void *dw,
                                           const void *x, const void *dy,
                                           const long long *xdims,
                                           const long long *wdims,
                                           const long long *ydims,
                                           const long long *stride,
                                           const long long *pad,
                                           const long long *dil,
                                           void *workspace,
                                           long long workspace_bytes,
                                           void *stream);

/* Ragged batch of FP32 2-D spatial convolutions (afhip_ragged_plan_* in
 * include/af/hip.h): problems is a device array of per-problem descriptors
 * {const float *in; float *out; const float *filt; int S0, S1, O0, O1, FX,
 * FY, sx, sy, nbx, nby;} and items a device array of nitems (problem, tile)
 * int pairs; grid persistent workgroups process items round-robin. Filters
 * up to 15x15. afhip_k_ragged_tile returns the output-tile edges of a
 * variant (nbx/nby are the problem's tile counts for that variant). */
AFHIP_KAPI int afhip_k_ragged_tile(int variant, int *bx, int *by);
AFHIP_KAPI int afhip_k_convolve2_ragged(int variant, const void *problems,
                                        const void *items, long long nitems,
                                        int grid, void *stream);

/* The two launches of afhip_k_conv2_nn_patch, separately: _pack writes the
 * packed weights into the workspace, _run convolves with them (the workspace
 * may be reused across _run calls while the weights are unchanged). */
AFHIP_KAPI int afhip_k_conv2_nn_patch_pack(int pass, int in_type, int precision,
                                           const void *w, const long long *xdims,
                                           const long long *wdims,
                                           const long long *ydims,
                                           const long long *stride,
                                           const long long *pad,
                                           const long long *dil, void *workspace,
                                           long long workspace_bytes,
                                           void *stream);
AFHIP_KAPI int afhip_k_conv2_nn_patch_run(int pass, int in_type, int out_type,
                                          int precision, void *out,
                                          const void *in,
                                          const long long *xdims,
                                          const long long *wdims,
                                          const long long *ydims,
                                          const long long *stride,
                                          const long long *pad,
                                          const long long *dil,
                                          const void *workspace,
                                          long long workspace_bytes, void *stream);

/* The two launches of afhip_k_conv2_nn_wgrad_rows, separately: _partial
 * writes the FP32 partial tiles into the workspace, _reduce sums them into
 * dw in a fixed order. */
AFHIP_KAPI int afhip_k_conv2_nn_wgrad_rows_partial(
    int in_type, int precision, const void *x, const void *dy,
    const long long *xdims, const long long *wdims, const long long *ydims,
    const long long *stride, const long long *pad, const long long *dil,
    void *workspace, long long workspace_bytes, void *stream);
AFHIP_KAPI int afhip_k_conv2_nn_wgrad_rows_reduce(
    int in_type, int out_type, int precision, void *dw, const long long *xdims,
    const long long *wdims, const long long *ydims, const long long *stride,
    const long long *pad, const long long *dil, const void *workspace,
    long long workspace_bytes, void *stream);

/* unwrap<T> (src/backend/cpu/unwrap.cpp:20-40, kernel/unwrap.hpp:20-80):
 * im2col with strides, zero padding and dilation (dx, dy); odims must be
 * [wx*wy, nx*ny, C, N] (is_column) or [nx*ny, wx*wy, C, N]. Any dtype (the
 * elements are moved as bytes). */
AFHIP_KAPI int afhip_k_unwrap(int dtype, void *out, const long long *odims,
                              const void *in, const long long *idims,
                              long long wx, long long wy, long long sx,
                              long long sy, long long px, long long py,
                              long long dx, long long dy, int is_column,
                              void *stream);

/* wrap<T> (src/backend/cpu/wrap.cpp:20-40, kernel/wrap.hpp:20-80): col2im.
 * out [ox, oy, C, N] receives (accumulate ? out : 0) plus every patch
 * element that lands on it, added in the reference's patch order in the
 * element type. Every dtype but f16 / bf16. */
AFHIP_KAPI int afhip_k_wrap(int dtype, void *out, const long long *odims,
                            const void *in, const long long *idims,
                            long long wx, long long wy, long long sx,
                            long long sy, long long px, long long py,
                            int is_column, int accumulate, void *stream);

/* af_index of a region (src/api/c/index.cpp), as af_fir uses it: copies the
 * odims-sized block at `offsets` of a contiguous idims array into a
 * contiguous out. esize is the element size in bytes (1, 2, 4, 8 or 16). */
AFHIP_KAPI int afhip_k_region_copy(int esize, void *out, const long long *odims,
                                   const void *in, const long long *idims,
                                   const long long *offsets, void *stream);

/* out = a - b over n elements of f32 or f64 (arithOp<af_sub_t>, as af_dog
 * uses it, src/api/c/dog.cpp:50). */
AFHIP_KAPI int afhip_k_sub(int dtype, void *out, const void *a, const void *b,
                           long long n, void *stream);

/* out = b / a with a broadcast along b's columns: a holds one element per
 * column (per_col) or one in total (af_div(b, a, true) in af_iir,
 * src/api/c/iir.cpp:76-81). f32, f64, c32, c64. */
AFHIP_KAPI int afhip_k_div_cols(int dtype, void *out, const void *b,
                                const long long *bdims, const void *a,
                                int per_col, void *stream);

/* kernel::iir<T> (src/backend/cpu/kernel/iir.hpp:20-55): the feedback
 * recurrence over the first ydims[0] rows of each column of c (the EXPAND
 * convolution of x with b, dims cdims); a is [na] (one for all columns) or
 * [na, ydims[1], ydims[2], ydims[3]]. f32, f64, c32, c64. More than 16
 * feedback taps need a workspace of afhip_k_iir_workspace bytes. */
AFHIP_KAPI long long afhip_k_iir_workspace(int dtype, long long na,
                                           const long long *ydims);
AFHIP_KAPI int afhip_k_iir(int dtype, void *y, const long long *ydims,
                           const void *c, const long long *cdims, const void *a,
                           const long long *adims, void *workspace,
                           long long workspace_bytes, void *stream);

/* conv2DataGradient<T> for stride > 1 and dilation 1
 * (src/backend/cpu/convolve.cpp:179-211) by polyphase decomposition: one
 * stride-1 convolution of dy per residue pair (u mod s0, v mod s1) with the
 * filter taps that reach those positions (channel dims transposed), on the
 * patch-staged forward kernels, each scattered into the interleaved dx. The
 * workspace holds one phase's sub-filter, packed weights and output; the
 * query returns 0 when the geometry is not served (stride 1, dilation > 1,
 * a phase the patch kernels do not take), and the call then returns 301. */
AFHIP_KAPI long long afhip_k_conv2_nn_dgrad_poly_workspace(
    int in_type, int out_type, int precision, const long long *xdims,
    const long long *wdims, const long long *ydims, const long long *stride,
    const long long *pad, const long long *dil);
AFHIP_KAPI int afhip_k_conv2_nn_dgrad_poly(int in_type, int out_type,
                                           int precision, void *dx,
                                           const void *dy, const void *w,
                                           const long long *xdims,
                                           const long long *wdims,
                                           const long long *ydims,
                                           const long long *stride,
                                           const long long *pad,
                                           const long long *dil,
                                           void *workspace,
                                           long long workspace_bytes,
                                           void *stream);

/* Harris corner detector kernels (af_harris, src/api/c/harris.cpp:52-95;
 * src/backend/cpu/harris.cpp:26-148; src/backend/cpu/kernel/harris.hpp).
 * dtype f32 or f64; W = dim0, H = dim1 of the image.
 * deriv: d3 [W,H,3] <- (ix*ix, ix*iy, iy*iy), iy/ix the gradients along
 * dim0/dim1 (src/backend/cpu/kernel/gradient.hpp:18-80). */
AFHIP_KAPI int afhip_k_harris_deriv(int dtype, void *d3, const void *in,
                                    long long W, long long H, void *stream);
/* nms: flags[i] = 1 where the response det - k*tr^2 of the smoothed s3
 * products exceeds its 8 neighbours and is >= min_r, for pixels at least
 * `border` from every edge; resp[i] = the response as float there. */
AFHIP_KAPI int afhip_k_harris_nms(int dtype, unsigned char *flags,
                                  float *resp, const void *s3, long long W,
                                  long long H, float k_thr, float min_r,
                                  int border, void *stream);
/* Stable compaction of the flagged indices (ascending) into out_idx, their
 * count into *d_num (device); the workspace query returns its bytes. */
AFHIP_KAPI long long afhip_k_select_flagged_workspace(long long n);
AFHIP_KAPI int afhip_k_select_flagged(const unsigned char *flags, long long n,
                                      unsigned *out_idx, unsigned *d_num,
                                      void *temp, long long temp_bytes,
                                      void *stream);
/* Stable descending sort of resp[sel[i]], i < m (keys/vals: m-entry
 * buffers; the sorted keys and positions go to keys_sorted/vals_sorted). */
AFHIP_KAPI long long afhip_k_sort_desc_workspace(long long m);
AFHIP_KAPI int afhip_k_harris_sort(const unsigned *sel, const float *resp,
                                   long long m, float *keys, unsigned *vals,
                                   float *keys_sorted, unsigned *vals_sorted,
                                   void *temp, long long temp_bytes,
                                   void *stream);
/* The first n features: x = sel[i] / W, y = sel[i] % W, score, orientation
 * 0, size 1, with i = order[f] (or f when order is NULL). */
AFHIP_KAPI int afhip_k_harris_out(float *x, float *y, float *score,
                                  float *orient, float *size,
                                  const unsigned *sel, const unsigned *order,
                                  const float *sorted_keys, const float *resp,
                                  long long W, long long n, void *stream);

/* Canny edge detector kernels (af_canny, src/api/c/canny.cpp:198-285;
 * src/backend/cpu/kernel/sobel.hpp, canny.hpp, histogram.hpp). Images are
 * ns slices of d0 x d1 f32 values. grad: Sobel dx (dim0) / dy (dim1), reflect101
 * borders, and the L1 (is_fast) or L2 magnitude. nms: the suppressed map. */
AFHIP_KAPI int afhip_k_canny_grad(float *gx, float *gy, float *mag,
                                  const float *smooth, long long d0,
                                  long long d1, long long ns, int is_fast,
                                  void *stream);
AFHIP_KAPI int afhip_k_canny_nms(float *out, const float *mag, const float *gx,
                                 const float *gy, long long d0, long long d1,
                                 long long ns, void *stream);
/* mm[0] = min, mm[1] = max of n values; part: 2048 floats of scratch. */
AFHIP_KAPI int afhip_k_canny_minmax(float *mm, const float *v, long long n,
                                    float *part, void *stream);
/* Per slice: the histogram of nbins over [0, maxval] into hist (zeroed by
 * the caller) and Otsu's threshold (a bin index) into thr. */
AFHIP_KAPI int afhip_k_canny_otsu(float *thr, unsigned *hist, const float *v,
                                  long long per, long long ns, int nbins,
                                  float maxval, void *stream);
/* Strong / weak maps: mode 0 normalises by mm and compares with t1, t2;
 * mode 1 compares with thr[slice] * t1 and thr[slice]. */
AFHIP_KAPI int afhip_k_canny_candidates(unsigned char *strong,
                                        unsigned char *weak, const float *v,
                                        const float *mm, const float *thr,
                                        long long per, long long ns, float t1,
                                        float t2, int mode, void *stream);
/* Hysteresis: seeds (strong pixels at the reference's scan positions), then
 * sweeps marking weak pixels next to marked ones; *changed is set by a
 * sweep that marks anything. */
AFHIP_KAPI int afhip_k_canny_seed(unsigned char *out,
                                  const unsigned char *strong, long long d0,
                                  long long d1, long long ns, void *stream);
AFHIP_KAPI int afhip_k_canny_sweep(unsigned char *out,
                                   const unsigned char *weak, long long d0,
                                   long long d1, long long ns,
                                   unsigned *changed, void *stream);

/* ORB kernels (af_orb, src/api/c/orb.cpp:58-89; src/backend/cpu/orb.cpp:44-283;
 * src/backend/cpu/kernel/orb.hpp, fast.hpp, resize.hpp). f32 or f64 images
 * of d0 x d1. */
/* Bilinear resize of one id0 x id1 slice to od0 x od1 (resize.hpp:86-130). */
AFHIP_KAPI int afhip_k_orb_resize(int dtype, void *out, const void *in,
                                  long long id0, long long id1, long long od0,
                                  long long od1, void *stream);
/* FAST test of every pixel at least `edge` from the borders: flags and
 * scores at the pixel's scan rank (y - edge) * (d1 - 2 edge) + (x - edge). */
AFHIP_KAPI int afhip_k_orb_fast(int dtype, unsigned char *flags, float *score,
                                const void *img, long long d0, long long d1,
                                int edge, float thr, int arc, void *stream);
/* Suppression of the first min(n, *count) corners sel[] (scan ranks; count
 * in device memory, or null for n) on the score map V (d0 x d1 floats,
 * cleared here): flags[j] keeps corner j; flags past the count are 0. */
AFHIP_KAPI int afhip_k_orb_fast_nonmax(unsigned char *flags, float *V,
                                       const unsigned *sel, const float *score,
                                       long long n, const unsigned *count,
                                       long long d0, long long d1, int edge,
                                       void *stream);
/* Harris response of the first min(n, *count) corners sel1[sel2[j]] (count
 * in device memory, or null for n) over a 7 x 7 block window; flags[j] = 0
 * for corners within patch_r of a border and past the count. */
AFHIP_KAPI int afhip_k_orb_harris(int dtype, unsigned char *flags, float *resp,
                                  const void *img, const unsigned *sel1,
                                  const unsigned *sel2, long long n,
                                  const unsigned *count, long long d0,
                                  long long d1, int edge, int patch_r,
                                  int block, float k_thr, void *stream);
/* The n kept features sel1[sel2[sel3[order[f]]]] with scores keys[f]:
 * positions and intensity-centroid angle over a patch x patch window. */
AFHIP_KAPI int afhip_k_orb_angle(int dtype, float *x, float *y, float *score,
                                 float *ori, const void *img,
                                 const unsigned *sel1, const unsigned *sel2,
                                 const unsigned *sel3, const unsigned *order,
                                 const float *keys, long long n, long long d0,
                                 long long d1, int edge, int patch,
                                 void *stream);
/* Rotated-BRIEF descriptors (8 words per feature) sampled from img, and the
 * positions and sizes scaled by scl. */
AFHIP_KAPI int afhip_k_orb_describe(int dtype, unsigned *desc, float *ox,
                                    float *oy, float *osize, const float *x,
                                    const float *y, const float *ori,
                                    const void *img, long long n, long long d0,
                                    long long d1, int patch, int r, float scl,
                                    void *stream);
