/*
 * MI355X-specific extensions of the ArrayFire C API.
 *
 * afhip_get_stream mirrors afcu_get_stream (include/af/cuda.h:50). The
 * remaining entry points have no counterpart in the reference: ArrayFire has
 * no bfloat16 dtype (include/af/defines.h:210-233) and no reduced-precision
 * or ragged convolution calls. They sit beside the af_* calls and share their
 * handles, error codes and af_get_last_error.
 */
#pragma once
#include <af/defines.h>

/* Storage type for bfloat16 arrays (2 bytes, the upper half of an IEEE
 * binary32). Accepted by af_create_array, af_get_data_ptr, af_cast and the
 * convolve2_nn entry points. The value lies outside af_dtype's range so it
 * never aliases a reference dtype. */
#define afhip_bf16 ((af_dtype)64)

/* Arithmetic used by the channel-reduction GEMM of convolve2_nn and its two
 * adjoints. */
typedef enum {
    /* The dtype of the inputs decides: f32 -> FP32, f16 -> FP16 MFMA with
     * FP32 accumulation, afhip_bf16 -> BF16 MFMA with FP32 accumulation,
     * f64 -> FP64. This is what af_convolve2_nn does. */
    AFHIP_PRECISION_DEFAULT = 0,
    /* Operands rounded to bfloat16 (round to nearest even) as they are
     * loaded, products on BF16 MFMA, FP32 accumulation. */
    AFHIP_PRECISION_BF16 = 1,
    /* Same with IEEE binary16 operands on FP16 MFMA. */
    AFHIP_PRECISION_F16 = 2,
    /* Each FP32 operand split as hi = bf16(x), lo = bf16(x - hi); three BF16
     * MFMA products hi*hi' + hi*lo' + lo*hi' accumulated in FP32. */
    AFHIP_PRECISION_SPLIT_BF16 = 3,
    /* FP32 operands, FP32 products and accumulation. */
    AFHIP_PRECISION_F32 = 4
} afhip_precision;

/* include/af/cuda.h:50 (afcu_get_stream): the HIP stream (hipStream_t,
 * returned as void*) on which the library queues work for device id. */
AFAPI af_err afhip_get_stream(void **stream, int id);

/* include/af/arith.h (af_cast): element type conversion. Supported among
 * f32, f64, f16, afhip_bf16 and the integer types; floating to bfloat16 and
 * binary16 rounds to nearest even. */
AFAPI af_err af_cast(af_array *out, const af_array in, const af_dtype type);

/* af_convolve2_nn with an explicit arithmetic. signal and filter must have
 * the same dtype (f32, f16 or afhip_bf16; f64 only with
 * AFHIP_PRECISION_DEFAULT). out_type selects the output dtype (f32, or the
 * input dtype when it is f16/afhip_bf16); (af_dtype)-1 means the input
 * dtype, as af_convolve2_nn does. */
AFAPI af_err afhip_convolve2_nn_ex(af_array *out, const af_array signal,
                                   const af_array filter,
                                   const unsigned stride_dims,
                                   const dim_t *strides,
                                   const unsigned padding_dims,
                                   const dim_t *paddings,
                                   const unsigned dilation_dims,
                                   const dim_t *dilations,
                                   afhip_precision precision, af_dtype out_type);

/* af_convolve2_gradient_nn with an explicit arithmetic (see above). The
 * incoming gradient, signal and filter share one dtype; convolved_output is
 * read for its dims only (it may be the FP32 output of a BF16 forward). The
 * gradient is returned in out_type ((af_dtype)-1: the input dtype). */
AFAPI af_err afhip_convolve2_gradient_nn_ex(
    af_array *out, const af_array incoming_gradient,
    const af_array original_signal, const af_array original_filter,
    const af_array convolved_output, const unsigned stride_dims,
    const dim_t *strides, const unsigned padding_dims, const dim_t *paddings,
    const unsigned dilation_dims, const dim_t *dilations,
    af_conv_gradient_type grad_type, afhip_precision precision,
    af_dtype out_type);

/* A convolve2_nn training step as one captured hipGraph: the forward
 * af_convolve2_nn(y, x, w, ...) and both adjoints
 * af_convolve2_gradient_nn(dx | dw, dy, x, w, y, ..., DATA | FILTER), with
 * the arithmetic of afhip_convolve2_nn_ex / afhip_convolve2_gradient_nn_ex
 * (precision, out_type). The plan runs the three passes once eagerly (the
 * same argument checks and error codes as those calls), then captures them
 * on the library stream with the kernels, tiles and workspaces the dispatch
 * picks, and allocates y, dx and dw (new handles, owned by the caller; the
 * plan keeps their storage and that of x, w and dy alive until it is
 * released). afhip_nn_step_plan_run replays the graph: it reads the current
 * contents of x, w and dy (weight packing included) and rewrites y, dx and
 * dw. No counterpart in the reference (its CUDA backend launches each pass
 * separately, src/backend/cuda/convolveNN.cpp). */
typedef struct afhip_nn_step_plan_t *afhip_nn_step_plan;
AFAPI af_err afhip_nn_step_plan_create(
    afhip_nn_step_plan *plan, af_array *y, af_array *dx, af_array *dw,
    const af_array x, const af_array w, const af_array dy,
    const unsigned stride_dims, const dim_t *strides,
    const unsigned padding_dims, const dim_t *paddings,
    const unsigned dilation_dims, const dim_t *dilations,
    afhip_precision precision, af_dtype out_type);
AFAPI af_err afhip_nn_step_plan_run(afhip_nn_step_plan plan);
AFAPI af_err afhip_nn_step_plan_release(afhip_nn_step_plan plan);

/* Ragged batches: count independent 2-D spatial convolutions, each exactly
 * af_convolve2(&outputs[i], signals[i], filters[i], modes[i],
 * AF_CONV_SPATIAL) for an f32 signal with exactly 2 dims (dim1 > 1) and a
 * 2-D filter of at most 15x15
 * (any dtype, converted to f32 as the reference's castArray does), run as one
 * persistent kernel over a cost-sorted (problem, tile) work list. The plan
 * allocates the outputs (new handles in outputs[], owned by the caller),
 * picks the tile shape with the autotuner (cached per problem list), and
 * captures the launch in a hipGraph; afhip_ragged_plan_run replays it on the
 * library stream (inputs are read again, so new signal values are picked
 * up). No counterpart in the reference. */
typedef struct afhip_ragged_plan_t *afhip_ragged_plan;
AFAPI af_err afhip_ragged_plan_create(afhip_ragged_plan *plan, af_array *outputs,
                                      const af_array *signals,
                                      const af_array *filters,
                                      const af_conv_mode *modes, unsigned count);
AFAPI af_err afhip_ragged_plan_run(afhip_ragged_plan plan);
/* The tile variant the autotuner chose (0: 128x64, 1: 64x64, 2: 64x32). */
AFAPI af_err afhip_ragged_plan_variant(int *variant, const afhip_ragged_plan plan);
AFAPI af_err afhip_ragged_plan_release(afhip_ragged_plan plan);
/* One-shot: create, run once, release (outputs stay valid). */
AFAPI af_err afhip_convolve2_ragged(af_array *outputs, const af_array *signals,
                                    const af_array *filters,
                                    const af_conv_mode *modes, unsigned count);
