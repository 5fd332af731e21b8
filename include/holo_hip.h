/*
 * holo_hip.h -- C-ABI of the MI355X-native Holovibes per-batch compute path.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an hv_status.  Device pointers are caller-owned (hipMalloc'd, or a
 * torch tensor's data_ptr()); `stream` is a hipStream_t passed as void*
 * (NULL = the default stream).  All step functions are asynchronous with
 * respect to the host unless stated otherwise; none of them allocates device
 * memory unless stated otherwise.
 *
 * Complex buffers are interleaved float2 (re, im) -- the layout of cuComplex
 * in the reference -- and hold frames frame-major, rows row-major:
 * element (frame f, row y, column x) is at f*W*H + y*W + x.
 *
 * Supported geometry: W and H in [16, 4096], powers of two or other even
 * products 2^a 3^b 5^c (e.g. 1536, 1080, 1920), with W*H a multiple of 256;
 * time-window sizes T in [1, 256] (PCA: powers of two, or below 32).  Other
 * sizes return HV_ERR_UNSUPPORTED.
 *
 * Each declaration cites the reference interface it replaces
 * (paths relative to the Holovibes 14.8.2 source tree).
 */
#ifndef HOLO_HIP_H
#define HOLO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes.  The reference returns nothing and exit()s on CUDA errors   */
/* (Backend/cuda_includes/common.cuh:53-64); this boundary reports instead.  */
/* ------------------------------------------------------------------------ */
typedef int hv_status;
#define HV_OK 0
#define HV_ERR_INVALID_ARG 1  /* null pointer, zero size, range out of bounds */
#define HV_ERR_UNSUPPORTED 2  /* valid but not implemented (e.g. non power-of-two) */
#define HV_ERR_OUT_OF_MEMORY 3
#define HV_ERR_HIP 4          /* a HIP runtime call failed */
#define HV_ERR_BUDGET 5       /* admission refused: HBM budget exhausted */
#define HV_ERR_NOT_READY 6    /* no result available yet */
#define HV_ERR_IO 7           /* a file could not be opened, read or written */
#define HV_ERR_FORMAT 8       /* not a valid .holo file (header checks of the reference) */

/* camera::PixelDepth (Camera/include/frame_desc.hh:13-24), subset on the path */
#define HV_DEPTH_8 1
#define HV_DEPTH_16 2
#define HV_DEPTH_32F 4

/* holovibes::SpaceTransformation (Backend/includes/enum/enum_space_transformation.hh:17-22) */
#define HV_SPACE_NONE 0
#define HV_SPACE_FRESNELTR 1
#define HV_SPACE_ANGULARSP 2

/* holovibes::TimeTransformation (Backend/includes/enum/enum_time_transformation.hh:16-22) */
#define HV_TIME_NONE 0
#define HV_TIME_STFT 1
#define HV_TIME_PCA 2

/* holovibes::ImgType (Backend/includes/enum/enum_img_type.hh:15-25), subset on the path */
#define HV_IMG_MODULUS 0
#define HV_IMG_SQUARED_MODULUS 1
#define HV_IMG_ARGUMENT 2
#define HV_IMG_MOMENTS_0 5
#define HV_IMG_MOMENTS_1 6
#define HV_IMG_MOMENTS_2 7

/* Lens (transfer function / chirp) arithmetic.  REFERENCE_FP32 evaluates the
 * reference kernels' float expressions in the same order (masks.cu:7-59);
 * FP64 evaluates the phase in double and rounds the lens to float once. */
#define HV_LENS_FP64 0
#define HV_LENS_REFERENCE_FP32 1

/* PCA covariance arithmetic. */
#define HV_PCA_FP32 0          /* plain FP32 FMA reduction */
#define HV_PCA_BF16_SPLIT 1    /* MFMA bf16, x = hi + lo, hi*hi + hi*lo + lo*hi, FP32 accumulate */
#define HV_PCA_FP16_SPLIT 2    /* MFMA fp16 (scaled), same split */
#define HV_PCA_BF16 3          /* MFMA bf16, single product (no split) */
#define HV_PCA_FP16 4          /* MFMA fp16 (scaled), single product */

/* Contrast range policy. */
#define HV_CONTRAST_REFERENCE 0   /* auto-contrast by percentiles on the first full window, then frozen (rendering.cc:239-258) */
#define HV_CONTRAST_PERCENTILE 1  /* percentiles recomputed on every image, device-resident */
#define HV_CONTRAST_MINMAX 2      /* min/max of every image, device-resident */
#define HV_CONTRAST_MANUAL 3      /* fixed contrast_min / contrast_max */

/* 2-D phase unwrapping variant. */
#define HV_UNWRAP_REFERENCE 0     /* periodic least squares with 7 FFT2 (tools_unwrap.cu:269-369) */

/* ------------------------------------------------------------------------ */
/* Library information                                                       */
/* ------------------------------------------------------------------------ */

/* No reference counterpart: version string of this library. */
const char* hv_version(void);

/* No reference counterpart: last error text of the calling thread. */
const char* hv_last_error(void);

/* No reference counterpart: number of visible GPUs (0 without a GPU). */
hv_status hv_device_count(int* count);

/* ------------------------------------------------------------------------ */
/* Step functions (one per reference step)                                   */
/* ------------------------------------------------------------------------ */

/* Replaces input_queue_to_input_buffer (Backend/cuda_sources/tools_conversion.cu:120-171):
 * u8 -> float(x)*257, u16 -> float(x), f32 -> x; imaginary part 0.
 * in: batch*frame_res pixels of `depth` bytes; out: batch*frame_res complex. */
hv_status hv_convert_to_complex(void* out_c64, const void* in, size_t frame_res, int batch, int depth,
                                void* stream);

/* Replaces angular_spectrum_lens (Backend/cuda_sources/angular_spectrum.cu:14-31,
 * kernel_spectral_lens masks.cu:31-59): centred transfer function
 * H = exp(i*2*pi*z/lambda*sqrt(max(0, 1 - lambda^2 (u^2+v^2)))), Nx*Ny complex.
 * x_step, y_step in metres (the caller passes PixelSize*1e-6). */
hv_status hv_angular_spectrum_lens(void* lens_c64, int nx, int ny, float z, float lambda, float x_step,
                                   float y_step, int precision, void* stream);

/* Replaces shift_corners (Backend/cuda_sources/shift_corners.cu:8-117): in-place
 * quadrant swap of `batch` complex frames (fftshift for even sizes). */
hv_status hv_shift_corners_c64(void* buf_c64, int batch, int size_x, int size_y, void* stream);

/* Same as hv_shift_corners_c64 for float frames (rendering.cc:21-40). */
hv_status hv_shift_corners_f32(float* buf, int batch, int size_x, int size_y, void* stream);

/* Replaces fresnel_transform_lens (Backend/cuda_sources/fresnel_transform.cu:14-63,
 * kernel_quadratic_lens masks.cu:7-29): chirp exp(i*pi/(lambda z)*(x^2+y^2)),
 * x = (i - side/2)*pixel_size*1e-6, built on a square of side lens_side and
 * cropped centrally to width x height.  pixel_size in micrometres. */
hv_status hv_fresnel_transform_lens(void* lens_c64, int lens_side, int height, int width, float lambda,
                                    float z, float pixel_size, int precision, void* stream);

/* Replaces angular_spectrum (Backend/cuda_sources/angular_spectrum.cu:33-63):
 * in place on `batch` frames: x = IFFT2(FFT2(x) * lens) / (W*H).
 * lens_shifted is the corner-shifted transfer function. */
hv_status hv_angular_spectrum(void* buf_c64, int batch, const void* lens_shifted, int width, int height,
                              void* stream);

/* Replaces fresnel_transform (Backend/cuda_sources/fresnel_transform.cu:65-83):
 * in place on `batch` frames: x = FFT2(x * lens), unnormalised. */
hv_status hv_fresnel_transform(void* buf_c64, int batch, const void* lens, int width, int height,
                               void* stream);

/* Batched 2-D FFT of `batch` frames, in place, unnormalised; direction -1
 * forward, +1 inverse.  Replaces the cufftXtExec calls on the spatial plan
 * (Backend/sources/core/icompute.cc:151-175). */
hv_status hv_fft2(void* buf_c64, int batch, int width, int height, int direction, void* stream);

/* Fused form of steps a2-a8 (conversion, space transform, transfer into the
 * time-transformation queue; converts.cc:369-396, fourier_transform.cc:30-137,
 * pipe.cc:457-467, queue.cc:332-376): raw frames -> complex -> space transform,
 * computed in place in ring slots (slot0 + b) mod ring_size, b < batch (of an
 * oversized batch only the last ring_size frames, as the queue keeps them).
 * `lens` is the shifted angular-spectrum lens pre-multiplied by 1/(W*H) (see
 * hv_prepare_lens) or the Fresnel chirp; ignored for HV_SPACE_NONE. */
hv_status hv_propagate_to_ring(const void* raw, int depth, int batch, int width, int height, int space,
                               const void* lens, void* ring_c64, int ring_size, int slot0, void* stream);

/* Builds the lens hv_propagate_to_ring expects (refresh time, not per batch):
 * angular spectrum -> corner-shifted H / (W*H); Fresnel -> chirp; NONE -> no-op. */
hv_status hv_prepare_lens(void* lens_c64, int space, int width, int height, float lambda, float z,
                          float pixel_size_um, int precision, void* stream);

/* Replaces stft (Backend/cuda_sources/stft.cu:9-15): for each pixel p,
 * out[k*P+p] = sum_t in[t*P+p] exp(-2 pi i k t / T), unnormalised. */
hv_status hv_stft(void* out_c64, const void* in_c64, int T, size_t frame_res, void* stream);

/* Replace complex_to_modulus / complex_to_squared_modulus / complex_to_argument
 * (Backend/cuda_sources/tools_conversion.cu:11-30, 84-103, 199-232): mean over
 * the inclusive range [pmin, pmax] of |x|, |x|^2 or atan2(im, re). */
hv_status hv_complex_to_modulus(float* out, const void* in_c64, int pmin, int pmax, size_t frame_res,
                                void* stream);
hv_status hv_complex_to_squared_modulus(float* out, const void* in_c64, int pmin, int pmax,
                                        size_t frame_res, void* stream);
hv_status hv_complex_to_argument(float* out, const void* in_c64, int pmin, int pmax, size_t frame_res,
                                 void* stream);

/* Fused STFT + image type: the ring (T frames, physical slot order) ->
 * mean over k in [pmin, pmax] of the chosen image type of S[k].  Equivalent
 * to hv_stft followed by hv_complex_to_* without the T*P spectrum in HBM.
 * If minmax (device, 2 x uint32) is non-NULL the image's min and max are
 * folded into it (see hv_minmax_init). */
hv_status hv_stft_image(float* out, const void* ring_c64, int T, size_t frame_res, int img_type, int pmin,
                        int pmax, uint32_t* minmax, void* stream);

/* Replaces the moments path (tools_conversion.cu:33-66, tools_compute.cu:38-97,
 * icompute.cc:26-69, pipe.cc:429-455): from the ring, S = STFT, then
 * M_k = sum_{f=f_start}^{f_end} freq_k[f] |S_f|^2 with the Nyquist bin added
 * again for k = 0, 2 and subtracted for k = 1 when T is even and it is in range.
 * freq_1 = fftfreq(T) * camera_fps.  Any of m0/m1/m2 may be NULL. */
hv_status hv_stft_moments(float* m0, float* m1, float* m2, const void* ring_c64, int T, size_t frame_res,
                          int f_start, int f_end, float camera_fps, void* stream);

/* time_transformation_cuts_begin (Backend/cuda_sources/stft.cu:75-115): from a
 * time-transformed window (T frames of W x H, bin-major), XZ (T x W) = mean
 * over rows [ymin, ymax] and YZ (H x T) = mean over columns [xmin, xmax] of
 * the image value (modulus, squared modulus, atan(im/re) + pi/2; 0 for other
 * types).  xmax < W and ymax < H (the reference reads past the frame
 * otherwise: HV_ERR_INVALID_ARG here). */
hv_status hv_stft_cuts(float* out_xz, float* out_yz, const void* stft_c64, int width, int height, int T, int xmin,
                       int ymin, int xmax, int ymax, int img_type, void* stream);

/* Replaces cov_matrix (Backend/sources/compute/matrix_operations.cc:9-26):
 * cov = H^H H, H = ring as P x T column-major; cov is T x T complex
 * column-major.  Partial sums go to `workspace` (hv_pca_workspace_size bytes). */
hv_status hv_pca_workspace_size(int T, size_t frame_res, int precision, size_t* bytes);
hv_status hv_pca_cov(void* cov_c64, const void* ring_c64, int T, size_t frame_res, int precision,
                     void* workspace, void* stream);

/* Replaces eigen_values_vectors (matrix_operations.cc:28-65, cusolverDnCheevd
 * lower, ascending): cyclic Jacobi in FP64 on the GPU.  cov: T x T complex
 * float (column-major); evals: T doubles ascending; evecs: T x T complex
 * double column-major (column j = eigenvector j, phase normalised so its
 * largest-magnitude component is real positive).  Only the lower triangle
 * of cov is read.  T <= 64 runs in LDS; larger T needs `workspace`
 * (hv_eig_workspace_size bytes, may be NULL when that size is 0). */
hv_status hv_eig_workspace_size(int T, size_t* bytes);
hv_status hv_eig_hermitian_f64(double* evals, void* evecs_c128, const void* cov_c64, int T, void* workspace,
                               void* stream);

/* Replaces matrix_multiply_complex (matrix_operations.cc:77-102) followed by the
 * image-type conversion and the PCA x131071 scale (converts.cc:22-63):
 * C = H V, image = 131071 * mean over t in [pmin, pmax] of type(C[:, t]). */
hv_status hv_pca_image(float* out, const void* ring_c64, const void* evecs_c128, int T, size_t frame_res,
                       int img_type, int pmin, int pmax, uint32_t* minmax, void* stream);

/* Replaces, for a PCA time transform, matrix_multiply_complex (matrix_operations.cc:77-102)
 * followed by the moments path over its output gpu_p_acc_buffer (pipe.cc:429-455,
 * tools_conversion.cu:33-66, tools_compute.cu:38-97): C = H V, then
 * out = sum_{f=f_start}^{f_end} freq_k[f] |C[:, f]|^2 for k = moment (0, 1, 2), with the
 * Nyquist component counted again for k = 0, 2 and cancelled for k = 1 when T is even;
 * freq_1 = fftfreq(T) * camera_fps.  No x131071 scale (the reference's scale acts on the
 * frame that the moment copy then overwrites).  minmax as in hv_stft_image. */
hv_status hv_pca_moments(float* out, const void* ring_c64, const void* evecs_c128, int T, size_t frame_res,
                         int moment, int f_start, int f_end, float camera_fps, uint32_t* minmax, void* stream);

/* Replaces unwrap_2d (Backend/cuda_sources/tools_unwrap.cu:269-369):
 * least-squares 2-D phase unwrapping of a W x H phase image (the reference's
 * FFT2 sequence, computing FFT2(z) once instead of twice);
 * out = Im(IFFT2(...)) / (-2 pi).  scratch: hv_unwrap_scratch_size bytes. */
hv_status hv_unwrap_scratch_size(int width, int height, size_t* bytes);
hv_status hv_unwrap_2d(float* out, const float* phase, int width, int height, void* scratch, void* stream);

/* Replaces rescale_float_unwrap2d (tools_conversion.cu:315-335) with a
 * device-side min/max: out = (in - min) / (max - min) * 65535.
 * minmax_scratch: 2 x uint32 of device memory. */
hv_status hv_rescale_unwrap2d(float* out, const float* in, size_t frame_res, uint32_t* minmax_scratch,
                              void* stream);

/* Replaces subsample_frame (Backend/cuda_sources/frame_reshape.cu:278-330) for
 * float frames: out(x,y) = in[(y*s + (x odd ? s/2 : 0))*W + x*s + (y odd ? s/2 : 0)]. */
hv_status hv_subsample_frame_f32(float* out, const float* in, int width, int height, int step, void* stream);

/* Device-side min/max accumulators: two uint32 holding the order-preserving
 * encoding of a float (sign bit set for non-negative values, all bits
 * inverted for negative ones), so that one integer atomicMin/atomicMax per
 * workgroup updates them exactly and in any order.
 * hv_minmax_init sets them to (+inf, -inf); hv_minmax folds n floats in;
 * hv_minmax_decode writes the two floats (min, max) to out. */
hv_status hv_minmax_init(uint32_t* minmax, void* stream);
hv_status hv_minmax(uint32_t* minmax, const float* in, size_t n, void* stream);
hv_status hv_minmax_decode(float* out, const uint32_t* minmax, void* stream);

/* Replaces compute_percentile_xy_view + update_contrast
 * (Backend/cuda_sources/percentile.cu:75-222, contrast_api.cc:137-146):
 * exact order statistics of the (optionally subsampled, step 0/1 = none)
 * image by radix select on the float bits; value at index
 * uint(pct/100 * N) of the ascending order, clamped to >= 1.0 when
 * clamp_ge_one.  range_out: 2 floats (device).  scratch: hv_percentile_scratch_size. */
hv_status hv_percentile_scratch_size(size_t n, size_t* bytes);
hv_status hv_percentile_range(float* range_out, const float* img, int width, int height, int subsample_step,
                              float lower_pct, float upper_pct, int clamp_ge_one, void* scratch, void* stream);

/* No counterpart in the reference: hv_percentile_range without the scratch
 * argument, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  It allocates nothing: its 16 KiB
 * of histograms live in a static array of the library, one per GPU, held
 * until the process ends.  Calls that overlap on one GPU race on that array
 * (serialise them, e.g. on one stream).  Images of more than 2^32 - 1
 * selected pixels are refused (HV_ERR_INVALID_ARG). */
hv_status hv_percentile_range_noalloc(float* range_out, const float* img, int width, int height, int subsample_step,
                                      float lower_pct, float upper_pct, int clamp_ge_one, void* stream);

/* Replaces compute_percentile_xy_view with the contrast reticle on
 * (percentile.cu:15-43 copy_circle_kernel, 150-222): the same order
 * statistics over the pixels (x, y) of the W x H image with
 * (x - W/2)^2 + (y - H/2)^2 <= r^2, r = uint(reticle_scale * min(W, H) / 2);
 * N = the number of such pixels.  reticle_scale in [0, 1]. */
hv_status hv_percentile_range_reticle(float* range_out, const float* img, int width, int height, float reticle_scale,
                                      float lower_pct, float upper_pct, int clamp_ge_one, void* scratch,
                                      void* stream);

/* Replaces apply_contrast_correction (Backend/cuda_sources/contrast_correction.cu:5-17):
 * x <- 65535 / (max - min + FLT_EPSILON) * (x - min), in place, host range. */
hv_status hv_apply_contrast(float* img, size_t n, float min, float max, void* stream);

/* Replaces float_to_ushort (tools_conversion.cu:350-387): x <= 0 -> 0,
 * x > 65535 -> 65535, else truncation. */
hv_status hv_float_to_ushort(uint16_t* out, const float* in, size_t n, void* stream);

/* Fused apply_contrast_correction + float_to_ushort with the range read from
 * device memory (range_dev[0] = min, range_dev[1] = max); invert swaps them
 * (rendering.cc:185-237). */
hv_status hv_contrast_to_ushort(uint16_t* out, const float* in, size_t n, const float* range_dev, int invert,
                                void* stream);

/* ------------------------------------------------------------------------ */
/* Pipe: the per-batch orchestrator (Backend/sources/core/pipe.cc:255-387)   */
/* ------------------------------------------------------------------------ */

/* The compute settings the path reads (Backend/includes/settings/settings.hh:94-206). */
typedef struct hv_compute_settings {
    uint32_t width, height;          /* FrameDescriptor width/height */
    uint32_t depth;                  /* HV_DEPTH_8 / HV_DEPTH_16 / HV_DEPTH_32F */
    uint32_t batch_size;             /* BatchSize */
    uint32_t time_stride;            /* TimeStride (multiple of batch_size) */
    uint32_t time_transformation_size; /* TimeTransformationSize, T */
    int32_t space_transformation;    /* HV_SPACE_* */
    int32_t time_transformation;     /* HV_TIME_* */
    int32_t img_type;                /* HV_IMG_* */
    float lambda;                    /* Lambda, metres */
    float z_distance;                /* ZDistance, metres */
    float pixel_size;                /* PixelSize, micrometres */
    uint32_t p_start, p_width;       /* P view (accumulation range) */
    int32_t unwrap_2d;               /* Unwrap2d (argument images) */
    int32_t fft_shift;               /* FftShiftEnabled */
    int32_t contrast_enabled;        /* XY.contrast.enabled */
    int32_t contrast_invert;         /* XY.contrast.invert */
    int32_t contrast_mode;           /* HV_CONTRAST_* */
    float contrast_min, contrast_max;      /* XYContrastRange (HV_CONTRAST_MANUAL) */
    float contrast_lower, contrast_upper;  /* ContrastLower/UpperThreshold, percent */
    uint32_t contrast_subsample;     /* percentile on an s x s strided subsample (0/1 = full) */
    float camera_fps;                /* CameraFps (moments frequency scale) */
    int32_t lens_precision;          /* HV_LENS_* */
    int32_t pca_precision;           /* HV_PCA_* */
    /* optional post-processing of the float image, in the reference's order
     * (Backend/sources/core/pipe.cc:353-370): renormalise, fft shift,
     * accumulate, log10, then contrast */
    int32_t renorm_enabled;          /* RenormEnabled: x * 2^renorm_constant / mean(x) */
    uint32_t renorm_constant;        /* RenormConstant, 0..31 */
    uint32_t img_accu_level;         /* XY.output_image_accumulation: mean of the last n images (0, 1: off) */
    int32_t log_enabled;             /* XY.log_enabled: log10 */
    /* CutsViewEnabled with the X / Y accumulation views (fourier_transform.cc:466-511):
     * XZ (T x W) = mean over rows [y_start, y_start + y_width], YZ (H x T) =
     * mean over columns [x_start, x_start + x_width] of the STFT image values */
    int32_t cuts_enabled;
    uint32_t x_start, x_width, y_start, y_width;
    /* RegistrationEnabled / RegistrationZone (registration.cc:9-122): shift
     * each image onto the first one by the peak of their cross-correlation
     * over a centred disc of radius zone * min(W, H) / 2 */
    int32_t registration_enabled;
    float registration_zone;
    /* ContrastReticleDisplayEnabled / ContrastReticleScale (Backend/includes/core/holovibes.hh:400, 403;
     * Backend/sources/compute/rendering.cc:310-327): the autocontrast percentiles
     * are taken over the centred disc of radius uint(scale * min(W, H) / 2)
     * instead of the whole image (percentile.cu:15-43, 150-190).  scale in [0, 1];
     * exclusive with contrast_subsample > 1. */
    int32_t contrast_reticle_enabled;
    float contrast_reticle_scale;
    /* OutputBufferSize (Backend/includes/core/holovibes.hh:375): u16 images the
     * output queue holds (gpu_output_queue, Backend/sources/core/icompute.cc:238-246);
     * 0: no queue, only the latest image per lane. */
    uint32_t output_buffer_size;
} hv_compute_settings;

/* Fills the reference defaults (Backend/includes/core/holovibes.hh:359-447). */
void hv_default_settings(hv_compute_settings* s);

/* Validates settings the way the reference setters clamp/refuse them
 * (Backend/sources/api/transform_api.cc:16-206). */
hv_status hv_validate_settings(const hv_compute_settings* s);

typedef struct hv_pipe hv_pipe;

/* Replaces Pipe construction + refresh (pipe.hh:70-125, icompute.hh:122-171,
 * pipe.cc:255-387): allocates every buffer once (ring, lens, scratch, output)
 * on the current device and builds the lens.  No allocation after this. */
hv_status hv_pipe_create(const hv_compute_settings* s, void* stream, hv_pipe** out);

/* Device bytes a pipe with these settings holds (admission accounting). */
hv_status hv_pipe_memory_bytes(const hv_compute_settings* s, size_t* bytes);

/* One Pipe::exec cycle (pipe.cc:650-676) for one batch already in device
 * memory (batch_size raw frames): stride discard, transfer into the ring,
 * wait for T frames, time transform, image type, contrast, u16.  Sets
 * *produced to 1 when an output image was written (hv_pipe_output), 0 when
 * the cycle exited early.  Asynchronous: work is queued on the pipe stream. */
hv_status hv_pipe_push_batch(hv_pipe* p, const void* frames, int* produced);

/* hv_pipe_push_batch for n batches in order, in one call: the loop of the
 * compute worker (Pipe::exec's cycle loop, pipe.cc:650-676, run by
 * compute_worker.cc:30-32) for a caller that holds several batches in device
 * memory.  *produced (may be NULL) counts the output images written.  Stops
 * at the first failing push and returns its status. */
hv_status hv_pipe_push_batches(hv_pipe* p, const void* const* frames, int n, int* produced);

/* Device pointers of the latest output: u16 image, float image after contrast
 * (before the u16 cast), and the float image before contrast. */
hv_status hv_pipe_output(hv_pipe* p, const uint16_t** out_u16, const float** out_f32_contrast,
                         const float** out_f32_image);

/* The output queue (gpu_output_queue, OutputBufferSize frames): every
 * produced u16 image is written by the cycle's last kernel straight into the
 * next slot, (index of the produced image) mod max_size, and enqueued as in
 * Queue::enqueue (Backend/sources/data_struct/queue.cc:149-189, called from
 * pipe.cc:491-524): when the queue is full the oldest image is overwritten
 * and has_overwritten set.  base: slot k at base + k * W * H.  A consumer
 * must read an image before max_size more images are produced. */
hv_status hv_pipe_output_queue(hv_pipe* p, const uint16_t** base, uint32_t* max_size, uint32_t* start_index,
                               uint32_t* size, int* has_overwritten);
/* Queue::dequeue(dest, stream, kind, nb_elts) (queue.cc:385-407): copies the
 * nb oldest images (memcpy_kind 2 = device to host, 3 = device to device) to
 * dst on `stream` once they are complete, waits for the copy, and removes
 * them (a range that wraps around the end is copied in two parts).
 * HV_ERR_NOT_READY when fewer than nb images are queued. */
hv_status hv_pipe_output_dequeue(hv_pipe* p, void* dst, int memcpy_kind, uint32_t nb, void* stream);
/* No reference counterpart: the u16 image of the index-th image the pipe
 * produced (0-based since creation), while the queue still holds it
 * (HV_ERR_NOT_READY once overwritten or not yet produced).  `stream` (may be
 * NULL) waits for the image to be complete. */
hv_status hv_pipe_output_frame(hv_pipe* p, uint64_t index, void* stream, const uint16_t** frame);
/* Images produced since creation. */
hv_status hv_pipe_output_count(hv_pipe* p, uint64_t* count);

/* Current contrast range (device pointer to 2 floats). */
hv_status hv_pipe_contrast_range(hv_pipe* p, const float** range_dev);

/* Captures the steady-state per-batch work (ring full, no discard) into a
 * hipGraph per ring phase and replays it on later pushes; 0 disables. */
hv_status hv_pipe_set_graph(hv_pipe* p, int enabled);

/* No reference counterpart: the pipe's own input buffer (batch_size frames).
 * Pushed batches are read in place; with graphs enabled each distinct input
 * pointer gets its own captured graph (up to 64, then batches are copied
 * into this buffer first), so producers should cycle through a fixed set of
 * landing buffers -- or write here. */
hv_status hv_pipe_input_buffer(hv_pipe* p, void** frames);

/* No reference counterpart (the reference's Chrono / benchmark worker time
 * whole cycles only): with timing enabled every steady-state cycle runs
 * eagerly with events between its stages and waits for them.
 * hv_pipe_stage_times writes the mean milliseconds of the 5 stages -- row
 * FFT pass, column pass, inverse row pass, image kernel(s), contrast + u16 --
 * over the timed cycles. */
hv_status hv_pipe_set_timing(hv_pipe* p, int enabled);
hv_status hv_pipe_stage_times(hv_pipe* p, float* ms5, int* cycles);

/* No reference counterpart.  Two shorter cycle paths than the three passes
 * (level 1 needs B = S = T, every batch replacing the whole window):
 *   level 1 -- angular spectrum, T <= 32, a T x W tile within LDS: the
 *              inverse row FFT pass is fused with the STFT and image
 *              epilogue; the ring holds each frame after the column pass
 *              instead of the propagated frame;
 *   level 2 -- angular spectrum or Fresnel, image types and moments: the
 *              STFT bins the image reads ([pmin, pmax], or [f_start, f_end]
 *              for the moments) of the raw window are taken first and only
 *              those K planes are propagated (the STFT and the propagation
 *              are linear and commute); the ring's first K frames hold the
 *              planes.  PCA likewise: the propagators are unitary up to a
 *              constant, so the propagated window's covariance is the raw
 *              window's up to that constant (same eigenvectors); it is taken
 *              from the raw frames and the K selected components of the raw
 *              window are propagated.  W <= 1024 and T <= 64 fuse the bins
 *              with the row FFTs (and, for images, the inverse row FFTs with
 *              the image term); larger frames and windows run them as passes
 *              of their own (T up to 256).  B = S = T: the bins are taken
 *              from the batch itself.  B < T (a sliding window): the pipe
 *              keeps the window's raw frames in a raw ring (T frames of
 *              `depth` bytes, slot n holding frame n mod T as the reference's
 *              ring) and takes the bins from it; up to 4 STFT bins are summed
 *              directly, the batch written into the raw ring as the window
 *              is read.
 * hv_pipe_set_fused picks the highest eligible path at or below `level`
 * (default: 2); 0 forces the three-pass + STFT sequence.  hv_pipe_path
 * reports the path in use (0, 1 or 2).  Leaving the sliding-window level 2
 * path rebuilds the three-pass ring from the raw window; entering it, the
 * window refills first (T / B batches without an image). */
hv_status hv_pipe_set_fused(hv_pipe* p, int level);
hv_status hv_pipe_path(const hv_pipe* p, int* path);

/* No reference counterpart.  Lets up to `lanes` (1..4) consecutive cycles
 * run on the GPU at once, each on its own stream with its own image,
 * contrast and output buffers (allocated here, outside the per-batch path);
 * cycle n runs on lane n mod lanes.  When every batch replaces the whole
 * window (B = S = T) cycles share no device state and each lane also has its
 * own bins-first planes or time-window ring and PCA buffers; otherwise the
 * lanes share the ring (the three-pass ring, or the raw ring of the
 * bins-first path with their own planes): a cycle's write into it waits
 * until the previous cycle's time transform has read it, and the rest of
 * the cycle (the image post-processing; on the bins-first path also the
 * bins' propagation) overlaps the next cycle (eager launches, no graphs).
 * Only for a pipe that owns its streams (created without a stream).  Outputs (hv_pipe_output, hv_pipe_contrast_range) are the latest
 * cycle's; call hv_pipe_sync before reading them.  A frozen autocontrast
 * range reaches every lane in stream order.  Changing the path
 * (hv_pipe_set_fused) returns to one lane; the frames a lane reads must stay
 * unchanged until hv_pipe_sync. */
hv_status hv_pipe_set_lanes(hv_pipe* p, int lanes);

/* No reference counterpart (measurement).  Mean milliseconds of one launch of
 * a bins-first stage -- 0 bins (or PCA components) + row FFTs (cycling over
 * the n_frames input batches; a sliding window: from the raw ring), 1 column pass, 2 (PCA) the raw window's
 * covariance + eigensolver, 3 inverse row FFT + image, 4 contrast + u16 --
 * over `iters` launches back to back on lane 0's stream, between two HIP
 * events.  The pipe's buffers are overwritten; the next cycle rewrites them. */
hv_status hv_pipe_time_stage(hv_pipe* p, int stage, const void* const* frames, int n_frames, int iters, float* ms);

/* No reference counterpart (measurement).  Batch -> image latency under a
 * fixed offered load: batch i (frames[i mod n_frames]) arrives at host time
 * t0 + i * period_ns and is pushed then; the completion of each produced
 * image is polled on the host (an event after its cycle) while waiting for
 * the next arrival.  Writes one latency (ns, completion seen - arrival) per
 * produced image into latency_ns (room for n_batches) and their count. */
hv_status hv_pipe_measure_latency(hv_pipe* p, const void* const* frames, int n_frames, int n_batches,
                                  double period_ns, int64_t* latency_ns, int* n_out);

/* The time-transformation queue: ring base (T frames, complex), its
 * start index and size (Queue::get_start_index / get_size, queue.hh).  On
 * the bins-first path (level 2) the ring's first frames hold the bins'
 * planes instead of the propagated window. */
hv_status hv_pipe_ring(hv_pipe* p, const void** ring, uint32_t* start, uint32_t* size);

/* The pipe's stream (hipStream_t). */
hv_status hv_pipe_stream(hv_pipe* p, void** stream);

/* Synchronises the pipe stream (end-of-pipe cudaStreamSynchronize, pipe.cc:386). */
hv_status hv_pipe_sync(hv_pipe* p);

/* The latest cycle's XZ (T x W) and YZ (H x T) cut images (device floats,
 * valid until the next push); settings.cuts_enabled. */
hv_status hv_pipe_cuts(hv_pipe* p, const float** xz, const float** yz);

/* The (x, y) circular shift registration applied to the latest image
 * (blocking read; settings.registration_enabled). */
hv_status hv_pipe_registration_shift(hv_pipe* p, int* shift_x, int* shift_y);

/* RecordMode::MOMENTS (Backend/sources/core/pipe.cc:429-455, 583-598): every
 * cycle also writes the three moment frames M0, M1, M2 of the window (the
 * STFT's, or the PCA output's) as float frames, whatever the image type.
 * Allocates 3 frames of floats here and keeps the three-pass path (the
 * moments read the whole window).  One lane only.  enabled 0 releases. */
hv_status hv_pipe_set_record_moments(hv_pipe* p, int enabled);
/* The latest cycle's M0, M1, M2 (device, 3 * W * H floats, valid until the
 * next push), for hv_holo_store_frames; HV_ERR_NOT_READY before a cycle. */
hv_status hv_pipe_record_frames(hv_pipe* p, const float** frames);

hv_status hv_pipe_destroy(hv_pipe* p);

/* ------------------------------------------------------------------------ */
/* Synthetic camera source (host, deterministic)                              */
/* ------------------------------------------------------------------------ */

/* Seeded interferogram generator standing in for ICamera::get_frames
 * (Camera/include/icamera.hh:152): I = |R + O_t|^2, R = exp(2 pi i (fx x + fy y)),
 * O_t = O_0 exp(i q r^2/r_max^2) exp(2 pi i f_d(x,y) t / period), then
 * Poisson(K I) clipped to the bit depth.  Counter-based: frame t of a source
 * is the same whatever order frames are generated in. */
typedef struct hv_synth_params {
    uint32_t width, height;
    uint32_t depth;          /* HV_DEPTH_8 or HV_DEPTH_16 */
    uint32_t bits;           /* significant bits (8 or 12) */
    uint32_t period;         /* Doppler period in frames (T of the config) */
    float sigma;             /* speckle amplitude */
    float tilt_x, tilt_y;    /* reference-wave tilt, cycles/px */
    float quad_phase;        /* quadratic phase amplitude (rad) at r_max */
    float mean_counts;       /* mean of K*I */
    uint32_t n_discs;        /* Doppler discs (<= 4) */
    float disc_cx[4], disc_cy[4], disc_r[4];   /* pixels */
    int32_t fd_lo[4], fd_hi[4];                /* integer Doppler range per disc */
    uint64_t seed;
} hv_synth_params;

/* Writes frames [first_frame, first_frame + count) to host memory `out`
 * (count * W * H * depth bytes).  Runs on host threads. */
hv_status hv_synth_frames(const hv_synth_params* p, uint64_t first_frame, uint32_t count, void* out);

/* ------------------------------------------------------------------------ */
/* Input side: packed-pixel unpack and the batch input queue                */
/* ------------------------------------------------------------------------ */

/* unpack_12_to_16bit / unpack_10_to_16bit (Backend/cuda_sources/unpack.cu:34-60):
 * output_size u16 pixels from a stream packed MSB first; each output holds
 * the value in its top bits (12-bit: v << 4, 10-bit: v << 6), as the
 * reference writes it.  output_size must be a multiple of 2 (12-bit) or 4
 * (10-bit) and input_size must cover output_size * bits / 8 bytes
 * (HV_ERR_INVALID_ARG otherwise; the reference leaves such pixels
 * unwritten).  input and output must not overlap.  Any number of frames in
 * one call (the frames' packed streams are contiguous). */
hv_status hv_unpack_12_to_16bit(uint16_t* output, size_t output_size, const uint8_t* input, size_t input_size,
                                void* stream);
hv_status hv_unpack_10_to_16bit(uint16_t* output, size_t output_size, const uint8_t* input, size_t input_size,
                                void* stream);

/* A queue outlives the readers (hv_reader_*) that feed it: stop them first. */
typedef struct hv_input_queue hv_input_queue;

/* BatchInputQueue (Backend/sources/data_struct/batch_input_queue.cc:8-59):
 * room for frame_capacity - frame_capacity % batch_size frames of
 * width * height * depth bytes in one HBM allocation (made here). */
hv_status hv_input_queue_create(uint32_t frame_capacity, uint32_t batch_size, uint32_t width, uint32_t height,
                                uint32_t depth, hv_input_queue** out);

/* BatchInputQueue::enqueue (batch_input_queue.cc:124-187): appends nb_frames
 * frames (host memory: memcpy_kind 1 = hipMemcpyHostToDevice, pinned for an
 * asynchronous copy; device memory: 3 = hipMemcpyDeviceToDevice) on `stream`.
 * The frames must stay valid until the copy has run on `stream`, host
 * memory included (a pageable copy may still be reading it on return).
 * A batch becomes visible when complete; when the queue is full the oldest
 * batch is overwritten (has_overwritten).  Waits on the host only while the
 * consumer holds the slot to overwrite (HV_ERR_INVALID_ARG if the caller's
 * own thread holds it). */
hv_status hv_input_queue_enqueue(hv_input_queue* q, const void* frames, int memcpy_kind, uint32_t nb_frames,
                                 void* stream);

/* The packed-file read (Backend/sources/thread/file_frame_read_worker.cc:212-240)
 * as an enqueue: nb_frames frames packed on `bits` (10 or 12) bits, from the
 * host (copied into a one-batch device staging buffer, allocated on first
 * use) or the device, unpacked straight into the queue's 16-bit slots. */
hv_status hv_input_queue_enqueue_packed(hv_input_queue* q, const void* packed_frames, int memcpy_kind, int bits,
                                        uint32_t nb_frames, void* stream);

/* The oldest whole batch, in place (batch_size frames in HBM), for
 * hv_pipe_push_batch: `stream` waits for the producer's copy.  Waits up to
 * timeout_ms on the host for one (0: no wait); HV_ERR_NOT_READY when there
 * is none.  The batch stays the consumer's until hv_input_queue_pop. */
hv_status hv_input_queue_front(hv_input_queue* q, void* stream, int timeout_ms, const void** batch);

/* Releases the batch front returned once the work queued on `stream` so far
 * has run (the producer's overwrite of the slot waits for it); without a
 * held batch, drops the oldest one like BatchInputQueue::dequeue()
 * (batch_input_queue.cc:205-213). */
hv_status hv_input_queue_pop(hv_input_queue* q, void* stream);

/* BatchInputQueue::dequeue(dest, ...) (batch_input_queue.cc:189-203) with a
 * plain copy: front + copy (memcpy_kind 3 device, 2 host) + pop, then waits
 * for `stream`.  HV_ERR_NOT_READY when the queue has no whole batch. */
hv_status hv_input_queue_dequeue(hv_input_queue* q, void* dst, int memcpy_kind, void* stream);

/* get_size (whole batches), max_size, start index, frames held, has_overwritten. */
hv_status hv_input_queue_state(hv_input_queue* q, uint32_t* size, uint32_t* max_size, uint32_t* start_index,
                               uint32_t* curr_nb_frames, int* has_overwritten);
/* get_data: the HBM allocation (slot k at k * batch_size * frame size). */
hv_status hv_input_queue_data(hv_input_queue* q, const void** data);

/* BatchInputQueue::resize (batch_input_queue.cc:241-266): new batch size,
 * same frame capacity, queue emptied. */
hv_status hv_input_queue_resize(hv_input_queue* q, uint32_t new_batch_size);
/* make_empty / reset_override (batch_input_queue.cc:88-106). */
hv_status hv_input_queue_reset(hv_input_queue* q);
hv_status hv_input_queue_destroy(hv_input_queue* q);

/* CapturedFramesDescriptor (Camera/include/icamera.hh:8-58): up to two
 * regions of frames (the end and the start of a camera's ring buffer). */
typedef struct hv_captured_frames {
    void* region1;
    uint32_t count1;
    void* region2;
    uint32_t count2;
    int on_gpu;  /* 1: device memory */
} hv_captured_frames;

/* ICamera (Camera/include/icamera.hh:91-166) as a C function table: any
 * camera plugin with a C ABI fills one.  get_frames' regions stay valid
 * until the next get_frames call. */
typedef struct hv_camera {
    void* self;
    hv_status (*init_camera)(void* self);
    hv_status (*start_acquisition)(void* self);
    hv_status (*stop_acquisition)(void* self);
    hv_status (*shutdown_camera)(void* self);  /* also frees the source */
    hv_status (*get_frames)(void* self, hv_captured_frames* out);
    hv_status (*get_fd)(void* self, uint32_t* width, uint32_t* height, uint32_t* depth);
    int (*get_camera_fps)(void* self);         /* -1 when unknown */
} hv_camera;

/* A synthetic camera: ring_frames frames of hv_synth_frames generated into
 * pinned memory at init_camera, replayed frames_per_call at a time. */
hv_status hv_camera_synthetic(const hv_synth_params* p, uint32_t ring_frames, uint32_t frames_per_call,
                              hv_camera* out);

/* A .holo recording as a source (FileFrameReadWorker,
 * Backend/sources/thread/file_frame_read_worker.cc:35-312): frames
 * [first_frame, end_frame) (end 0: to the end), frames_per_call at a time
 * (the reference's file buffer size), from first_frame again at the end. */
hv_status hv_camera_holo_file(const char* path, uint64_t first_frame, uint64_t end_frame, uint32_t frames_per_call,
                              hv_camera* out);

typedef struct hv_reader hv_reader;

/* The read worker (Backend/sources/thread/camera_frame_read_worker.cc:20-72):
 * init_camera here, then a host thread that starts the acquisition and
 * enqueues every get_frames into q on its own stream until stopped, at most
 * fps_limit frames/s when fps_limit > 0.  The consumer must keep popping
 * (a full queue waits for the batch the consumer holds). */
hv_status hv_reader_start(const hv_camera* cam, hv_input_queue* q, float fps_limit, hv_reader** out);
/* Frames enqueued so far. */
hv_status hv_reader_frames(hv_reader* r, uint64_t* frames);
/* Stops the thread (stop_acquisition), reports the frames enqueued and the
 * first error the thread met.  The camera stays the caller's (shutdown_camera). */
hv_status hv_reader_stop(hv_reader* r, uint64_t* frames);

/* ------------------------------------------------------------------------ */
/* .holo recordings (Backend/includes/io/holo_file.hh:19-102,               */
/* Backend/sources/io/input_file/input_holo_file.cc:19-167,                 */
/* Backend/sources/io/output_file/output_holo_file.cc:10-122)               */
/* ------------------------------------------------------------------------ */

/* holovibes::RecordedDataType (Backend/includes/enum/enum_recorded_data_type.hh:15-21) */
#define HV_HOLO_RAW 0
#define HV_HOLO_PROCESSED 1
#define HV_HOLO_MOMENTS 2

/* HoloFile::HoloFileHeader (holo_file.hh:58-83): 64 bytes, packed on 2
 * bytes, little-endian; the frames follow it, then the JSON footer. */
#pragma pack(push, 2)
typedef struct hv_holo_header {
    char magic_number[4];     /* "HOLO" */
    uint16_t version;         /* 7 when written here */
    uint16_t bits_per_pixel;  /* 8 * bytes per pixel */
    uint32_t img_width;
    uint32_t img_height;
    uint32_t img_nb;
    uint64_t total_data_size; /* img_width * img_height * img_nb * bits_per_pixel / 8 */
    uint8_t endianness;       /* 0 little, 1 big (recorded, not applied, as in the reference) */
    uint8_t data_type;        /* HV_HOLO_* (version 7) */
    char padding[34];
} hv_holo_header;
#pragma pack(pop)

/* What InputHoloFile derives from the header (input_holo_file.cc:19-43, 82-90). */
typedef struct hv_holo_info {
    uint32_t width, height;
    uint32_t depth;            /* bytes per pixel = bits_per_pixel / 8 */
    uint32_t img_nb;
    uint16_t version, bits_per_pixel;
    uint8_t endianness;
    uint8_t data_type;         /* header data_type for version 7, HV_HOLO_RAW before */
    uint8_t truncated;         /* the file ends before the frames the header announces */
    uint8_t reserved;
    uint64_t frame_size;       /* width * height * depth */
    uint64_t total_data_size;
    uint64_t frames_in_file;   /* whole frames present (img_nb unless truncated) */
    uint64_t footer_offset;    /* 64 + total_data_size */
    uint64_t footer_size;      /* bytes after the frames (0: no footer) */
} hv_holo_info;

typedef struct hv_holo_reader hv_holo_reader;
typedef struct hv_holo_writer hv_holo_writer;

/* Fills a version-7 header as OutputHoloFile's constructor does
 * (output_holo_file.cc:10-34): depth in bytes, data_type HV_HOLO_*. */
hv_status hv_holo_make_header(uint32_t width, uint32_t height, uint32_t depth, uint64_t img_nb, int data_type,
                              hv_holo_header* out);

/* InputHoloFile::InputHoloFile + load_header + load_fd (input_holo_file.cc:19-90).
 * HV_ERR_IO when the file cannot be opened; HV_ERR_FORMAT with the
 * reference's messages for a short header or a wrong magic ("invalid
 * bytecode size") and for total_data_size != frame_size * img_nb ("invalid
 * total data size"). */
hv_status hv_holo_open(const char* path, hv_holo_reader** out);
hv_status hv_holo_get_info(const hv_holo_reader* r, hv_holo_info* info);
hv_status hv_holo_get_header(const hv_holo_reader* r, hv_holo_header* h);

/* set_pos_to_frame + read_frames (input_holo_file.cc:45-51,
 * Backend/sources/io/input_file/input_frame_file.cc:7-17) into host memory:
 * reads min(count, frames_in_file - first) frames, never footer bytes;
 * first > frames_in_file is HV_ERR_INVALID_ARG.  Blocking. */
hv_status hv_holo_read_frames(hv_holo_reader* r, uint64_t first, uint64_t count, void* host_dst,
                              uint64_t* frames_read);

/* FileFrameReadWorker::read_copy_file (Backend/sources/thread/file_frame_read_worker.cc:197-271):
 * the same frames straight into device memory through two pinned 32 MiB
 * staging chunks (file reads overlap the DMA of the previous chunk).
 * Returns when the frames are in device memory.  Allocates the reader's
 * pinned staging on first use; no device allocation. */
hv_status hv_holo_load_frames(hv_holo_reader* r, uint64_t first, uint64_t count, void* dev_dst, void* stream,
                              uint64_t* frames_read);

/* The raw footer bytes (input_holo_file.cc:92-133; JSON parsing is the
 * caller's).  dst NULL: *size = footer size only.  HV_ERR_FORMAT "invalid
 * footer" for a truncated file. */
hv_status hv_holo_read_footer(hv_holo_reader* r, char* dst, size_t capacity, size_t* size);
hv_status hv_holo_close(hv_holo_reader* r);

/* OutputHoloFile (output_holo_file.cc:10-122).  create truncates the file and
 * fills the header (not yet written); each write_* appends at the cursor. */
hv_status hv_holo_create(const char* path, uint32_t width, uint32_t height, uint32_t depth, uint64_t img_nb,
                         int data_type, hv_holo_writer** out);
hv_status hv_holo_write_header(hv_holo_writer* w);
/* write_frame for count frames in host memory. */
hv_status hv_holo_write_frames(hv_holo_writer* w, const void* host_src, uint64_t count);
/* The record path's device -> file copy: count frames in device memory,
 * through two pinned staging chunks (the file write of one chunk overlaps
 * the DMA of the next); writes of 256 MiB or more reserve the file's blocks
 * and copy through a shared mapping from several threads.  Returns when the
 * bytes are in the file (page cache). */
hv_status hv_holo_store_frames(hv_holo_writer* w, const void* dev_src, uint64_t count, void* stream);
hv_status hv_holo_write_footer(hv_holo_writer* w, const char* json, size_t size);
/* Rewrites img_nb and total_data_size in the header at offset 0; the cursor stays. */
hv_status hv_holo_correct_number_of_frames(hv_holo_writer* w, uint64_t frames_written);
hv_status hv_holo_writer_close(hv_holo_writer* w);

/* ------------------------------------------------------------------------ */
/* Serving: several sources, admission by HBM budget, latency accounting     */
/* ------------------------------------------------------------------------ */

typedef struct hv_server hv_server;

/* No single reference counterpart (the reference serves one source per
 * process through BatchInputQueue, batch_input_queue.cc:124-235): a server
 * owns several pipes, each with its own HBM ring and lens, admitted while
 * their memory fits in budget_bytes. */
hv_status hv_server_create(size_t budget_bytes, hv_server** out);

/* Admits a source; HV_ERR_BUDGET when its pipe does not fit the budget. */
hv_status hv_server_add_source(hv_server* s, const hv_compute_settings* settings, int* source_id);

/* Releases a source and its memory. */
hv_status hv_server_remove_source(hv_server* s, int source_id);

/* No reference counterpart (continuous batching).  Every source admits at
 * most `slots` requests to the GPU at once (their work overlaps other
 * sources' on the device); later requests wait in a host queue of
 * `pending` entries, admitted in arrival order as slots free (at submit and
 * poll); a request arriving at a full queue sheds the oldest waiting one --
 * newest data wins, as the reference's input queue overwrites its oldest
 * batch when the consumer lags (Backend/sources/data_struct/batch_input_queue.cc:173-178).
 * Defaults: 2 slots, 16 pending.  Applies to sources added later. */
hv_status hv_server_configure(hv_server* s, int slots, int pending);

/* Submits one request: a batch (device memory, batch_size frames, already in
 * HBM) that became available at host time t_arrival_ns (steady clock).
 * Returns immediately with its id; `frames` must stay valid until the
 * request completes or is shed. */
hv_status hv_server_submit_request(hv_server* s, int source_id, const void* frames, int64_t t_arrival_ns,
                                   int64_t* request_id);
/* hv_server_submit_request without the id. */
hv_status hv_server_submit(hv_server* s, int source_id, const void* frames, int64_t t_arrival_ns);

#define HV_REQ_DONE 0       /* completed; output_index names its image */
#define HV_REQ_NO_OUTPUT 1  /* completed without an image (window filling, stride discard) */
#define HV_REQ_SHED 2       /* dropped from a full pending queue, never run */

typedef struct hv_request_result {
    int64_t request_id;
    int32_t source;
    int32_t status;        /* HV_REQ_* */
    int64_t latency_ns;    /* completion seen - t_arrival (0 when shed) */
    int64_t output_index;  /* HV_REQ_DONE: index for hv_server_output_frame; -1 otherwise */
} hv_request_result;

/* Reports finished (and shed) requests, up to max_n, and admits waiting
 * requests into the slots that freed.  Latency = the time the completion
 * of the request's u16 image is seen here - t_arrival. */
hv_status hv_server_poll_requests(hv_server* s, hv_request_result* out, int max_n, int* n_out);

/* hv_server_poll_requests reduced to (source, latency_ns) of the completed
 * requests that produced an image. */
hv_status hv_server_poll(hv_server* s, int* sources, int64_t* latency_ns, int max_n, int* n_out);

/* Blocks until every submitted request has completed or been shed (waiting
 * requests are admitted as slots free); results stay for the next poll. */
hv_status hv_server_drain(hv_server* s);

/* Latest output of a source (device pointers, see hv_pipe_output). */
hv_status hv_server_output(hv_server* s, int source_id, const uint16_t** out_u16, const float** out_f32_image);

/* The u16 image of a request (its output_index), from the source's output
 * queue (hv_pipe_output_frame); HV_ERR_NOT_READY once overwritten. */
hv_status hv_server_output_frame(hv_server* s, int source_id, int64_t output_index, void* stream,
                                 const uint16_t** frame);

/* Requests shed so far and requests waiting now, over all sources. */
hv_status hv_server_stats(hv_server* s, uint64_t* shed, uint32_t* pending);

/* Bytes currently admitted. */
hv_status hv_server_memory_bytes(hv_server* s, size_t* bytes);

hv_status hv_server_destroy(hv_server* s);

#ifdef __cplusplus
}
#endif

#endif /* HOLO_HIP_H */
