// hv_kernels.hpp -- host launchers of the HIP kernels (internal C++ API).
// Arguments are validated by the C-ABI layer (hv_capi.cpp) before these run.
#pragma once

#include "hv_common.hpp"

namespace hv {

// hv_spatial.hip
hv_status launch_convert(float2* out, const void* in, size_t n, int depth, hipStream_t st);
hv_status launch_lens_as(float2* out, int nx, int ny, float z, float lambda, float xs, float ys, int precision,
                         int shift, float scale, hipStream_t st);
hv_status launch_lens_fresnel(float2* out, int side, int width, int height, float lambda, float z, float px,
                              int precision, hipStream_t st);
hv_status launch_shift_corners_c64(float2* buf, int batch, int sx, int sy, hipStream_t st);
hv_status launch_shift_corners_f32(float* buf, int batch, int sx, int sy, hipStream_t st);
hv_status launch_fft2(float2* buf, int batch, int width, int height, int direction, hipStream_t st);
// x = IFFT2(FFT2(x) * lens) * post_scale
hv_status launch_angular_spectrum(float2* buf, int batch, const float2* lens, int width, int height,
                                  float post_scale, hipStream_t st);
hv_status launch_fresnel(float2* buf, int batch, const float2* chirp, int width, int height, hipStream_t st);
// One column pass over `planes` frames of `buf` (in place): forward, inverse,
// or forward x lens then inverse.
enum { COLPASS_FWD = 0, COLPASS_INV = 1, COLPASS_LENS = 2 };
hv_status launch_col_pass(float2* buf, const float2* lens, int width, int height, int planes, int mode,
                          hipStream_t st);
// One row pass over `planes` frames of `buf` (in place, unnormalised): direction -1 forward, +1 inverse.
hv_status launch_row_pass(float2* buf, int width, int height, int planes, int direction, hipStream_t st);
// Column-tiled planes for 2048- and 4096-long columns (power-of-two widths):
// element (y, x) at ((x / CW) H + y) CW + x mod CW with CW the column tile of
// the column pass, so that every column tile is one contiguous block.
bool col_tiled_supported(int width, int height);
// Forward row FFT (direction -1), row-major in, tiled out (out of place).
hv_status launch_row_pass_tiled(const float2* in, float2* out, int width, int height, int planes, int direction,
                                hipStream_t st);
// Inverse row FFT of K planes (column-tiled, or row-major) with the moments
// (k_bins_moments' outputs, img = moment sel) or the image term mean in the
// same kernel; widths row_ifft_epi_supported.
bool row_ifft_epi_supported(int width);
hv_status launch_row_ifft_moments(float* m0, float* m1, float* m2, float* img, const float2* planes, int width,
                                  int height, int T, int fs, int K, float fps, int sel, uint32_t* mm, int mm_slots,
                                  hipStream_t st, bool tiled);
hv_status launch_row_ifft_image(float* out, const float2* planes, int width, int height, int K, int img_type,
                                float count, float scale, uint32_t* mm, int mm_slots, hipStream_t st, bool tiled);
// column FFT x lens x inverse column FFT of tiled planes (in place), lens tiled too
hv_status launch_col_lens_tiled(float2* buf, const float2* lens_t, int width, int height, int planes,
                                hipStream_t st);
hv_status launch_tile_frame(float2* out, const float2* in, int width, int height, hipStream_t st);
hv_status launch_propagate_to_ring(const void* raw, int depth, int batch, int width, int height, int space,
                                   const float2* lens, float2* ring, int ring_size, int slot0, hipStream_t st,
                                   hipEvent_t* marks = nullptr, bool spectral_rows = false,
                                   uint32_t* mm_reset = nullptr, int mm_slots = 1);

// hv_time.hip
hv_status launch_stft(float2* out, const float2* in, int T, size_t P, hipStream_t st);
hv_status launch_complex_to_image(float* out, const float2* in, int img_type, int pmin, int pmax, size_t P,
                                  hipStream_t st);
// Fused STFT + image type.  mm: encoded min/max (may be null).
hv_status launch_stft_image(float* out, const float2* ring, int T, size_t P, int img_type, int pmin, int pmax,
                            uint32_t* mm, hipStream_t st, int mm_slots = 1);
// Fused inverse row FFT + STFT + image type for a window whose T frames hold
// the angular spectrum after its column pass (B = S = T); HV_ERR_UNSUPPORTED
// when (width, T) has no fused kernel (its LDS tile would not fit).
bool row_ifft_stft_supported(int width, int T);
hv_status launch_row_ifft_stft_image(float* out, float2* ring, int width, int height, int T, int img_type, int pmin,
                                     int pmax, uint32_t* mm, hipStream_t st, int mm_slots = 1);
hv_status launch_stft_moments(float* m0, float* m1, float* m2, const float2* ring, int T, size_t P, int f_start,
                              int f_end, float fps, uint32_t* mm, int mm_moment, hipStream_t st, int mm_slots = 1);

// hv_bins.hip: B = S = T STFT images computed bins-first (the STFT bins
// [pmin, pmin + K) of the raw window are propagated instead of its T frames).
// bins_first_supported: the fused row kernels fit (W <= 1024, T <= 64, the
// K plane rows of an image row in LDS); bins_split_supported: the path as
// separate passes (time DFT, row, column, row, image), any W, H, T.
bool bins_first_supported(int width, int height, int T, int K, int depth, bool pca = false);
bool bins_split_supported(int width, int height, int T, int K, int depth);
// Moments M0, M1, M2 (any may be null; `out` receives moment `moment` too, for
// the image) of the propagated planes holding bins [fs, fs + K) of a length-T
// window (inverse row FFT first for the angular spectrum).
hv_status launch_bins_moments(float* m0, float* m1, float* m2, const float2* planes, int width, int height, int T,
                              int fs, int K, int space, float fps, int mm_moment, uint32_t* mm, hipStream_t st,
                              int mm_slots = 1, float2* tiles = nullptr);
// The split path with column-tiled planes (angular spectrum, 2048- and
// 4096-long columns): `tiles` (K x P, apart from `planes`) holds the planes
// between the forward row pass and the inverse row pass, `lens_t` the lens in
// the same layout; the inverse row pass writes row-major planes back.
bool bins_tiled_supported(int width, int height, int T, int K, int depth, int space, bool pca = false);
// The cycle's batch on its way into a sliding window's raw ring: `n` frames
// of `batch` (0 < n < T) go to slots (slot0 + b) mod T of `ring`, the window
// the bins read (in slot order); the few-bin time DFT writes them as it reads
// the window, the other kernels after copies.  Empty (batch null): nothing.
struct RawInsert {
    void* ring = nullptr;
    const void* batch = nullptr;
    int n = 0;
    int slot0 = 0;
};
// Bins + row FFTs (x chirp for Fresnel) into planes[0 .. K) (K x P complex),
// then the column pass over them; launch_bins_propagate runs both, with
// `mark` (optional) recorded in between.
// wts (bins-first PCA): plane k = sum_t wts[t T + k] x_t (the projection onto
// eigenvectors pmin .. pmin + K - 1: the row-major float eigenvectors from
// column pmin, readable 8 columns past the last) instead of DFT bin pmin + k.
hv_status launch_bins_rows(const void* raw, int depth, int width, int height, int T, int slot0, int pmin, int K,
                           int space, const float2* lens, float2* planes, hipStream_t st, uint32_t* mm_reset,
                           int mm_slots = 1, const float2* wts = nullptr, float2* tiles = nullptr,
                           RawInsert ins = {});
hv_status launch_bins_columns(float2* planes, int width, int height, int K, int space, const float2* lens,
                              hipStream_t st, float2* tiles = nullptr, const float2* lens_t = nullptr);
hv_status launch_bins_propagate(const void* raw, int depth, int width, int height, int T, int slot0, int pmin,
                                int K, int space, const float2* lens, float2* planes, hipStream_t st,
                                hipEvent_t mark, uint32_t* mm_reset, int mm_slots = 1, const float2* wts = nullptr,
                                float2* tiles = nullptr, const float2* lens_t = nullptr, RawInsert ins = {});
// Image of the propagated planes: inverse row FFT (angular spectrum) + image
// term, mean over [pmin, pmax], x scale (131071 for PCA images); mm: encoded
// min/max (may be null).
hv_status launch_bins_image(float* out, const float2* planes, int width, int height, int K, int space, int img_type,
                            int pmin, int pmax, uint32_t* mm, hipStream_t st, int mm_slots = 1, float scale = 1.f,
                            float2* tiles = nullptr);

// hv_render.hip
// mm: `slots` accumulator pairs (block_minmax_atomic); the C-ABI uses one.
hv_status launch_minmax_init(uint32_t* mm, hipStream_t st, int slots = 1);
hv_status launch_minmax(uint32_t* mm, const float* in, size_t n, hipStream_t st, int slots = 1);
hv_status launch_minmax_decode(float* out, const uint32_t* mm, hipStream_t st, int slots = 1);
hv_status launch_subsample_f32(float* out, const float* in, int width, int height, int step, hipStream_t st);
hv_status launch_apply_contrast(float* img, size_t n, float mn, float mx, hipStream_t st);
// Where a cycle's u16 image goes: `out` (ring null), or slot pos[0] of a ring
// of `cap` images; the kernel's last workgroup to finish advances pos[0] by
// `step` (the lanes' interleave) modulo cap and clears the count pos[1].
struct OutSlot {
    uint16_t* ring = nullptr;
    uint32_t cap = 0, step = 1;
    uint32_t* pos = nullptr;
};
hv_status launch_float_to_ushort(uint16_t* out, const float* in, size_t n, hipStream_t st, OutSlot slot = OutSlot{});
// range: 2 floats (encoded = 0) or `encoded` pairs of ordered-uint min/max
// accumulators (encoded >= 1).
// range_out (optional): the decoded range actually applied (before invert).
hv_status launch_contrast_to_ushort(uint16_t* out, float* out_f32, const float* in, size_t n, const void* range,
                                    int encoded, int invert, hipStream_t st, float* range_out = nullptr,
                                    OutSlot slot = OutSlot{});
hv_status launch_scale(float* img, size_t n, float factor, hipStream_t st);
size_t percentile_scratch_bytes();
// Writes range[0..1] = values of rank k0, k1 (ascending) of the image (or its
// subsample), optionally clamped to >= 1.
// reticle_scale >= 0 (with step 1): only the pixels of the contrast reticle's
// centred disc (their count is the N of the index rule); < 0: off.
hv_status launch_percentile(float* range, const float* img, int width, int height, int step, float lo, float hi,
                            int clamp, void* scratch, hipStream_t st, float reticle_scale = -1.f);
// launch_percentile with its scratch in a static device array of the library
// (no allocation; calls on one GPU share it).
hv_status launch_percentile_static(float* range, const float* img, int width, int height, int step, float lo, float hi,
                                   int clamp, hipStream_t st);
uint32_t reticle_radius(int width, int height, float scale);
size_t reticle_pixels(int width, int height, float scale);
// have_mm: mm already holds the input's min/max (launch_unwrap_2d with mm)
hv_status launch_rescale_unwrap(float* out, const float* in, size_t n, uint32_t* mm, hipStream_t st, int slots = 1,
                                bool have_mm = false);
hv_status launch_unwrap_2d(float* out, const float* phase, int width, int height, void* scratch, hipStream_t st,
                           uint32_t* mm = nullptr, int slots = 1);
// Power-of-two sides in [16, 4096]: four fused passes (hv_spatial.hip)
bool unwrap_fused_supported(int width, int height);
hv_status launch_unwrap_2d_fused(float* out, const float* phase, int width, int height, void* scratch,
                                 hipStream_t st, uint32_t* mm, int slots);
size_t unwrap_scratch_bytes(int width, int height);

// hv_pca.hip
size_t pca_workspace_bytes(int T, size_t P, int precision);
// cov: T x T column-major, double2 when cov_double (the pipe's, for the FP64
// eigensolver), float2 otherwise (the C-ABI step's).
hv_status launch_pca_cov(void* cov, int cov_double, const float2* ring, int T, size_t P, int precision, void* ws,
                         hipStream_t st);
// Bins-first PCA: the covariance (double2, T x T) of the converted raw
// window, frames in physical ring-slot order (slot n = frame (n - slot0) mod T),
// equal to the propagated ring's up to a constant factor.
bool pca_raw_cov_supported(int T, size_t P);
hv_status launch_pca_cov_raw(double2* cov, const void* raw, int depth, int T, size_t P, int slot0, int precision,
                             void* ws, hipStream_t st);
size_t eig_workspace_bytes(int T);
// evecs32 (optional): the eigenvectors again as float2, row-major (component
// k of eigenvector j at k T + j).  real: cov is real symmetric (its imaginary
// parts are ignored) and the rotations real (T <= 128).
hv_status launch_eig_jacobi(double* evals, double2* evecs, const void* cov, int cov_double, int T, void* work,
                            hipStream_t st, float2* evecs32 = nullptr, int real = 0);
// img_type: an image type (mean over [pmin, pmax], x131071) or HV_IMG_MOMENTS_k
// (weights from fps over [pmin, pmax] = [f_start, f_end])
hv_status launch_pca_image(float* out, const float2* ring, const double2* evecs, int T, size_t P, int img_type,
                           int pmin, int pmax, uint32_t* mm, hipStream_t st, int mm_slots = 1, float fps = 0.f);

// M0, M1, M2 of the PCA output in one projection (MFMA path; three otherwise)
hv_status launch_pca_moments3(float* m0, float* m1, float* m2, const float2* ring, const double2* evecs, int T,
                              size_t P, int fs, int fe, float fps, hipStream_t st);

// hv_post.hip: post-processing of the float image (P a multiple of 4)
size_t renorm_scratch_bytes();
hv_status launch_renormalize(float* img, size_t P, unsigned norm_constant, void* scratch, hipStream_t st);
// state: {next slot, images held}, zeroed when the ring is (re)started
hv_status launch_accumulate(float* img, float* ring, uint32_t* state, int level, size_t P, hipStream_t st);
hv_status launch_log10(float* img, size_t P, hipStream_t st);
// registration: scratch of registration_scratch_bytes(P); mask once, reference
// once, then one shift per cycle (img shifted in place through tmp)
size_t registration_scratch_bytes(size_t P);
hv_status launch_registration_mask(void* scratch, int W, int H, float zone, int swap, hipStream_t st);
hv_status launch_registration_set_reference(void* scratch, const float* img, int W, int H, hipStream_t st);
hv_status launch_registration(void* scratch, float* img, float* tmp, int W, int H, hipStream_t st);
hv_status registration_shift(void* scratch, size_t P, int* shift_host);
// XZ (T x W) and YZ (H x T) cuts of a time-transformed window (bin-major)
hv_status launch_cuts(float* xz, float* yz, const float2* stft, int W, int H, int T, int xmin, int ymin, int xmax,
                      int ymax, int img_type, hipStream_t st);

}  // namespace hv
