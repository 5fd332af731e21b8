// hv_capi.cpp -- the extern "C" step functions of include/holo_hip.h:
// argument validation, then dispatch to the kernel launchers.
#include <cstdio>
#include <cstring>
#include <string>

#include "hv_kernels.hpp"

namespace hv {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

hv_status fail(hv_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

hv_status hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? HV_ERR_OUT_OF_MEMORY : HV_ERR_HIP;
}

}  // namespace hv

using namespace hv;

namespace {

hv_status check_frame(int width, int height) {
    if (width <= 0 || height <= 0) return fail(HV_ERR_INVALID_ARG, "width and height must be positive");
    if (!fft_len_ok(width) || !fft_len_ok(height))
        return fail(HV_ERR_UNSUPPORTED, "width and height must be powers of two or even 2^a 3^b 5^c in [16, 4096]");
    return HV_OK;
}

hv_status check_T(int T) {
    if (T <= 0) return fail(HV_ERR_INVALID_ARG, "T must be positive");
    if (!time_len_ok(T)) return fail(HV_ERR_UNSUPPORTED, "T must be in [1, 256]");
    return HV_OK;
}

hv_status check_range(int T, int pmin, int pmax) {
    if (pmin < 0 || pmax < pmin || pmax >= T)
        return fail(HV_ERR_INVALID_ARG, "accumulation range must satisfy 0 <= pmin <= pmax < T");
    return HV_OK;
}

hv_status check_img(int img_type) {
    if (img_type != HV_IMG_MODULUS && img_type != HV_IMG_SQUARED_MODULUS && img_type != HV_IMG_ARGUMENT)
        return fail(HV_ERR_INVALID_ARG, "img_type must be MODULUS, SQUARED_MODULUS or ARGUMENT");
    return HV_OK;
}

hv_status check_res(size_t frame_res) {
    if (frame_res == 0) return fail(HV_ERR_INVALID_ARG, "frame_res must be positive");
    if (frame_res % 256) return fail(HV_ERR_UNSUPPORTED, "frame_res must be a multiple of 256");
    return HV_OK;
}

#define HV_NOT_NULL(p) \
    if (!(p)) return fail(HV_ERR_INVALID_ARG, #p " is null")

}  // namespace

extern "C" {

const char* hv_version(void) { return "holovibes_amd 0.1 (gfx950)"; }

const char* hv_last_error(void) { return g_last_error.c_str(); }

hv_status hv_device_count(int* count) {
    HV_NOT_NULL(count);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return HV_OK;
}

hv_status hv_convert_to_complex(void* out_c64, const void* in, size_t frame_res, int batch, int depth,
                                void* stream) {
    HV_NOT_NULL(out_c64);
    HV_NOT_NULL(in);
    if (frame_res == 0 || batch <= 0) return fail(HV_ERR_INVALID_ARG, "empty input");
    return launch_convert((float2*)out_c64, in, frame_res * (size_t)batch, depth, as_stream(stream));
}

hv_status hv_angular_spectrum_lens(void* lens_c64, int nx, int ny, float z, float lambda, float x_step,
                                   float y_step, int precision, void* stream) {
    HV_NOT_NULL(lens_c64);
    if (nx <= 0 || ny <= 0) return fail(HV_ERR_INVALID_ARG, "lens size must be positive");
    if (lambda <= 0.f || x_step <= 0.f || y_step <= 0.f)
        return fail(HV_ERR_INVALID_ARG, "lambda and steps must be positive");
    return launch_lens_as((float2*)lens_c64, nx, ny, z, lambda, x_step, y_step, precision, 0, 1.0f,
                          as_stream(stream));
}

hv_status hv_shift_corners_c64(void* buf_c64, int batch, int size_x, int size_y, void* stream) {
    HV_NOT_NULL(buf_c64);
    if (batch <= 0 || size_x <= 0 || size_y <= 0) return fail(HV_ERR_INVALID_ARG, "empty input");
    if (size_x % 2 || size_y % 2) return fail(HV_ERR_UNSUPPORTED, "shift_corners needs even sizes");
    return launch_shift_corners_c64((float2*)buf_c64, batch, size_x, size_y, as_stream(stream));
}

hv_status hv_shift_corners_f32(float* buf, int batch, int size_x, int size_y, void* stream) {
    HV_NOT_NULL(buf);
    if (batch <= 0 || size_x <= 0 || size_y <= 0) return fail(HV_ERR_INVALID_ARG, "empty input");
    if (size_x % 2 || size_y % 2) return fail(HV_ERR_UNSUPPORTED, "shift_corners needs even sizes");
    return launch_shift_corners_f32(buf, batch, size_x, size_y, as_stream(stream));
}

hv_status hv_fresnel_transform_lens(void* lens_c64, int lens_side, int height, int width, float lambda, float z,
                                    float pixel_size, int precision, void* stream) {
    HV_NOT_NULL(lens_c64);
    if (width <= 0 || height <= 0 || lens_side < width || lens_side < height)
        return fail(HV_ERR_INVALID_ARG, "lens_side must cover the frame");
    if (lambda <= 0.f || z == 0.f || pixel_size <= 0.f)
        return fail(HV_ERR_INVALID_ARG, "lambda, z and pixel_size must be non-zero");
    return launch_lens_fresnel((float2*)lens_c64, lens_side, width, height, lambda, z, pixel_size, precision,
                               as_stream(stream));
}

hv_status hv_prepare_lens(void* lens_c64, int space, int width, int height, float lambda, float z,
                          float pixel_size_um, int precision, void* stream) {
    HV_TRY(check_frame(width, height));
    if (space == HV_SPACE_NONE) return HV_OK;
    HV_NOT_NULL(lens_c64);
    if (lambda <= 0.f || pixel_size_um <= 0.f)
        return fail(HV_ERR_INVALID_ARG, "lambda and pixel size must be positive");
    if (space == HV_SPACE_ANGULARSP) {
        const float step = pixel_size_um * 1e-6f;  // fourier_transform.cc:128-129
        return launch_lens_as((float2*)lens_c64, width, height, z, lambda, step, step, precision, 1,
                              1.0f / (float)((size_t)width * height), as_stream(stream));
    }
    if (space == HV_SPACE_FRESNELTR) {
        if (z == 0.f) return fail(HV_ERR_INVALID_ARG, "Fresnel needs z != 0");
        const int side = width > height ? width : height;
        return launch_lens_fresnel((float2*)lens_c64, side, width, height, lambda, z, pixel_size_um, precision,
                                   as_stream(stream));
    }
    return fail(HV_ERR_INVALID_ARG, "unknown space transformation");
}

hv_status hv_angular_spectrum(void* buf_c64, int batch, const void* lens_shifted, int width, int height,
                              void* stream) {
    HV_NOT_NULL(buf_c64);
    HV_NOT_NULL(lens_shifted);
    if (batch <= 0) return fail(HV_ERR_INVALID_ARG, "batch must be positive");
    HV_TRY(check_frame(width, height));
    // 1/N on the inverse row pass: N is a power of two, so this equals the
    // division after the inverse FFT (kernel_complex_divide) exactly.
    return launch_angular_spectrum((float2*)buf_c64, batch, (const float2*)lens_shifted, width, height,
                                   1.0f / (float)((size_t)width * height), as_stream(stream));
}

hv_status hv_fresnel_transform(void* buf_c64, int batch, const void* lens, int width, int height, void* stream) {
    HV_NOT_NULL(buf_c64);
    HV_NOT_NULL(lens);
    if (batch <= 0) return fail(HV_ERR_INVALID_ARG, "batch must be positive");
    HV_TRY(check_frame(width, height));
    return launch_fresnel((float2*)buf_c64, batch, (const float2*)lens, width, height, as_stream(stream));
}

hv_status hv_fft2(void* buf_c64, int batch, int width, int height, int direction, void* stream) {
    HV_NOT_NULL(buf_c64);
    if (batch <= 0) return fail(HV_ERR_INVALID_ARG, "batch must be positive");
    if (direction != -1 && direction != 1) return fail(HV_ERR_INVALID_ARG, "direction must be -1 or +1");
    HV_TRY(check_frame(width, height));
    return launch_fft2((float2*)buf_c64, batch, width, height, direction, as_stream(stream));
}

hv_status hv_propagate_to_ring(const void* raw, int depth, int batch, int width, int height, int space,
                               const void* lens, void* ring_c64, int ring_size, int slot0, void* stream) {
    HV_NOT_NULL(raw);
    HV_NOT_NULL(ring_c64);
    if (batch <= 0 || ring_size <= 0) return fail(HV_ERR_INVALID_ARG, "batch and ring_size must be positive");
    if (slot0 < 0 || slot0 >= ring_size) return fail(HV_ERR_INVALID_ARG, "slot0 out of range");
    if (depth != HV_DEPTH_8 && depth != HV_DEPTH_16 && depth != HV_DEPTH_32F)
        return fail(HV_ERR_INVALID_ARG, "depth must be 1, 2 or 4");
    if (space != HV_SPACE_NONE && space != HV_SPACE_FRESNELTR && space != HV_SPACE_ANGULARSP)
        return fail(HV_ERR_INVALID_ARG, "unknown space transformation");
    if (space != HV_SPACE_NONE) {
        HV_NOT_NULL(lens);
        HV_TRY(check_frame(width, height));
    } else if (width <= 0 || height <= 0) {
        return fail(HV_ERR_INVALID_ARG, "width and height must be positive");
    }
    return launch_propagate_to_ring(raw, depth, batch, width, height, space, (const float2*)lens,
                                    (float2*)ring_c64, ring_size, slot0, as_stream(stream));
}

hv_status hv_stft(void* out_c64, const void* in_c64, int T, size_t frame_res, void* stream) {
    HV_NOT_NULL(out_c64);
    HV_NOT_NULL(in_c64);
    HV_TRY(check_T(T));
    HV_TRY(check_res(frame_res));
    if (out_c64 == in_c64) return fail(HV_ERR_INVALID_ARG, "hv_stft is out of place");
    return launch_stft((float2*)out_c64, (const float2*)in_c64, T, frame_res, as_stream(stream));
}

static hv_status complex_to(int img_type, float* out, const void* in_c64, int pmin, int pmax, size_t frame_res,
                            void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(in_c64);
    if (frame_res == 0) return fail(HV_ERR_INVALID_ARG, "frame_res must be positive");
    if (pmin < 0 || pmax < pmin) return fail(HV_ERR_INVALID_ARG, "range must satisfy 0 <= pmin <= pmax");
    return launch_complex_to_image(out, (const float2*)in_c64, img_type, pmin, pmax, frame_res, as_stream(stream));
}

hv_status hv_complex_to_modulus(float* out, const void* in_c64, int pmin, int pmax, size_t frame_res,
                                void* stream) {
    return complex_to(HV_IMG_MODULUS, out, in_c64, pmin, pmax, frame_res, stream);
}
hv_status hv_complex_to_squared_modulus(float* out, const void* in_c64, int pmin, int pmax, size_t frame_res,
                                        void* stream) {
    return complex_to(HV_IMG_SQUARED_MODULUS, out, in_c64, pmin, pmax, frame_res, stream);
}
hv_status hv_complex_to_argument(float* out, const void* in_c64, int pmin, int pmax, size_t frame_res,
                                 void* stream) {
    return complex_to(HV_IMG_ARGUMENT, out, in_c64, pmin, pmax, frame_res, stream);
}

hv_status hv_stft_image(float* out, const void* ring_c64, int T, size_t frame_res, int img_type, int pmin,
                        int pmax, uint32_t* minmax, void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(ring_c64);
    HV_TRY(check_T(T));
    HV_TRY(check_res(frame_res));
    HV_TRY(check_img(img_type));
    HV_TRY(check_range(T, pmin, pmax));
    return launch_stft_image(out, (const float2*)ring_c64, T, frame_res, img_type, pmin, pmax, minmax,
                             as_stream(stream));
}

hv_status hv_stft_moments(float* m0, float* m1, float* m2, const void* ring_c64, int T, size_t frame_res,
                          int f_start, int f_end, float camera_fps, void* stream) {
    HV_NOT_NULL(ring_c64);
    if (!m0 && !m1 && !m2) return fail(HV_ERR_INVALID_ARG, "no moment requested");
    HV_TRY(check_T(T));
    HV_TRY(check_res(frame_res));
    HV_TRY(check_range(T, f_start, f_end));
    return launch_stft_moments(m0, m1, m2, (const float2*)ring_c64, T, frame_res, f_start, f_end, camera_fps,
                               nullptr, 0, as_stream(stream));
}

hv_status hv_stft_cuts(float* out_xz, float* out_yz, const void* stft_c64, int width, int height, int T, int xmin,
                       int ymin, int xmax, int ymax, int img_type, void* stream) {
    HV_NOT_NULL(out_xz);
    HV_NOT_NULL(out_yz);
    HV_NOT_NULL(stft_c64);
    HV_TRY(check_frame(width, height));
    HV_TRY(check_T(T));
    if (xmin < 0 || ymin < 0 || xmin > xmax || ymin > ymax || xmax >= width || ymax >= height)
        return fail(HV_ERR_INVALID_ARG, "cuts: need 0 <= min <= max < frame side");
    return launch_cuts(out_xz, out_yz, (const float2*)stft_c64, width, height, T, xmin, ymin, xmax, ymax, img_type,
                       as_stream(stream));
}

hv_status hv_pca_workspace_size(int T, size_t frame_res, int precision, size_t* bytes) {
    HV_NOT_NULL(bytes);
    HV_TRY(check_T(T));
    HV_TRY(check_res(frame_res));
    *bytes = pca_workspace_bytes(T, frame_res, precision);
    return HV_OK;
}

hv_status hv_pca_cov(void* cov_c64, const void* ring_c64, int T, size_t frame_res, int precision, void* workspace,
                     void* stream) {
    HV_NOT_NULL(cov_c64);
    HV_NOT_NULL(ring_c64);
    HV_NOT_NULL(workspace);
    HV_TRY(check_T(T));
    HV_TRY(check_res(frame_res));
    if (precision < HV_PCA_FP32 || precision > HV_PCA_FP16) return fail(HV_ERR_INVALID_ARG, "unknown PCA precision");
    return launch_pca_cov(cov_c64, 0, (const float2*)ring_c64, T, frame_res, precision, workspace,
                          as_stream(stream));
}

hv_status hv_eig_workspace_size(int T, size_t* bytes) {
    HV_NOT_NULL(bytes);
    HV_TRY(check_T(T));
    *bytes = eig_workspace_bytes(T);
    return HV_OK;
}

hv_status hv_eig_hermitian_f64(double* evals, void* evecs_c128, const void* cov_c64, int T, void* workspace,
                               void* stream) {
    HV_NOT_NULL(evals);
    HV_NOT_NULL(evecs_c128);
    HV_NOT_NULL(cov_c64);
    HV_TRY(check_T(T));
    if (eig_workspace_bytes(T) && !workspace) return fail(HV_ERR_INVALID_ARG, "T > 64 needs a workspace");
    return launch_eig_jacobi(evals, (double2*)evecs_c128, cov_c64, 0, T, workspace, as_stream(stream));
}

hv_status hv_pca_image(float* out, const void* ring_c64, const void* evecs_c128, int T, size_t frame_res,
                       int img_type, int pmin, int pmax, uint32_t* minmax, void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(ring_c64);
    HV_NOT_NULL(evecs_c128);
    HV_TRY(check_T(T));
    HV_TRY(check_res(frame_res));
    HV_TRY(check_img(img_type));
    HV_TRY(check_range(T, pmin, pmax));
    return launch_pca_image(out, (const float2*)ring_c64, (const double2*)evecs_c128, T, frame_res, img_type, pmin,
                            pmax, minmax, as_stream(stream));
}

hv_status hv_pca_moments(float* out, const void* ring_c64, const void* evecs_c128, int T, size_t frame_res,
                         int moment, int f_start, int f_end, float camera_fps, uint32_t* minmax, void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(ring_c64);
    HV_NOT_NULL(evecs_c128);
    HV_TRY(check_T(T));
    HV_TRY(check_res(frame_res));
    if (moment < 0 || moment > 2) return fail(HV_ERR_INVALID_ARG, "moment must be 0, 1 or 2");
    HV_TRY(check_range(T, f_start, f_end));
    return launch_pca_image(out, (const float2*)ring_c64, (const double2*)evecs_c128, T, frame_res,
                            HV_IMG_MOMENTS_0 + moment, f_start, f_end, minmax, as_stream(stream), 1, camera_fps);
}

hv_status hv_unwrap_scratch_size(int width, int height, size_t* bytes) {
    HV_NOT_NULL(bytes);
    HV_TRY(check_frame(width, height));
    *bytes = unwrap_scratch_bytes(width, height);
    return HV_OK;
}

hv_status hv_unwrap_2d(float* out, const float* phase, int width, int height, void* scratch, void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(phase);
    HV_NOT_NULL(scratch);
    HV_TRY(check_frame(width, height));
    return launch_unwrap_2d(out, phase, width, height, scratch, as_stream(stream));
}

hv_status hv_rescale_unwrap2d(float* out, const float* in, size_t frame_res, uint32_t* minmax_scratch,
                              void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(in);
    HV_NOT_NULL(minmax_scratch);
    if (frame_res == 0) return fail(HV_ERR_INVALID_ARG, "frame_res must be positive");
    return launch_rescale_unwrap(out, in, frame_res, minmax_scratch, as_stream(stream));
}

hv_status hv_subsample_frame_f32(float* out, const float* in, int width, int height, int step, void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(in);
    if (width <= 0 || height <= 0 || step <= 0) return fail(HV_ERR_INVALID_ARG, "sizes must be positive");
    if (width % step || height % step) return fail(HV_ERR_INVALID_ARG, "step must divide width and height");
    return launch_subsample_f32(out, in, width, height, step, as_stream(stream));
}

hv_status hv_minmax_init(uint32_t* minmax, void* stream) {
    HV_NOT_NULL(minmax);
    return launch_minmax_init(minmax, as_stream(stream));
}

hv_status hv_minmax(uint32_t* minmax, const float* in, size_t n, void* stream) {
    HV_NOT_NULL(minmax);
    HV_NOT_NULL(in);
    if (n == 0) return fail(HV_ERR_INVALID_ARG, "empty input");
    return launch_minmax(minmax, in, n, as_stream(stream));
}

hv_status hv_minmax_decode(float* out, const uint32_t* minmax, void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(minmax);
    return launch_minmax_decode(out, minmax, as_stream(stream));
}

hv_status hv_percentile_scratch_size(size_t n, size_t* bytes) {
    HV_NOT_NULL(bytes);
    (void)n;
    *bytes = percentile_scratch_bytes();
    return HV_OK;
}

hv_status hv_percentile_range(float* range_out, const float* img, int width, int height, int subsample_step,
                              float lower_pct, float upper_pct, int clamp_ge_one, void* scratch, void* stream) {
    HV_NOT_NULL(range_out);
    HV_NOT_NULL(img);
    HV_NOT_NULL(scratch);
    if (width <= 0 || height <= 0) return fail(HV_ERR_INVALID_ARG, "sizes must be positive");
    const int s = subsample_step > 1 ? subsample_step : 1;
    if (width % s || height % s) return fail(HV_ERR_INVALID_ARG, "step must divide width and height");
    if (!(lower_pct >= 0.f && lower_pct <= 100.f && upper_pct >= 0.f && upper_pct <= 100.f))
        return fail(HV_ERR_INVALID_ARG, "percentiles must lie in [0, 100]");
    return launch_percentile(range_out, img, width, height, s, lower_pct, upper_pct, clamp_ge_one, scratch,
                             as_stream(stream));
}

hv_status hv_percentile_range_noalloc(float* range_out, const float* img, int width, int height, int subsample_step,
                                      float lower_pct, float upper_pct, int clamp_ge_one, void* stream) {
    HV_NOT_NULL(range_out);
    HV_NOT_NULL(img);
    if (width <= 0 || height <= 0) return fail(HV_ERR_INVALID_ARG, "sizes must be positive");
    const int s = subsample_step > 1 ? subsample_step : 1;
    if (width % s || height % s) return fail(HV_ERR_INVALID_ARG, "step must divide width and height");
    if (!(lower_pct >= 0.f && lower_pct <= 100.f && upper_pct >= 0.f && upper_pct <= 100.f))
        return fail(HV_ERR_INVALID_ARG, "percentiles must lie in [0, 100]");
    return launch_percentile_static(range_out, img, width, height, s, lower_pct, upper_pct, clamp_ge_one,
                                    as_stream(stream));
}

hv_status hv_percentile_range_reticle(float* range_out, const float* img, int width, int height, float reticle_scale,
                                      float lower_pct, float upper_pct, int clamp_ge_one, void* scratch,
                                      void* stream) {
    HV_NOT_NULL(range_out);
    HV_NOT_NULL(img);
    HV_NOT_NULL(scratch);
    if (width <= 0 || height <= 0) return fail(HV_ERR_INVALID_ARG, "sizes must be positive");
    // ContrastApi::set_contrast_reticle_scale (contrast_api.cc:278-297)
    if (!(reticle_scale >= 0.f && reticle_scale <= 1.f))
        return fail(HV_ERR_INVALID_ARG, "Reticle scale must be in range [0., 1.]");
    if (!(lower_pct >= 0.f && lower_pct <= 100.f && upper_pct >= 0.f && upper_pct <= 100.f))
        return fail(HV_ERR_INVALID_ARG, "percentiles must lie in [0, 100]");
    return launch_percentile(range_out, img, width, height, 1, lower_pct, upper_pct, clamp_ge_one, scratch,
                             as_stream(stream), reticle_scale);
}

hv_status hv_apply_contrast(float* img, size_t n, float min, float max, void* stream) {
    HV_NOT_NULL(img);
    if (n == 0) return fail(HV_ERR_INVALID_ARG, "empty input");
    return launch_apply_contrast(img, n, min, max, as_stream(stream));
}

hv_status hv_float_to_ushort(uint16_t* out, const float* in, size_t n, void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(in);
    if (n == 0) return fail(HV_ERR_INVALID_ARG, "empty input");
    return launch_float_to_ushort(out, in, n, as_stream(stream));
}

hv_status hv_contrast_to_ushort(uint16_t* out, const float* in, size_t n, const float* range_dev, int invert,
                                void* stream) {
    HV_NOT_NULL(out);
    HV_NOT_NULL(in);
    HV_NOT_NULL(range_dev);
    if (n == 0) return fail(HV_ERR_INVALID_ARG, "empty input");
    return launch_contrast_to_ushort(out, nullptr, in, n, range_dev, 0, invert, as_stream(stream));
}

}  // extern "C"
