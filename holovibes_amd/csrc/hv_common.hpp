// hv_common.hpp -- shared host/device helpers of the HIP path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/holo_hip.h"

namespace hv {

// ---------------------------------------------------------------------------
// host-side error reporting (thread-local text behind hv_last_error)
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
hv_status fail(hv_status code, const std::string& msg);
hv_status hip_fail(hipError_t e, const char* what);

#define HV_HIP(call)                                      \
    do {                                                  \
        hipError_t hv_e_ = (call);                        \
        if (hv_e_ != hipSuccess) return ::hv::hip_fail(hv_e_, #call); \
    } while (0)

#define HV_CHECK_LAUNCH(name)                                  \
    do {                                                       \
        hipError_t hv_e_ = hipGetLastError();                  \
        if (hv_e_ != hipSuccess) return ::hv::hip_fail(hv_e_, name); \
    } while (0)

#define HV_TRY(expr)                      \
    do {                                  \
        hv_status hv_s_ = (expr);         \
        if (hv_s_ != HV_OK) return hv_s_; \
    } while (0)

inline bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }
inline int ilog2(long n) {
    int l = 0;
    while ((1L << l) < n) ++l;
    return l;
}
inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Geometry accepted by the FFT kernels.
inline bool fft_len_ok(long n) { return is_pow2(n) && n >= 16 && n <= 4096; }
inline bool time_len_ok(long n) { return is_pow2(n) && n >= 1 && n <= 256; }

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// Order-preserving float <-> uint mapping for atomic min/max on floats.
__device__ __forceinline__ uint32_t float_to_ordered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
inline uint32_t host_float_to_ordered(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// 64-lane wavefront reductions (DPP/permute via __shfl_xor).
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Pixel conversion of input_queue_to_input_buffer (tools_conversion.cu:134-141).
template <typename TIn>
__device__ __forceinline__ float convert_px(TIn v);
template <>
__device__ __forceinline__ float convert_px<uint8_t>(uint8_t v) { return static_cast<float>(v) * 257; }
template <>
__device__ __forceinline__ float convert_px<uint16_t>(uint16_t v) { return static_cast<float>(v); }
template <>
__device__ __forceinline__ float convert_px<float>(float v) { return v; }

// Image-type contribution of one spectrum value (tools_conversion.cu:20-24, 96-100, 222-226).
__device__ __forceinline__ float image_term(float2 z, int img_type) {
    if (img_type == HV_IMG_MODULUS) return hypotf(z.x, z.y);
    if (img_type == HV_IMG_SQUARED_MODULUS) {
        const float h = hypotf(z.x, z.y);
        return h * h;
    }
    return atan2f(z.y, z.x);
}

// Compile-time image type with a short modulus: sqrt(x^2 + y^2) with the
// hardware square root (1 ulp) instead of hypotf's scaled evaluation; the
// operands here stay far from overflow (|z| < 1e19).
template <int IMG>
__device__ __forceinline__ float image_term_fast(float2 z) {
    if constexpr (IMG == HV_IMG_MODULUS) return __builtin_amdgcn_sqrtf(fmaf(z.x, z.x, z.y * z.y));
    else if constexpr (IMG == HV_IMG_SQUARED_MODULUS) return fmaf(z.x, z.x, z.y * z.y);
    else return atan2f(z.y, z.x);
}

// Block min/max of v, then one atomic per block on the ordered encodings in
// mm[0] (min) and mm[1] (max).  Every thread of the block must call it.
// Deterministic: min and max do not depend on the order of updates.
__device__ __forceinline__ void block_minmax_atomic(float vmin, float vmax, uint32_t* mm) {
    __shared__ float s_min[16], s_max[16];
    vmin = wave_min(vmin);
    vmax = wave_max(vmax);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = (blockDim.x + 63) >> 6;
    if (lane == 0) {
        s_min[wave] = vmin;
        s_max[wave] = vmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = s_min[0], b = s_max[0];
        for (int w = 1; w < nwaves; ++w) {
            a = fminf(a, s_min[w]);
            b = fmaxf(b, s_max[w]);
        }
        atomicMin(&mm[0], float_to_ordered(a));
        atomicMax(&mm[1], float_to_ordered(b));
    }
}

}  // namespace hv
