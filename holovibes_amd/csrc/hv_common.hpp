// hv_common.hpp -- shared host/device helpers of the HIP path.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
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

// Batch input queue, for the read worker (hv_source.cpp): an enqueue whose
// wait for the slot the consumer holds ends with HV_ERR_NOT_READY once
// `stop` is set, and a wake-up of such waits.
hv_status input_queue_enqueue_until(hv_input_queue* q, const void* frames, int memcpy_kind, uint32_t nb_frames,
                                    void* stream, const std::atomic<bool>* stop);
void input_queue_wake(hv_input_queue* q);

// Geometry accepted by the FFT kernels: powers of two in [16, 4096] (the
// templated kernels) and the other even 2^a 3^b 5^c in that range (the
// runtime-length kernels, hv_spatial.hip).
bool fft_len_any_ok(long n);
inline bool fft_len_ok(long n) { return n >= 16 && n <= 4096 && (is_pow2(n) || fft_len_any_ok(n)); }
// Time windows: any length in [1, 256] (powers of two on the FFT kernels,
// the others by direct DFT of the bins used).
inline bool time_len_ok(long n) { return n >= 1 && n <= 256; }

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// 2-D unwrap helpers (tools_unwrap.cu): the FFT frequency index of the
// circularly shifted fx / fy grids (kernel_init_unwrap_2d + circ_shift by
// (-W/2, -H/2)), and eq_unwrap_2d's product for one element, g 2 pi i conj(i z).
__device__ __forceinline__ float kfreq(int i, int n) { return (float)(i < n / 2 ? i : i - n); }
__device__ __forceinline__ float2 unwrap_eq_term(float2 g, float2 zz) {
    const float2 tpi = make_float2(0.f, (float)(2.0 * M_PI));
    const float2 zi = make_float2(zz.x * 0.f - zz.y * 1.f, -(zz.x * 1.f + zz.y * 0.f));
    return cmul(cmul(g, tpi), zi);
}

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

// Image min/max accumulation, by the `slots` argument of the producers:
//   slots >= 1   -- `slots` pairs of ordered encodings, mm[2 s] (min) and
//                   mm[2 s + 1] (max), updated with atomics (the C-ABI step
//                   functions use one pair);
//   kMmPartials  -- no atomics: block b writes its (min, max) as floats at
//                   ((float2*) mm)[1 + b] and block 0 the block count at mm[0];
//                   the consumer reduces the partials.  Device-scope atomics
//                   from hundreds of blocks finishing together serialise at
//                   the memory side for microseconds; the pipe uses this mode.
constexpr int kMmPartials = -1;

// Block min/max of v, then its contribution to mm (see above).  Every thread
// of the block must call it.  Deterministic: the result does not depend on
// the order of the blocks.  In atomic mode a block first reads its pair and
// skips an update that cannot change it (a stale read only costs an atomic).
__device__ __forceinline__ void block_minmax_atomic(float vmin, float vmax, uint32_t* mm, int slots = 1) {
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
        const unsigned bid = blockIdx.x + blockIdx.y * gridDim.x;
        if (slots == kMmPartials) {
            reinterpret_cast<float2*>(mm)[1 + bid] = make_float2(a, b);
            if (bid == 0) mm[0] = gridDim.x * gridDim.y;
            return;
        }
        const uint32_t ea = float_to_ordered(a), eb = float_to_ordered(b);
        const int slot = slots > 1 ? (int)(bid % (unsigned)slots) : 0;
        uint32_t* p = mm + 2 * slot;
        if (ea < __hip_atomic_load(&p[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&p[0], ea);
        if (eb > __hip_atomic_load(&p[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&p[1], eb);
    }
}

// Min and max of the accumulators, for every thread of the block (all must
// call it: the partials are reduced by the whole block).
__device__ __forceinline__ void mm_decode_block(const uint32_t* mm, int slots, float& mn, float& mx) {
    __shared__ float d_min[16], d_max[16];
    float a = INFINITY, b = -INFINITY;
    if (slots == kMmPartials) {
        const unsigned n = mm[0];
        const float2* parts = reinterpret_cast<const float2*>(mm) + 1;
        for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
            const float2 v = parts[i];
            a = fminf(a, v.x);
            b = fmaxf(b, v.y);
        }
    } else {
        for (int i = threadIdx.x; i < slots; i += blockDim.x) {
            a = fminf(a, ordered_to_float(mm[2 * i]));
            b = fmaxf(b, ordered_to_float(mm[2 * i + 1]));
        }
    }
    a = wave_min(a);
    b = wave_max(b);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = (blockDim.x + 63) >> 6;
    if (lane == 0) {
        d_min[wave] = a;
        d_max[wave] = b;
    }
    __syncthreads();
    a = d_min[0];
    b = d_max[0];
    for (int w = 1; w < nwaves; ++w) {
        a = fminf(a, d_min[w]);
        b = fmaxf(b, d_max[w]);
    }
    mn = a;
    mx = b;
}

// Re-arms atomic accumulator pairs (threads of one block, any count); the
// partials mode needs no re-arming.
__device__ __forceinline__ void mm_reset_slots(uint32_t* mm, int slots) {
    for (int s = threadIdx.x; s < slots; s += blockDim.x) {
        mm[2 * s] = float_to_ordered(INFINITY);
        mm[2 * s + 1] = float_to_ordered(-INFINITY);
    }
}

}  // namespace hv
