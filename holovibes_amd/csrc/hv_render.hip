// hv_render.hip -- contrast, float -> u16, subsample, percentile, 2-D unwrap.
//
// Reductions are device-resident: the min/max of an image is kept as the
// order-preserving uint32 encoding of two floats and updated with one
// atomicMin/atomicMax per workgroup after a 64-lane shuffle reduction, so no
// stage reads anything back to the host (the reference's rescale_float,
// rescale_float_unwrap2d and compute_percentile all synchronise the stream).
// Percentiles are exact order statistics found by a 3-pass radix select
// (11 + 11 + 10 bits of the ordered key) instead of a full sort.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <mutex>

#include "hv_fft.hpp"
#include "hv_kernels.hpp"

namespace hv {

static unsigned blocks_for(size_t n, int threads, size_t cap = 16384) {
    size_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// ---------------------------------------------------------------------------
// min / max
// ---------------------------------------------------------------------------
__global__ void k_minmax_init(uint32_t* mm, int slots) { mm_reset_slots(mm, slots); }

__global__ void __launch_bounds__(256) k_minmax(uint32_t* mm, const float* __restrict__ in, size_t n, int slots) {
    float a = INFINITY, b = -INFINITY;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = in[i];
        a = fminf(a, v);
        b = fmaxf(b, v);
    }
    block_minmax_atomic(a, b, mm, slots);
}

__global__ void k_minmax_decode(float* out, const uint32_t* mm, int slots) {
    float a, b;
    mm_decode_block(mm, slots, a, b);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

// Reduces per-block partials in place to one entry (count 1).  Consumers
// with many blocks each reduce every partial; past a few hundred consumer
// blocks one reduction up front is cheaper than that quadratic read traffic.
__global__ void __launch_bounds__(1024) k_mm_collapse(uint32_t* mm) {
    float a, b;
    mm_decode_block(mm, kMmPartials, a, b);  // ends with a barrier after every read of the partials
    if (threadIdx.x == 0) {
        reinterpret_cast<float2*>(mm)[1] = make_float2(a, b);
        mm[0] = 1;
    }
}

constexpr unsigned kDecodeBlocksDirect = 256;

static hv_status collapse_partials(uint32_t* mm, unsigned consumer_blocks, int slots, hipStream_t st) {
    if (slots != kMmPartials || consumer_blocks <= kDecodeBlocksDirect) return HV_OK;
    k_mm_collapse<<<1, 1024, 0, st>>>(mm);
    HV_CHECK_LAUNCH("k_mm_collapse");
    return HV_OK;
}

hv_status launch_minmax_init(uint32_t* mm, hipStream_t st, int slots) {
    k_minmax_init<<<1, 64, 0, st>>>(mm, slots);
    HV_CHECK_LAUNCH("k_minmax_init");
    return HV_OK;
}
hv_status launch_minmax(uint32_t* mm, const float* in, size_t n, hipStream_t st, int slots) {
    k_minmax<<<blocks_for(n, 256, 2048), 256, 0, st>>>(mm, in, n, slots);
    HV_CHECK_LAUNCH("k_minmax");
    return HV_OK;
}
hv_status launch_minmax_decode(float* out, const uint32_t* mm, hipStream_t st, int slots) {
    k_minmax_decode<<<1, 64, 0, st>>>(out, mm, slots);
    HV_CHECK_LAUNCH("k_minmax_decode");
    return HV_OK;
}

// ---------------------------------------------------------------------------
// subsample_frame (frame_reshape.cu:278-305)
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t subsample_index(size_t index, int in_w, int out_w, int step) {
    const uint32_t x = index % out_w, y = index / out_w;
    const uint32_t x_off = y % 2 == 0 ? 0 : step / 2;
    const uint32_t y_off = x % 2 == 0 ? 0 : step / 2;
    return (size_t)y * step * in_w + (size_t)y_off * in_w + (size_t)x * step + x_off;
}

__global__ void k_subsample(float* __restrict__ out, const float* __restrict__ in, int in_w, int out_w, int step,
                            size_t n_out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_out) out[i] = in[subsample_index(i, in_w, out_w, step)];
}

hv_status launch_subsample_f32(float* out, const float* in, int width, int height, int step, hipStream_t st) {
    const size_t n = (size_t)(width / step) * (height / step);
    k_subsample<<<blocks_for(n, 256, 1u << 30), 256, 0, st>>>(out, in, width, width / step, step, n);
    HV_CHECK_LAUNCH("k_subsample");
    return HV_OK;
}

// ---------------------------------------------------------------------------
// contrast and float -> u16 (contrast_correction.cu:5-17, tools_conversion.cu:350-359)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f2u16(float x) {
    if (x <= 0.0f) return 0;
    if (x >= 65536.0f) return 65535;  // uint(x) > 65535 (the cast saturates on the GPU)
    return (uint16_t)(uint32_t)x;
}

__global__ void __launch_bounds__(256) k_apply_contrast(float* img, size_t n, float mn, float mx) {
    const float factor = (unsigned short)65535 / (mx - mn + FLT_EPSILON);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        img[i] = factor * (img[i] - mn);
}

// The image's destination: `out`, or the output queue slot pos[0].  Every
// block reads the slot before it counts itself done (out_slot_done), so the
// last block's advance cannot reach a block still to write.  One thread per
// block reads it (device-scope accesses to one address are served one after
// another: a read per wavefront cost ~10 us over 4096 wavefronts).
__device__ __forceinline__ uint16_t* out_slot_target(uint16_t* out, const OutSlot& o, size_t n) {
    if (!o.ring) return out;
    __shared__ uint32_t slot;
    if (threadIdx.x == 0) slot = __hip_atomic_load(&o.pos[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return o.ring + (size_t)slot * n;
}
__device__ __forceinline__ void out_slot_done(const OutSlot& o) {
    if (!o.ring) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t prev = atomicAdd(&o.pos[1], 1u);
        if (prev == gridDim.x * gridDim.y - 1) {
            const uint32_t slot = __hip_atomic_load(&o.pos[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&o.pos[0], (slot + o.step) % o.cap, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&o.pos[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ void __launch_bounds__(256) k_float_to_ushort(uint16_t* out, const float* in, size_t n, OutSlot slot) {
    uint16_t* dst = out_slot_target(out, slot, n);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = f2u16(in[i]);
    out_slot_done(slot);
}

// V float4 per thread, their loads issued together, V chosen so that an
// image of 1024^2 or more still gets 256 blocks: fewer blocks mean fewer of
// the output slot's device-scope accesses and of the blocks' min/max
// decodes; smaller images keep one float4 per thread and 256 blocks or less.
static int float4_per_thread(size_t n) {
    const size_t n4 = n / 4;
    return n4 >= (size_t)4 * 256 * 256 ? 4 : n4 >= (size_t)2 * 256 * 256 ? 2 : 1;
}

template <int V>
__global__ void __launch_bounds__(256)
    k_contrast_u16(uint16_t* __restrict__ out_arg, float* __restrict__ out_f32, const float* __restrict__ in, size_t n,
                   const void* range, int encoded, int invert, float* range_out, OutSlot slot, int vec) {
    uint16_t* __restrict__ out = out_slot_target(out_arg, slot, n);
    float mn, mx;
    if (encoded) {
        mm_decode_block((const uint32_t*)range, encoded, mn, mx);
    } else {
        mn = ((const float*)range)[0];
        mx = ((const float*)range)[1];
    }
    if (range_out && blockIdx.x == 0 && threadIdx.x == 0) {
        range_out[0] = mn;
        range_out[1] = mx;
    }
    if (invert) {
        const float t = mn;
        mn = mx;
        mx = t;
    }
    const float factor = (unsigned short)65535 / (mx - mn + FLT_EPSILON);
    const size_t n4 = vec ? n / 4 : 0;  // unaligned buffers: the element loop below takes every pixel
    const float4* in4 = reinterpret_cast<const float4*>(in);
    const size_t step = (size_t)gridDim.x * blockDim.x * V;
    for (size_t i0 = (size_t)blockIdx.x * blockDim.x * V + threadIdx.x; i0 < n4; i0 += step) {
        float4 v[V];
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const size_t i = i0 + (size_t)j * blockDim.x;
            if (i < n4) v[j] = in4[i];
        }
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const size_t i = i0 + (size_t)j * blockDim.x;
            if (i >= n4) break;
            float4 w = v[j];
            w.x = factor * (w.x - mn);
            w.y = factor * (w.y - mn);
            w.z = factor * (w.z - mn);
            w.w = factor * (w.w - mn);
            if (out_f32) reinterpret_cast<float4*>(out_f32)[i] = w;
            ushort4 u;
            u.x = f2u16(w.x);
            u.y = f2u16(w.y);
            u.z = f2u16(w.z);
            u.w = f2u16(w.w);
            reinterpret_cast<ushort4*>(out)[i] = u;
        }
    }
    for (size_t i = n4 * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = factor * (in[i] - mn);
        if (out_f32) out_f32[i] = v;
        out[i] = f2u16(v);
    }
    out_slot_done(slot);
}

__global__ void __launch_bounds__(256) k_scale(float* img, size_t n, float f) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        img[i] = img[i] * f;
}

hv_status launch_apply_contrast(float* img, size_t n, float mn, float mx, hipStream_t st) {
    k_apply_contrast<<<blocks_for(n, 256), 256, 0, st>>>(img, n, mn, mx);
    HV_CHECK_LAUNCH("k_apply_contrast");
    return HV_OK;
}
hv_status launch_float_to_ushort(uint16_t* out, const float* in, size_t n, hipStream_t st, OutSlot slot) {
    k_float_to_ushort<<<blocks_for(n, 256), 256, 0, st>>>(out, in, n, slot);
    HV_CHECK_LAUNCH("k_float_to_ushort");
    return HV_OK;
}
hv_status launch_contrast_to_ushort(uint16_t* out, float* out_f32, const float* in, size_t n, const void* range,
                                    int encoded, int invert, hipStream_t st, float* range_out, OutSlot slot) {
    const int V = float4_per_thread(n);
    const unsigned blocks = blocks_for((n + 3) / 4, 256 * V);
    if (encoded == kMmPartials) HV_TRY(collapse_partials((uint32_t*)range, blocks, encoded, st));
    const int vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out_f32)) & 15) == 0 &&
                    ((reinterpret_cast<uintptr_t>(out) | reinterpret_cast<uintptr_t>(slot.ring)) & 7) == 0 &&
                    (!slot.ring || n % 4 == 0);  // queue slots of n pixels stay 8-byte aligned
    if (V == 4) k_contrast_u16<4><<<blocks, 256, 0, st>>>(out, out_f32, in, n, range, encoded, invert, range_out, slot, vec);
    else if (V == 2) k_contrast_u16<2><<<blocks, 256, 0, st>>>(out, out_f32, in, n, range, encoded, invert, range_out, slot, vec);
    else k_contrast_u16<1><<<blocks, 256, 0, st>>>(out, out_f32, in, n, range, encoded, invert, range_out, slot, vec);
    HV_CHECK_LAUNCH("k_contrast_u16");
    return HV_OK;
}
hv_status launch_scale(float* img, size_t n, float factor, hipStream_t st) {
    k_scale<<<blocks_for(n, 256), 256, 0, st>>>(img, n, factor);
    HV_CHECK_LAUNCH("k_scale");
    return HV_OK;
}

// ---------------------------------------------------------------------------
// percentiles: exact radix select on ordered keys
// ---------------------------------------------------------------------------
constexpr int kPctBins = 2048;
struct PctScratch {
    uint32_t hist[2][kPctBins];
    uint32_t prefix[2];
    uint32_t rank[2];
};
size_t percentile_scratch_bytes() { return sizeof(PctScratch); }

__global__ void k_pct_init(PctScratch* s, uint32_t r0, uint32_t r1) {
    for (int i = threadIdx.x; i < 2 * kPctBins; i += blockDim.x) (&s->hist[0][0])[i] = 0;
    if (threadIdx.x == 0) {
        s->prefix[0] = s->prefix[1] = 0;
        s->rank[0] = r0;
        s->rank[1] = r1;
    }
}

// The pixels the percentiles are taken over: every pixel of the (optionally
// subsampled) image, or those of the contrast reticle's centred disc.
struct PctZone {
    int step = 1;
    int disc = 0;          // 1: only (x - cx)^2 + (y - cy)^2 <= r2
    int cx = 0, cy = 0;
    uint32_t r2 = 0;
};

// Histogram of digit [shift, shift + width) over keys whose higher bits equal
// the running prefix of each of the two selections.
__global__ void __launch_bounds__(256)
    k_pct_hist(PctScratch* s, const float* __restrict__ img, int in_w, int out_w, PctZone z, size_t n, int shift,
               int width) {
    __shared__ uint32_t h[2][kPctBins];
    for (int i = threadIdx.x; i < 2 * kPctBins; i += blockDim.x) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t p0 = s->prefix[0], p1 = s->prefix[1];
    const int hs = shift + width;
    const uint32_t mask = (1u << width) - 1;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        size_t src = i;
        if (z.step > 1) {
            src = subsample_index(i, in_w, out_w, z.step);
        } else if (z.disc) {
            // copy_circle_kernel (percentile.cu:15-43)
            const int dx = (int)(i % (size_t)in_w) - z.cx, dy = (int)(i / (size_t)in_w) - z.cy;
            if ((uint32_t)(dx * dx + dy * dy) > z.r2) continue;
        }
        const uint32_t key = float_to_ordered(img[src]);
        const uint32_t digit = (key >> shift) & mask;
        const uint32_t high = hs >= 32 ? 0u : (key >> hs);
        if (high == p0) atomicAdd(&h[0][digit], 1u);
        if (high == p1) atomicAdd(&h[1][digit], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * kPctBins; i += blockDim.x) {
        const uint32_t v = (&h[0][0])[i];
        if (v) atomicAdd(&(&s->hist[0][0])[i], v);
    }
}

// One block per selection q: find the bin holding rank[q], extend the prefix,
// reduce the rank, clear the histogram for the next pass.  On the last pass
// write the selected value (clamped to >= 1 when requested).
__global__ void __launch_bounds__(256)
    k_pct_select(PctScratch* s, int width, int last, float* out, int clamp) {
    const int q = blockIdx.x;
    constexpr int PER = kPctBins / 256;
    __shared__ uint32_t part[256];
    __shared__ uint32_t found_bin, found_before;
    const int t = threadIdx.x;
    uint32_t local[PER];
    uint32_t sum = 0;
    for (int k = 0; k < PER; ++k) {
        local[k] = (t * PER + k) < (1 << width) ? s->hist[q][t * PER + k] : 0u;
        sum += local[k];
    }
    part[t] = sum;
    __syncthreads();
    // inclusive Hillis-Steele scan of the 256 partial sums
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const uint32_t rank = s->rank[q];
    const uint32_t before = t > 0 ? part[t - 1] : 0u;
    if (rank >= before && rank < part[t]) {
        uint32_t c = before;
        for (int k = 0; k < PER; ++k) {
            if (rank < c + local[k]) {
                found_bin = t * PER + k;
                found_before = c;
                break;
            }
            c += local[k];
        }
    }
    __syncthreads();
    for (int k = 0; k < PER; ++k) s->hist[q][t * PER + k] = 0;
    if (t == 0) {
        const uint32_t prefix = (s->prefix[q] << width) | found_bin;
        s->prefix[q] = prefix;
        s->rank[q] = rank - found_before;
        if (last) {
            float v = ordered_to_float(prefix);
            if (clamp) v = v > 1.0f ? v : 1.0f;  // contrast_api.cc:139-140
            out[q] = v;
        }
    }
}

// The same radix select in one workgroup when the visited pixels fit in LDS
// (subsampled images: the 8 x 8 subsample of 1024^2 is 16384 keys): the
// keys are read once, the three digit passes and both selections run on
// them in LDS, and the range is written -- one launch instead of seven,
// with the same result (the selection is exact).
constexpr int kPctSmallMax = 16384;
constexpr int kPctSmallThreads = 1024;

__global__ void __launch_bounds__(kPctSmallThreads)
    k_pct_small(const float* __restrict__ img, int in_w, int out_w, int step, int n, uint32_t r0, uint32_t r1,
                float* out, int clamp) {
    __shared__ uint32_t keys[kPctSmallMax];
    __shared__ uint32_t h[2][kPctBins];
    __shared__ uint32_t wsum[2][kPctSmallThreads / 2 / 64];
    __shared__ uint32_t sel_bin[2], sel_before[2];
    constexpr int PER = 2 * kPctBins / kPctSmallThreads;  // bins per thread of its selection's half
    const int t = threadIdx.x;
    // every key's load in flight at once (a loop of dependent loads took 32 us for 16384 keys)
    constexpr int KPT = kPctSmallMax / kPctSmallThreads;
    float v[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const int i = t + j * kPctSmallThreads;
        if (i < n) v[j] = img[step > 1 ? subsample_index((size_t)i, in_w, out_w, step) : (size_t)i];
    }
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const int i = t + j * kPctSmallThreads;
        if (i < n) keys[i] = float_to_ordered(v[j]);
    }
    uint32_t prefix0 = 0, prefix1 = 0, rank0 = r0, rank1 = r1;
    const int shifts[3] = {21, 10, 0}, widths[3] = {11, 11, 10};
    for (int d = 0; d < 3; ++d) {
        for (int i = t; i < 2 * kPctBins; i += kPctSmallThreads) (&h[0][0])[i] = 0;
        __syncthreads();
        const int shift = shifts[d], width = widths[d], hs = shift + width;
        const uint32_t mask = (1u << width) - 1;
        // both selections share the first digit's histogram (both prefixes are empty)
        const bool shared = d == 0;
        uint32_t kv[KPT];
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int i = t + j * kPctSmallThreads;
            kv[j] = i < n ? keys[i] : 0u;
        }
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            if (t + j * kPctSmallThreads >= n) break;
            const uint32_t key = kv[j];
            const uint32_t digit = (key >> shift) & mask;
            const uint32_t high = hs >= 32 ? 0u : (key >> hs);
            if (high == prefix0) atomicAdd(&h[0][digit], 1u);
            if (!shared && high == prefix1) atomicAdd(&h[1][digit], 1u);
        }
        __syncthreads();
        // threads [0, 512) select for the lower percentile, [512, 1024) for the upper
        const int q = t / (kPctSmallThreads / 2), u = t % (kPctSmallThreads / 2);
        const int qh = shared ? 0 : q;
        const uint32_t rank = q ? rank1 : rank0;
        uint32_t local[PER], sum = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int b = u * PER + k;
            local[k] = b < (1 << width) ? h[qh][b] : 0u;
            sum += local[k];
        }
        // inclusive scan of each half's partial sums: within each wavefront by
        // shuffles, then the half's wavefront totals
        const int lane = u & 63, wv = u >> 6;
        uint32_t incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (lane == 63) wsum[q][wv] = incl;
        __syncthreads();
        for (int w = 0; w < wv; ++w) incl += wsum[q][w];
        const uint32_t before = incl - sum;
        if (rank >= before && rank < incl) {
            uint32_t c = before;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                if (rank < c + local[k]) {
                    sel_bin[q] = u * PER + k;
                    sel_before[q] = c;
                    break;
                }
                c += local[k];
            }
        }
        __syncthreads();
        prefix0 = (prefix0 << width) | sel_bin[0];
        prefix1 = (prefix1 << width) | sel_bin[1];
        rank0 -= sel_before[0];
        rank1 -= sel_before[1];
        __syncthreads();  // sel_* and h are rewritten by the next pass
    }
    if (t < 2) {
        float v = ordered_to_float(t ? prefix1 : prefix0);
        if (clamp) v = v > 1.0f ? v : 1.0f;  // contrast_api.cc:139-140
        out[t] = v;
    }
}

uint32_t reticle_radius(int width, int height, float scale) {
    // compute_percentile_xy_view (percentile.cu:166-170): float product, truncated
    return (uint32_t)(scale * (float)std::min(width, height) / 2);
}

size_t reticle_pixels(int width, int height, float scale) {
    const long r = (long)reticle_radius(width, height, scale), r2 = r * r;
    const long cx = width / 2, cy = height / 2;
    size_t count = 0;
    for (long y = 0; y < height; ++y) {
        const long dy2 = (y - cy) * (y - cy);
        if (dy2 > r2) continue;
        // largest m with m^2 <= r2 - dy2
        long m = (long)std::sqrt((double)(r2 - dy2));
        while (m * m > r2 - dy2) --m;
        while ((m + 1) * (m + 1) <= r2 - dy2) ++m;
        const long x0 = std::max(0L, cx - m), x1 = std::min((long)width - 1, cx + m);
        if (x1 >= x0) count += (size_t)(x1 - x0 + 1);
    }
    return count;
}

hv_status launch_percentile(float* range, const float* img, int width, int height, int step, float lo, float hi,
                            int clamp, void* scratch, hipStream_t st, float reticle_scale) {
    if (step < 1) step = 1;
    const int out_w = width / step, out_h = height / step;
    PctZone z;
    z.step = step;
    size_t n = (size_t)out_w * out_h;  // pixels visited
    size_t count = n;                  // pixels selected (N of the index rule)
    if (reticle_scale >= 0.f && step == 1) {
        const uint32_t r = reticle_radius(width, height, reticle_scale);
        z.disc = 1;
        z.cx = width / 2;
        z.cy = height / 2;
        z.r2 = r * r;
        count = reticle_pixels(width, height, reticle_scale);
    }
    if (count == 0) return fail(HV_ERR_INVALID_ARG, "empty image");
    // compute_percentile (percentile.cu:84-91): index = uint(p / 100 * N), float arithmetic
    auto rank_of = [count](float pct) {
        const float f = pct / 100 * (float)count;
        uint32_t r = f <= 0.f ? 0u : (uint32_t)f;
        if (r >= count) r = (uint32_t)count - 1;  // p = 100 would read one past the end
        return r;
    };
    if (!z.disc && n <= (size_t)kPctSmallMax) {
        k_pct_small<<<1, kPctSmallThreads, 0, st>>>(img, width, out_w, step, (int)n, rank_of(lo), rank_of(hi), range,
                                                    clamp);
        HV_CHECK_LAUNCH("k_pct_small");
        return HV_OK;
    }
    PctScratch* s = reinterpret_cast<PctScratch*>(scratch);
    k_pct_init<<<1, 256, 0, st>>>(s, rank_of(lo), rank_of(hi));
    HV_CHECK_LAUNCH("k_pct_init");
    const int shifts[3] = {21, 10, 0}, widths[3] = {11, 11, 10};
    const unsigned g = blocks_for(n, 256 * 8, 1024);
    for (int d = 0; d < 3; ++d) {
        k_pct_hist<<<g, 256, 0, st>>>(s, img, width, out_w, z, n, shifts[d], widths[d]);
        HV_CHECK_LAUNCH("k_pct_hist");
        k_pct_select<<<2, 256, 0, st>>>(s, widths[d], d == 2, range, clamp);
        HV_CHECK_LAUNCH("k_pct_select");
    }
    return HV_OK;
}

// ---------------------------------------------------------------------------
// rescale_float_unwrap2d (tools_conversion.cu:315-335), device min/max
// ---------------------------------------------------------------------------
// V float4 per thread like k_contrast_u16: 256 blocks for a 1024^2 image,
// which decode the unwrap's min/max partials directly.
template <int V>
__global__ void __launch_bounds__(256) k_rescale_unwrap(float* out, const float* in, size_t n, const uint32_t* mm,
                                                        int slots, int vec) {
    float mn, mx;
    mm_decode_block(mm, slots, mn, mx);
    const float m = (unsigned short)65535;
    auto f = [&](float v) { return mn < 0.f ? (v + fabsf(mn)) / (fabsf(mn) + mx) * m : (v - mn) / (mx - mn) * m; };
    const size_t n4 = vec ? n / 4 : 0;
    const size_t step = (size_t)gridDim.x * blockDim.x * V;
    for (size_t i0 = (size_t)blockIdx.x * blockDim.x * V + threadIdx.x; i0 < n4; i0 += step) {
        float4 v[V];
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const size_t i = i0 + (size_t)j * blockDim.x;
            if (i < n4) v[j] = reinterpret_cast<const float4*>(in)[i];
        }
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const size_t i = i0 + (size_t)j * blockDim.x;
            if (i >= n4) break;
            reinterpret_cast<float4*>(out)[i] = make_float4(f(v[j].x), f(v[j].y), f(v[j].z), f(v[j].w));
        }
    }
    for (size_t i = n4 * 4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = f(in[i]);
}

hv_status launch_rescale_unwrap(float* out, const float* in, size_t n, uint32_t* mm, hipStream_t st, int slots,
                                bool have_mm) {
    if (!have_mm) {
        HV_TRY(launch_minmax_init(mm, st, slots));
        HV_TRY(launch_minmax(mm, in, n, st, slots));
    }
    const int V = float4_per_thread(n);
    const unsigned blocks = blocks_for((n + 3) / 4, 256 * V);
    HV_TRY(collapse_partials(mm, blocks, slots, st));
    const int vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    if (V == 4) k_rescale_unwrap<4><<<blocks, 256, 0, st>>>(out, in, n, mm, slots, vec);
    else if (V == 2) k_rescale_unwrap<2><<<blocks, 256, 0, st>>>(out, in, n, mm, slots, vec);
    else k_rescale_unwrap<1><<<blocks, 256, 0, st>>>(out, in, n, mm, slots, vec);
    HV_CHECK_LAUNCH("k_rescale_unwrap");
    return HV_OK;
}

// ---------------------------------------------------------------------------
// unwrap_2d (tools_unwrap.cu:269-369): periodic least-squares unwrapping
// ---------------------------------------------------------------------------
size_t unwrap_scratch_bytes(int width, int height) { return (size_t)width * height * sizeof(float2) * 3; }

__global__ void k_unwrap_init(float2* z, float2* g, const float* phase, size_t P) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float2 v = make_float2(cosf(phase[i]), sinf(phase[i]));
    z[i] = v;
    g[i] = v;
}

// gradient_unwrap_2d (:298-322): Gx = FFT(z) fx, Gy = FFT(z) fy
__global__ void k_unwrap_grad(float2* gx, float2* gy, int width, int height) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= width) return;
    const size_t i = (size_t)y * width + x;
    const float2 v = gx[i];
    const float fx = kfreq(x, width), fy = kfreq(y, height);
    gx[i] = make_float2(v.x * fx, v.y * fx);
    gy[i] = make_float2(v.x * fy, v.y * fy);
}

// (gx, gy) * 2 pi i, then eq_unwrap_2d (:324-351): z <- conj(z i), g *= z
__global__ void k_unwrap_eq(float2* gx, float2* gy, const float2* z, size_t P) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float2 zz = z[i];
    gx[i] = unwrap_eq_term(gx[i], zz);
    gy[i] = unwrap_eq_term(gy[i], zz);
}

// kernel_norm_ratio + kernel_add_complex_frames: G = Gx fx/|f|^2 + Gy fy/|f|^2, 0 at DC
__global__ void k_unwrap_ratio(float2* gx, const float2* gy, int width, int height) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= width) return;
    const size_t i = (size_t)y * width + x;
    const float fx = kfreq(x, width), fy = kfreq(y, height);
    const float norm = fx * fx + fy * fy;
    float2 a = make_float2(0.f, 0.f), b = make_float2(0.f, 0.f);
    if (norm != 0) {
        const float cx = fx / norm, cy = fy / norm;
        a = make_float2(gx[i].x * cx, gx[i].y * cx);
        b = make_float2(gy[i].x * cy, gy[i].y * cy);
    }
    gx[i] = cadd(a, b);
}

// kernel_unwrap2d_last_step (:197-203)
__global__ void k_unwrap_last(float* out, const float2* g, size_t P) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P) out[i] = (float)((double)g[i].y / -(2.0 * M_PI));
}

// The steps above one after another (frame sides that are not powers of
// two); powers of two run the fused passes of launch_unwrap_2d_fused.
// mm: the output's min/max accumulated there (re-armed first), for the
// rescale; nullptr: none.
hv_status launch_unwrap_2d(float* out, const float* phase, int width, int height, void* scratch, hipStream_t st,
                           uint32_t* mm, int slots) {
    if (unwrap_fused_supported(width, height))
        return launch_unwrap_2d_fused(out, phase, width, height, scratch, st, mm, slots);
    const size_t P = (size_t)width * height;
    float2* z = reinterpret_cast<float2*>(scratch);
    float2* gx = z + P;
    float2* gy = gx + P;  // gx, gy adjacent: batched FFT2 of 2 frames
    const unsigned g1 = (unsigned)((P + 255) / 256);
    const dim3 g2((width + 255) / 256, height);
    k_unwrap_init<<<g1, 256, 0, st>>>(z, gx, phase, P);
    HV_CHECK_LAUNCH("k_unwrap_init");
    HV_TRY(launch_fft2(gx, 1, width, height, -1, st));
    k_unwrap_grad<<<g2, 256, 0, st>>>(gx, gy, width, height);
    HV_CHECK_LAUNCH("k_unwrap_grad");
    HV_TRY(launch_fft2(gx, 2, width, height, +1, st));
    k_unwrap_eq<<<g1, 256, 0, st>>>(gx, gy, z, P);
    HV_CHECK_LAUNCH("k_unwrap_eq");
    HV_TRY(launch_fft2(gx, 2, width, height, -1, st));
    k_unwrap_ratio<<<g2, 256, 0, st>>>(gx, gy, width, height);
    HV_CHECK_LAUNCH("k_unwrap_ratio");
    HV_TRY(launch_fft2(gx, 1, width, height, +1, st));
    k_unwrap_last<<<g1, 256, 0, st>>>(out, gx, P);
    HV_CHECK_LAUNCH("k_unwrap_last");
    if (mm) {
        HV_TRY(launch_minmax_init(mm, st, slots));
        HV_TRY(launch_minmax(mm, out, P, st, slots));
    }
    return HV_OK;
}

// ---------------------------------------------------------------------------
// percentiles without any allocation: the radix-select histograms in a
// file-scope device array, one per GPU, resolved once per device
// ---------------------------------------------------------------------------
constexpr size_t kPctStaticBytes = 16 * 1024 + 256;  // the array's fixed size
static_assert(sizeof(PctScratch) <= kPctStaticBytes, "percentile scratch outgrew its static array");

// In host code this name is a host-side stand-in, not a device address:
// only the per-device pointer from hipGetSymbolAddress is passed to kernels.
static __device__ __align__(256) unsigned char scratch[kPctStaticBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The current device's copy of `scratch`, looked up on its first use there.
static hv_status static_scratch(void** out) {
    int dev = 0;
    HV_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(HV_ERR_UNSUPPORTED, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(HV_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return HV_OK;
}

hv_status launch_percentile_static(float* range, const float* img, int width, int height, int step, float lo, float hi,
                                   int clamp, hipStream_t st) {
    if (step < 1) step = 1;
    // the histograms count in 32 bits
    const size_t n = (size_t)(width / step) * (size_t)(height / step);
    if (n > (size_t)UINT32_MAX)
        return fail(HV_ERR_INVALID_ARG, "hv_percentile_range_noalloc: more than 2^32 - 1 pixels to select from");
    void* ws = nullptr;
    HV_TRY(static_scratch(&ws));
    // k_pct_init clears the histograms at the start of every call
    return launch_percentile(range, img, width, height, step, lo, hi, clamp, ws, st);
}

}  // namespace hv
