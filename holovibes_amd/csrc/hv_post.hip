// hv_post.hip -- optional post-processing of the float image, between the
// image kernel and the contrast (Backend/sources/core/pipe.cc:353-370):
// renormalisation, image accumulation, log10; and the XZ / YZ cuts of the
// time-transformed window.  Each is one HBM pass over the image (float4 per
// thread); accumulation also reads the level - 1 previous images.  The
// accumulation ring's position lives in device memory, so a captured cycle
// replays correctly.
#include "hv_common.hpp"
#include "hv_kernels.hpp"

namespace {

constexpr int kRenormParts = 1024;

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
    return t;  // valid in thread 0
}

// per-block double sums of the image (reduce_add, Backend/cuda_includes/reduce.cuh:207-221)
__global__ void __launch_bounds__(256) k_renorm_partials(const float4* __restrict__ x, size_t n4,
                                                         double* __restrict__ parts) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) {
        float4 v = x[i];
        acc += (double)v.x + (double)v.y + (double)v.z + (double)v.w;
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = t;
}

// x = x * 2^c * (frame_res / sum), evaluated left to right in float as the
// reference does (tools_compute.cu:99-123); every block adds the partials in
// the same order, so all get the same total
__global__ void __launch_bounds__(256) k_renorm_apply(float4* __restrict__ x, size_t n4, const double* __restrict__ parts,
                                                      int nparts, float frame_res, float multiplier) {
    __shared__ double sh[4];
    __shared__ float ratio;
    double acc = 0.0;
    for (int k = threadIdx.x; k < nparts; k += 256) acc += parts[k];
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) ratio = frame_res / (float)t;
    __syncthreads();
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n4) return;
    const float r = ratio;
    float4 v = x[i];
    v.x = (v.x * multiplier) * r;
    v.y = (v.y * multiplier) * r;
    v.z = (v.z * multiplier) * r;
    v.w = (v.w * multiplier) * r;
    x[i] = v;
}

// The image joins the accumulation ring (Queue::enqueue, queue.cc:149-189)
// and becomes the mean of the ring's images, summed from the oldest in float
// (kernel_accumulate_images, tools_conversion.cu:418-441).  state = {next
// slot, images held} before this image.
__global__ void __launch_bounds__(256) k_accumulate(float4* __restrict__ img, float4* __restrict__ ring,
                                                    const uint32_t* __restrict__ state, int level, size_t n4) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n4) return;
    const int slot = (int)state[0];
    const int held = min((int)state[1] + 1, level);
    const float4 cur = img[i];
    ring[(size_t)slot * n4 + i] = cur;
    int pos = (slot + 1 - held + level) % level;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < held; ++k) {
        const float4 v = pos == slot ? cur : ring[(size_t)pos * n4 + i];
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
        pos = pos + 1 == level ? 0 : pos + 1;
    }
    const float d = (float)held;
    img[i] = make_float4(s.x / d, s.y / d, s.z / d, s.w / d);
}

__global__ void k_accumulate_advance(uint32_t* state, int level) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        state[0] = (state[0] + 1) % (uint32_t)level;
        state[1] = min(state[1] + 1, (uint32_t)level);
    }
}

__global__ void __launch_bounds__(256) k_log10(float4* __restrict__ x, size_t n4) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n4) return;
    float4 v = x[i];
    x[i] = make_float4(log10f(v.x), log10f(v.y), log10f(v.z), log10f(v.w));
}

inline unsigned grid_for(size_t n4) { return (unsigned)((n4 + 255) / 256); }

}  // namespace

namespace hv {

size_t renorm_scratch_bytes() { return kRenormParts * sizeof(double) + 256; }

hv_status launch_renormalize(float* img, size_t P, unsigned norm_constant, void* scratch, hipStream_t st) {
    if (P % 4) return fail(HV_ERR_UNSUPPORTED, "renormalize: frame_res must be a multiple of 4");
    const size_t n4 = P / 4;
    double* parts = (double*)scratch;
    const int nparts = (int)std::min<size_t>(kRenormParts, std::max<size_t>(1, (n4 + 2047) / 2048));
    hipLaunchKernelGGL(k_renorm_partials, dim3(nparts), dim3(256), 0, st, (const float4*)img, n4, parts);
    hipLaunchKernelGGL(k_renorm_apply, dim3(grid_for(n4)), dim3(256), 0, st, (float4*)img, n4, parts, nparts, (float)P,
                       (float)(1u << norm_constant));
    HV_CHECK_LAUNCH("renormalize");
    return HV_OK;
}

hv_status launch_accumulate(float* img, float* ring, uint32_t* state, int level, size_t P, hipStream_t st) {
    if (P % 4) return fail(HV_ERR_UNSUPPORTED, "accumulation: frame_res must be a multiple of 4");
    const size_t n4 = P / 4;
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for(n4)), dim3(256), 0, st, (float4*)img, (float4*)ring, state, level,
                       n4);
    hipLaunchKernelGGL(k_accumulate_advance, dim3(1), dim3(64), 0, st, state, level);
    HV_CHECK_LAUNCH("accumulate");
    return HV_OK;
}

hv_status launch_log10(float* img, size_t P, hipStream_t st) {
    if (P % 4) return fail(HV_ERR_UNSUPPORTED, "log10: frame_res must be a multiple of 4");
    hipLaunchKernelGGL(k_log10, dim3(grid_for(P / 4)), dim3(256), 0, st, (float4*)img, P / 4);
    HV_CHECK_LAUNCH("log10");
    return HV_OK;
}

}  // namespace hv

// ---------------------------------------------------------------------------
// XZ / YZ cuts (fill_32bit_slices, Backend/cuda_sources/stft.cu:17-115): the
// time-transformed window seen from the side.  XZ (T rows x W) holds, per
// frequency bin and column, the mean over rows [ymin, ymax]; YZ (H rows x T)
// per row and bin the mean over columns [xmin, xmax].  Values as the
// reference computes them: hypotf, its square, or atanf(im / re) + pi/2
// (atan, not atan2), summed in float in index order.
// ---------------------------------------------------------------------------
namespace {

// other image types (the moments) leave the value at 0, as in the reference
__device__ __forceinline__ float cut_value(float2 z, int img_type) {
    if (img_type == HV_IMG_MODULUS) return hypotf(z.x, z.y);
    if (img_type == HV_IMG_SQUARED_MODULUS) {
        const float m = hypotf(z.x, z.y);
        return m * m;
    }
    if (img_type == HV_IMG_ARGUMENT) return (float)((double)atanf(z.y / z.x) + 1.5707963267948966);
    return 0.f;
}

// one thread per (bin, x): rows ymin..ymax of the bin's plane, coalesced over x
__global__ void __launch_bounds__(256) k_cut_xz(float* __restrict__ xz, const float2* __restrict__ stft, int W,
                                                size_t P, int T, int ymin, int ymax, int img_type) {
    const int x = blockIdx.x * 256 + threadIdx.x, bin = blockIdx.y;
    if (x >= W || bin >= T) return;
    const float2* plane = stft + (size_t)bin * P;
    float sum = 0.f;
    for (int y = ymin; y <= ymax; ++y) sum += cut_value(plane[(size_t)y * W + x], img_type);
    xz[(size_t)bin * W + x] = sum / (float)(ymax - ymin + 1);
}

// one thread per (y, bin): columns xmin..xmax of row y in the bin's plane
__global__ void __launch_bounds__(256) k_cut_yz(float* __restrict__ yz, const float2* __restrict__ stft, int W, int H,
                                                size_t P, int T, int xmin, int xmax, int img_type) {
    const size_t id = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (id >= (size_t)H * T) return;
    const int y = (int)(id / T), bin = (int)(id % T);
    const float2* row = stft + (size_t)bin * P + (size_t)y * W;
    float sum = 0.f;
    for (int x = xmin; x <= xmax; ++x) sum += cut_value(row[x], img_type);
    yz[id] = sum / (float)(xmax - xmin + 1);
}

}  // namespace

namespace hv {

hv_status launch_cuts(float* xz, float* yz, const float2* stft, int W, int H, int T, int xmin, int ymin, int xmax,
                      int ymax, int img_type, hipStream_t st) {
    const size_t P = (size_t)W * H;
    hipLaunchKernelGGL(k_cut_xz, dim3((W + 255) / 256, T), dim3(256), 0, st, xz, stft, W, P, T, ymin, ymax, img_type);
    hipLaunchKernelGGL(k_cut_yz, dim3((unsigned)(((size_t)H * T + 255) / 256)), dim3(256), 0, st, yz, stft, W, H, P, T,
                       xmin, xmax, img_type);
    HV_CHECK_LAUNCH("cuts");
    return HV_OK;
}

}  // namespace hv

// ---------------------------------------------------------------------------
// Registration (Backend/sources/compute/registration.cc:9-122): the image is
// shifted circularly so that its cross-correlation with a reference image
// peaks at the origin.  Both images are prepared the same way: the mean over
// a centred disc is removed inside the disc and everything outside is set to
// 0.  The correlation is IFFT2(FFT2(current) * conj(FFT2(reference))) (real
// part; the reference's spectrum is kept from the cycle that set it), its
// peak is the first index of the largest |value| (cublasIsamax), and the
// shift is computed and applied on the device (the reference reads the peak
// back to the host).
// ---------------------------------------------------------------------------
namespace {

constexpr int kRegParts = 1024;

// get_circular_mask (masks.cu:61-77) with the quadrant swap of registration.cc:100-110
__global__ void k_reg_mask(float* mask, int W, int H, float cx, float cy, float radius, int swap) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const float d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
    const float v = d2 <= radius * radius ? 1.f : 0.f;
    int ox = x, oy = y;
    if (swap) {
        ox = (x + W / 2) % W;
        oy = (y + H / 2) % H;
    }
    mask[(size_t)oy * W + ox] = v;
}

// per-block sums of input * mask and of mask (get_mean_in_mask, apply_mask.cu:45-202), in double
__global__ void __launch_bounds__(256) k_reg_mean_parts(const float* __restrict__ in, const float* __restrict__ mask,
                                                        size_t n, double* __restrict__ parts) {
    __shared__ double sh[4];
    double s = 0.0, c = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        s += (double)(in[i] * mask[i]);
        c += (double)mask[i];
    }
    const double ts = block_sum(s, sh);
    __syncthreads();
    const double tc = block_sum(c, sh);
    if (threadIdx.x == 0) {
        parts[2 * blockIdx.x] = ts;
        parts[2 * blockIdx.x + 1] = tc;
    }
}

// out = (in - mean) * mask (rescale_in_mask + apply_mask), as the real part
// of a complex plane; the mean inside the mask is sum / count in float, 0 for
// an empty mask (apply_mask.cu:198-202), every block adding the partials in
// the same order
__global__ void __launch_bounds__(256) k_reg_prepare(float2* __restrict__ out, const float* __restrict__ in,
                                                     const float* __restrict__ mask, size_t n,
                                                     const double* __restrict__ parts, int nparts) {
    __shared__ double sh[4];
    __shared__ float mean_s;
    double s = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < nparts; k += 256) {
        s += parts[2 * k];
        c += parts[2 * k + 1];
    }
    const double ts = block_sum(s, sh);
    __syncthreads();
    const double tc = block_sum(c, sh);
    if (threadIdx.x == 0) mean_s = tc > 0.0 ? (float)ts / (float)tc : 0.f;
    __syncthreads();
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[i] = make_float2((in[i] - mean_s) * mask[i], 0.f);
}

// cur *= conj(ref)
__global__ void __launch_bounds__(256) k_reg_cross(float2* __restrict__ cur, const float2* __restrict__ ref, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 a = cur[i], b = ref[i];
    cur[i] = make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// (|value|, index) maxima per block, first index on ties
__global__ void __launch_bounds__(256) k_reg_peak_parts(const float2* __restrict__ xc, size_t n,
                                                        float* __restrict__ pv, uint32_t* __restrict__ pi) {
    __shared__ float sv[256];
    __shared__ uint32_t si[256];
    float best = -1.f;
    uint32_t bi = 0xFFFFFFFFu;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float v = fabsf(xc[i].x);
        if (v > best) {  // increasing i per thread: the first maximum stays
            best = v;
            bi = (uint32_t)i;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const float v2 = sv[threadIdx.x + o];
            const uint32_t i2 = si[threadIdx.x + o];
            if (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x])) {
                sv[threadIdx.x] = v2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pv[blockIdx.x] = sv[0];
        pi[blockIdx.x] = si[0];
    }
}

// the peak's (x, y), wrapped to the signed shift and negated (registration.cc:46-55);
// one block reduces the per-block maxima, first index on ties
__global__ void __launch_bounds__(256) k_reg_peak_final(const float* __restrict__ pv, const uint32_t* __restrict__ pi,
                                                        int nparts, int W, int H, int* __restrict__ shift) {
    __shared__ float sv[256];
    __shared__ uint32_t si[256];
    float best = -1.f;
    uint32_t bi = 0xFFFFFFFFu;
    for (int k = threadIdx.x; k < nparts; k += 256)
        if (pv[k] > best || (pv[k] == best && pi[k] < bi)) {
            best = pv[k];
            bi = pi[k];
        }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const float v2 = sv[threadIdx.x + o];
            const uint32_t i2 = si[threadIdx.x + o];
            if (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x])) {
                sv[threadIdx.x] = v2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t b = si[0] == 0xFFFFFFFFu ? 0u : si[0];
        int x = (int)(b % (uint32_t)W), y = (int)(b / (uint32_t)W);
        if (x > W / 2) x -= W;
        if (y > H / 2) y -= H;
        shift[0] = -x;
        shift[1] = -y;
    }
}

// circ_shift (tools.cu:40-79): out[(y + sy) mod H][(x + sx) mod W] = in[y][x]
__global__ void __launch_bounds__(256) k_reg_shift(float* __restrict__ out, const float* __restrict__ in, int W, int H,
                                                   const int* __restrict__ shift) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    int nx = (x + shift[0]) % W, ny = (y + shift[1]) % H;
    nx += W * (nx < 0);
    ny += H * (ny < 0);
    out[(size_t)ny * W + nx] = in[(size_t)y * W + x];
}

}  // namespace

namespace hv {

// mask (P floats) + reference spectrum and work plane (P complex each) + partials + shift
size_t registration_scratch_bytes(size_t P) {
    return P * sizeof(float) + 2 * P * sizeof(float2) + 2 * kRegParts * sizeof(double) + kRegParts * 8 + 256;
}

struct RegScratch {
    float* mask;
    float2* ref;
    float2* work;
    double* parts;
    float* pv;
    uint32_t* pi;
    int* shift;
};

static RegScratch reg_layout(void* scratch, size_t P) {
    char* c = (char*)scratch;
    RegScratch r;
    r.mask = (float*)c;
    c += P * sizeof(float);
    r.ref = (float2*)c;
    c += P * sizeof(float2);
    r.work = (float2*)c;
    c += P * sizeof(float2);
    r.parts = (double*)c;
    c += 2 * kRegParts * sizeof(double);
    r.pv = (float*)c;
    c += kRegParts * sizeof(float);
    r.pi = (uint32_t*)c;
    c += kRegParts * sizeof(uint32_t);
    r.shift = (int*)c;
    return r;
}

hv_status launch_registration_mask(void* scratch, int W, int H, float zone, int swap, hipStream_t st) {
    RegScratch r = reg_layout(scratch, (size_t)W * H);
    const float radius = (float)(std::min(W, H) / 2) * zone;
    hipLaunchKernelGGL(k_reg_mask, dim3((W + 255) / 256, H), dim3(256), 0, st, r.mask, W, H, W / 2.0f, H / 2.0f, radius,
                       swap);
    HV_CHECK_LAUNCH("registration mask");
    return HV_OK;
}

static hv_status reg_prepare(const RegScratch& r, float2* out, const float* img, size_t P, hipStream_t st) {
    const int nparts = (int)std::min<size_t>(kRegParts, std::max<size_t>(1, (P + 8191) / 8192));
    hipLaunchKernelGGL(k_reg_mean_parts, dim3(nparts), dim3(256), 0, st, img, r.mask, P, r.parts);
    hipLaunchKernelGGL(k_reg_prepare, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, out, img, r.mask, P, r.parts,
                       nparts);
    HV_CHECK_LAUNCH("registration prepare");
    return HV_OK;
}

// the reference image: prepared, its spectrum kept
hv_status launch_registration_set_reference(void* scratch, const float* img, int W, int H, hipStream_t st) {
    const size_t P = (size_t)W * H;
    RegScratch r = reg_layout(scratch, P);
    HV_TRY(reg_prepare(r, r.ref, img, P, st));
    return launch_fft2(r.ref, 1, W, H, -1, st);
}

// shifts img (in place, through tmp) onto the reference
hv_status launch_registration(void* scratch, float* img, float* tmp, int W, int H, hipStream_t st) {
    const size_t P = (size_t)W * H;
    RegScratch r = reg_layout(scratch, P);
    HV_TRY(reg_prepare(r, r.work, img, P, st));
    HV_TRY(launch_fft2(r.work, 1, W, H, -1, st));
    hipLaunchKernelGGL(k_reg_cross, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, r.work, r.ref, P);
    HV_TRY(launch_fft2(r.work, 1, W, H, 1, st));
    const int nparts = (int)std::min<size_t>(kRegParts, std::max<size_t>(1, (P + 4095) / 4096));
    hipLaunchKernelGGL(k_reg_peak_parts, dim3(nparts), dim3(256), 0, st, r.work, P, r.pv, r.pi);
    hipLaunchKernelGGL(k_reg_peak_final, dim3(1), dim3(256), 0, st, r.pv, r.pi, nparts, W, H, r.shift);
    hipLaunchKernelGGL(k_reg_shift, dim3((W + 255) / 256, H), dim3(256), 0, st, tmp, img, W, H, r.shift);
    HV_CHECK_LAUNCH("registration");
    HV_HIP(hipMemcpyAsync(img, tmp, P * sizeof(float), hipMemcpyDeviceToDevice, st));
    return HV_OK;
}

hv_status registration_shift(void* scratch, size_t P, int* shift_host) {
    RegScratch r = reg_layout(scratch, P);
    HV_HIP(hipMemcpy(shift_host, r.shift, 2 * sizeof(int), hipMemcpyDeviceToHost));
    return HV_OK;
}

}  // namespace hv
