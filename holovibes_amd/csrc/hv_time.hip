// hv_time.hip -- time transforms (STFT) fused with the image-type epilogue.
//
// The time window is the ring of T frames (T x P complex, frame-major, slots
// in physical order as the reference's Queue stores them, queue.cc:332-376).
// The STFT (stft.cu:9-15: a batched length-T FFT with stride P) runs per
// pixel: lanes take consecutive pixels, so every load of slot t is a
// contiguous run.  For T <= 64 one thread holds the pixel's whole window in
// registers; for T = 128, 256 a group of threads shares it through LDS.
// The epilogue turns the spectrum into the image without writing the T x P
// spectrum to HBM: modulus / squared modulus / argument mean over [pmin, pmax]
// (tools_conversion.cu:11-30, 84-103, 199-232) or the moments
// (tools_conversion.cu:33-66, tools_compute.cu:38-97), plus the image's
// min/max for the contrast stage.
#include "hv_fft.hpp"
#include "hv_kernels.hpp"

namespace hv {

enum { TM_IMAGE = 0, TM_FULL = 1, TM_MOMENTS = 2 };

#ifndef HV_TIME_DIRECT_BINS
#define HV_TIME_DIRECT_BINS 1   // STFT images of <= 4 bins by direct accumulation
#endif

struct TimeParams {
    float* out;        // image (TM_IMAGE) or M0 (TM_MOMENTS)
    float* m1;         // TM_MOMENTS
    float* m2;         // TM_MOMENTS
    float2* spec;      // TM_FULL
    int img_type;
    int pmin, pmax;    // inclusive range (image) or [f_start, f_end] (moments)
    float fps;         // moments frequency scale (CameraFps)
    uint32_t* mm;      // encoded min/max of the image, may be null
    int mm_moment;     // which moment feeds mm (0, 1, 2)
    int mm_slots;      // accumulation mode of mm (block_minmax_atomic; 0 counts as 1)
};

// fft_freqs (icompute.cc:26-69): f1[i] = i d for i < T/2, -(T - i) d above
// (even T), d = fps / T, in float as the reference computes it.
template <int T>
__device__ __forceinline__ float freq1(int i, float d) {
    if constexpr (T % 2 == 0) return i < T / 2 ? i * d : -((float)T - i) * d;
    else return i < (T + 1) / 2 ? i * d : (i - (float)T) * d;
}

// Moments of one pixel from its spectrum (tools_compute.cu:38-97 applied to
// |S_f|^2 from tools_conversion.cu:33-66).
template <int T, class Spec>
__device__ __forceinline__ void pixel_moments(Spec&& spec, int fs, int fe, float fps, float& a0, float& a1,
                                              float& a2) {
    const float d = fps / T;
    a0 = a1 = a2 = 0.f;
    for (int i = fs; i <= fe; ++i) {
        const float2 z = spec(i);
        const float s = z.x * z.x + z.y * z.y;
        const float f = freq1<T>(i, d);
        a0 += s * 1.0f;
        a1 += s * f;
        a2 += s * (f * f);
    }
    constexpr int nyq = T / 2;
    if (T % 2 == 0 && nyq >= fs && nyq <= fe) {
        const float2 z = spec(nyq);
        const float s = z.x * z.x + z.y * z.y;
        const float f = freq1<T>(nyq, d);
        a0 += s * 1.0f;
        a1 -= s * f;
        a2 += s * (f * f);
    }
}

// ---------------------------------------------------------------------------
// T <= 64: one thread per pixel, whole window in registers
// ---------------------------------------------------------------------------
template <int T, int MODE>
__global__ void __launch_bounds__(256) k_time_reg(const float2* __restrict__ ring, size_t P, TimeParams prm) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < P;
    float2 v[T];
#pragma unroll
    for (int t = 0; t < T; ++t) v[t] = valid ? ring[(size_t)t * P + p] : make_float2(0.f, 0.f);
    if constexpr (T > 1) fft_reg<T, -1>(v);
    float vmin = INFINITY, vmax = -INFINITY;
    if constexpr (MODE == TM_FULL) {
        if (valid) {
#pragma unroll
            for (int t = 0; t < T; ++t) prm.spec[(size_t)t * P + p] = v[t];
        }
    } else if constexpr (MODE == TM_IMAGE) {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < T; ++k)
            if (k >= prm.pmin && k <= prm.pmax) acc += image_term(v[k], prm.img_type);
        const float val = acc / (prm.pmax - prm.pmin + 1);
        if (valid) {
            prm.out[p] = val;
            vmin = vmax = val;
        }
    } else {
        // registers are indexed with compile-time k only: select from the unrolled loop
        auto spec = [&](int i) -> float2 {
            float2 r = make_float2(0.f, 0.f);
#pragma unroll
            for (int k = 0; k < T; ++k)
                if (k == i) r = v[k];
            return r;
        };
        float a0, a1, a2;
        pixel_moments<T>(spec, prm.pmin, prm.pmax, prm.fps, a0, a1, a2);
        if (valid) {
            if (prm.out) prm.out[p] = a0;
            if (prm.m1) prm.m1[p] = a1;
            if (prm.m2) prm.m2[p] = a2;
            const float sel = prm.mm_moment == 0 ? a0 : prm.mm_moment == 1 ? a1 : a2;
            vmin = vmax = sel;
        }
    }
    if (MODE != TM_FULL && prm.mm) block_minmax_atomic(vmin, vmax, prm.mm, prm.mm_slots != 0 ? prm.mm_slots : 1);
}

// ---------------------------------------------------------------------------
// T = 128, 256: CW pixels per workgroup, NT threads per pixel, LDS exchange
// ---------------------------------------------------------------------------
// T = 128, 256 moments: the bins' sums shared by the pixel's threads (1) or
// taken by one thread per pixel (0)
#ifndef HV_TIME_MOMENTS_SPLIT
#define HV_TIME_MOMENTS_SPLIT 1
#endif

template <int T>
struct TimeLdsCfg {
    static constexpr int E = 16;
    static constexpr int NT = FftShape<T, E>::NT;
    static constexpr int CW = 64;
    static constexpr int THREADS = NT * CW;
};

template <int T, int MODE>
__global__ void __launch_bounds__(TimeLdsCfg<T>::THREADS)
    k_time_lds(const float2* __restrict__ ring, size_t P, TimeParams prm) {
    using C = TimeLdsCfg<T>;
    __shared__ float2 lds[T * C::CW];
    const int c = threadIdx.x % C::CW, t = threadIdx.x / C::CW;
    const size_t p = (size_t)blockIdx.x * C::CW + c;
    const bool valid = p < P;
    auto lds_at = [&](int n) -> float2& { return lds[n * C::CW + c]; };
    auto load = [&](int n) -> float2 { return valid ? ring[(size_t)n * P + p] : make_float2(0.f, 0.f); };
    float vmin = INFINITY, vmax = -INFINITY;
    if constexpr (MODE == TM_FULL) {
        auto store = [&](int n, float2 z) {
            if (valid) prm.spec[(size_t)n * P + p] = z;
        };
        fft_run<T, C::E, -1>(t, load, store, lds_at);
    } else if constexpr (MODE == TM_MOMENTS && HV_TIME_MOMENTS_SPLIT) {
        // the pixel's NT threads each sum the bins fs + t, fs + t + NT, ...;
        // thread 0 adds the NT partial sums in order (and the Nyquist rule)
        __shared__ float red[3][C::NT][C::CW];
        auto store = [&](int n, float2 z) { lds_at(n) = z; };
        fft_run<T, C::E, -1, true>(t, load, store, lds_at);
        __syncthreads();
        const float d = prm.fps / T;
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        for (int i = prm.pmin + t; i <= prm.pmax; i += C::NT) {
            const float2 z = lds_at(i);
            const float sq = z.x * z.x + z.y * z.y;
            const float f = freq1<T>(i, d);
            a0 += sq * 1.0f;
            a1 += sq * f;
            a2 += sq * (f * f);
        }
        red[0][t][c] = a0;
        red[1][t][c] = a1;
        red[2][t][c] = a2;
        __syncthreads();
        if (t == 0 && valid) {
            a0 = a1 = a2 = 0.f;
#pragma unroll
            for (int k = 0; k < C::NT; ++k) {
                a0 += red[0][k][c];
                a1 += red[1][k][c];
                a2 += red[2][k][c];
            }
            constexpr int nyq = T / 2;
            if (nyq >= prm.pmin && nyq <= prm.pmax) {
                const float2 z = lds_at(nyq);
                const float sq = z.x * z.x + z.y * z.y;
                const float f = freq1<T>(nyq, d);
                a0 += sq * 1.0f;
                a1 -= sq * f;
                a2 += sq * (f * f);
            }
            if (prm.out) prm.out[p] = a0;
            if (prm.m1) prm.m1[p] = a1;
            if (prm.m2) prm.m2[p] = a2;
            const float sel = prm.mm_moment == 0 ? a0 : prm.mm_moment == 1 ? a1 : a2;
            vmin = vmax = sel;
        }
        if (prm.mm) block_minmax_atomic(vmin, vmax, prm.mm, prm.mm_slots != 0 ? prm.mm_slots : 1);
    } else {
        auto store = [&](int n, float2 z) { lds_at(n) = z; };
        fft_run<T, C::E, -1, true>(t, load, store, lds_at);
        __syncthreads();
        if (t == 0 && valid) {
            if constexpr (MODE == TM_IMAGE) {
                float acc = 0.f;
                for (int k = prm.pmin; k <= prm.pmax; ++k) acc += image_term(lds_at(k), prm.img_type);
                const float val = acc / (prm.pmax - prm.pmin + 1);
                prm.out[p] = val;
                vmin = vmax = val;
            } else {
                float a0, a1, a2;
                pixel_moments<T>([&](int i) { return lds_at(i); }, prm.pmin, prm.pmax, prm.fps, a0, a1, a2);
                if (prm.out) prm.out[p] = a0;
                if (prm.m1) prm.m1[p] = a1;
                if (prm.m2) prm.m2[p] = a2;
                const float sel = prm.mm_moment == 0 ? a0 : prm.mm_moment == 1 ? a1 : a2;
                vmin = vmax = sel;
            }
        }
        if (prm.mm) block_minmax_atomic(vmin, vmax, prm.mm, prm.mm_slots != 0 ? prm.mm_slots : 1);
    }
}

template <int T, int MODE>
static hv_status time_launch_t(const float2* ring, size_t P, const TimeParams& prm, hipStream_t st) {
    if constexpr (T <= 64) {
        const unsigned g = (unsigned)((P + 255) / 256);
        k_time_reg<T, MODE><<<g, 256, 0, st>>>(ring, P, prm);
    } else {
        using C = TimeLdsCfg<T>;
        const unsigned g = (unsigned)((P + C::CW - 1) / C::CW);
        k_time_lds<T, MODE><<<g, C::THREADS, 0, st>>>(ring, P, prm);
    }
    HV_CHECK_LAUNCH("k_time");
    return HV_OK;
}

// STFT image of a few bins (K = pmax - pmin + 1 <= kDirectBins): the bins
// accumulated directly as the window streams in, X_k = sum_t x_t w^{k t}
// (w = exp(-2 pi i / T), the ring's physical slot order as in the FFT
// kernels) -- K T complex products and 2 K accumulators per pixel instead of
// the whole window in registers and a length-T FFT, so the pass runs at the
// ring's read rate.
constexpr int kDirectBins = 4;

template <int T>
__global__ void __launch_bounds__(256) k_time_bins(const float2* __restrict__ ring, size_t P, TimeParams prm) {
    constexpr int LOGT = clog2(T);
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < P;
    const float2* xp = ring + (valid ? p : 0);
    const int K = prm.pmax - prm.pmin + 1;
    float2 acc[kDirectBins];
#pragma unroll
    for (int j = 0; j < kDirectBins; ++j) acc[j] = make_float2(0.f, 0.f);
#pragma unroll 8
    for (int t = 0; t < T; ++t) {
        const float2 x = xp[(size_t)t * P];
#pragma unroll
        for (int j = 0; j < kDirectBins; ++j) {
            if (j >= K) break;  // uniform
            const float2 w = twiddle<-1>(((prm.pmin + j) * t) & (T - 1), LOGT);
            acc[j].x = fmaf(x.x, w.x, fmaf(-x.y, w.y, acc[j].x));
            acc[j].y = fmaf(x.x, w.y, fmaf(x.y, w.x, acc[j].y));
        }
    }
    float a = 0.f;
#pragma unroll
    for (int j = 0; j < kDirectBins; ++j)
        if (j < K) a += image_term(acc[j], prm.img_type);
    const float val = a / K;
    float vmin = INFINITY, vmax = -INFINITY;
    if (valid) {
        prm.out[p] = val;
        vmin = vmax = val;
    }
    if (prm.mm) block_minmax_atomic(vmin, vmax, prm.mm, prm.mm_slots != 0 ? prm.mm_slots : 1);
}

template <int T>
static hv_status bins_launch_t(const float2* ring, size_t P, const TimeParams& prm, hipStream_t st) {
    k_time_bins<T><<<(unsigned)((P + 255) / 256), 256, 0, st>>>(ring, P, prm);
    HV_CHECK_LAUNCH("k_time_bins");
    return HV_OK;
}

// Windows whose length is not a power of two (2 <= T <= 256; cufftPlanMany
// takes any n, icompute.cc:109-115): the bins the output needs by direct DFT,
// X_k = sum_t x_t exp(-2 pi i k t / T) over the ring's physical slot order,
// eight bins at a time as the window streams in, twiddles from an LDS table
// of exp(-2 pi i m / T) built with sincospif.  All T bins for the full
// spectrum, [pmin, pmax] for an image, [f_start, f_end] for the moments.
__device__ __forceinline__ float freq1_rt(int i, int T, float d) {
    if (T % 2 == 0) return i < T / 2 ? i * d : -((float)T - i) * d;
    return i < (T + 1) / 2 ? i * d : (i - (float)T) * d;
}

constexpr int kAnyBins = 8;

template <int MODE>
__global__ void __launch_bounds__(256) k_time_any(const float2* __restrict__ ring, size_t P, int T, TimeParams prm) {
    __shared__ float2 tw[256];
    for (int i = threadIdx.x; i < T; i += blockDim.x) {
        float sn, cs;
        sincospif(2.0f * (float)i / (float)T, &sn, &cs);
        tw[i] = make_float2(cs, -sn);
    }
    __syncthreads();
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < P;
    const float2* xp = ring + (valid ? p : 0);
    const int k0 = MODE == TM_FULL ? 0 : prm.pmin, k1 = MODE == TM_FULL ? T - 1 : prm.pmax;
    const float d = prm.fps / T;
    float acc = 0.f, a0 = 0.f, a1 = 0.f, a2 = 0.f;
    float2 nyq = make_float2(0.f, 0.f);
    for (int kb = k0; kb <= k1; kb += kAnyBins) {
        const int nk = min(kAnyBins, k1 - kb + 1);
        float2 sb[kAnyBins];
        int idx[kAnyBins];  // (k t) mod T of each bin
#pragma unroll
        for (int j = 0; j < kAnyBins; ++j) {
            sb[j] = make_float2(0.f, 0.f);
            idx[j] = 0;
        }
        for (int t = 0; t < T; ++t) {
            const float2 x = xp[(size_t)t * P];
#pragma unroll
            for (int j = 0; j < kAnyBins; ++j) {
                if (j >= nk) break;  // uniform
                const float2 w = tw[idx[j]];
                sb[j].x = fmaf(x.x, w.x, fmaf(-x.y, w.y, sb[j].x));
                sb[j].y = fmaf(x.x, w.y, fmaf(x.y, w.x, sb[j].y));
                idx[j] += kb + j;
                if (idx[j] >= T) idx[j] -= T;
            }
        }
#pragma unroll
        for (int j = 0; j < kAnyBins; ++j) {
            if (j >= nk) break;
            const int k = kb + j;
            if constexpr (MODE == TM_FULL) {
                if (valid) prm.spec[(size_t)k * P + p] = sb[j];
            } else if constexpr (MODE == TM_IMAGE) {
                acc += image_term(sb[j], prm.img_type);
            } else {
                const float sq = sb[j].x * sb[j].x + sb[j].y * sb[j].y;
                const float f = freq1_rt(k, T, d);
                a0 += sq * 1.0f;
                a1 += sq * f;
                a2 += sq * (f * f);
                if (2 * k == T) nyq = sb[j];
            }
        }
    }
    float vmin = INFINITY, vmax = -INFINITY;
    if constexpr (MODE == TM_IMAGE) {
        const float val = acc / (prm.pmax - prm.pmin + 1);
        if (valid) {
            prm.out[p] = val;
            vmin = vmax = val;
        }
    } else if constexpr (MODE == TM_MOMENTS) {
        // the Nyquist bin of an even window counted again for M0, M2 and cancelled for M1 (tools_compute.cu:61-69)
        if (T % 2 == 0 && T / 2 >= k0 && T / 2 <= k1) {
            const float sq = nyq.x * nyq.x + nyq.y * nyq.y;
            const float f = freq1_rt(T / 2, T, d);
            a0 += sq * 1.0f;
            a1 -= sq * f;
            a2 += sq * (f * f);
        }
        if (valid) {
            if (prm.out) prm.out[p] = a0;
            if (prm.m1) prm.m1[p] = a1;
            if (prm.m2) prm.m2[p] = a2;
            const float sel = prm.mm_moment == 0 ? a0 : prm.mm_moment == 1 ? a1 : a2;
            vmin = vmax = sel;
        }
    }
    if (MODE != TM_FULL && prm.mm) block_minmax_atomic(vmin, vmax, prm.mm, prm.mm_slots != 0 ? prm.mm_slots : 1);
}

template <int MODE>
static hv_status time_launch(const float2* ring, int T, size_t P, const TimeParams& prm, hipStream_t st) {
    if (!is_pow2(T)) {
        if (T < 2 || T > 256) return fail(HV_ERR_UNSUPPORTED, "T must be in [1, 256]");
        k_time_any<MODE><<<(unsigned)((P + 255) / 256), 256, 0, st>>>(ring, P, T, prm);
        HV_CHECK_LAUNCH("k_time_any");
        return HV_OK;
    }
    if (MODE == TM_IMAGE && T >= 16 && prm.pmax - prm.pmin + 1 <= kDirectBins && HV_TIME_DIRECT_BINS) {
        switch (T) {
            case 16: return bins_launch_t<16>(ring, P, prm, st);
            case 32: return bins_launch_t<32>(ring, P, prm, st);
            case 64: return bins_launch_t<64>(ring, P, prm, st);
            case 128: return bins_launch_t<128>(ring, P, prm, st);
            default: return bins_launch_t<256>(ring, P, prm, st);
        }
    }
    switch (T) {
        case 1: return time_launch_t<1, MODE>(ring, P, prm, st);
        case 2: return time_launch_t<2, MODE>(ring, P, prm, st);
        case 4: return time_launch_t<4, MODE>(ring, P, prm, st);
        case 8: return time_launch_t<8, MODE>(ring, P, prm, st);
        case 16: return time_launch_t<16, MODE>(ring, P, prm, st);
        case 32: return time_launch_t<32, MODE>(ring, P, prm, st);
        case 64: return time_launch_t<64, MODE>(ring, P, prm, st);
        case 128: return time_launch_t<128, MODE>(ring, P, prm, st);
        case 256: return time_launch_t<256, MODE>(ring, P, prm, st);
        default: return fail(HV_ERR_UNSUPPORTED, "T must be in [1, 256]");
    }
}

hv_status launch_stft(float2* out, const float2* in, int T, size_t P, hipStream_t st) {
    TimeParams prm{};
    prm.spec = out;
    return time_launch<TM_FULL>(in, T, P, prm, st);
}

hv_status launch_stft_image(float* out, const float2* ring, int T, size_t P, int img_type, int pmin, int pmax,
                            uint32_t* mm, hipStream_t st, int mm_slots) {
    TimeParams prm{};
    prm.out = out;
    prm.img_type = img_type;
    prm.pmin = pmin;
    prm.pmax = pmax;
    prm.mm = mm;
    prm.mm_slots = mm_slots;
    return time_launch<TM_IMAGE>(ring, T, P, prm, st);
}

hv_status launch_stft_moments(float* m0, float* m1, float* m2, const float2* ring, int T, size_t P, int f_start,
                              int f_end, float fps, uint32_t* mm, int mm_moment, hipStream_t st, int mm_slots) {
    TimeParams prm{};
    prm.out = m0;
    prm.m1 = m1;
    prm.m2 = m2;
    prm.pmin = f_start;
    prm.pmax = f_end;
    prm.fps = fps;
    prm.mm = mm;
    prm.mm_slots = mm_slots;
    prm.mm_moment = mm_moment;
    return time_launch<TM_MOMENTS>(ring, T, P, prm, st);
}

// ---------------------------------------------------------------------------
// B = S = T: inverse row FFT of every frame of one image row, then the STFT
// of that row's pixels, in one workgroup.  The window is replaced whole by
// every batch, so its frames never need to reach HBM in the spatial domain:
// this kernel reads the column pass's output rows and writes one image row,
// instead of writing the T frames back and reading them again.
// ---------------------------------------------------------------------------
#ifndef HV_FUSED_E_SMALL
#define HV_FUSED_E_SMALL 8      // values per thread of the row transforms, W <= 512
#endif

template <int W, int T>
struct RowStftCfg {
    static constexpr int E = W >= 1024 ? 16 : HV_FUSED_E_SMALL;
    static constexpr int NT = FftShape<W, E>::NT;                  // threads per row transform
    static constexpr int THREADS = 512;                             // 256 VGPRs for the T-point FFT
    static constexpr int G = THREADS / NT < T ? THREADS / NT : T;   // rows transformed at once
    static constexpr bool PAD = (size_t)T * (W + W / 8) * 8 <= 160 * 1024;
    static constexpr int PITCH = PAD ? W + W / 8 : W;
    // threads beyond G row transforms work on a spare LDS row
    static constexpr int ROWS = G * NT == THREADS ? T : T + 1;
    static constexpr bool FITS = (size_t)ROWS * PITCH * 8 <= 160 * 1024 && T <= 32 && THREADS % NT == 0 &&
                                 T % G == 0;
};

template <int W, int T>
__global__ void __launch_bounds__(512)
    k_row_ifft_stft(const float2* __restrict__ ring, size_t P, TimeParams prm) {
    using C = RowStftCfg<W, T>;
    constexpr int kE = FftShape<W, C::E>::kE;
    constexpr int ROUNDS = T / C::G;
    __shared__ float2 tile[C::ROWS * C::PITCH];
    const int g = threadIdx.x / C::NT, t = threadIdx.x % C::NT;
    const bool active = g < C::G;
    auto pad = [](int n) { return C::PAD ? n + (n >> 3) : n; };
    const int y = blockIdx.x;
    // all ROUNDS x kE inputs of the row are loaded before any is transformed
    float2 v[ROUNDS][kE];
#pragma unroll
    for (int rd = 0; rd < ROUNDS; ++rd) {
        const float2* src = ring + (size_t)(active ? rd * C::G + g : 0) * P + (size_t)y * W;
        fft_load<W, C::E>(t, [&](int n) -> float2 { return active ? src[n] : make_float2(0.f, 0.f); }, v[rd]);
    }
    // 1. inverse row FFTs, G frames at a time, results in the tile
    float2 tw[FftTw<W, C::E>::COUNT];
    fft_twiddles<W, C::E>(t, tw);
#pragma unroll
    for (int rd = 0; rd < ROUNDS; ++rd) {
        float2* row = tile + (active ? rd * C::G + g : T) * C::PITCH;
        auto lds_at = [&](int n) -> float2& { return row[pad(n)]; };
        auto store = [&](int n, float2 x) { row[pad(n)] = x; };
        fft_compute_tw<W, C::E, +1, true, (C::NT <= 64)>(t, v[rd], tw, store, lds_at);
        if constexpr (C::NT > 64) __syncthreads();
    }
    __syncthreads();
    // 2. STFT of each pixel of the row across the T frames + image epilogue
    float vmin = INFINITY, vmax = -INFINITY;
#pragma unroll 1
    for (int x = threadIdx.x; x < W; x += blockDim.x) {
        float2 s[T];
#pragma unroll
        for (int f = 0; f < T; ++f) s[f] = tile[f * C::PITCH + pad(x)];
        fft_reg<T, -1>(s);
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < T; ++k)
            if (k >= prm.pmin && k <= prm.pmax) acc += image_term(s[k], prm.img_type);
        const float val = acc / (prm.pmax - prm.pmin + 1);
        prm.out[(size_t)y * W + x] = val;
        vmin = fminf(vmin, val);
        vmax = fmaxf(vmax, val);
    }
    if (prm.mm) block_minmax_atomic(vmin, vmax, prm.mm, prm.mm_slots != 0 ? prm.mm_slots : 1);
}


template <int W, int T>
static hv_status row_stft_wt(float* out, float2* ring, int height, size_t P, const TimeParams& prm, hipStream_t st) {
    if constexpr (RowStftCfg<W, T>::FITS) {
        k_row_ifft_stft<W, T><<<height, RowStftCfg<W, T>::THREADS, 0, st>>>(ring, P, prm);
        HV_CHECK_LAUNCH("k_row_ifft_stft");
        return HV_OK;
    } else {
        return fail(HV_ERR_UNSUPPORTED, "no fused row-IFFT + STFT kernel for this width and T");
    }
}

template <int W>
static hv_status row_stft_w(float* out, float2* ring, int height, int T, size_t P, const TimeParams& prm,
                            hipStream_t st, bool query) {
    switch (T) {
        case 8: return query ? (RowStftCfg<W, 8>::FITS ? HV_OK : HV_ERR_UNSUPPORTED) : row_stft_wt<W, 8>(out, ring, height, P, prm, st);
        case 16: return query ? (RowStftCfg<W, 16>::FITS ? HV_OK : HV_ERR_UNSUPPORTED) : row_stft_wt<W, 16>(out, ring, height, P, prm, st);
        case 32: return query ? (RowStftCfg<W, 32>::FITS ? HV_OK : HV_ERR_UNSUPPORTED) : row_stft_wt<W, 32>(out, ring, height, P, prm, st);
        case 64: return query ? (RowStftCfg<W, 64>::FITS ? HV_OK : HV_ERR_UNSUPPORTED) : row_stft_wt<W, 64>(out, ring, height, P, prm, st);
        default: return HV_ERR_UNSUPPORTED;
    }
}

static hv_status row_stft_dispatch(float* out, float2* ring, int width, int height, int T, const TimeParams& prm,
                                   hipStream_t st, bool query) {
    const size_t P = (size_t)width * height;
    switch (width) {
        case 64: return row_stft_w<64>(out, ring, height, T, P, prm, st, query);
        case 128: return row_stft_w<128>(out, ring, height, T, P, prm, st, query);
        case 256: return row_stft_w<256>(out, ring, height, T, P, prm, st, query);
        case 512: return row_stft_w<512>(out, ring, height, T, P, prm, st, query);
        case 1024: return row_stft_w<1024>(out, ring, height, T, P, prm, st, query);
        default: return HV_ERR_UNSUPPORTED;
    }
}

bool row_ifft_stft_supported(int width, int T) {
    TimeParams prm{};
    return row_stft_dispatch(nullptr, nullptr, width, 1, T, prm, nullptr, true) == HV_OK;
}

hv_status launch_row_ifft_stft_image(float* out, float2* ring, int width, int height, int T, int img_type, int pmin,
                                     int pmax, uint32_t* mm, hipStream_t st, int mm_slots) {
    TimeParams prm{};
    prm.out = out;
    prm.img_type = img_type;
    prm.pmin = pmin;
    prm.pmax = pmax;
    prm.mm = mm;
    prm.mm_slots = mm_slots;
    return row_stft_dispatch(out, ring, width, height, T, prm, st, false);
}

// complex_to_modulus / squared_modulus / argument on a stored spectrum
// (tools_conversion.cu:11-30, 84-103, 199-232).
__global__ void __launch_bounds__(256)
    k_complex_to_image(float* __restrict__ out, const float2* __restrict__ in, int img_type, int pmin, int pmax,
                       size_t P) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    float acc = 0.f;
    for (int k = pmin; k <= pmax; ++k) acc += image_term(in[(size_t)k * P + p], img_type);
    out[p] = acc / (pmax - pmin + 1);
}

hv_status launch_complex_to_image(float* out, const float2* in, int img_type, int pmin, int pmax, size_t P,
                                  hipStream_t st) {
    k_complex_to_image<<<(unsigned)((P + 255) / 256), 256, 0, st>>>(out, in, img_type, pmin, pmax, P);
    HV_CHECK_LAUNCH("k_complex_to_image");
    return HV_OK;
}

}  // namespace hv
