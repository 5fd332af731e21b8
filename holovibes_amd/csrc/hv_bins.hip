// hv_bins.hip -- the window-replacing STFT path computed bins-first.
//
// When B = S = T every batch replaces the whole time window
// (pipe.cc:400-427, queue.cc:332-376), and the image reads only the STFT bins
// k in [pmin, pmax] (converts.cc:76-92, tools_conversion.cu:11-30, 84-103,
// 199-232).  The propagation L (angular spectrum, angular_spectrum.cu:33-63,
// or Fresnel, fresnel_transform.cu:65-83) is one linear operator applied to
// every frame, and the STFT (stft.cu:9-15) is a linear combination of frames:
//     S_k = sum_t w^{k t} L(x_t) = L( sum_t w^{k t} x_t ),  w = exp(-2 pi i / T).
// So this path takes the K = pmax - pmin + 1 needed bins of the converted
// frames first and propagates K planes instead of T frames:
//   1. k_bins_rows: per image row, the length-T DFT of every pixel of the T
//      raw frames (physical ring-slot order), bins [pmin, pmax] kept in LDS,
//      x chirp (Fresnel), then their forward row FFTs -> plane rows in HBM;
//   2. the column pass over the K planes (x lens, inverse column FFT for the
//      angular spectrum; forward only for Fresnel);
//   3. k_bins_row_ifft_image (angular spectrum): inverse row FFT of each
//      plane row + image term, summed over the planes in registers and LDS;
//      k_bins_image (Fresnel): the image term summed over the planes.
// The image equals the reference's up to FP32 rounding; bytes moved per batch
// drop from about 3 x 16 B x T P (three FFT passes over the window) to
// 2 B T P + 3 x 16 B K P.
#include <algorithm>

#include "hv_fft.hpp"
#include "hv_kernels.hpp"

namespace hv {

namespace {

// image rows per workgroup of the two row kernels (pipelined: the next
// row's loads overlap this row's transforms)
#ifndef HV_BINS_RPW_ROWS
#define HV_BINS_RPW_ROWS 1
#endif
#ifndef HV_BINS_IMAGE_EPI_K
#define HV_BINS_IMAGE_EPI_K 4  // up to this many planes: the image through the epilogue kernel
#endif
#ifndef HV_BINS_RPW_IMAGE
#define HV_BINS_RPW_IMAGE 1
#endif

#ifndef HV_BINS_IMG_THREADS
#define HV_BINS_IMG_THREADS 512
#endif
#ifndef HV_BINS_IMG_MT
#define HV_BINS_IMG_MT 1        // plane rows each wavefront transforms side by side in the image kernel
#endif
#ifndef HV_BINS_IMG_DEPTH
#define HV_BINS_IMG_DEPTH 1     // plane rows loaded ahead by the image kernel
#endif

constexpr int kBinThreads = 512;  // time DFT + row FFT kernel

#ifndef HV_BINS_ROW_E
#define HV_BINS_ROW_E 8         // values per thread of the plane-row FFTs, rows up to 512
#endif

// inverse row FFT + image kernel: its exchange rows and partial sums stay
// within LDS
template <int W>
constexpr int img_threads() {
    return W >= 1024 ? 512 : HV_BINS_IMG_THREADS;
}

template <int W, int THREADS = kBinThreads>
struct BinRowCfg {
    static constexpr int E = W >= 1024 ? 16 : HV_BINS_ROW_E;
    static constexpr int NT = FftShape<W, E>::NT;          // threads per row transform
    static constexpr int G = THREADS / NT;                  // row transforms at once
    static constexpr int PITCH = W + W / 8;                  // padded LDS row
    static constexpr int PPT = (W + THREADS - 1) / THREADS;  // pixels per thread
    static __device__ __forceinline__ int pad(int n) { return n + (n >> 3); }
};

// LDS bytes of k_bins_rows: K plane rows + one spare row for idle groups
// (+ the T x K8 projection weights for PCA, T > 0).
template <int W>
constexpr size_t bins_rows_lds(int K, int pca_T = 0) {
    return (size_t)(K + 1) * BinRowCfg<W>::PITCH * sizeof(float2) +
           (size_t)pca_T * (pca_T < 8 ? 8 : pca_T) * sizeof(float2);
}

// The next row's window is prefetched into registers while the current
// row's plane rows are transformed, when it fits beside the DFT registers.
template <int W, int T>
constexpr bool bins_rows_prefetch() {
    return HV_BINS_RPW_ROWS > 1 && T * BinRowCfg<W>::PPT <= 32;
}

// Bins k in [pmin, pmin + K) of the length-T DFT of a real sequence x, from
// one complex FFT of half the length: z_n = x_2n + i x_2n+1, Z = FFT_M(z)
// (M = T/2), then X_k = E_k + w^k O_k with E_k = (Z_k + conj Z_{M-k}) / 2,
// O_k = (Z_k - conj Z_{M-k}) / 2i, w = exp(-2 pi i / T) (k <= M), and
// X_k = conj X_{T-k} above M.  emit(k, X_k) for each wanted bin.
template <int T, class Emit>
__device__ __forceinline__ void real_dft_bins(const float* x, int pmin, int K, Emit&& emit) {
    constexpr int M = T / 2;
    float2 z[M];
#pragma unroll
    for (int n = 0; n < M; ++n) z[n] = make_float2(x[2 * n], x[2 * n + 1]);
    if constexpr (M > 1) fft_reg<M, -1>(z);
#pragma unroll
    for (int k = 0; k < T; ++k) {
        if (k < pmin || k >= pmin + K) continue;  // uniform: the bin range is a kernel argument
        const int kk = k <= M ? k : T - k;
        const float2 a = z[kk % M], bc = z[(M - kk) % M];
        const float2 b = make_float2(bc.x, -bc.y);
        const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
        const float2 d = make_float2(0.5f * (a.x - b.x), 0.5f * (a.y - b.y));
        const float2 o = make_float2(d.y, -d.x);  // d / i
        const float2 w = twiddle<-1>(kk, clog2(T));
        float2 v = make_float2(e.x + (w.x * o.x - w.y * o.y), e.y + (w.x * o.y + w.y * o.x));
        if (k > M) v.y = -v.y;
        emit(k, v);
    }
}

// PCA: plane k = sum_t wts[t T + k] x_t over the window in physical slot
// order (the projection C = H V onto eigenvectors pmin .. pmin + K - 1 of the
// ring covariance, wts = their row-major float copy from column pmin),
// instead of DFT bin pmin + k.
template <int W, int T, typename TIn, bool PCA>
__global__ void __launch_bounds__(kBinThreads)
    k_bins_rows(const TIn* __restrict__ raw, float2* __restrict__ planes, const float2* __restrict__ chirp,
                int height, int slot0, int pmin, int K, uint32_t* mm_reset, int mm_slots,
                const float2* __restrict__ wts) {
    constexpr int rpw = HV_BINS_RPW_ROWS;
    using C = BinRowCfg<W>;
    constexpr int kE = FftShape<W, C::E>::kE;
    constexpr bool PF = bins_rows_prefetch<W, T>();
    extern __shared__ float2 rows[];
    if (mm_reset && blockIdx.x == 0) mm_reset_slots(mm_reset, mm_slots);
    const size_t P = (size_t)W * height;
    const int y0 = blockIdx.x * rpw;
    // PCA: the K x T weights behind the plane rows, slot-major with a row pitch
    // of KP >= K8 columns (a compile-time stride: constant LDS offsets), zero padded
    constexpr int KP = T < 8 ? 8 : T;
    const int K8 = (K + 7) & ~7;
    float2* lw = rows + (size_t)(K + 1) * C::PITCH;
    if constexpr (PCA) {
        for (int i = threadIdx.x; i < T * KP; i += kBinThreads) {
            const int f = i / KP, k = i % KP;
            lw[i] = k < K ? wts[(size_t)f * T + k] : make_float2(0.f, 0.f);
        }
        __syncthreads();
    }
    const int g = threadIdx.x / C::NT, t = threadIdx.x % C::NT;
    // Slot f of the ring holds frame (f - slot0) mod T of the batch.
    auto load_window = [&](int y, int x, float* xr) {
#pragma unroll
        for (int f = 0; f < T; ++f) {
            const int j = (f - slot0) & (T - 1);
            xr[f] = convert_px<TIn>(raw[(size_t)j * P + (size_t)y * W + x]);
        }
    };
    float xr[PF ? C::PPT : 1][T];
    if constexpr (PF) {
#pragma unroll
        for (int q = 0; q < C::PPT; ++q)
            if (threadIdx.x + q * kBinThreads < W) load_window(y0, threadIdx.x + q * kBinThreads, xr[q]);
    }
#pragma unroll
    for (int i = 0; i < rpw; ++i) {
        const int y = y0 + i;
        if (y >= height) break;  // uniform across the workgroup
        // 1. DFT over the window of each pixel; bins [pmin, pmin + K) to LDS
#pragma unroll 1
        for (int q = 0; q < C::PPT; ++q) {
            const int x = threadIdx.x + q * kBinThreads;
            if (x >= W) continue;
            if constexpr (!PF) load_window(y, x, xr[0]);
            const float* xs = xr[PF ? q : 0];
            const float2 c = chirp ? chirp[(size_t)y * W + x] : make_float2(1.f, 0.f);
            if constexpr (PCA) {
                // eight components at a time (sixteen independent FMA chains), the
                // weights of a slot read from LDS by the whole wave at one address
#pragma unroll 1
                for (int kb = 0; kb < K; kb += 8) {
                    float2 v[8];
#pragma unroll
                    for (int b = 0; b < 8; ++b) v[b] = make_float2(0.f, 0.f);
                    // FG slots per window: their weights are loaded, then used; the
                    // empty asm statements pin the accumulators and memory at the end
                    // of each window, so that the compiler neither hoists every slot's
                    // loads nor sinks the FMAs (either needs more registers than a
                    // wave has)
                    constexpr int FG = T < 4 ? T : 4;
#pragma unroll
                    for (int f0 = 0; f0 < T; f0 += FG) {
                        float4 w[FG][4];
#pragma unroll
                        for (int u = 0; u < FG; ++u)
#pragma unroll
                            for (int b2 = 0; b2 < 4; ++b2)
                                w[u][b2] = reinterpret_cast<const float4*>(lw + kb + (f0 + u) * KP)[b2];
#pragma unroll
                        for (int u = 0; u < FG; ++u) {
                            const float x = xs[f0 + u];
#pragma unroll
                            for (int b2 = 0; b2 < 4; ++b2) {
                                v[2 * b2].x = fmaf(w[u][b2].x, x, v[2 * b2].x);
                                v[2 * b2].y = fmaf(w[u][b2].y, x, v[2 * b2].y);
                                v[2 * b2 + 1].x = fmaf(w[u][b2].z, x, v[2 * b2 + 1].x);
                                v[2 * b2 + 1].y = fmaf(w[u][b2].w, x, v[2 * b2 + 1].y);
                            }
                        }
                        // the window's FMAs complete here and the next window's loads stay after them
#pragma unroll
                        for (int b = 0; b < 8; ++b) asm volatile("" : "+v"(v[b].x), "+v"(v[b].y)::"memory");
                    }
#pragma unroll
                    for (int b = 0; b < 8; ++b)
                        if (kb + b < K) rows[(kb + b) * C::PITCH + C::pad(x)] = chirp ? cmul(v[b], c) : v[b];
                }
            } else {
                real_dft_bins<T>(xs, pmin, K, [&](int k, float2 v) {
                    rows[(k - pmin) * C::PITCH + C::pad(x)] = chirp ? cmul(v, c) : v;
                });
            }
        }
        if constexpr (PF) {
            if (i + 1 < rpw && y + 1 < height) {
#pragma unroll
                for (int q = 0; q < C::PPT; ++q)
                    if (threadIdx.x + q * kBinThreads < W) load_window(y + 1, threadIdx.x + q * kBinThreads, xr[q]);
            }
        }
        __syncthreads();
        // 2. forward row FFT of each plane row, G rows at a time, in place in LDS
        float2 tw[FftTw<W, C::E>::COUNT];
        fft_twiddles<W, C::E>(t, tw);
        const int rounds = (K + C::G - 1) / C::G;
#pragma unroll 1
        for (int r = 0; r < rounds; ++r) {
            const int row = r * C::G + g;
            const bool active = row < K;
            float2* lrow = rows + (active ? row : K) * C::PITCH;
            auto lds_at = [&](int n) -> float2& { return lrow[C::pad(n)]; };
            float2 v[kE];
            fft_load<W, C::E>(t, [&](int n) -> float2 { return lds_at(n); }, v);
            float2* dst = planes + (size_t)(active ? row : 0) * P + (size_t)y * W;
            fft_compute_tw<W, C::E, -1, false, (C::NT <= 64)>(t, v, tw, [&](int n, float2 z) {
                if (active) dst[n] = z;
            }, lds_at);
        }
        __syncthreads();  // the plane rows are rewritten by the next row's DFT
    }
}

template <int W, int IMG>
__global__ void __launch_bounds__(img_threads<W>())
    k_bins_row_ifft_image(const float2* __restrict__ planes, float* __restrict__ out, int height, int K, int rpw,
                          float count, uint32_t* mm, int mm_slots, float scale) {
    constexpr int kImgThreads = img_threads<W>();
    using C = BinRowCfg<W, kImgThreads>;
    constexpr int kE = FftShape<W, C::E>::kE;
    constexpr int MT = W >= 1024 ? 1 : HV_BINS_IMG_MT;  // plane rows per group per round, interleaved
    constexpr int GR = C::G * MT;      // plane rows per round
    __shared__ float2 xch[GR * C::PITCH];
    __shared__ float part[C::G * W];
    const int g = threadIdx.x / C::NT, t = threadIdx.x % C::NT;
    const size_t P = (size_t)W * height;
    const int y0 = blockIdx.x * rpw;
    const int nrows = min(rpw, height - y0);
    const int rounds = (K + GR - 1) / GR;
    const int items = nrows * rounds;  // (row, round) pairs, loaded ahead
    // plane row of transform m of this group in round r: r GR + m G + g
    auto lds_at = [&](int m, int n) -> float2& { return xch[(m * C::G + g) * C::PITCH + C::pad(n)]; };
    auto load_item = [&](int it, float2* v) {
        if (it >= items) return;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int row = (it % rounds) * GR + m * C::G + g;
            const float2* src = planes + (size_t)(row < K ? row : 0) * P + (size_t)(y0 + it / rounds) * W;
            fft_load<W, C::E>(t, [&](int n) -> float2 { return src[n]; }, v + m * kE);
        }
    };
    float2 tw[FftTw<W, C::E>::COUNT];
    fft_twiddles<W, C::E>(t, tw);
    float acc[kE];
    // D items in flight beyond the one being transformed
    constexpr int D = HV_BINS_IMG_DEPTH;
    float2 buf[D + 1][MT * kE];
#pragma unroll
    for (int d = 0; d < D; ++d) load_item(d, buf[d]);
    float vmin = INFINITY, vmax = -INFINITY;
#pragma unroll 1
    for (int it = 0; it < items; ++it) {
        const int r = it % rounds;
        load_item(it + D, buf[D]);
        if (r == 0) {
#pragma unroll
            for (int i = 0; i < kE; ++i) acc[i] = 0.f;
        }
        fft_compute_multi<W, C::E, +1, MT, (C::NT <= 64)>(t, buf[0], tw, [&](int m, int, float2 z, int slot) {
            if (r * GR + m * C::G + g < K) acc[slot] += image_term_fast<IMG>(z);
        }, lds_at);
        if (r == rounds - 1) {
            // sum the G groups' partial sums of each pixel of this row
            const int y = y0 + it / rounds;
            __syncthreads();  // the previous row's sums have been read
#pragma unroll
            for (int i = 0; i < kE; ++i) part[g * W + fft_out_index<W, C::E>(t, i)] = acc[i];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < C::PPT; ++q) {
                const int x = threadIdx.x + q * kImgThreads;
                if (x >= W) continue;
                float a = 0.f;
#pragma unroll
                for (int j2 = 0; j2 < C::G; ++j2) a += part[j2 * W + x];
                const float val = a / count * scale;
                out[(size_t)y * W + x] = val;
                vmin = fminf(vmin, val);
                vmax = fmaxf(vmax, val);
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int i = 0; i < MT * kE; ++i) buf[d][i] = buf[d + 1][i];
    }
    if (mm) block_minmax_atomic(vmin, vmax, mm, mm_slots);
}

// Fresnel: the planes already hold the propagated bins.
template <int IMG>
__global__ void __launch_bounds__(256)
    k_bins_image(const float2* __restrict__ planes, float* __restrict__ out, size_t P, int K, float count,
                 uint32_t* mm, int mm_slots, float scale) {
    float vmin = INFINITY, vmax = -INFINITY;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        float a = 0.f;
        for (int k = 0; k < K; ++k) a += image_term_fast<IMG>(planes[(size_t)k * P + p]);
        const float val = a / count * scale;
        out[p] = val;
        vmin = fminf(vmin, val);
        vmax = fmaxf(vmax, val);
    }
    if (mm) block_minmax_atomic(vmin, vmax, mm, mm_slots);
}

// ---------------------------------------------------------------------------
// Split path (frames wider than 1024, windows longer than 64, moments): the
// bins-first steps as separate passes over the K planes.
#ifndef HV_BINS_TIME_REAL
#define HV_BINS_TIME_REAL 1     // T >= 32: the real window's DFT through a half-length complex FFT
#endif
#ifndef HV_BINS_TIME_LDS_KB
#define HV_BINS_TIME_LDS_KB 64  // LDS tile of the real time DFT (sets its pixels per workgroup)
#endif
// ---------------------------------------------------------------------------
// Time DFT: CW pixels per workgroup, NT threads per pixel, the length-T
// transform of each pixel's window through LDS (the window read in physical
// ring-slot order, slot n holding frame (n - slot0) mod T of the batch);
// only bins [pmin, pmin + K) are written, x chirp for Fresnel.
template <int T>
struct BinTimeCfg {
    static constexpr int E = T >= 16 ? 16 : T;
    static constexpr int NT = FftShape<T, E>::NT;
    static constexpr int CW = 64;
    static constexpr int THREADS = NT * CW;
};

template <int T, typename TIn>
__global__ void __launch_bounds__(BinTimeCfg<T>::THREADS)
    k_bins_time(const TIn* __restrict__ raw, float2* __restrict__ planes, const float2* __restrict__ chirp, size_t P,
                int slot0, int pmin, int K, uint32_t* mm_reset, int mm_slots, int vec) {
    using C = BinTimeCfg<T>;
    __shared__ float2 lds[T * C::CW];
    // the cycle's first kernel re-arms the image min/max accumulators of its last kernels
    if (mm_reset && blockIdx.x == 0) mm_reset_slots(mm_reset, mm_slots);
    const int c = threadIdx.x % C::CW, t = threadIdx.x / C::CW;
    const size_t p0 = (size_t)blockIdx.x * C::CW, p = p0 + c;
    const bool valid = p < P;  // P is a multiple of CW: every pixel of the tile exists
    auto lds_at = [&](int n) -> float2& { return lds[n * C::CW + c]; };
    // the tile (T frames of CW pixels) through LDS with 16-byte loads: the
    // per-pixel loads of the transform would move CW * sizeof(TIn) bytes per
    // wave instruction
    constexpr int PX = 16 / sizeof(TIn);  // pixels per 16-byte load
    constexpr int CHUNKS = T * C::CW / PX;
    static_assert(C::CW % PX == 0, "a tile row holds whole 16-byte loads");
    for (int idx = threadIdx.x; idx < CHUNKS; idx += C::THREADS) {
        const int n = idx / (C::CW / PX), q = idx % (C::CW / PX);
        const int j = (n - slot0) & (T - 1);  // physical slot n holds frame (n - slot0) mod T
        const TIn* src = raw + (size_t)j * P + p0 + (size_t)q * PX;
        float2* dst = lds + n * C::CW + q * PX;
        if (vec) {
            const uint4 w = *reinterpret_cast<const uint4*>(src);
            const TIn* px = reinterpret_cast<const TIn*>(&w);
#pragma unroll
            for (int e = 0; e < PX; ++e) dst[e] = make_float2(convert_px<TIn>(px[e]), 0.f);
        } else {  // a caller's buffer that is not 16-byte aligned
#pragma unroll
            for (int e = 0; e < PX; ++e) dst[e] = make_float2(convert_px<TIn>(src[e]), 0.f);
        }
    }
    __syncthreads();
    auto load = [&](int n) -> float2 { return lds_at(n); };
    const float2 cf = (chirp && valid) ? chirp[p] : make_float2(1.f, 0.f);
    auto store = [&](int n, float2 z) {
        const int k = n - pmin;
        if (valid && k >= 0 && k < K) planes[(size_t)k * P + p] = chirp ? cmul(z, cf) : z;
    };
    fft_run<T, C::E, -1>(t, load, store, lds_at);
}

// T >= 32: the window is real, so each pixel's length-T DFT is one complex
// FFT of half the length, z_m = x_2m + i x_2m+1, then the split of
// real_dft_bins (X_k = E_k + w^k O_k, conjugate symmetry above T / 2): half
// the LDS tile (T / 2 x 64 pixels: two or more workgroups per CU at T = 256,
// four at T = 128) and half the butterflies.  Thread t of a pixel writes bins
// pmin + t, pmin + t + NT, ..., so a wavefront (64 pixels, one t) writes 512
// contiguous bytes of one plane per store.
// CW pixels per workgroup: the widest tile of HV_BINS_TIME_LDS_KB KiB of LDS
// (each frame of the window read as CW-pixel runs: 64 pixels of u16 frames,
// 128-byte runs, took 1.14 ms for 57 bins of 128 frames of 2048^2 u16 --
// against 0.92 ms with 128 pixels), or 64 pixels when the frame is not a
// multiple of that width.
template <int T>
constexpr int bins_time_cw() {
    return HV_BINS_TIME_LDS_KB * 1024 / (T / 2 * (int)sizeof(float2));
}

template <int T, int CWT>
struct BinTimeRealCfg {
    static constexpr int M = T / 2;  // complex points
    static constexpr int E = M >= 16 ? 16 : M;
    static constexpr int NT = FftShape<M, E>::NT;  // threads per pixel
    static constexpr int CW = CWT;
    static constexpr int THREADS = NT * CW;
};

template <int T, typename TIn, int CWT>
__global__ void __launch_bounds__((BinTimeRealCfg<T, CWT>::THREADS))
    k_bins_time_real(const TIn* __restrict__ raw, float2* __restrict__ planes, const float2* __restrict__ chirp,
                     size_t P, int slot0, int pmin, int K, uint32_t* mm_reset, int mm_slots, int vec) {
    using C = BinTimeRealCfg<T, CWT>;
    __shared__ float2 lds[C::M * C::CW];
    if (mm_reset && blockIdx.x == 0) mm_reset_slots(mm_reset, mm_slots);
    const int c = threadIdx.x % C::CW, t = threadIdx.x / C::CW;
    const size_t p0 = (size_t)blockIdx.x * C::CW, p = p0 + c;
    auto lds_at = [&](int m) -> float2& { return lds[m * C::CW + c]; };
    // sample n of the window (physical slot n holds frame (n - slot0) mod T,
    // as in k_bins_time) into z_{n/2}, real part for even n
    float* ldsf = reinterpret_cast<float*>(lds);
    constexpr int PX = 16 / sizeof(TIn);  // pixels per 16-byte load
    constexpr int CHUNKS = T * C::CW / PX;
    static_assert(C::CW % PX == 0, "a tile row holds whole 16-byte loads");
    for (int idx = threadIdx.x; idx < CHUNKS; idx += C::THREADS) {
        const int n = idx / (C::CW / PX), q = idx % (C::CW / PX);
        const int j = (n - slot0) & (T - 1);
        const TIn* src = raw + (size_t)j * P + p0 + (size_t)q * PX;
        float* dst = ldsf + ((size_t)(n >> 1) * C::CW + q * PX) * 2 + (n & 1);
        if (vec) {
            const uint4 w = *reinterpret_cast<const uint4*>(src);
            const TIn* px = reinterpret_cast<const TIn*>(&w);
#pragma unroll
            for (int e = 0; e < PX; ++e) dst[2 * e] = convert_px<TIn>(px[e]);
        } else {  // a caller's buffer that is not 16-byte aligned
#pragma unroll
            for (int e = 0; e < PX; ++e) dst[2 * e] = convert_px<TIn>(src[e]);
        }
    }
    __syncthreads();
    // Z = FFT_M(z), left in the tile in natural order
    fft_run<C::M, C::E, -1, true>(t, [&](int m) -> float2 { return lds_at(m); },
                                  [&](int m, float2 z) { lds_at(m) = z; }, lds_at);
    __syncthreads();
    const float2 cf = chirp ? chirp[p] : make_float2(1.f, 0.f);
    for (int k = pmin + t; k < pmin + K; k += C::NT) {
        const int kk = k <= C::M ? k : T - k;
        const float2 a = lds_at(kk % C::M), bc = lds_at((C::M - kk) % C::M);
        const float2 b = make_float2(bc.x, -bc.y);
        const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
        const float2 d = make_float2(0.5f * (a.x - b.x), 0.5f * (a.y - b.y));
        const float2 o = make_float2(d.y, -d.x);  // d / i
        const float2 w = twiddle<-1>(kk, clog2(T));
        float2 v = make_float2(e.x + (w.x * o.x - w.y * o.y), e.y + (w.x * o.y + w.y * o.x));
        if (k > C::M) v.y = -v.y;
        planes[(size_t)(k - pmin) * P + p] = chirp ? cmul(v, cf) : v;
    }
}

// Few bins (K <= HV_BINS_DIRECT_K) in their own pass: each thread sums the
// window of 8 pixels directly, X_k = sum_n x_n e^{-2 pi i k n / T} over the
// physical slots n (slot n holding frame (n - slot0) mod T), each frame read
// once with 8-pixel loads, kDirectG frames' loads in flight, no LDS and no barrier
// (the transform kernels above stage a tile of every pixel's window in LDS
// and compute all T bins: latency-bound when only a few are kept).  The sums
// run over x_n - x_0: for k != 0 mod T the constant drops out exactly, so the
// partial sums stay at the scale of the window's variation, not of its mean
// (T x_0 is added back to bin 0).  With `ins_src`, the cycle's batch is
// inserted on the way: slots (ws0 + b) mod T, b < nb, are read from the batch
// and written to ins_dst (the raw ring being read), so the ring's update and
// the transform's read are one pass.
#ifndef HV_BINS_DIRECT_K
#define HV_BINS_DIRECT_K 4
#endif
constexpr int kDirectThreads = 256;
// frames whose loads are issued together: 1024^2 u16, T = 64, one bin, 3 lanes ran at
// 327 k / 324 k / 340 k / 358 k / 328 k frames/s with 16 / 8 / 4 / 2 / 1 (the
// fewer registers, the more of the other lanes' kernels beside it)
#ifndef HV_BINS_DIRECT_G
#define HV_BINS_DIRECT_G 2
#endif
constexpr int kDirectG = HV_BINS_DIRECT_G;

template <typename TIn>
struct Px8 {  // 8 pixels as loaded: 8, 16 or 32 bytes
    uint4 a, b;
};

template <typename TIn>
__device__ __forceinline__ void px8_load(const TIn* src, int vec, Px8<TIn>& v) {
    if (vec) {
        if constexpr (sizeof(TIn) == 1) {
            const uint2 w = *reinterpret_cast<const uint2*>(src);
            v.a = make_uint4(w.x, w.y, 0u, 0u);
        } else {
            v.a = *reinterpret_cast<const uint4*>(src);
            if constexpr (sizeof(TIn) == 4) v.b = *reinterpret_cast<const uint4*>(src + 4);
        }
    } else {  // a caller's buffer that is not 16-byte aligned
        TIn* px = reinterpret_cast<TIn*>(&v);
#pragma unroll
        for (int e = 0; e < 8; ++e) px[e] = src[e];
    }
}

template <typename TIn>
__device__ __forceinline__ void px8_store(TIn* dst, const Px8<TIn>& v) {  // dst 16-byte aligned (the raw ring)
    if constexpr (sizeof(TIn) == 1) {
        *reinterpret_cast<uint2*>(dst) = make_uint2(v.a.x, v.a.y);
    } else {
        *reinterpret_cast<uint4*>(dst) = v.a;
        if constexpr (sizeof(TIn) == 4) *reinterpret_cast<uint4*>(dst + 4) = v.b;
    }
}

template <int T, typename TIn, int KT>
__global__ void __launch_bounds__(kDirectThreads)
    k_bins_time_direct(const TIn* raw, int slot0, TIn* ins_dst, const TIn* __restrict__ ins_src, int nb, int ws0,
                       float2* __restrict__ planes, const float2* __restrict__ chirp, size_t P, int pmin, int K,
                       uint32_t* mm_reset, int mm_slots, int vec) {
    if (mm_reset && blockIdx.x == 0) mm_reset_slots(mm_reset, mm_slots);
    const size_t p = ((size_t)blockIdx.x * kDirectThreads + threadIdx.x) * 8;
    if (p >= P) return;  // P is a multiple of 8: a thread's 8 pixels exist together
    constexpr int LG = clog2(T);
    constexpr int G = T < kDirectG ? T : kDirectG;
    float2 acc[KT][8];
#pragma unroll
    for (int k = 0; k < KT; ++k)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[k][e] = make_float2(0.f, 0.f);
    float x0[8];
    for (int n0 = 0; n0 < T; n0 += G) {
        Px8<TIn> v[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int n = n0 + g;
            const int r = (n - ws0) & (T - 1);
            const bool ins = ins_src && r < nb;  // uniform
            const TIn* src = ins ? ins_src + (size_t)r * P + p : raw + (size_t)((n - slot0) & (T - 1)) * P + p;
            px8_load<TIn>(src, ins ? 1 : vec, v[g]);
        }
        if (ins_src) {
#pragma unroll
            for (int g = 0; g < G; ++g)
                if (((n0 + g - ws0) & (T - 1)) < nb) px8_store<TIn>(ins_dst + (size_t)(n0 + g) * P + p, v[g]);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const TIn* px = reinterpret_cast<const TIn*>(&v[g]);
            float x[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = convert_px<TIn>(px[e]);
            if (n0 == 0 && g == 0) {
#pragma unroll
                for (int e = 0; e < 8; ++e) x0[e] = x[e];
            }
            const int n = n0 + g;
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                if (k >= K) break;
                const float2 w = twiddle<-1>(((pmin + k) * n) & (T - 1), LG);
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float d = x[e] - x0[e];
                    acc[k][e].x = fmaf(d, w.x, acc[k][e].x);
                    acc[k][e].y = fmaf(d, w.y, acc[k][e].y);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        if (k >= K) break;
        const bool dc = ((pmin + k) & (T - 1)) == 0;
        float2 out[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float2 z = acc[k][e];
            if (dc) z.x += (float)T * x0[e];
            out[e] = chirp ? cmul(z, chirp[p + e]) : z;
        }
        float4* dst = reinterpret_cast<float4*>(planes + (size_t)k * P + p);
#pragma unroll
        for (int e = 0; e < 4; ++e) dst[e] = make_float4(out[2 * e].x, out[2 * e].y, out[2 * e + 1].x, out[2 * e + 1].y);
    }
}

// The cycle's batch into the raw ring by copies (the kernels that stage the
// window in LDS read the ring after it): at most two contiguous runs.
hv_status raw_insert_copy(const RawInsert& ins, size_t frame_bytes, int T, hipStream_t st) {
    if (!ins.batch) return HV_OK;
    if (ins.n < 1 || ins.n >= T || ins.slot0 < 0 || ins.slot0 >= T)
        return fail(HV_ERR_INVALID_ARG, "raw ring insert: 0 < n < T and 0 <= slot0 < T");
    const int n1 = std::min(ins.n, T - ins.slot0);
    HV_HIP(hipMemcpyAsync((char*)ins.ring + (size_t)ins.slot0 * frame_bytes, ins.batch, (size_t)n1 * frame_bytes,
                          hipMemcpyDeviceToDevice, st));
    if (ins.n > n1)
        HV_HIP(hipMemcpyAsync(ins.ring, (const char*)ins.batch + (size_t)n1 * frame_bytes,
                              (size_t)(ins.n - n1) * frame_bytes, hipMemcpyDeviceToDevice, st));
    return HV_OK;
}

template <int T, typename TIn, int KT>
hv_status time_direct_t(const void* raw, float2* planes, const float2* chirp, size_t P, int slot0, int pmin, int K,
                        uint32_t* mm_reset, int mm_slots, hipStream_t st, const RawInsert& ins) {
    if (P % 8) return fail(HV_ERR_UNSUPPORTED, "bins-first time DFT: frames of a multiple of 8 pixels");
    if (ins.batch) {
        if (ins.ring != raw || slot0 != 0 || ins.n < 1 || ins.n >= T || ins.slot0 < 0 || ins.slot0 >= T)
            return fail(HV_ERR_INVALID_ARG, "raw ring insert: the ring read in slot order, 0 < n < T");
        if ((reinterpret_cast<uintptr_t>(ins.ring) & 15) || (reinterpret_cast<uintptr_t>(ins.batch) & 15)) {
            HV_TRY(raw_insert_copy(ins, P * sizeof(TIn), T, st));
            return time_direct_t<T, TIn, KT>(raw, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, RawInsert{});
        }
    }
    const int vec = (reinterpret_cast<uintptr_t>(raw) & 15) == 0;
    const unsigned g = (unsigned)((P / 8 + kDirectThreads - 1) / kDirectThreads);
    k_bins_time_direct<T, TIn, KT><<<g, kDirectThreads, 0, st>>>(
        (const TIn*)raw, slot0, (TIn*)ins.ring, (const TIn*)ins.batch, ins.n, ins.slot0, planes, chirp, P, pmin, K,
        mm_reset, mm_slots, vec);
    HV_CHECK_LAUNCH("k_bins_time_direct");
    return HV_OK;
}

template <int T, typename TIn>
hv_status time_launch_t(const void* raw, float2* planes, const float2* chirp, size_t P, int slot0, int pmin, int K,
                        uint32_t* mm_reset, int mm_slots, hipStream_t st, const RawInsert& ins) {
    if (K <= HV_BINS_DIRECT_K) {
        if (K == 1) return time_direct_t<T, TIn, 1>(raw, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        if (K == 2) return time_direct_t<T, TIn, 2>(raw, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        return time_direct_t<T, TIn, 4>(raw, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
    }
    HV_TRY(raw_insert_copy(ins, P * sizeof(TIn), T, st));
    if constexpr (T >= 32 && HV_BINS_TIME_REAL) {
        constexpr int CWB = bins_time_cw<T>();
        static_assert(CWB >= 64 && BinTimeRealCfg<T, CWB>::THREADS <= 1024, "time DFT tile");
        if (P % 64) return fail(HV_ERR_UNSUPPORTED, "bins-first time DFT: frames of a multiple of 64 pixels");
        const int vec = (reinterpret_cast<uintptr_t>(raw) & 15) == 0;
        if (P % CWB == 0) {
            k_bins_time_real<T, TIn, CWB><<<(unsigned)(P / CWB), BinTimeRealCfg<T, CWB>::THREADS, 0, st>>>(
                (const TIn*)raw, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, vec);
        } else {
            k_bins_time_real<T, TIn, 64><<<(unsigned)(P / 64), BinTimeRealCfg<T, 64>::THREADS, 0, st>>>(
                (const TIn*)raw, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, vec);
        }
        HV_CHECK_LAUNCH("k_bins_time_real");
        return HV_OK;
    }
    using C = BinTimeCfg<T>;
    if (P % C::CW) return fail(HV_ERR_UNSUPPORTED, "bins-first time DFT: frames of a multiple of 64 pixels");
    const int vec = (reinterpret_cast<uintptr_t>(raw) & 15) == 0;
    const unsigned g = (unsigned)(P / C::CW);
    k_bins_time<T, TIn><<<g, C::THREADS, 0, st>>>((const TIn*)raw, planes, chirp, P, slot0, pmin, K, mm_reset,
                                                  mm_slots, vec);
    HV_CHECK_LAUNCH("k_bins_time");
    return HV_OK;
}

template <int T>
hv_status time_launch_d(const void* raw, int depth, float2* planes, const float2* chirp, size_t P, int slot0,
                        int pmin, int K, uint32_t* mm_reset, int mm_slots, hipStream_t st, const RawInsert& ins) {
    switch (depth) {
        case HV_DEPTH_8: return time_launch_t<T, uint8_t>(raw, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        case HV_DEPTH_16: return time_launch_t<T, uint16_t>(raw, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        default: return time_launch_t<T, float>(raw, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
    }
}

hv_status bins_time(const void* raw, int depth, float2* planes, const float2* chirp, size_t P, int T, int slot0,
                    int pmin, int K, uint32_t* mm_reset, int mm_slots, hipStream_t st, const RawInsert& ins) {
    switch (T) {
        case 2: return time_launch_d<2>(raw, depth, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        case 4: return time_launch_d<4>(raw, depth, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        case 8: return time_launch_d<8>(raw, depth, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        case 16: return time_launch_d<16>(raw, depth, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        case 32: return time_launch_d<32>(raw, depth, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        case 64: return time_launch_d<64>(raw, depth, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        case 128: return time_launch_d<128>(raw, depth, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        case 256: return time_launch_d<256>(raw, depth, planes, chirp, P, slot0, pmin, K, mm_reset, mm_slots, st, ins);
        default: return fail(HV_ERR_UNSUPPORTED, "bins-first path: T must be a power of two in [2, 256]");
    }
}

// PCA projection for the split path: the tile (T slots x 64 pixels, converted,
// physical slot order) through LDS; wave w computes components
// [8 w + 64 i, + 8) of every pixel of the tile (lane = pixel), the weights of
// a component and slot uniform across the wave (scalar loads).
constexpr int kPcaTimeThreads = 512;
constexpr int kPcaTimeCW = 64;

template <int T, typename TIn>
__global__ void __launch_bounds__(kPcaTimeThreads)
    k_pca_bins_time(const TIn* __restrict__ raw, float2* __restrict__ planes, const float2* __restrict__ chirp,
                    size_t P, int slot0, int K, const float2* __restrict__ wts, uint32_t* mm_reset, int mm_slots,
                    int vec) {
    constexpr int CW = kPcaTimeCW;
    __shared__ float lds[T * CW];
    if (mm_reset && blockIdx.x == 0) mm_reset_slots(mm_reset, mm_slots);
    const size_t p0 = (size_t)blockIdx.x * CW;
    constexpr int PX = 16 / sizeof(TIn);  // pixels per 16-byte load
    constexpr int CHUNKS = T * CW / PX;
    for (int idx = threadIdx.x; idx < CHUNKS; idx += kPcaTimeThreads) {
        const int n = idx / (CW / PX), q = idx % (CW / PX);
        const int j = (n - slot0) & (T - 1);  // physical slot n holds frame (n - slot0) mod T
        const TIn* src = raw + (size_t)j * P + p0 + (size_t)q * PX;
        float* dst = lds + n * CW + q * PX;
        if (vec) {
            const uint4 w = *reinterpret_cast<const uint4*>(src);
            const TIn* px = reinterpret_cast<const TIn*>(&w);
#pragma unroll
            for (int e = 0; e < PX; ++e) dst[e] = convert_px<TIn>(px[e]);
        } else {
#pragma unroll
            for (int e = 0; e < PX; ++e) dst[e] = convert_px<TIn>(src[e]);
        }
    }
    __syncthreads();
    // the wave index made uniform for the compiler: the weights are then scalar loads
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t p = p0 + lane;
    const float2 cf = chirp ? chirp[p] : make_float2(1.f, 0.f);
    constexpr int NB = 8;
    for (int kb = wave * NB; kb < K; kb += NB * (kPcaTimeThreads / 64)) {
        float2 acc[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = make_float2(0.f, 0.f);
#pragma unroll 4
        for (int f = 0; f < T; ++f) {
            const float x = lds[f * CW + lane];
            const float2* wf = wts + (size_t)f * T + kb;  // 8 consecutive columns: one scalar load
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const float2 w = wf[b];
                acc[b].x = fmaf(w.x, x, acc[b].x);
                acc[b].y = fmaf(w.y, x, acc[b].y);
            }
        }
#pragma unroll
        for (int b = 0; b < NB; ++b)
            if (kb + b < K) planes[(size_t)(kb + b) * P + p] = chirp ? cmul(acc[b], cf) : acc[b];
    }
}

// The projection on MFMA (integer frames, T = 32, 64, 128, at most 64
// components): the planes are the product of the window (pixels x T frames)
// and the T x 2K real weight matrix (component k's real and imaginary
// weights in columns 2k, 2k + 1), as v_mfma_f32_32x32x16_f16 steps on split
// operands: frames x / 16 = hi + lo exactly (converted pixels have at most 16
// significant bits), weights 256 w = hi + lo (22 bits), three products per
// step (lo hi, hi lo, hi hi) accumulated in FP32; the powers of two are
// divided out exactly.  Each workgroup walks 64-pixel chunks (persistent,
// one grid of resident workgroups): the chunk's T x 64 window is staged in
// LDS as pixel rows of FP16 hi / lo (8 x 8 blocks transposed in registers),
// wave w runs columns 32 w .. 32 w + 31 with their B fragments held in
// registers for every chunk, and writes its 16 components through a
// wave-private LDS tile as 256-byte plane runs.
#ifndef HV_PCA_TIME_MFMA
#define HV_PCA_TIME_MFMA 1
#endif
constexpr int kPcaMfmaWaves = 4;
constexpr int kPcaMfmaCW = 64;

typedef _Float16 v8h_p __attribute__((ext_vector_type(8)));
typedef float v16f_p __attribute__((ext_vector_type(16)));

template <int T>
struct PcaMfmaCfg {
    static constexpr int KS = T / 16;                          // k-steps of 16 frames
    static constexpr int PITCH = T + 8;                        // halfs per pixel row (16-byte pad)
    static constexpr int TILE = kPcaMfmaCW * PITCH;            // halfs of the hi (or lo) tile
};

template <typename TIn>
constexpr bool pca_time_mfma_ok(int T, int K) {
    return HV_PCA_TIME_MFMA && !std::is_same<TIn, float>::value && (T == 32 || T == 64 || T == 128) && K <= 64;
}

// Eight pixels of one frame, converted.
template <typename TIn>
__device__ __forceinline__ void load8_px(const TIn* src, int vec, float* x) {
    if constexpr (sizeof(TIn) == 2) {
        if (vec) {
            const uint4 w = *reinterpret_cast<const uint4*>(src);
            const TIn* px = reinterpret_cast<const TIn*>(&w);
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = convert_px<TIn>(px[e]);
            return;
        }
    } else {
        if (vec) {
            const uint2 w = *reinterpret_cast<const uint2*>(src);
            const TIn* px = reinterpret_cast<const TIn*>(&w);
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = convert_px<TIn>(px[e]);
            return;
        }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = convert_px<TIn>(src[e]);
}

template <int T, typename TIn>
__global__ void __launch_bounds__(64 * kPcaMfmaWaves)
    k_pca_time_mfma(const TIn* __restrict__ raw, float2* __restrict__ planes, const float2* __restrict__ chirp,
                    size_t P, int slot0, int K, const float2* __restrict__ wts, uint32_t* mm_reset, int mm_slots,
                    int vec) {
    using C = PcaMfmaCfg<T>;
    constexpr float XS = 1.f / 16, WS = 256.f, OS = 1.f / (XS * WS);
    __shared__ __align__(16) _Float16 th[C::TILE], tl[C::TILE];
    __shared__ float dl[kPcaMfmaWaves][32 * 33];
    if (mm_reset && blockIdx.x == 0) mm_reset_slots(mm_reset, mm_slots);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5, l32 = lane & 31;
    const bool active = 32 * wave < 2 * K;  // uniform across the wave
    // B fragments: lane l holds columns 32 w + l % 32, frames 16 s + 8 (l / 32) + e
    v8h_p bh[C::KS], bl[C::KS];
    {
        const int n = 32 * wave + l32, j = n >> 1;
#pragma unroll
        for (int s = 0; s < C::KS; ++s)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float v = 0.f;
                if (j < K) {
                    const float2 w = wts[(size_t)(16 * s + 8 * half + e) * T + j];
                    v = ((n & 1) ? w.y : w.x) * WS;
                }
                const _Float16 h = (_Float16)v;
                bh[s][e] = h;
                bl[s][e] = (_Float16)(v - (float)h);
            }
    }
    float* d = dl[wave];
    const size_t nchunks = P / kPcaMfmaCW;
    for (size_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const size_t p0 = chunk * kPcaMfmaCW;
        __syncthreads();  // the previous chunk's tile is read
        // T blocks of 8 frames x 8 pixels (window frame f from physical slot (f - slot0) mod T)
        for (int b = threadIdx.x; b < T; b += 64 * kPcaMfmaWaves) {
            const int fg = b >> 3, pg = b & 7;
            float x[8][8];
#pragma unroll
            for (int fi = 0; fi < 8; ++fi) {
                const int slot = (8 * fg + fi - slot0) & (T - 1);
                load8_px<TIn>(raw + (size_t)slot * P + p0 + 8 * pg, vec, x[fi]);
            }
#pragma unroll
            for (int pi = 0; pi < 8; ++pi) {
                v8h_p hi, lo;
#pragma unroll
                for (int fi = 0; fi < 8; ++fi) {
                    const float v = x[fi][pi] * XS;
                    const _Float16 h = (_Float16)v;
                    hi[fi] = h;
                    lo[fi] = (_Float16)(v - (float)h);
                }
                *reinterpret_cast<v8h_p*>(th + (8 * pg + pi) * C::PITCH + 8 * fg) = hi;
                *reinterpret_cast<v8h_p*>(tl + (8 * pg + pi) * C::PITCH + 8 * fg) = lo;
            }
        }
        __syncthreads();
        if (!active) continue;
#pragma unroll
        for (int mt = 0; mt < kPcaMfmaCW / 32; ++mt) {
            v16f_p acc = {};
            const int row = (32 * mt + l32) * C::PITCH + 8 * half;
#pragma unroll
            for (int s = 0; s < C::KS; ++s) {
                const v8h_p xh = *reinterpret_cast<const v8h_p*>(th + row + 16 * s);
                const v8h_p xl = *reinterpret_cast<const v8h_p*>(tl + row + 16 * s);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(xl, bh[s], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh, bl[s], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh, bh[s], acc, 0, 0, 0);
            }
            // register r of lane l: pixel 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
#pragma unroll
            for (int r = 0; r < 16; ++r) d[l32 * 33 + 8 * (r / 4) + 4 * half + (r % 4)] = acc[r];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int cl = 2 * q + half, k = 16 * wave + cl;
                if (k < K) {
                    const size_t p = p0 + 32 * mt + l32;
                    float2 v = make_float2(d[(2 * cl) * 33 + l32] * OS, d[(2 * cl + 1) * 33 + l32] * OS);
                    if (chirp) v = cmul(v, chirp[p]);
                    planes[(size_t)k * P + p] = v;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// Resident workgroups of k_pca_time_mfma on the current device.
template <int T, typename TIn>
static hv_status pca_mfma_grid(unsigned& grid) {
    static int per_cu = 0, cus = 0;
    if (!per_cu) {
        int dev = 0;
        HV_HIP(hipGetDevice(&dev));
        HV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_pca_time_mfma<T, TIn>), 64 * kPcaMfmaWaves, 0));
        if (per_cu < 1) per_cu = 1;
    }
    grid = (unsigned)(cus * per_cu);
    return HV_OK;
}

template <int T, typename TIn>
hv_status pca_time_t(const void* raw, float2* planes, const float2* chirp, size_t P, int slot0, int K,
                     const float2* wts, uint32_t* mm_reset, int mm_slots, hipStream_t st) {
    if (P % kPcaTimeCW) return fail(HV_ERR_UNSUPPORTED, "bins-first PCA: frames of a multiple of 64 pixels");
    const int vec = (reinterpret_cast<uintptr_t>(raw) & 15) == 0;
    if constexpr (T == 32 || T == 64 || T == 128) {
        if (pca_time_mfma_ok<TIn>(T, K)) {
            unsigned grid = 0;
            HV_TRY((pca_mfma_grid<T, TIn>(grid)));
            const size_t nchunks = P / kPcaMfmaCW;
            if (grid > nchunks) grid = (unsigned)nchunks;
            k_pca_time_mfma<T, TIn><<<grid, 64 * kPcaMfmaWaves, 0, st>>>((const TIn*)raw, planes, chirp, P, slot0,
                                                                         K, wts, mm_reset, mm_slots, vec);
            HV_CHECK_LAUNCH("k_pca_time_mfma");
            return HV_OK;
        }
    }
    k_pca_bins_time<T, TIn><<<(unsigned)(P / kPcaTimeCW), kPcaTimeThreads, 0, st>>>(
        (const TIn*)raw, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, vec);
    HV_CHECK_LAUNCH("k_pca_bins_time");
    return HV_OK;
}

template <int T>
hv_status pca_time_d(const void* raw, int depth, float2* planes, const float2* chirp, size_t P, int slot0, int K,
                     const float2* wts, uint32_t* mm_reset, int mm_slots, hipStream_t st) {
    switch (depth) {
        case HV_DEPTH_8: return pca_time_t<T, uint8_t>(raw, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        case HV_DEPTH_16: return pca_time_t<T, uint16_t>(raw, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        default: return pca_time_t<T, float>(raw, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
    }
}

hv_status pca_time(const void* raw, int depth, float2* planes, const float2* chirp, size_t P, int T, int slot0,
                   int K, const float2* wts, uint32_t* mm_reset, int mm_slots, hipStream_t st) {
    switch (T) {
        case 2: return pca_time_d<2>(raw, depth, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        case 4: return pca_time_d<4>(raw, depth, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        case 8: return pca_time_d<8>(raw, depth, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        case 16: return pca_time_d<16>(raw, depth, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        case 32: return pca_time_d<32>(raw, depth, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        case 64: return pca_time_d<64>(raw, depth, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        case 128: return pca_time_d<128>(raw, depth, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        case 256: return pca_time_d<256>(raw, depth, planes, chirp, P, slot0, K, wts, mm_reset, mm_slots, st);
        default: return fail(HV_ERR_UNSUPPORTED, "bins-first PCA: T must be a power of two in [2, 256]");
    }
}

// Moments of the propagated bins (tools_conversion.cu:33-66, tools_compute.cu:38-97):
// plane k holds bin i = fs + k; f1 = fftfreq(T) * fps in float as the
// reference computes it; the Nyquist bin counted again for M0 and M2 and
// cancelled for M1.  `img` receives moment `sel`, whose min/max feeds mm.
__global__ void __launch_bounds__(256)
    k_bins_moments(const float2* __restrict__ planes, float* m0, float* m1, float* m2, float* img, size_t P, int T,
                   int fs, int K, float fps, int sel, uint32_t* mm, int mm_slots) {
    const float d = fps / T;
    const int nyq = T / 2;
    float vmin = INFINITY, vmax = -INFINITY;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        for (int k = 0; k < K; ++k) {
            const int i = fs + k;
            const float2 z = planes[(size_t)k * P + p];
            const float sq = z.x * z.x + z.y * z.y;
            const float f = i < nyq ? i * d : -((float)T - i) * d;
            a0 += sq * 1.0f;
            a1 += sq * f;
            a2 += sq * (f * f);
        }
        if (nyq >= fs && nyq < fs + K) {
            const float2 z = planes[(size_t)(nyq - fs) * P + p];
            const float sq = z.x * z.x + z.y * z.y;
            const float f = -((float)T - nyq) * d;
            a0 += sq * 1.0f;
            a1 -= sq * f;
            a2 += sq * (f * f);
        }
        if (m0) m0[p] = a0;
        if (m1) m1[p] = a1;
        if (m2) m2[p] = a2;
        const float v = sel == 0 ? a0 : sel == 1 ? a1 : a2;
        if (img) img[p] = v;
        vmin = fminf(vmin, v);
        vmax = fmaxf(vmax, v);
    }
    if (mm) block_minmax_atomic(vmin, vmax, mm, mm_slots);
}

template <int W, int T, typename TIn, bool PCA>
hv_status rows_launch(const void* raw, float2* planes, const float2* chirp, int height, int slot0, int pmin, int K,
                      uint32_t* mm_reset, int mm_slots, const float2* wts, hipStream_t st) {
    // the plane rows may take more than the default 64 KiB of dynamic LDS
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bins_rows<W, T, TIn, PCA>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (attr != hipSuccess) return hip_fail(attr, "hipFuncSetAttribute(k_bins_rows)");
    const int rpw = HV_BINS_RPW_ROWS;
    k_bins_rows<W, T, TIn, PCA><<<(height + rpw - 1) / rpw, kBinThreads, bins_rows_lds<W>(K, PCA ? T : 0), st>>>(
        (const TIn*)raw, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts);
    HV_CHECK_LAUNCH("k_bins_rows");
    return HV_OK;
}

template <int W, int T, typename TIn>
hv_status rows_wtd(const void* raw, float2* planes, const float2* chirp, int height, int slot0, int pmin, int K,
                   uint32_t* mm_reset, int mm_slots, const float2* wts, hipStream_t st) {
    return wts ? rows_launch<W, T, TIn, true>(raw, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st)
               : rows_launch<W, T, TIn, false>(raw, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
}

template <int W, int T>
hv_status rows_wt(const void* raw, int depth, float2* planes, const float2* chirp, int height, int slot0, int pmin,
                  int K, uint32_t* mm_reset, int mm_slots, const float2* wts, hipStream_t st) {
    switch (depth) {
        case HV_DEPTH_8: return rows_wtd<W, T, uint8_t>(raw, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        case HV_DEPTH_16: return rows_wtd<W, T, uint16_t>(raw, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        default: return rows_wtd<W, T, float>(raw, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
    }
}

template <int W>
hv_status rows_w(const void* raw, int depth, float2* planes, const float2* chirp, int height, int T, int slot0,
                 int pmin, int K, uint32_t* mm_reset, int mm_slots, const float2* wts, hipStream_t st) {
    switch (T) {
        case 2: return rows_wt<W, 2>(raw, depth, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        case 4: return rows_wt<W, 4>(raw, depth, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        case 8: return rows_wt<W, 8>(raw, depth, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        case 16: return rows_wt<W, 16>(raw, depth, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        case 32: return rows_wt<W, 32>(raw, depth, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        case 64: return rows_wt<W, 64>(raw, depth, planes, chirp, height, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        default: return fail(HV_ERR_UNSUPPORTED, "bins-first path: T must be a power of two in [2, 64]");
    }
}

size_t rows_lds_for(int width, int K, int pca_T) {
    switch (width) {
        case 64: return bins_rows_lds<64>(K, pca_T);
        case 128: return bins_rows_lds<128>(K, pca_T);
        case 256: return bins_rows_lds<256>(K, pca_T);
        case 512: return bins_rows_lds<512>(K, pca_T);
        case 1024: return bins_rows_lds<1024>(K, pca_T);
        default: return 0;
    }
}

}  // namespace

bool bins_first_supported(int width, int height, int T, int K, int depth, bool pca) {
    if (!(width == 64 || width == 128 || width == 256 || width == 512 || width == 1024)) return false;
    if (!is_pow2(height) || height < 16 || height > 4096) return false;
    if (!is_pow2(T) || T < 2 || T > 64 || K < 1 || K > T) return false;
    if (width == 1024 && T > 32) return false;  // two pixels' windows per thread exceed the registers
    if (depth != HV_DEPTH_8 && depth != HV_DEPTH_16 && depth != HV_DEPTH_32F) return false;
    return rows_lds_for(width, K, pca ? T : 0) <= 160 * 1024;
}

bool bins_split_supported(int width, int height, int T, int K, int depth) {
    if (!fft_len_ok(width) || !fft_len_ok(height)) return false;
    if (!is_pow2(T) || T < 2 || T > 256 || K < 1 || K > T) return false;
    return depth == HV_DEPTH_8 || depth == HV_DEPTH_16 || depth == HV_DEPTH_32F;
}

hv_status launch_bins_rows(const void* raw, int depth, int width, int height, int T, int slot0, int pmin, int K,
                           int space, const float2* lens, float2* planes, hipStream_t st, uint32_t* mm_reset,
                           int mm_slots, const float2* wts, float2* tiles, RawInsert ins) {
    if (space != HV_SPACE_ANGULARSP && space != HV_SPACE_FRESNELTR)
        return fail(HV_ERR_UNSUPPORTED, "bins-first path needs a space transform");
    const float2* chirp = space == HV_SPACE_FRESNELTR ? lens : nullptr;
    if (!bins_first_supported(width, height, T, K, depth, wts != nullptr)) {
        if (!bins_split_supported(width, height, T, K, depth))
            return fail(HV_ERR_UNSUPPORTED, "bins-first path: unsupported shape");
        // split: the window's bins (or PCA components), then their forward row FFTs
        const size_t P = (size_t)width * height;
        if (wts) {
            HV_TRY(raw_insert_copy(ins, P * depth, T, st));
            HV_TRY(pca_time(raw, depth, planes, chirp, P, T, slot0, K, wts, mm_reset, mm_slots, st));
        } else {
            HV_TRY(bins_time(raw, depth, planes, chirp, P, T, slot0, pmin, K, mm_reset, mm_slots, st, ins));
        }
        // column-tiled planes: the forward row pass writes the tiles the column pass reads
        if (tiles) return launch_row_pass_tiled(planes, tiles, width, height, K, -1, st);
        return launch_row_pass(planes, width, height, K, -1, st);
    }
    HV_TRY(raw_insert_copy(ins, (size_t)width * height * depth, T, st));
    switch (width) {
        case 64: return rows_w<64>(raw, depth, planes, chirp, height, T, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        case 128: return rows_w<128>(raw, depth, planes, chirp, height, T, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        case 256: return rows_w<256>(raw, depth, planes, chirp, height, T, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        case 512: return rows_w<512>(raw, depth, planes, chirp, height, T, slot0, pmin, K, mm_reset, mm_slots, wts, st);
        default: return rows_w<1024>(raw, depth, planes, chirp, height, T, slot0, pmin, K, mm_reset, mm_slots, wts, st);
    }
}

hv_status launch_bins_columns(float2* planes, int width, int height, int K, int space, const float2* lens,
                              hipStream_t st, float2* tiles, const float2* lens_t) {
    if (tiles) return launch_col_lens_tiled(tiles, lens_t, width, height, K, st);
    return launch_col_pass(planes, space == HV_SPACE_ANGULARSP ? lens : nullptr, width, height, K,
                           space == HV_SPACE_ANGULARSP ? COLPASS_LENS : COLPASS_FWD, st);
}

hv_status launch_bins_propagate(const void* raw, int depth, int width, int height, int T, int slot0, int pmin,
                                int K, int space, const float2* lens, float2* planes, hipStream_t st,
                                hipEvent_t mark, uint32_t* mm_reset, int mm_slots, const float2* wts, float2* tiles,
                                const float2* lens_t, RawInsert ins) {
    HV_TRY(launch_bins_rows(raw, depth, width, height, T, slot0, pmin, K, space, lens, planes, st, mm_reset,
                            mm_slots, wts, tiles, ins));
    if (mark) HV_HIP(hipEventRecord(mark, st));
    return launch_bins_columns(planes, width, height, K, space, lens, st, tiles, lens_t);
}

bool bins_tiled_supported(int width, int height, int T, int K, int depth, int space, bool pca) {
    return space == HV_SPACE_ANGULARSP && !bins_first_supported(width, height, T, K, depth, pca) &&
           bins_split_supported(width, height, T, K, depth) && col_tiled_supported(width, height);
}

// The inverse row FFT and the image term in one kernel (image rows of
// W <= 1024 within LDS).
bool bins_image_fused(int width, int space) {
    return space == HV_SPACE_ANGULARSP && is_pow2(width) && width >= 64 && width <= 1024;
}

template <int IMG>
hv_status bins_image_t(float* out, float2* planes, int width, int height, int K, int space, float count,
                       uint32_t* mm, int mm_slots, float scale, hipStream_t st, float2* tiles) {
    // column-tiled planes, or row-major ones too long for the image rows in
    // LDS: the inverse row FFT and the image term in one kernel
    if (tiles) return launch_row_ifft_image(out, tiles, width, height, K, IMG, count, scale, mm, mm_slots, st, true);
    // a few planes: the fused kernel transforms G plane rows per round (8 at
    // W = 1024), padding rows included; the epilogue kernel takes the planes
    // one after another (one plane of 1024^2: 21.0 -> 7 us)
    if (space == HV_SPACE_ANGULARSP && K <= HV_BINS_IMAGE_EPI_K && row_ifft_epi_supported(width))
        return launch_row_ifft_image(out, planes, width, height, K, IMG, count, scale, mm, mm_slots, st, false);
    if (!bins_image_fused(width, space)) {
        if (space == HV_SPACE_ANGULARSP && row_ifft_epi_supported(width))
            return launch_row_ifft_image(out, planes, width, height, K, IMG, count, scale, mm, mm_slots, st, false);
        // the planes' inverse row FFT (angular spectrum, other widths), then the image term
        if (space == HV_SPACE_ANGULARSP) HV_TRY(launch_row_pass(planes, width, height, K, +1, st));
        const size_t P = (size_t)width * height;
        const int blocks = (int)std::min<size_t>((P + 255) / 256, 2048);
        k_bins_image<IMG><<<blocks, 256, 0, st>>>(planes, out, P, K, count, mm, mm_slots, scale);
        HV_CHECK_LAUNCH("k_bins_image");
        return HV_OK;
    }
    const int rpw = HV_BINS_RPW_IMAGE;
    const dim3 grid((height + rpw - 1) / rpw);
    switch (width) {
        case 64: k_bins_row_ifft_image<64, IMG><<<grid, img_threads<64>(), 0, st>>>(planes, out, height, K, rpw, count, mm, mm_slots, scale); break;
        case 128: k_bins_row_ifft_image<128, IMG><<<grid, img_threads<128>(), 0, st>>>(planes, out, height, K, rpw, count, mm, mm_slots, scale); break;
        case 256: k_bins_row_ifft_image<256, IMG><<<grid, img_threads<256>(), 0, st>>>(planes, out, height, K, rpw, count, mm, mm_slots, scale); break;
        case 512: k_bins_row_ifft_image<512, IMG><<<grid, img_threads<512>(), 0, st>>>(planes, out, height, K, rpw, count, mm, mm_slots, scale); break;
        case 1024: k_bins_row_ifft_image<1024, IMG><<<grid, img_threads<1024>(), 0, st>>>(planes, out, height, K, rpw, count, mm, mm_slots, scale); break;
        default: return fail(HV_ERR_UNSUPPORTED, "bins-first path: unsupported width");
    }
    HV_CHECK_LAUNCH("k_bins_row_ifft_image");
    return HV_OK;
}

hv_status launch_bins_moments(float* m0, float* m1, float* m2, const float2* planes_c, int width, int height, int T,
                              int fs, int K, int space, float fps, int mm_moment, uint32_t* mm, hipStream_t st,
                              int mm_slots, float2* tiles) {
    float2* planes = const_cast<float2*>(planes_c);
    // angular spectrum: the inverse row FFT and the moments in one kernel (column-tiled or row-major planes)
    if (tiles)
        return launch_row_ifft_moments(m0, m1, m2, nullptr, tiles, width, height, T, fs, K, fps, mm_moment, mm,
                                       mm_slots, st, true);
    if (space == HV_SPACE_ANGULARSP && row_ifft_epi_supported(width))
        return launch_row_ifft_moments(m0, m1, m2, nullptr, planes, width, height, T, fs, K, fps, mm_moment, mm,
                                       mm_slots, st, false);
    if (space == HV_SPACE_ANGULARSP) HV_TRY(launch_row_pass(planes, width, height, K, +1, st));
    const size_t P = (size_t)width * height;
    const int blocks = (int)std::min<size_t>((P + 255) / 256, 2048);
    k_bins_moments<<<blocks, 256, 0, st>>>(planes, m0, m1, m2, nullptr, P, T, fs, K, fps, mm_moment, mm, mm_slots);
    HV_CHECK_LAUNCH("k_bins_moments");
    return HV_OK;
}

hv_status launch_bins_image(float* out, const float2* planes_c, int width, int height, int K, int space, int img_type,
                            int pmin, int pmax, uint32_t* mm, hipStream_t st, int mm_slots, float scale,
                            float2* tiles) {
    float2* planes = const_cast<float2*>(planes_c);
    const float count = (float)(pmax - pmin + 1);
    switch (img_type) {
        case HV_IMG_MODULUS:
            return bins_image_t<HV_IMG_MODULUS>(out, planes, width, height, K, space, count, mm, mm_slots, scale, st, tiles);
        case HV_IMG_SQUARED_MODULUS:
            return bins_image_t<HV_IMG_SQUARED_MODULUS>(out, planes, width, height, K, space, count, mm, mm_slots, scale, st,
                                                        tiles);
        case HV_IMG_ARGUMENT:
            return bins_image_t<HV_IMG_ARGUMENT>(out, planes, width, height, K, space, count, mm, mm_slots, scale, st,
                                                 tiles);
        default: return fail(HV_ERR_UNSUPPORTED, "bins-first path: image type must be modulus, squared or argument");
    }
}

}  // namespace hv
