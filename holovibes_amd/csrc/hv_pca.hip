// hv_pca.hip -- PCA time transform (fourier_transform.cc:340-377,
// matrix_operations.cc:9-102) on MFMA + FP64 Jacobi.
//
// Covariance.  cov = H^H H with H the ring seen as P x T complex.  Writing
// x = re + i im per frame and X = [Re | Im] (P x 2T real), the real Gram
// matrix X^T X holds everything:
//   Re cov[i][j] = Re_i . Re_j + Im_i . Im_j,  Im cov[i][j] = Re_i . Im_j - Im_i . Re_j.
// Frames are taken in groups of 32; for each group pair (F >= G) one wave
// accumulates two 32 x 32 MFMA tiles over its slice of pixels:
//   acc_re += Re_F^T Re_G + Im_F^T Im_G,   acc_im += Re_F^T Im_G + (-Im_F)^T Re_G.
// The A operand of group F and the B operand of group G are the same
// register fragment shape (lane l <-> frame 32F + l%32, k <-> pixels), so a
// lane loads 8 consecutive complex pixels (64 bytes) of one frame per step.
// Precisions: FP32 (v_mfma_f32_32x32x2_f32, exact products), BF16 / FP16
// (v_mfma_f32_32x32x16_{bf16,f16}), and their hi+lo splits
// (x = hi + lo, x y ~ hi hi + hi lo + lo hi, FP32 accumulation).  FP16 runs
// on the ring scaled by a power of two chosen from its max |component| so the
// values sit in FP16's normal range.  Partial tiles of every workgroup go to a
// workspace; a second kernel sums them and writes the Hermitian T x T matrix.
//
// Centring.  Every pixel's time series is dominated by its mean (the
// interferogram's constant term), so every entry of H^H H is about
// lambda_max / T, and FP32 accumulation leaves errors of ~1e-7 lambda_max on
// all of them -- far above the 1e-9 lambda_max gaps between the small
// eigenvalues that decide which eigenvectors a component range selects.  So
// a first pass takes each pixel's mean mu_p over the window, the MFMA kernels
// accumulate the centred Y = H - mu 1^T, and
//   H^H H = Y^H Y + a 1^T + 1 a^H + m 1 1^T,  a_t = sum_p conj(y_pt) mu_p,
//   m = sum_p |mu_p|^2
// is assembled in FP64 (the identity holds for any mu; the mean makes Y^H Y
// small).  Y^H Y's accumulation errors scale with the centred data, ~1e-4
// of lambda_max here, and the FP64 matrix goes to the FP64 eigensolver.
//
// Eigendecomposition.  cusolverDnCheevd (lower, ascending) is replaced by a
// cyclic Jacobi in FP64 on one workgroup with the round-robin (circle)
// ordering: T/2 disjoint rotations per round, T - 1 rounds per sweep.
//
// Projection.  C = H V restricted to the columns [pmin, pmax] the image uses,
// fused with the image type, the mean, the x131071 PCA scale
// (converts.cc:50-61) and the image min/max.
#include <cfloat>

#include "hv_fft.hpp"
#include "hv_kernels.hpp"

namespace hv {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// covariance
// ---------------------------------------------------------------------------
__host__ __device__ inline int pca_groups(int T) { return T < 32 ? 1 : T / 32; }
__host__ __device__ inline int pca_pairs(int T) {
    const int g = pca_groups(T);
    return g * (g + 1) / 2;
}
static constexpr int kCovWaves = 4;
static constexpr int kTile = 2 * 32 * 32;  // floats of one wave's two accumulators

#ifndef HV_PCA_COV_GROUPED
#define HV_PCA_COV_GROUPED 1    // T >= 64, half precision: every pair per workgroup, fragments shared in LDS
#endif
#ifndef HV_PCA_COV_WG
#define HV_PCA_COV_WG 2048      // workgroups to aim for over all pairs and slices
#endif

// Pixel slices per frame-group pair: about HV_PCA_COV_WG workgroups in all
// (several per CU: a wave waits on its next pixels' loads), each wave
// covering at least 256 pixels.
static int pca_slices_pairs(int T, size_t P) {
    const int np = pca_pairs(T);
    int ns = 1;
    // every wave's pixel range a multiple of 16 pixels (frames whose pixel count is not a power of two)
    while (np * ns * 2 <= HV_PCA_COV_WG && (size_t)(ns * 2) * kCovWaves * 256 <= P &&
           P % ((size_t)(ns * 2) * kCovWaves * 16) == 0)
        ns *= 2;
    return ns;
}

// Half-precision covariance with T >= 64 (two or more frame groups): one
// workgroup per pixel slice computes every pair (k_pca_cov_groups).
static bool pca_cov_grouped(int T, int precision) {
    return T >= 64 && precision != HV_PCA_FP32 && HV_PCA_COV_GROUPED;
}
constexpr int kCovGroupWaves = 8;
constexpr int kCovGroupSlices = 512;  // about two workgroups per CU

// Slices of the grouped kernel: at most kCovGroupSlices, at least four
// 16-pixel steps per workgroup.
static int pca_slices_grouped(size_t P) {
    int ns = 1;
    while (ns * 2 <= kCovGroupSlices && (size_t)(ns * 2) * 64 <= P && P % ((size_t)(ns * 2) * 16) == 0) ns *= 2;
    return ns;
}

static int pca_slices(int T, size_t P, int precision) {
    return pca_cov_grouped(T, precision) ? pca_slices_grouped(P) : pca_slices_pairs(T, P);
}

static size_t align_up(size_t n) { return (n + 255) & ~size_t(255); }
constexpr int kMeanBlocks = 1024;  // most workgroups of the mean pass

// Workspace: MFMA partial tiles, a_t partials (one row of T per slice), the
// pixel means, the FP64 partials of sum |mu|^2, the assembled FP64 terms
// (a_0 .. a_{T-1}, then m) and the scale words.
struct CovLayout {
    size_t partial_floats;  // slices * pairs * kTile
    size_t off_apart, off_mean, off_mpart, off_terms;
    size_t offset_scale;    // byte offset of the scale words
    size_t bytes;
};
static CovLayout cov_layout(int T, size_t P, int precision) {
    CovLayout l;
    const int slices = pca_slices(T, P, precision);
    l.partial_floats = (size_t)slices * pca_pairs(T) * kTile;
    l.off_apart = align_up(l.partial_floats * sizeof(float));
    l.off_mean = l.off_apart + align_up((size_t)slices * std::max(T, 32) * sizeof(float2));
    l.off_mpart = l.off_mean + align_up(P * sizeof(float2));
    l.off_terms = l.off_mpart + align_up(kMeanBlocks * sizeof(double));
    l.offset_scale = l.off_terms + align_up((size_t)(T + 1) * sizeof(double2));
    l.bytes = l.offset_scale + 256;
    return l;
}

static CovLayout raw_cov_layout(int T, size_t P);
// room for the ring covariance and for the raw-window covariance (bins-first PCA)
size_t pca_workspace_bytes(int T, size_t P, int precision) {
    return std::max(cov_layout(T, P, precision).bytes, raw_cov_layout(T, P).bytes);
}

__device__ __forceinline__ void pair_of(int idx, int& F, int& G) {
    // pairs enumerated (0,0), (1,0), (1,1), (2,0), ... with F >= G
    F = 0;
    while ((F + 1) * (F + 2) / 2 <= idx) ++F;
    G = idx - F * (F + 1) / 2;
}

// Per-pixel window mean mu_p (float), max |component| of the ring and of mu
// (ordered uints in mm[1], mm[3]: the FP16 scale bounds |y| <= |x| + |mu|),
// and one FP64 partial of sum_p |mu_p|^2 per workgroup.  Two pixels (one
// 16-byte load) per thread and frame, eight frames' loads in flight.
static int pca_mean_blocks(size_t P) { return (int)std::min<size_t>(kMeanBlocks, (P / 2 + 255) / 256); }

__global__ void __launch_bounds__(256)
    k_pca_mean(float2* __restrict__ mean, double* __restrict__ mpart, uint32_t* mm, const float2* __restrict__ ring,
               int T, size_t P) {
    const size_t P2 = P / 2;
    const float4* r4 = reinterpret_cast<const float4*>(ring);
    float4* m4 = reinterpret_cast<float4*>(mean);
    const float inv = 1.0f / (float)T;  // exact for power-of-two windows
    float mx = 0.f, mmu = 0.f;
    double msum = 0.0;
    auto amax = [](float4 v) { return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))); };
    for (size_t p2 = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p2 < P2; p2 += (size_t)gridDim.x * blockDim.x) {
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        int t = 0;
        for (; t + 8 <= T; t += 8) {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = r4[(size_t)(t + u) * P2 + p2];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                acc.x += v[u].x;
                acc.y += v[u].y;
                acc.z += v[u].z;
                acc.w += v[u].w;
                mx = fmaxf(mx, amax(v[u]));
            }
        }
        for (; t < T; ++t) {
            const float4 v = r4[(size_t)t * P2 + p2];
            acc.x += v.x;
            acc.y += v.y;
            acc.z += v.z;
            acc.w += v.w;
            mx = fmaxf(mx, amax(v));
        }
        acc = make_float4(acc.x * inv, acc.y * inv, acc.z * inv, acc.w * inv);
        m4[p2] = acc;
        mmu = fmaxf(mmu, amax(acc));
        msum += (double)acc.x * acc.x + (double)acc.y * acc.y + (double)acc.z * acc.z + (double)acc.w * acc.w;
    }
    block_minmax_atomic(mx, mx, mm);
    __syncthreads();  // the helper's LDS is reused by the second call
    block_minmax_atomic(mmu, mmu, mm + 2);
    __shared__ double red[256];
    red[threadIdx.x] = msum;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) mpart[blockIdx.x] = red[0];
}

__global__ void k_scale_init(uint32_t* mm) {
    if (threadIdx.x == 0) {
        mm[0] = mm[2] = float_to_ordered(INFINITY);
        mm[1] = mm[3] = float_to_ordered(0.f);
    }
}

// Power of two bringing the bound max |x| + max |mu| of the centred values
// to about 2^12 (FP16 max 65504; products of 2^12-sized values stay far from
// the FP32 accumulator's limits).
__device__ __forceinline__ float fp16_scale(const uint32_t* mm) {
    const float mx = ordered_to_float(mm[1]) + ordered_to_float(mm[3]);
    if (!(mx > 0.f) || !isfinite(mx)) return 1.0f;
    const int e = ilogbf(mx);
    return ldexpf(1.0f, 12 - e);
}

// conj(y) mu
__device__ __forceinline__ float2 conj_mul(float2 y, float2 mu) {
    return make_float2(y.x * mu.x + y.y * mu.y, y.x * mu.y - y.y * mu.x);
}

template <int PREC>
struct CovStep {
    static constexpr int K = (PREC == HV_PCA_FP32) ? 2 : 16;      // pixels per MFMA
    static constexpr int KPL = (PREC == HV_PCA_FP32) ? 1 : 8;     // pixels per lane
};

template <int PREC>
__global__ void __launch_bounds__(64 * kCovWaves)
    k_pca_cov(float* __restrict__ partial, float2* __restrict__ apart, const float2* __restrict__ ring,
              const float2* __restrict__ mean, int T, size_t P, int slices, const uint32_t* __restrict__ scale_mm) {
    using S = CovStep<PREC>;
    // Workgroups go round-robin over the 8 XCDs (linear id mod 8), each with
    // its own L2.  With slices a multiple of 8, the pairs of one slice run as
    // consecutive workgroups of one XCD and read the same frame groups at
    // about the same time, so that XCD's L2 fetches them once.
    const int npairs = pca_pairs(T);
    int pair, slice;
    if (slices % 8 == 0) {
        const int idx = blockIdx.x >> 3;
        pair = idx % npairs;
        slice = (idx / npairs) * 8 + (blockIdx.x & 7);
    } else {
        pair = blockIdx.x % npairs;
        slice = blockIdx.x / npairs;
    }
    int F, G;
    pair_of(pair, F, G);
    const bool diag = F == G;  // diagonal pairs also accumulate a_t for their group's frames
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t per_wave = P / ((size_t)slices * kCovWaves);
    const size_t p_begin = ((size_t)slice * kCovWaves + wave) * per_wave;
    const size_t p_end = p_begin + per_wave;
    const int fF = 32 * F + (lane & 31), fG = 32 * G + (lane & 31);
    const bool okF = fF < T, okG = fG < T;
    const float2* rowF = ring + (size_t)(okF ? fF : 0) * P;
    const float2* rowG = ring + (size_t)(okG ? fG : 0) * P;
    const int koff = (lane >> 5) * S::KPL;
    float scale = 1.0f;
    if constexpr (PREC == HV_PCA_FP16 || PREC == HV_PCA_FP16_SPLIT) scale = fp16_scale(scale_mm);
    // centred value of a frame that exists (frames past T stay zero)
    auto centre = [](float2 x, float2 mu, bool ok) { return ok ? make_float2(x.x - mu.x, x.y - mu.y) : make_float2(0.f, 0.f); };
    float2 aacc = make_float2(0.f, 0.f);

    v16f acc_re = {}, acc_im = {};
    if constexpr (PREC == HV_PCA_FP32) {
        // 4 MFMA steps (2 pixels each) per iteration, the next iteration's
        // pixels loaded before this one's products
        constexpr int U = 4;
        float2 na[U], nb[U], nm[U];
        auto load = [&](size_t p0) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                na[u] = okF ? rowF[p0 + u * S::K + koff] : make_float2(0.f, 0.f);
                nb[u] = okG ? rowG[p0 + u * S::K + koff] : make_float2(0.f, 0.f);
                nm[u] = mean[p0 + u * S::K + koff];
            }
        };
        load(p_begin);
        for (size_t p0 = p_begin; p0 < p_end; p0 += U * S::K) {
            float2 ca[U], cb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ca[u] = centre(na[u], nm[u], okF);
                cb[u] = centre(nb[u], nm[u], okG);
                if (diag) {
                    const float2 t = conj_mul(ca[u], nm[u]);
                    aacc.x += t.x;
                    aacc.y += t.y;
                }
            }
            if (p0 + U * S::K < p_end) load(p0 + U * S::K);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float2 a = ca[u], b = cb[u];
                acc_re = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b.x, acc_re, 0, 0, 0);
                acc_re = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b.y, acc_re, 0, 0, 0);
                acc_im = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b.y, acc_im, 0, 0, 0);
                acc_im = __builtin_amdgcn_mfma_f32_32x32x2f32(-a.y, b.x, acc_im, 0, 0, 0);
            }
        }
    } else {
        // the next 16 pixels' loads are issued before this step's conversions and products
        float4 na[4], nb[4], nm[4];
        auto load = [&](size_t p0) {
            const float4* pa = reinterpret_cast<const float4*>(rowF + p0 + koff);
            const float4* pb = reinterpret_cast<const float4*>(rowG + p0 + koff);
            const float4* pm = reinterpret_cast<const float4*>(mean + p0 + koff);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                na[q] = okF ? pa[q] : make_float4(0.f, 0.f, 0.f, 0.f);
                nb[q] = okG ? pb[q] : make_float4(0.f, 0.f, 0.f, 0.f);
                nm[q] = pm[q];
            }
        };
        load(p_begin);
        for (size_t p0 = p_begin; p0 < p_end; p0 += S::K) {
            float2 xa[8], xb[8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 va = na[q], vb = nb[q], vm = nm[q];
                const float2 m0 = make_float2(vm.x, vm.y), m1 = make_float2(vm.z, vm.w);
                const float2 ya0 = centre(make_float2(va.x, va.y), m0, okF), ya1 = centre(make_float2(va.z, va.w), m1, okF);
                const float2 yb0 = centre(make_float2(vb.x, vb.y), m0, okG), yb1 = centre(make_float2(vb.z, vb.w), m1, okG);
                if (diag) {
                    const float2 t0 = conj_mul(ya0, m0), t1 = conj_mul(ya1, m1);
                    aacc.x += t0.x + t1.x;
                    aacc.y += t0.y + t1.y;
                }
                xa[2 * q] = make_float2(ya0.x * scale, ya0.y * scale);
                xa[2 * q + 1] = make_float2(ya1.x * scale, ya1.y * scale);
                xb[2 * q] = make_float2(yb0.x * scale, yb0.y * scale);
                xb[2 * q + 1] = make_float2(yb1.x * scale, yb1.y * scale);
            }
            if (p0 + S::K < p_end) load(p0 + S::K);
            constexpr bool split = (PREC == HV_PCA_BF16_SPLIT || PREC == HV_PCA_FP16_SPLIT);
            constexpr bool bf = (PREC == HV_PCA_BF16_SPLIT || PREC == HV_PCA_BF16);
            using V8 = typename std::conditional<bf, v8bf, v8h>::type;
            using E = typename std::conditional<bf, __bf16, _Float16>::type;
            V8 ar_hi, ai_hi, ain_hi, br_hi, bi_hi, ar_lo, ai_lo, ain_lo, br_lo, bi_lo;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const E arh = (E)xa[e].x, aih = (E)xa[e].y, brh = (E)xb[e].x, bih = (E)xb[e].y;
                ar_hi[e] = arh;
                ai_hi[e] = aih;
                ain_hi[e] = -aih;
                br_hi[e] = brh;
                bi_hi[e] = bih;
                if constexpr (split) {
                    const E arl = (E)(xa[e].x - (float)arh), ail = (E)(xa[e].y - (float)aih);
                    ar_lo[e] = arl;
                    ai_lo[e] = ail;
                    ain_lo[e] = -ail;
                    br_lo[e] = (E)(xb[e].x - (float)brh);
                    bi_lo[e] = (E)(xb[e].y - (float)bih);
                }
            }
            auto mfma = [](V8 a, V8 b, v16f c) {
                if constexpr (bf) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
                else return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
            };
            if constexpr (split) {
                // small cross terms first, then hi*hi
                acc_re = mfma(ar_lo, br_hi, acc_re);
                acc_re = mfma(ar_hi, br_lo, acc_re);
                acc_re = mfma(ai_lo, bi_hi, acc_re);
                acc_re = mfma(ai_hi, bi_lo, acc_re);
                acc_im = mfma(ar_lo, bi_hi, acc_im);
                acc_im = mfma(ar_hi, bi_lo, acc_im);
                acc_im = mfma(ain_lo, br_hi, acc_im);
                acc_im = mfma(ain_hi, br_lo, acc_im);
            }
            acc_re = mfma(ar_hi, br_hi, acc_re);
            acc_re = mfma(ai_hi, bi_hi, acc_re);
            acc_im = mfma(ar_hi, bi_hi, acc_im);
            acc_im = mfma(ain_hi, br_hi, acc_im);
        }
    }
    // reduce the workgroup's waves in LDS, one partial tile per workgroup
    __shared__ float red[kCovWaves][kTile];
    __shared__ float2 ared[kCovWaves][64];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        red[wave][r * 64 + lane] = acc_re[r];
        red[wave][1024 + r * 64 + lane] = acc_im[r];
    }
    ared[wave][lane] = aacc;
    __syncthreads();
    float* out = partial + ((size_t)slice * npairs + pair) * kTile;
    for (int i = threadIdx.x; i < kTile; i += blockDim.x) {
        float s = 0.f;
#pragma unroll
        for (int w = 0; w < kCovWaves; ++w) s += red[w][i];
        out[i] = s;
    }
    // a_t of group F's frames over this slice: lanes l and l + 32 of every wave (a fixed order)
    if (diag && threadIdx.x < 32 && 32 * F + (int)threadIdx.x < T) {
        float2 a = make_float2(0.f, 0.f);
#pragma unroll
        for (int w = 0; w < kCovWaves; ++w) {
            a.x += ared[w][threadIdx.x].x + ared[w][threadIdx.x + 32].x;
            a.y += ared[w][threadIdx.x].y + ared[w][threadIdx.x + 32].y;
        }
        apart[(size_t)slice * T + 32 * F + threadIdx.x] = a;
    }
}

// T = 32 G frames (G >= 2), half precision.  Per 16-pixel step, wave g < G
// loads group g's 32 frames x 16 pixels (lane l: frame 32 g + l % 32,
// pixels 8 (l / 32) .. + 7, the MFMA operand layout of both A and B), converts
// them once to the operand fragments -- real, imaginary and negated imaginary
// parts, hi and (split) lo -- and stores them in LDS (double-buffered: one
// barrier per step); then every wave runs the products of its pairs
// (w, w + 8, ...) from the fragments.  The ring is read once; the
// pair-per-workgroup kernel reads each group once per pair it belongs to.
template <int PREC, int G>
__global__ void __launch_bounds__(64 * kCovGroupWaves)
    k_pca_cov_groups(float* __restrict__ partial, float2* __restrict__ apart, const float2* __restrict__ ring,
                     const float2* __restrict__ mean, size_t P, int slices, const uint32_t* __restrict__ scale_mm) {
    constexpr int NP = G * (G + 1) / 2;
    constexpr int W = kCovGroupWaves;
    constexpr int PPW = (NP + W - 1) / W;
    constexpr bool split = (PREC == HV_PCA_BF16_SPLIT || PREC == HV_PCA_FP16_SPLIT);
    constexpr bool bf = (PREC == HV_PCA_BF16_SPLIT || PREC == HV_PCA_BF16);
    using V8 = typename std::conditional<bf, v8bf, v8h>::type;
    using E = typename std::conditional<bf, __bf16, _Float16>::type;
    constexpr int NF = split ? 6 : 3;  // re, im, -im (hi), then the same (lo)
    __shared__ V8 frag[2][G][NF][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t per_slice = P / slices;
    const size_t p_begin = (size_t)blockIdx.x * per_slice, p_end = p_begin + per_slice;
    float scale = 1.0f;
    if constexpr (PREC == HV_PCA_FP16 || PREC == HV_PCA_FP16_SPLIT) scale = fp16_scale(scale_mm);
    // converting waves: group `wave`, frame 32 wave + lane % 32, pixel offset 8 (lane / 32)
    const bool conv = wave < G;
    const float2* src = ring + (size_t)(32 * (conv ? wave : 0) + (lane & 31)) * P + (lane >> 5) * 8;
    const float2* msrc = mean + (lane >> 5) * 8;
    float4 nx[4], nm[4];
    float2 aacc = make_float2(0.f, 0.f);  // a_t of this lane's frame over its pixels
    auto load = [&](size_t p0) {
        const float4* q = reinterpret_cast<const float4*>(src + p0);
        const float4* qm = reinterpret_cast<const float4*>(msrc + p0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            nx[k] = q[k];
            nm[k] = qm[k];
        }
    };
    if (conv) load(p_begin);
    v16f acc_re[PPW], acc_im[PPW];
#pragma unroll
    for (int j = 0; j < PPW; ++j) acc_re[j] = acc_im[j] = v16f{};
    int buf = 0;
    for (size_t p0 = p_begin; p0 < p_end; p0 += 16, buf ^= 1) {
        if (conv) {
            float2 x[8];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float2 m0 = make_float2(nm[k].x, nm[k].y), m1 = make_float2(nm[k].z, nm[k].w);
                const float2 y0 = make_float2(nx[k].x - m0.x, nx[k].y - m0.y);
                const float2 y1 = make_float2(nx[k].z - m1.x, nx[k].w - m1.y);
                const float2 t0 = conj_mul(y0, m0), t1 = conj_mul(y1, m1);
                aacc.x += t0.x + t1.x;
                aacc.y += t0.y + t1.y;
                x[2 * k] = make_float2(y0.x * scale, y0.y * scale);
                x[2 * k + 1] = make_float2(y1.x * scale, y1.y * scale);
            }
            if (p0 + 16 < p_end) load(p0 + 16);
            V8 rh, ih, nh, rl, il, nl;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const E r = (E)x[e].x, i = (E)x[e].y;
                rh[e] = r;
                ih[e] = i;
                nh[e] = -i;
                if constexpr (split) {
                    const E rlo = (E)(x[e].x - (float)r), ilo = (E)(x[e].y - (float)i);
                    rl[e] = rlo;
                    il[e] = ilo;
                    nl[e] = -ilo;
                }
            }
            frag[buf][wave][0][lane] = rh;
            frag[buf][wave][1][lane] = ih;
            frag[buf][wave][2][lane] = nh;
            if constexpr (split) {
                frag[buf][wave][3][lane] = rl;
                frag[buf][wave][4][lane] = il;
                frag[buf][wave][5][lane] = nl;
            }
        }
        __syncthreads();
        auto mfma = [](V8 a, V8 b, v16f c) {
            if constexpr (bf) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
            else return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
        };
#pragma unroll
        for (int j = 0; j < PPW; ++j) {
            const int pair = wave + j * W;
            if (pair >= NP) break;
            int F, Gp;
            pair_of(pair, F, Gp);
            const V8 ar = frag[buf][F][0][lane], ai = frag[buf][F][1][lane], an = frag[buf][F][2][lane];
            const V8 br = frag[buf][Gp][0][lane], bi = frag[buf][Gp][1][lane];
            if constexpr (split) {
                const V8 arl = frag[buf][F][3][lane], ail = frag[buf][F][4][lane], anl = frag[buf][F][5][lane];
                const V8 brl = frag[buf][Gp][3][lane], bil = frag[buf][Gp][4][lane];
                // small cross terms first, then hi*hi
                acc_re[j] = mfma(arl, br, acc_re[j]);
                acc_re[j] = mfma(ar, brl, acc_re[j]);
                acc_re[j] = mfma(ail, bi, acc_re[j]);
                acc_re[j] = mfma(ai, bil, acc_re[j]);
                acc_im[j] = mfma(arl, bi, acc_im[j]);
                acc_im[j] = mfma(ar, bil, acc_im[j]);
                acc_im[j] = mfma(anl, br, acc_im[j]);
                acc_im[j] = mfma(an, brl, acc_im[j]);
            }
            acc_re[j] = mfma(ar, br, acc_re[j]);
            acc_re[j] = mfma(ai, bi, acc_re[j]);
            acc_im[j] = mfma(ar, bi, acc_im[j]);
            acc_im[j] = mfma(an, br, acc_im[j]);
        }
    }
    // a_t of the converting wave's 32 frames: lanes l and l + 32 hold the same frame
    if (conv) {
        const float ax = aacc.x + __shfl_xor(aacc.x, 32), ay = aacc.y + __shfl_xor(aacc.y, 32);
        if (lane < 32) apart[(size_t)blockIdx.x * (32 * G) + 32 * wave + lane] = make_float2(ax, ay);
    }
    // each pair's tile straight to its partial slot (one wave per pair: no cross-wave sum)
#pragma unroll
    for (int j = 0; j < PPW; ++j) {
        const int pair = wave + j * W;
        if (pair >= NP) break;
        float* out = partial + ((size_t)blockIdx.x * NP + pair) * kTile;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            out[r * 64 + lane] = acc_re[j][r];
            out[1024 + r * 64 + lane] = acc_im[j][r];
        }
    }
}

template <int PREC>
static hv_status cov_groups(float* partial, float2* apart, const float2* ring, const float2* mean, int T, size_t P,
                            int slices, const uint32_t* scale_mm, hipStream_t st) {
    constexpr int threads = 64 * kCovGroupWaves;
    switch (T) {
        case 64: k_pca_cov_groups<PREC, 2><<<slices, threads, 0, st>>>(partial, apart, ring, mean, P, slices, scale_mm); break;
        case 128: k_pca_cov_groups<PREC, 4><<<slices, threads, 0, st>>>(partial, apart, ring, mean, P, slices, scale_mm); break;
        case 256: k_pca_cov_groups<PREC, 8><<<slices, threads, 0, st>>>(partial, apart, ring, mean, P, slices, scale_mm); break;
        default: return fail(HV_ERR_UNSUPPORTED, "grouped covariance needs T in {64, 128, 256}");
    }
    HV_CHECK_LAUNCH("k_pca_cov_groups");
    return HV_OK;
}

// a_t = sum of its slice partials and m = sum of the mean pass's partials,
// in FP64 and a fixed order: one workgroup per frame (workgroup 0 also m).
__global__ void __launch_bounds__(256)
    k_pca_mean_terms(double2* __restrict__ terms, const float2* __restrict__ apart, int slices, int T,
                     const double* __restrict__ mpart, int mblocks) {
    const int t = blockIdx.x;
    __shared__ double rr[256], ri[256];
    double re = 0.0, im = 0.0;
    for (int sl = threadIdx.x; sl < slices; sl += 256) {
        const float2 a = apart[(size_t)sl * T + t];
        re += a.x;
        im += a.y;
    }
    rr[threadIdx.x] = re;
    ri[threadIdx.x] = im;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            rr[threadIdx.x] += rr[threadIdx.x + o];
            ri[threadIdx.x] += ri[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) terms[t] = make_double2(rr[0], ri[0]);
    if (t != 0) return;
    __syncthreads();
    double m = 0.0;
    for (int b = threadIdx.x; b < mblocks; b += 256) m += mpart[b];
    rr[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) rr[threadIdx.x] += rr[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) terms[T] = make_double2(rr[0], 0.0);
}

// Sum the slices in FP64, undo the FP16 scale (a power of two), add the mean
// terms -- cov(i, j) = (Y^H Y)(i, j) + a_i + conj(a_j) + m -- and write the
// full Hermitian matrix, column-major, as double2 (cov64) and / or float2
// (cov32).  Tile element (acc, r, lane) is row 8 (r / 4) + 4 (lane / 32) + r % 4
// and column lane % 32 of the 32 x 32 MFMA output.  Each workgroup owns 64
// elements of one pair's tile; its four waves sum every fourth slice, then
// one wave adds the four (a fixed order: deterministic).
__global__ void __launch_bounds__(256)
    k_pca_cov_reduce(double2* __restrict__ cov64, float2* __restrict__ cov32, const float* __restrict__ partial,
                     const double2* __restrict__ terms, int T, int slices, int npairs,
                     const uint32_t* __restrict__ scale_mm, int use_scale) {
    const int pair = blockIdx.x;
    int F, G;
    pair_of(pair, F, G);
    const int g = threadIdx.x >> 6, el = threadIdx.x & 63;
    const int e = blockIdx.y * 64 + el;  // 0 .. 1023
    double re = 0.0, im = 0.0;
    // unrolled so that several slices' loads are in flight (same summation order)
#pragma unroll 8
    for (int sl = g; sl < slices; sl += 4) {
        const float* t = partial + ((size_t)sl * npairs + pair) * kTile;
        re += (double)t[e];
        im += (double)t[1024 + e];
    }
    __shared__ double sr[4][64], si[4][64];
    sr[g][el] = re;
    si[g][el] = im;
    __syncthreads();
    if (g != 0) return;
    re = (sr[0][el] + sr[1][el]) + (sr[2][el] + sr[3][el]);
    im = (si[0][el] + si[1][el]) + (si[2][el] + si[3][el]);
    if (use_scale) {
        const double sc = (double)fp16_scale(scale_mm);
        const double inv = 1.0 / (sc * sc);  // exact: a power of two
        re *= inv;
        im *= inv;
    }
    const int r = e / 64, lane = e % 64;
    const int i = 8 * (r / 4) + 4 * (lane / 32) + (r % 4), jj = lane % 32;
    const int fi = 32 * F + i, fj = 32 * G + jj;
    if (fi >= T || fj >= T) return;
    const double2 ai = terms[fi], aj = terms[fj];
    const double m = terms[T].x;
    re += ai.x + aj.x + m;
    im += ai.y - aj.y;
    if (cov64) {
        cov64[fi + (size_t)fj * T] = make_double2(re, im);
        if (F != G) cov64[fj + (size_t)fi * T] = make_double2(re, -im);
    }
    if (cov32) {
        cov32[fi + (size_t)fj * T] = make_float2((float)re, (float)im);
        if (F != G) cov32[fj + (size_t)fi * T] = make_float2((float)re, (float)-im);
    }
}

hv_status launch_pca_cov(void* cov, int cov_double, const float2* ring, int T, size_t P, int precision, void* ws,
                         hipStream_t st) {
    // frame groups of 32: windows that are not powers of two fit one group
    if (!is_pow2(T) && T > 32) return fail(HV_ERR_UNSUPPORTED, "PCA covariance: T must be a power of two or below 32");
    const CovLayout l = cov_layout(T, P, precision);
    const int slices = pca_slices(T, P, precision), np = pca_pairs(T);
    const bool grouped = pca_cov_grouped(T, precision);
    char* w = reinterpret_cast<char*>(ws);
    float* partial = reinterpret_cast<float*>(w);
    float2* apart = reinterpret_cast<float2*>(w + l.off_apart);
    float2* mean = reinterpret_cast<float2*>(w + l.off_mean);
    double* mpart = reinterpret_cast<double*>(w + l.off_mpart);
    double2* terms = reinterpret_cast<double2*>(w + l.off_terms);
    uint32_t* scale_mm = reinterpret_cast<uint32_t*>(w + l.offset_scale);
    const bool fp16 = precision == HV_PCA_FP16 || precision == HV_PCA_FP16_SPLIT;
    // every wave's pixel range: a multiple of the 8 pixels per FP32 iteration and the 16 per half-precision step
    if (P % ((size_t)slices * (grouped ? 1 : kCovWaves) * 16))
        return fail(HV_ERR_UNSUPPORTED, "frame_res not a multiple of the PCA tile");
    if (reinterpret_cast<uintptr_t>(ring) & 15) return fail(HV_ERR_INVALID_ARG, "the ring must be 16-byte aligned");
    switch (precision) {
        case HV_PCA_FP32: case HV_PCA_BF16_SPLIT: case HV_PCA_FP16_SPLIT: case HV_PCA_BF16: case HV_PCA_FP16: break;
        default: return fail(HV_ERR_INVALID_ARG, "unknown PCA precision");
    }
    // window means, max |x|, max |mu| and the partials of m
    k_scale_init<<<1, 64, 0, st>>>(scale_mm);
    HV_CHECK_LAUNCH("k_scale_init");
    const int mblocks = pca_mean_blocks(P);
    k_pca_mean<<<mblocks, 256, 0, st>>>(mean, mpart, scale_mm, ring, T, P);
    HV_CHECK_LAUNCH("k_pca_mean");
    if (grouped) {
        switch (precision) {
            case HV_PCA_BF16_SPLIT: HV_TRY(cov_groups<HV_PCA_BF16_SPLIT>(partial, apart, ring, mean, T, P, slices, scale_mm, st)); break;
            case HV_PCA_FP16_SPLIT: HV_TRY(cov_groups<HV_PCA_FP16_SPLIT>(partial, apart, ring, mean, T, P, slices, scale_mm, st)); break;
            case HV_PCA_BF16: HV_TRY(cov_groups<HV_PCA_BF16>(partial, apart, ring, mean, T, P, slices, scale_mm, st)); break;
            default: HV_TRY(cov_groups<HV_PCA_FP16>(partial, apart, ring, mean, T, P, slices, scale_mm, st)); break;
        }
    } else {
        const unsigned grid = (unsigned)(np * slices);
        const dim3 b(64 * kCovWaves);
        switch (precision) {
            case HV_PCA_FP32: k_pca_cov<HV_PCA_FP32><<<grid, b, 0, st>>>(partial, apart, ring, mean, T, P, slices, scale_mm); break;
            case HV_PCA_BF16_SPLIT: k_pca_cov<HV_PCA_BF16_SPLIT><<<grid, b, 0, st>>>(partial, apart, ring, mean, T, P, slices, scale_mm); break;
            case HV_PCA_FP16_SPLIT: k_pca_cov<HV_PCA_FP16_SPLIT><<<grid, b, 0, st>>>(partial, apart, ring, mean, T, P, slices, scale_mm); break;
            case HV_PCA_BF16: k_pca_cov<HV_PCA_BF16><<<grid, b, 0, st>>>(partial, apart, ring, mean, T, P, slices, scale_mm); break;
            default: k_pca_cov<HV_PCA_FP16><<<grid, b, 0, st>>>(partial, apart, ring, mean, T, P, slices, scale_mm); break;
        }
        HV_CHECK_LAUNCH("k_pca_cov");
    }
    k_pca_mean_terms<<<T, 256, 0, st>>>(terms, apart, slices, T, mpart, mblocks);
    HV_CHECK_LAUNCH("k_pca_mean_terms");
    k_pca_cov_reduce<<<dim3(np, 1024 / 64), 256, 0, st>>>(cov_double ? (double2*)cov : nullptr,
                                                           cov_double ? nullptr : (float2*)cov, partial, terms, T,
                                                           slices, np, scale_mm, fp16 ? 1 : 0);
    HV_CHECK_LAUNCH("k_pca_cov_reduce");
    return HV_OK;
}

// ---------------------------------------------------------------------------
// covariance of the raw window (bins-first PCA)
// ---------------------------------------------------------------------------
// Both propagators are unitary up to a constant: the angular spectrum is
// L = F^H diag(H) F / P with |H| = 1 everywhere (L^H L = I), Fresnel is
// L = F diag(c) with a unit-modulus chirp (L^H L = P I).  The propagated
// window X = L H (frames as columns) therefore has X^H X = alpha H^T H: the
// covariance the reference takes of the ring (matrix_operations.cc:9-37) is
// the covariance of the converted raw window, a real symmetric matrix, up to
// a factor that changes no eigenvector and no eigenvalue order.  It is taken
// here straight from the raw u8 / u16 / f32 frames, in physical ring-slot
// order (slot n holds frame (n - slot0) mod T of the batch), centred as in
// launch_pca_cov: real products, a quarter of the complex MFMA work, and
// d bytes per value read instead of 8.

// N consecutive pixels of a raw frame, loaded as whole words (16-byte aligned
// buffers) or byte by byte (a caller's buffer that is not), converted on use.
template <typename TIn, int N>
struct RawPx {
    static constexpr int kBytes = N * (int)sizeof(TIn);
    static constexpr int kWords = kBytes / 4;
    uint32_t w[kWords];
    __device__ __forceinline__ void load(const TIn* src, int vec) {
        if (vec) {
            if constexpr (kBytes == 4) {
                w[0] = *reinterpret_cast<const uint32_t*>(src);
            } else if constexpr (kBytes == 8) {
                const uint2 v = *reinterpret_cast<const uint2*>(src);
                w[0] = v.x;
                w[1] = v.y;
            } else {
#pragma unroll
                for (int i = 0; i < kBytes / 16; ++i) {
                    const uint4 v = reinterpret_cast<const uint4*>(src)[i];
                    w[4 * i] = v.x;
                    w[4 * i + 1] = v.y;
                    w[4 * i + 2] = v.z;
                    w[4 * i + 3] = v.w;
                }
            }
        } else {
            const uint8_t* b = reinterpret_cast<const uint8_t*>(src);
#pragma unroll
            for (int i = 0; i < kWords; ++i)
                w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
                       ((uint32_t)b[4 * i + 3] << 24);
        }
    }
    __device__ __forceinline__ float get(int e) const {
        if constexpr (sizeof(TIn) == 1) return convert_px<uint8_t>((uint8_t)(w[e >> 2] >> (8 * (e & 3))));
        else if constexpr (sizeof(TIn) == 2) return convert_px<uint16_t>((uint16_t)(w[e >> 1] >> (16 * (e & 1))));
        else return __uint_as_float(w[e]);
    }
};

// Per-pixel window mean (float: integer frames sum exactly) and FP64 partials
// of m = sum mu^2 (one per workgroup), and -- for the FP16 scale -- max |x| and
// max |mu|.  A workgroup takes 256 pixels at a time: thread (g, q) sums frames
// g, g + 4, ... of pixels 4q .. 4q + 3 (one word per frame, the frames' loads
// independent), and the four frame groups are added in LDS in a fixed order.
constexpr int kRawMeanThreads = 256;
template <typename TIn>
__global__ void __launch_bounds__(kRawMeanThreads)
    k_raw_mean(float* __restrict__ mean, double* __restrict__ mpart, uint32_t* mm, const TIn* __restrict__ raw,
               int T, size_t P, int vec) {
    constexpr int kGroups = 4, kQuads = kRawMeanThreads / kGroups;  // 64 quads = 256 pixels
    const int q = threadIdx.x % kQuads, g = threadIdx.x / kQuads;
    const float inv = 1.0f / (float)T;  // exact for power-of-two windows
    __shared__ float4 part[kGroups][kQuads];
    float mx = 0.f, mmu = 0.f;
    double msum = 0.0;
    for (size_t c0 = (size_t)blockIdx.x * 256; c0 < P; c0 += (size_t)gridDim.x * 256) {
        const size_t p = c0 + 4 * q;
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
        for (int t = g; t < T; t += kGroups) {
            RawPx<TIn, 4> v;
            v.load(raw + (size_t)t * P + p, vec);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float x = v.get(e);
                acc[e] += x;
                if (mm) mx = fmaxf(mx, fabsf(x));
            }
        }
        __syncthreads();  // the previous chunk's partial sums have been read
        part[g][q] = make_float4(acc[0], acc[1], acc[2], acc[3]);
        __syncthreads();
        if (g == 0) {
            float4 s = part[0][q];
#pragma unroll
            for (int j = 1; j < kGroups; ++j) {
                const float4 u = part[j][q];
                s.x += u.x;
                s.y += u.y;
                s.z += u.z;
                s.w += u.w;
            }
            const float4 m = make_float4(s.x * inv, s.y * inv, s.z * inv, s.w * inv);
            reinterpret_cast<float4*>(mean)[p / 4] = m;
            mmu = fmaxf(mmu, fmaxf(fmaxf(fabsf(m.x), fabsf(m.y)), fmaxf(fabsf(m.z), fabsf(m.w))));
            msum += (double)m.x * m.x + (double)m.y * m.y + (double)m.z * m.z + (double)m.w * m.w;
        }
    }
    if (mm) {  // uniform: a kernel argument
        block_minmax_atomic(mx, mx, mm);
        __syncthreads();  // the helper's LDS is reused by the second call
        block_minmax_atomic(mmu, mmu, mm + 2);
    }
    __shared__ double red[kRawMeanThreads];
    red[threadIdx.x] = msum;
    __syncthreads();
    for (int o = kRawMeanThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) mpart[blockIdx.x] = red[0];
}

// Slice sums of the raw-window tiles in FP64 with the mean terms computed in
// place (k_pca_mean_terms + k_pca_cov_reduce in one kernel): a workgroup owns
// 64 elements of one pair's tile -- two rows i0, i0 + 4 and 32 columns -- and
// its 16 waves sum, each over every 16th slice, the tile elements and the a_t
// of those 34 frames, and m over the mean pass's partials; the waves' sums are
// added in a fixed order.  cov(i, j) = (Y^T Y)(i, j) + a_i + a_j + m.
constexpr int kRawReduceWaves = 16;
__global__ void __launch_bounds__(64 * kRawReduceWaves)
    k_raw_cov_reduce(double2* __restrict__ cov64, const float* __restrict__ partial, const float2* __restrict__ apart,
                     const double* __restrict__ mpart, int mblocks, int T, int slices, int npairs,
                     const uint32_t* __restrict__ scale_mm, int use_scale) {
    constexpr int NW = kRawReduceWaves;
    const int pair = blockIdx.x;
    int F, G;
    pair_of(pair, F, G);
    const int g = threadIdx.x >> 6, el = threadIdx.x & 63;
    const int r = blockIdx.y;  // MFMA output register: rows 8 (r / 4) + r % 4 (+ 4), all 32 columns
    const int e = r * 64 + el;
    const int i0 = 8 * (r / 4) + (r % 4);
    const int f = el < 32 ? 32 * G + el : 32 * F + i0 + 4 * (el - 32);
    double re = 0.0;
#pragma unroll 8
    for (int sl = g; sl < slices; sl += NW) re += (double)partial[((size_t)sl * npairs + pair) * kTile + e];
    double a = 0.0;
    if (el < 34 && f < T) {
#pragma unroll 8
        for (int sl = g; sl < slices; sl += NW) a += (double)apart[(size_t)sl * T + f].x;
    }
    double m = 0.0;
    for (int b = (int)threadIdx.x; b < mblocks; b += 64 * NW) m += mpart[b];
    for (int o = 32; o >= 1; o >>= 1) m += __shfl_xor(m, o, 64);
    __shared__ double sr[NW][64];
    __shared__ double ap[NW][34];
    __shared__ double mp[NW];
    sr[g][el] = re;
    if (el < 34) ap[g][el] = a;
    if (el == 0) mp[g] = m;
    __syncthreads();
    if (g != 0) return;
    re = 0.0;
    double msh = 0.0, ar = 0.0, ac = 0.0;
    const int half = el / 32, jj = el % 32;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        re += sr[w][el];
        msh += mp[w];
        ar += ap[w][32 + half];
        ac += ap[w][jj];
    }
    if (use_scale) {
        const double sc = (double)fp16_scale(scale_mm);
        re *= 1.0 / (sc * sc);  // exact: a power of two
    }
    const int fi = 32 * F + i0 + 4 * half, fj = 32 * G + jj;
    if (fi >= T || fj >= T) return;
    re += ar + ac + msh;
    cov64[fi + (size_t)fj * T] = make_double2(re, 0.0);
    if (F != G) cov64[fj + (size_t)fi * T] = make_double2(re, 0.0);
}

// The MFMA fragments of 8 centred values (pixels 8 (lane / 32) .. + 7 of a
// 16-pixel step, the operand layout of both A and B): FP32 values for the
// 32x32x2 f32 steps (value e in step e), or BF16 / FP16 hi and (split) lo.
template <int PREC>
struct RawFrag {
    static constexpr bool fp32 = PREC == HV_PCA_FP32;
    static constexpr bool split = (PREC == HV_PCA_BF16_SPLIT || PREC == HV_PCA_FP16_SPLIT);
    static constexpr bool bf = (PREC == HV_PCA_BF16_SPLIT || PREC == HV_PCA_BF16);
    using V8 = typename std::conditional<bf, v8bf, v8h>::type;
    using E = typename std::conditional<bf, __bf16, _Float16>::type;
};

// acc += A^T B over 16 pixels, A and B the fragments of two frame groups
// (the same registers for a diagonal pair).
template <int PREC>
__device__ __forceinline__ v16f raw_products(const float (&a)[8], const float (&b)[8], v16f acc) {
    using F = RawFrag<PREC>;
    if constexpr (F::fp32) {
#pragma unroll
        for (int e = 0; e < 8; ++e) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[e], b[e], acc, 0, 0, 0);
        return acc;
    } else {
        typename F::V8 ah, bh, al, bl;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const typename F::E xa = (typename F::E)a[e], xb = (typename F::E)b[e];
            ah[e] = xa;
            bh[e] = xb;
            if constexpr (F::split) {
                al[e] = (typename F::E)(a[e] - (float)xa);
                bl[e] = (typename F::E)(b[e] - (float)xb);
            }
        }
        auto mfma = [](typename F::V8 x, typename F::V8 y, v16f c) {
            if constexpr (F::bf) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, c, 0, 0, 0);
            else return __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y, c, 0, 0, 0);
        };
        if constexpr (F::split) {  // small cross terms first, then hi * hi
            acc = mfma(al, bh, acc);
            acc = mfma(ah, bl, acc);
        }
        return mfma(ah, bh, acc);
    }
}

// T <= 32 (one frame group): every wave streams its own pixel range, lane l
// holding frame slot l % 32 (zero past T); the four waves' tiles are summed
// in LDS into one partial tile per workgroup.
template <int PREC, typename TIn>
__global__ void __launch_bounds__(64 * kCovWaves)
    k_raw_cov(float* __restrict__ partial, float2* __restrict__ apart, const TIn* __restrict__ raw,
              const float* __restrict__ mean, int T, size_t P, int slices, int slot0, int vec,
              const uint32_t* __restrict__ scale_mm) {
    const int slice = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t per_wave = P / ((size_t)slices * kCovWaves);
    const size_t p_begin = ((size_t)slice * kCovWaves + wave) * per_wave, p_end = p_begin + per_wave;
    const int f = lane & 31;
    const bool ok = f < T;
    const TIn* row = raw + (size_t)(ok ? ((f - slot0) & (T - 1)) : 0) * P + (lane >> 5) * 8;
    const float* mrow = mean + (lane >> 5) * 8;
    float scale = 1.0f;
    if constexpr (PREC == HV_PCA_FP16 || PREC == HV_PCA_FP16_SPLIT) scale = fp16_scale(scale_mm);
    RawPx<TIn, 8> nx;
    float4 nm0, nm1;
    auto load = [&](size_t p0) {
        if (ok) nx.load(row + p0, vec);
        nm0 = *reinterpret_cast<const float4*>(mrow + p0);
        nm1 = *reinterpret_cast<const float4*>(mrow + p0 + 4);
    };
    load(p_begin);
    float aacc = 0.f;
    v16f acc = {};
    for (size_t p0 = p_begin; p0 < p_end; p0 += 16) {
        const float m[8] = {nm0.x, nm0.y, nm0.z, nm0.w, nm1.x, nm1.y, nm1.z, nm1.w};
        float y[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float c = ok ? nx.get(e) - m[e] : 0.f;
            aacc = fmaf(c, m[e], aacc);
            y[e] = c * scale;
        }
        if (p0 + 16 < p_end) load(p0 + 16);
        acc = raw_products<PREC>(y, y, acc);
    }
    __shared__ float red[kCovWaves][1024];
    __shared__ float ared[kCovWaves][64];
#pragma unroll
    for (int r = 0; r < 16; ++r) red[wave][r * 64 + lane] = acc[r];
    ared[wave][lane] = aacc;
    __syncthreads();
    float* out = partial + (size_t)slice * kTile;
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) {
        float s = 0.f;
#pragma unroll
        for (int w = 0; w < kCovWaves; ++w) s += red[w][i];
        out[i] = s;
    }
    if (threadIdx.x < 32 && (int)threadIdx.x < T) {
        float a = 0.f;
#pragma unroll
        for (int w = 0; w < kCovWaves; ++w) a += ared[w][threadIdx.x] + ared[w][threadIdx.x + 32];
        apart[(size_t)slice * T + threadIdx.x] = make_float2(a, 0.f);
    }
}

// T = 32 G (G >= 2): as k_pca_cov_groups -- wave g < G converts frame group
// g's 16-pixel step into LDS fragments once, every wave runs the products of
// its pairs -- with real fragments.  Half precisions store the fragments
// already converted (hi and the split lo), so a group's conversion is done
// once instead of once per pair that reads it.
template <int PREC, int G, typename TIn>
__global__ void __launch_bounds__(64 * kCovGroupWaves)
    k_raw_cov_groups(float* __restrict__ partial, float2* __restrict__ apart, const TIn* __restrict__ raw,
                     const float* __restrict__ mean, size_t P, int slices, int slot0, int vec,
                     const uint32_t* __restrict__ scale_mm) {
    constexpr int NP = G * (G + 1) / 2;
    constexpr int W = kCovGroupWaves;
    constexpr int PPW = (NP + W - 1) / W;
    constexpr int T = 32 * G;
    using FR = RawFrag<PREC>;
    // 8 centred, scaled values per lane: FP32, or converted once (hi, and the
    // split lo) for every pair that reads them
    constexpr int NH = FR::fp32 ? 1 : (FR::split ? 2 : 1);
    __shared__ float4 frag[2][FR::fp32 ? G : 1][2][64];
    __shared__ typename FR::V8 hfrag[2][FR::fp32 ? 1 : G][NH][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t per_slice = P / slices;
    const size_t p_begin = (size_t)blockIdx.x * per_slice, p_end = p_begin + per_slice;
    float scale = 1.0f;
    if constexpr (PREC == HV_PCA_FP16 || PREC == HV_PCA_FP16_SPLIT) scale = fp16_scale(scale_mm);
    const bool conv = wave < G;
    const int f = 32 * (conv ? wave : 0) + (lane & 31);
    const TIn* row = raw + (size_t)((f - slot0) & (T - 1)) * P + (lane >> 5) * 8;
    const float* mrow = mean + (lane >> 5) * 8;
    RawPx<TIn, 8> nx;
    float4 nm0, nm1;
    auto load = [&](size_t p0) {
        nx.load(row + p0, vec);
        nm0 = *reinterpret_cast<const float4*>(mrow + p0);
        nm1 = *reinterpret_cast<const float4*>(mrow + p0 + 4);
    };
    if (conv) load(p_begin);
    float aacc = 0.f;
    v16f acc[PPW];
#pragma unroll
    for (int j = 0; j < PPW; ++j) acc[j] = v16f{};
    int buf = 0;
    for (size_t p0 = p_begin; p0 < p_end; p0 += 16, buf ^= 1) {
        if (conv) {
            const float m[8] = {nm0.x, nm0.y, nm0.z, nm0.w, nm1.x, nm1.y, nm1.z, nm1.w};
            float y[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float c = nx.get(e) - m[e];
                aacc = fmaf(c, m[e], aacc);
                y[e] = c * scale;
            }
            if (p0 + 16 < p_end) load(p0 + 16);
            if constexpr (FR::fp32) {
                frag[buf][wave][0][lane] = make_float4(y[0], y[1], y[2], y[3]);
                frag[buf][wave][1][lane] = make_float4(y[4], y[5], y[6], y[7]);
            } else {
                typename FR::V8 h, l;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const typename FR::E x = (typename FR::E)y[e];
                    h[e] = x;
                    if constexpr (FR::split) l[e] = (typename FR::E)(y[e] - (float)x);
                }
                hfrag[buf][wave][0][lane] = h;
                if constexpr (FR::split) hfrag[buf][wave][1][lane] = l;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PPW; ++j) {
            const int pair = wave + j * W;
            if (pair >= NP) break;
            int F, Gp;
            pair_of(pair, F, Gp);
            if constexpr (FR::fp32) {
                const float4 a0 = frag[buf][F][0][lane], a1 = frag[buf][F][1][lane];
                const float4 b0 = frag[buf][Gp][0][lane], b1 = frag[buf][Gp][1][lane];
                const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
                const float b[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
                acc[j] = raw_products<PREC>(a, b, acc[j]);
            } else {
                auto mfma = [](typename FR::V8 x, typename FR::V8 y2, v16f c) {
                    if constexpr (FR::bf) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y2, c, 0, 0, 0);
                    else return __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y2, c, 0, 0, 0);
                };
                const typename FR::V8 ah = hfrag[buf][F][0][lane], bh = hfrag[buf][Gp][0][lane];
                if constexpr (FR::split) {  // small cross terms first, then hi * hi (raw_products' order)
                    acc[j] = mfma(hfrag[buf][F][1][lane], bh, acc[j]);
                    acc[j] = mfma(ah, hfrag[buf][Gp][1][lane], acc[j]);
                }
                acc[j] = mfma(ah, bh, acc[j]);
            }
        }
    }
    if (conv) {
        const float a = aacc + __shfl_xor(aacc, 32);
        if (lane < 32) apart[(size_t)blockIdx.x * T + 32 * wave + lane] = make_float2(a, 0.f);
    }
#pragma unroll
    for (int j = 0; j < PPW; ++j) {
        const int pair = wave + j * W;
        if (pair >= NP) break;
        float* out = partial + ((size_t)blockIdx.x * NP + pair) * kTile;
#pragma unroll
        for (int r = 0; r < 16; ++r) out[r * 64 + lane] = acc[j][r];
    }
}

// Half precisions, slices of a multiple of 64 pixels: the same products over
// 64-pixel steps, one barrier per step instead of four.  Lane (frame f, half h)
// of wave g < G loads pixels 32 h .. 32 h + 31 of the step (64-byte runs for
// u16, the next step's loads issued before this step's are converted) and
// gives k-step kk the 8 values 32 h + 8 kk + e, for A and B alike; the step's
// 64 pixel means are staged in LDS by the last wave one step ahead.
template <int PREC, int G, typename TIn>
__global__ void __launch_bounds__(64 * kCovGroupWaves)
    k_raw_cov_groups_h(float* __restrict__ partial, float2* __restrict__ apart, const TIn* __restrict__ raw,
                       const float* __restrict__ mean, size_t P, int slices, int slot0, int vec,
                       const uint32_t* __restrict__ scale_mm) {
    using FR = RawFrag<PREC>;
    static_assert(!FR::fp32, "half precisions");
    constexpr int NP = G * (G + 1) / 2;
    constexpr int W = kCovGroupWaves;
    constexpr int PPW = (NP + W - 1) / W;
    constexpr int T = 32 * G;
    constexpr int NH = FR::split ? 2 : 1;
    constexpr int SS = 64;  // pixels per step
    __shared__ typename FR::V8 hfrag[2][4][G][NH][64];
    __shared__ float4 mu[2][SS / 4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5;
    const size_t per_slice = P / slices;
    const size_t p_begin = (size_t)blockIdx.x * per_slice, p_end = p_begin + per_slice;
    float scale = 1.0f;
    if constexpr (PREC == HV_PCA_FP16 || PREC == HV_PCA_FP16_SPLIT) scale = fp16_scale(scale_mm);
    const bool conv = wave < G, stager = wave == W - 1 && lane < SS / 4;
    const int f = 32 * (conv ? wave : 0) + (lane & 31);
    const TIn* row = raw + (size_t)((f - slot0) & (T - 1)) * P + 32 * h;
    RawPx<TIn, 32> nx;
    if (conv) nx.load(row + p_begin, vec);
    if (stager) mu[0][lane] = reinterpret_cast<const float4*>(mean + p_begin)[lane];
    __syncthreads();
    float aacc = 0.f;
    v16f acc[PPW];
#pragma unroll
    for (int j = 0; j < PPW; ++j) acc[j] = v16f{};
    int buf = 0;
    for (size_t p0 = p_begin; p0 < p_end; p0 += SS, buf ^= 1) {
        const bool more = p0 + SS < p_end;
        if (conv) {
            const float* m = reinterpret_cast<const float*>(mu[buf]) + 32 * h;
            float xs[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) xs[i] = nx.get(i);
            if (more) nx.load(row + p0 + SS, vec);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                typename FR::V8 hv, lv;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float me = m[8 * kk + e];
                    const float c = xs[8 * kk + e] - me;
                    aacc = fmaf(c, me, aacc);
                    const float y = c * scale;
                    const typename FR::E x = (typename FR::E)y;
                    hv[e] = x;
                    if constexpr (FR::split) lv[e] = (typename FR::E)(y - (float)x);
                }
                hfrag[buf][kk][wave][0][lane] = hv;
                if constexpr (FR::split) hfrag[buf][kk][wave][1][lane] = lv;
            }
        }
        // the next step's means (read from the other buffer after the barrier)
        if (stager && more) mu[buf ^ 1][lane] = reinterpret_cast<const float4*>(mean + p0 + SS)[lane];
        __syncthreads();
        auto mfma = [](typename FR::V8 x, typename FR::V8 y2, v16f c) {
            if constexpr (FR::bf) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y2, c, 0, 0, 0);
            else return __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y2, c, 0, 0, 0);
        };
#pragma unroll
        for (int j = 0; j < PPW; ++j) {
            const int pair = wave + j * W;
            if (pair >= NP) break;
            int F, Gp;
            pair_of(pair, F, Gp);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const typename FR::V8 ah = hfrag[buf][kk][F][0][lane], bh = hfrag[buf][kk][Gp][0][lane];
                if constexpr (FR::split) {  // small cross terms first, then hi * hi
                    acc[j] = mfma(hfrag[buf][kk][F][1][lane], bh, acc[j]);
                    acc[j] = mfma(ah, hfrag[buf][kk][Gp][1][lane], acc[j]);
                }
                acc[j] = mfma(ah, bh, acc[j]);
            }
        }
    }
    if (conv) {
        const float a = aacc + __shfl_xor(aacc, 32);
        if (lane < 32) apart[(size_t)blockIdx.x * T + 32 * wave + lane] = make_float2(a, 0.f);
    }
#pragma unroll
    for (int j = 0; j < PPW; ++j) {
        const int pair = wave + j * W;
        if (pair >= NP) break;
        float* out = partial + ((size_t)blockIdx.x * NP + pair) * kTile;
#pragma unroll
        for (int r = 0; r < 16; ++r) out[r * 64 + lane] = acc[j][r];
    }
}

// Pixel slices of the raw covariance: T <= 32 as the pair kernel (one pair),
// larger windows as the grouped kernel.
// T >= 64: slices of a multiple of 64 pixels where the frame allows
// (k_raw_cov_groups_h), else of 16 (k_raw_cov_groups).
static int raw_cov_slices(int T, size_t P) {
    if (T <= 32) return pca_slices_pairs(T, P);
    if (P % 64) return pca_slices_grouped(P);
    int ns = 1;
    while (ns * 2 <= kCovGroupSlices && (size_t)(ns * 2) * 128 <= P && P % ((size_t)(ns * 2) * 64) == 0) ns *= 2;
    return ns;
}

static CovLayout raw_cov_layout(int T, size_t P) {
    CovLayout l;
    const int slices = raw_cov_slices(T, P);
    l.partial_floats = (size_t)slices * pca_pairs(T) * kTile;
    l.off_apart = align_up(l.partial_floats * sizeof(float));
    l.off_mean = l.off_apart + align_up((size_t)slices * std::max(T, 32) * sizeof(float2));
    l.off_mpart = l.off_mean + align_up(P * sizeof(float));
    l.off_terms = l.off_mpart + align_up(kMeanBlocks * sizeof(double));
    l.offset_scale = l.off_terms + align_up((size_t)(T + 1) * sizeof(double2));
    l.bytes = l.offset_scale + 256;
    return l;
}

template <int PREC, typename TIn>
static hv_status raw_cov_t(float* partial, float2* apart, const void* raw, const float* mean, int T, size_t P,
                           int slices, int slot0, int vec, const uint32_t* scale_mm, hipStream_t st) {
    const TIn* r = (const TIn*)raw;
    constexpr int tg = 64 * kCovGroupWaves;
    if constexpr (!RawFrag<PREC>::fp32) {
        // T = 256: eight groups' step fragments would take 128 KiB of LDS
        if ((T == 64 || T == 128) && (P / slices) % 64 == 0) {
            if (T == 64) k_raw_cov_groups_h<PREC, 2, TIn><<<slices, tg, 0, st>>>(partial, apart, r, mean, P, slices, slot0, vec, scale_mm);
            else k_raw_cov_groups_h<PREC, 4, TIn><<<slices, tg, 0, st>>>(partial, apart, r, mean, P, slices, slot0, vec, scale_mm);
            HV_CHECK_LAUNCH("k_raw_cov_groups_h");
            return HV_OK;
        }
    }
    switch (T) {
        case 64: k_raw_cov_groups<PREC, 2, TIn><<<slices, tg, 0, st>>>(partial, apart, r, mean, P, slices, slot0, vec, scale_mm); break;
        case 128: k_raw_cov_groups<PREC, 4, TIn><<<slices, tg, 0, st>>>(partial, apart, r, mean, P, slices, slot0, vec, scale_mm); break;
        case 256: k_raw_cov_groups<PREC, 8, TIn><<<slices, tg, 0, st>>>(partial, apart, r, mean, P, slices, slot0, vec, scale_mm); break;
        default:
            k_raw_cov<PREC, TIn><<<slices, 64 * kCovWaves, 0, st>>>(partial, apart, r, mean, T, P, slices, slot0, vec, scale_mm);
            break;
    }
    HV_CHECK_LAUNCH("k_raw_cov");
    return HV_OK;
}

template <typename TIn>
static hv_status raw_cov_p(int precision, float* partial, float2* apart, const void* raw, const float* mean, int T,
                           size_t P, int slices, int slot0, int vec, const uint32_t* scale_mm, hipStream_t st) {
    switch (precision) {
        case HV_PCA_FP32: return raw_cov_t<HV_PCA_FP32, TIn>(partial, apart, raw, mean, T, P, slices, slot0, vec, scale_mm, st);
        case HV_PCA_BF16_SPLIT: return raw_cov_t<HV_PCA_BF16_SPLIT, TIn>(partial, apart, raw, mean, T, P, slices, slot0, vec, scale_mm, st);
        case HV_PCA_FP16_SPLIT: return raw_cov_t<HV_PCA_FP16_SPLIT, TIn>(partial, apart, raw, mean, T, P, slices, slot0, vec, scale_mm, st);
        case HV_PCA_BF16: return raw_cov_t<HV_PCA_BF16, TIn>(partial, apart, raw, mean, T, P, slices, slot0, vec, scale_mm, st);
        default: return raw_cov_t<HV_PCA_FP16, TIn>(partial, apart, raw, mean, T, P, slices, slot0, vec, scale_mm, st);
    }
}

bool pca_raw_cov_supported(int T, size_t P) {
    if (!(T == 1 || T == 2 || T == 4 || T == 8 || T == 16 || T == 32 || T == 64 || T == 128 || T == 256)) return false;
    const int slices = raw_cov_slices(T, P);
    return P % ((size_t)slices * (T <= 32 ? kCovWaves : 1) * 16) == 0;
}

hv_status launch_pca_cov_raw(double2* cov, const void* raw, int depth, int T, size_t P, int slot0, int precision,
                             void* ws, hipStream_t st) {
    if (!pca_raw_cov_supported(T, P)) return fail(HV_ERR_UNSUPPORTED, "raw-window covariance: unsupported shape");
    switch (precision) {
        case HV_PCA_FP32: case HV_PCA_BF16_SPLIT: case HV_PCA_FP16_SPLIT: case HV_PCA_BF16: case HV_PCA_FP16: break;
        default: return fail(HV_ERR_INVALID_ARG, "unknown PCA precision");
    }
    const CovLayout l = raw_cov_layout(T, P);
    const int slices = raw_cov_slices(T, P), np = pca_pairs(T);
    char* w = reinterpret_cast<char*>(ws);
    float* partial = reinterpret_cast<float*>(w);
    float2* apart = reinterpret_cast<float2*>(w + l.off_apart);
    float* mean = reinterpret_cast<float*>(w + l.off_mean);
    double* mpart = reinterpret_cast<double*>(w + l.off_mpart);
    uint32_t* scale_mm = reinterpret_cast<uint32_t*>(w + l.offset_scale);
    const int vec = (reinterpret_cast<uintptr_t>(raw) & 15) == 0;
    const bool fp16 = precision == HV_PCA_FP16 || precision == HV_PCA_FP16_SPLIT;
    // the FP16 scale words (max |x|, max |mu|): only the FP16 products use them
    uint32_t* mm = fp16 ? scale_mm : nullptr;
    if (fp16) {
        k_scale_init<<<1, 64, 0, st>>>(scale_mm);
        HV_CHECK_LAUNCH("k_scale_init");
    }
    const int mblocks = (int)std::min<size_t>(kMeanBlocks, P / 256);
    switch (depth) {
        case HV_DEPTH_8: k_raw_mean<uint8_t><<<mblocks, kRawMeanThreads, 0, st>>>(mean, mpart, mm, (const uint8_t*)raw, T, P, vec); break;
        case HV_DEPTH_16: k_raw_mean<uint16_t><<<mblocks, kRawMeanThreads, 0, st>>>(mean, mpart, mm, (const uint16_t*)raw, T, P, vec); break;
        case HV_DEPTH_32F: k_raw_mean<float><<<mblocks, kRawMeanThreads, 0, st>>>(mean, mpart, mm, (const float*)raw, T, P, vec); break;
        default: return fail(HV_ERR_INVALID_ARG, "depth must be 1, 2 or 4");
    }
    HV_CHECK_LAUNCH("k_raw_mean");
    switch (depth) {
        case HV_DEPTH_8: HV_TRY(raw_cov_p<uint8_t>(precision, partial, apart, raw, mean, T, P, slices, slot0, vec, scale_mm, st)); break;
        case HV_DEPTH_16: HV_TRY(raw_cov_p<uint16_t>(precision, partial, apart, raw, mean, T, P, slices, slot0, vec, scale_mm, st)); break;
        default: HV_TRY(raw_cov_p<float>(precision, partial, apart, raw, mean, T, P, slices, slot0, vec, scale_mm, st)); break;
    }
    k_raw_cov_reduce<<<dim3(np, 1024 / 64), 64 * kRawReduceWaves, 0, st>>>(cov, partial, apart, mpart, mblocks, T,
                                                                          slices, np, scale_mm, fp16 ? 1 : 0);
    HV_CHECK_LAUNCH("k_raw_cov_reduce");
    return HV_OK;
}

// ---------------------------------------------------------------------------
// FP64 Jacobi
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }

// circle-method pairing: round r of n - 1, pair k of n / 2
__device__ __forceinline__ void rr_pair(int n, int r, int k, int& p, int& q) {
    int a, b;
    if (k == 0) {
        a = n - 1;
        b = r;
    } else {
        a = (r + k) % (n - 1);
        b = (r - k + (n - 1)) % (n - 1);
    }
    p = a < b ? a : b;
    q = a < b ? b : a;
}

#ifndef HV_JACOBI_RSQRT
#define HV_JACOBI_RSQRT 1       // rotation cosines by reciprocal square root
#endif
#ifndef HV_JACOBI_MAX_SWEEPS
#define HV_JACOBI_MAX_SWEEPS 40
#endif
#ifndef HV_JACOBI_NT32
#define HV_JACOBI_NT32 1024     // threads of the T <= 32 eigensolver (A's blocks and V's row pairs on separate threads)
#endif
#ifndef HV_JACOBI_NT64
#define HV_JACOBI_NT64 1024     // threads of the T <= 64 eigensolver
#endif

// One sweep round applies n/2 disjoint rotations J_k (pair k = (p_k, q_k)).
// A' = J^H A J acts on the 2 x 2 blocks A[k, k'] = rows (p_k, q_k) x columns
// (p_k', q_k') independently: A'[k, k'] = J_k^H A[k, k'] J_k'.  So a round is
// one rotation phase (n/2 threads) and one update phase in which each thread
// owns whole blocks of A and pairs of rows of V -- two barriers per round.
// A and V (n x n, column-major, leading dimension ld) live in LDS or global
// memory; n is a compile-time constant (the pairs of a round are computed
// with constant divisors).
struct Rot {
    double c, s;   // J = [[c, s], [-s conj(e), c conj(e)]]
    double2 ec;    // conj(e), e = a_pq / |a_pq|
};

// Rotation annihilating a_pq of the 2 x 2 Hermitian block [[app, apq], [conj apq, aqq]].
__device__ __forceinline__ Rot make_rot(double app, double aqq, double2 apq) {
    const double mag2 = apq.x * apq.x + apq.y * apq.y;
    Rot R;
    if (mag2 == 0.0) {
        R.c = 1.0;
        R.s = 0.0;
        R.ec = make_double2(1.0, 0.0);
    } else {
        // With h = aqq - app, g = 2 |apq|, r = sqrt(h^2 + g^2), the rotation
        // t = sgn(h) g / (|h| + r) (tau = h / g), c = 1 / sqrt(1 + t^2), s = t c
        // is c = (r + |h|) u, s = sgn(h) g u with u = 1 / sqrt(2 r (r + |h|)):
        // a square root and two reciprocal square roots, no division, and
        // no cancellation for small angles.
        const double inv = HV_JACOBI_RSQRT ? rsqrt(mag2) : 1.0 / sqrt(mag2);
        R.ec = make_double2(apq.x * inv, -apq.y * inv);
        const double h = aqq - app, ah = fabs(h);
        const double g = 2.0 * mag2 * inv;
        const double r = sqrt(fma(h, h, 4.0 * mag2));
        const double u = HV_JACOBI_RSQRT ? rsqrt(2.0 * r * (r + ah)) : 1.0 / sqrt(2.0 * r * (r + ah));
        R.c = (r + ah) * u;
        R.s = (h >= 0 ? g : -g) * u;
    }
    return R;
}

template <int NT>
__device__ double block_sum(double v, double* red) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < NT / 64; ++i) s += red[i];
    __syncthreads();
    return s;
}

// x J restricted to one pair: (x_p, x_q) -> (x_p c - x_q s conj(e), x_p s + x_q c conj(e))
__device__ __forceinline__ void rot_right(const Rot& R, double2& xp, double2& xq) {
    const double2 t = zmul(xq, R.ec);
    const double2 np_ = make_double2(xp.x * R.c - t.x * R.s, xp.y * R.c - t.y * R.s);
    const double2 nq_ = make_double2(xp.x * R.s + t.x * R.c, xp.y * R.s + t.y * R.c);
    xp = np_;
    xq = nq_;
}
// J^H y restricted to one pair: (y_p, y_q) -> (c y_p - s e y_q, s y_p + c e y_q)
__device__ __forceinline__ void rot_left(const Rot& R, double2& yp, double2& yq) {
    const double2 t = zmul(yq, zconj(R.ec));
    const double2 np_ = make_double2(R.c * yp.x - R.s * t.x, R.c * yp.y - R.s * t.y);
    const double2 nq_ = make_double2(R.s * yp.x + R.c * t.x, R.s * yp.y + R.c * t.y);
    yp = np_;
    yq = nq_;
}

// Real symmetric matrices (the raw-window covariance): the same rotation with
// e = sgn(a_pq) folded into s, J = [[c, s], [-s, c]]; every update real.
struct RotR {
    double c, s;
};
__device__ __forceinline__ RotR make_rot(double app, double aqq, double apq) {
    RotR R;
    if (apq == 0.0) {
        R.c = 1.0;
        R.s = 0.0;
    } else {
        const double h = aqq - app, ah = fabs(h);
        const double g = 2.0 * fabs(apq);
        const double r = sqrt(fma(h, h, g * g));
        const double u = HV_JACOBI_RSQRT ? rsqrt(2.0 * r * (r + ah)) : 1.0 / sqrt(2.0 * r * (r + ah));
        R.c = (r + ah) * u;
        const double s = g * u;
        R.s = ((h >= 0) == (apq >= 0)) ? s : -s;  // sgn(h) sgn(a_pq) g u
    }
    return R;
}
__device__ __forceinline__ void rot_right(const RotR& R, double& xp, double& xq) {
    const double np_ = xp * R.c - xq * R.s, nq_ = xp * R.s + xq * R.c;
    xp = np_;
    xq = nq_;
}
__device__ __forceinline__ void rot_left(const RotR& R, double& yp, double& yq) { rot_right(R, yp, yq); }

// Element helpers of the two matrix kinds (complex double2, real double).
__device__ __forceinline__ double el_re(double2 x) { return x.x; }
__device__ __forceinline__ double el_re(double x) { return x; }
__device__ __forceinline__ double el_norm2(double2 x) { return x.x * x.x + x.y * x.y; }
__device__ __forceinline__ double el_norm2(double x) { return x * x; }
__device__ __forceinline__ void el_zero(double2& x) { x = make_double2(0.0, 0.0); }
__device__ __forceinline__ void el_zero(double& x) { x = 0.0; }
__device__ __forceinline__ double2 el_conj(double2 x) { return make_double2(x.x, -x.y); }
__device__ __forceinline__ double el_conj(double x) { return x; }
__device__ __forceinline__ void el_one(double2& x) { x = make_double2(1.0, 0.0); }
__device__ __forceinline__ void el_one(double& x) { x = 1.0; }
__device__ __forceinline__ void el_set_real(double2& x, double v) { x = make_double2(v, 0.0); }
__device__ __forceinline__ void el_set_real(double& x, double v) { x = v; }
__device__ __forceinline__ void el_real(double2& x) { x.y = 0.0; }
__device__ __forceinline__ void el_real(double&) {}
__device__ __forceinline__ double2 el_c(double2 x) { return x; }
__device__ __forceinline__ double2 el_c(double x) { return make_double2(x, 0.0); }
template <class E>
using RotOf = decltype(make_rot(0.0, 0.0, E{}));
// x times the unit phase that makes ph real positive
__device__ __forceinline__ double2 el_phase_norm(double2 x, double2 ph) {
    const double inv = 1.0 / sqrt(ph.x * ph.x + ph.y * ph.y);
    return zmul(x, make_double2(ph.x * inv, -ph.y * inv));
}
__device__ __forceinline__ double2 el_phase_norm(double x, double ph) { return make_double2(ph < 0.0 ? -x : x, 0.0); }

// Eigenvalues ascending (stable ranks: ties keep their index order) and
// phase-normalised eigenvectors -- the largest component real positive, the
// first of equal magnitudes within a relative 1e-12 -- from the converged
// diagonal d(j) and the accumulated rotations v(k, j).
// nout < N: the matrix was padded (see k_jacobi_lds); only the first nout
// eigenpairs, of nout components, are written.
template <int NT, int N, class D, class VA>
__device__ void jacobi_finish(D&& d, VA&& v, double* evals, double2* evecs, int* order, float2* evecs32,
                              int nout = N) {
    constexpr int n = N;
    const int tid = threadIdx.x;
    for (int j = tid; j < n; j += NT) {
        const double dj = d(j);
        int rank = 0;
#pragma unroll 4
        for (int i = 0; i < n; ++i) {
            const double di = d(i);
            rank += (di < dj || (di == dj && i < j)) ? 1 : 0;
        }
        order[rank] = j;
    }
    __syncthreads();
    for (int jj = tid; jj < nout; jj += NT) {
        const int j = order[jj];
        evals[jj] = d(j);
        int kmax = 0;
        double best = -1.0;
#pragma unroll 4
        for (int k = 0; k < nout; ++k) {
            const double m2 = el_norm2(v(k, j));
            if (m2 > best * (1 + 2e-12)) {
                best = m2;
                kmax = k;
            }
        }
        const auto ph = v(kmax, j);
#pragma unroll 4
        for (int k = 0; k < nout; ++k) {
            const double2 z = el_phase_norm(v(k, j), ph);
            evecs[k + (size_t)jj * nout] = z;
            // row-major float copy: one component of consecutive eigenvectors is contiguous
            if (evecs32) evecs32[(size_t)k * nout + jj] = make_float2((float)z.x, (float)z.y);
        }
    }
}

// Rotation log of a deferred run: every round's n/2 rotations in order, then
// the round count; the eigenvectors are accumulated afterwards row by row
// (k_jacobi_apply), which many workgroups do in parallel.
template <class R = Rot>
struct JacobiLog {
    R* rot;         // HV_JACOBI_MAX_SWEEPS (n - 1) rounds of n / 2
    int* rounds;
    double* diag;   // converged diagonal, unordered
};

// nreal < N: a matrix padded to N (k_jacobi_lds); the convergence test
// measures the nreal x nreal block only, as for that matrix alone.
template <int NT, int N, bool DEFER = false, class E = double2>
__device__ void jacobi_run(E* A, E* V, int ld, double* evals, double2* evecs, RotOf<E>* rot, double* red,
                           int* order, float2* evecs32, JacobiLog<RotOf<E>> log = JacobiLog<RotOf<E>>{},
                           int nreal = N) {
    using R = RotOf<E>;
    constexpr int n = N, half = N / 2;
    const int tid = threadIdx.x;
    int g = 0;  // rounds done
    auto a = [&](int i, int j) -> E& { return A[i + (size_t)j * ld]; };
    auto v = [&](int i, int j) -> E& { return V[i + (size_t)j * ld]; };
    double fro = 0.0;
    for (int e = tid; e < n * n; e += NT)
        if (e % n < nreal && e / n < nreal) fro += el_norm2(a(e % n, e / n));
    fro = block_sum<NT>(fro, red);
    for (int sweep = 0; sweep < HV_JACOBI_MAX_SWEEPS && n > 1; ++sweep) {
        double off = 0.0;
        for (int e = tid; e < n * n; e += NT) {
            const int i = e % n, j = e / n;
            if (i != j) off += el_norm2(a(i, j));
        }
        off = block_sum<NT>(off, red);
        // relative off-diagonal norm below 1e-13: far under the FP32 input's own precision
        if (off <= 1e-26 * fro || off == 0.0) break;
        for (int r = 0; r < n - 1; ++r) {
            for (int k = tid; k < half; k += NT) {
                int p, q;
                rr_pair(n, r, k, p, q);
                const R rk = make_rot(el_re(a(p, p)), el_re(a(q, q)), a(p, q));
                rot[k] = rk;
                if constexpr (DEFER) log.rot[(size_t)g * half + k] = rk;
            }
            ++g;
            __syncthreads();
            // A <- J^H A J, block by block; V <- V J, row pair by row pair
            for (int task = tid; task < half * half; task += NT) {
                const int k = task % half, kk = task / half;  // row pair k, column pair kk
                int p, q, pp, qq;
                rr_pair(n, r, k, p, q);
                rr_pair(n, r, kk, pp, qq);
                const R Rr = rot[k], Rc = rot[kk];
                E b00 = a(p, pp), b01 = a(p, qq), b10 = a(q, pp), b11 = a(q, qq);
                rot_right(Rc, b00, b01);
                rot_right(Rc, b10, b11);
                rot_left(Rr, b00, b10);
                rot_left(Rr, b01, b11);
                if (k == kk) {  // the annihilated pair: exact zeros, real diagonal
                    el_zero(b01);
                    el_zero(b10);
                    el_real(b00);
                    el_real(b11);
                }
                a(p, pp) = b00;
                a(p, qq) = b01;
                a(q, pp) = b10;
                a(q, qq) = b11;
            }
            if constexpr (!DEFER) {
                // with threads for both, V's row pairs go to the threads past A's blocks, in parallel
                constexpr int v0 = NT >= half * half + half * n ? half * half : 0;
                for (int task = tid - v0; task >= 0 && task < half * n; task += NT) {
                    const int m = task % n, kk = task / n;
                    int pp, qq;
                    rr_pair(n, r, kk, pp, qq);
                    E x = v(m, pp), y = v(m, qq);
                    rot_right(rot[kk], x, y);
                    v(m, pp) = x;
                    v(m, qq) = y;
                }
            }
            __syncthreads();
        }
    }
    if constexpr (DEFER) {
        for (int j = tid; j < n; j += NT) log.diag[j] = el_re(a(j, j));
        if (tid == 0) *log.rounds = g;
        return;
    }
    jacobi_finish<NT, N>([&](int j) { return el_re(a(j, j)); }, v, evals, evecs, order, evecs32, nreal);
}

// A covariance element (float2 from the C-ABI step, double2 in the pipe) as double2.
__device__ __forceinline__ double2 cov_elem(const float2* c, size_t k) {
    const float2 v = c[k];
    return make_double2((double)v.x, (double)v.y);
}
__device__ __forceinline__ double2 cov_elem(const double2* c, size_t k) { return c[k]; }

// E = double: the real part of the covariance only (a real symmetric matrix).
__device__ __forceinline__ void set_cov_el(double2& dst, double2 c, int i, int j) {
    dst = make_double2(c.x, i == j ? 0.0 : (i > j ? c.y : -c.y));
}
__device__ __forceinline__ void set_cov_el(double& dst, double2 c, int, int) { dst = c.x; }

// nreal < N (windows that are not powers of two): the nreal x nreal matrix
// (leading dimension nreal) padded to N with a diagonal block pad * I, pad
// above every eigenvalue (1 + twice the largest absolute row sum, a
// Gershgorin bound) and zero coupling: the pad rows never rotate (their
// off-diagonal entries are exact zeros), the matrix's eigenpairs are the
// nreal smallest, and the first nreal components of their vectors.
template <int NT, int N, class CT, class E = double2>
__global__ void __launch_bounds__(NT)
    k_jacobi_lds(double* evals, double2* evecs, const CT* cov, float2* evecs32, int nreal) {
    constexpr int LD = N + 1;  // padded columns: row and column walks spread over the banks
    __shared__ E A[N * LD], V[N * LD];
    __shared__ RotOf<E> rot[N / 2 > 0 ? N / 2 : 1];
    __shared__ double red[NT / 64];
    __shared__ int order[N];
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        if (i < nreal && j < nreal) {
            // Hermitian from the lower triangle (CUBLAS_FILL_MODE_LOWER)
            const double2 c = i >= j ? cov_elem(cov, i + (size_t)j * nreal) : cov_elem(cov, j + (size_t)i * nreal);
            set_cov_el(A[i + j * LD], c, i, j);
        } else {
            el_zero(A[i + j * LD]);
        }
        el_zero(V[i + j * LD]);
        if (i == j) el_one(V[i + j * LD]);
    }
    __syncthreads();
    if (nreal < N) {
        if (threadIdx.x == 0) {
            double bound = 0.0;
            for (int i = 0; i < nreal; ++i) {
                double row = 0.0;
                for (int j = 0; j < nreal; ++j) row += sqrt(el_norm2(A[i + j * LD]));
                bound = fmax(bound, row);
            }
            for (int i = nreal; i < N; ++i) el_set_real(A[i + i * LD], 1.0 + 2.0 * bound);
        }
        __syncthreads();
    }
    jacobi_run<NT, N, false, E>(A, V, LD, evals, evecs, rot, red, order, evecs32, JacobiLog<RotOf<E>>{}, nreal);
}

// T >= 128: A in the workspace, the rotations logged (JacobiLog) instead of
// applied to V in the same workgroup: one workgroup streaming both matrices
// through L2 every round was bound by that one CU's L2 bandwidth.
template <int N>
struct JacobiWs {
    static constexpr size_t kLogRounds = (size_t)HV_JACOBI_MAX_SWEEPS * (N - 1);
    static constexpr size_t a_off = 0;                                             // N x N double2
    static constexpr size_t v_off = a_off + (size_t)N * N * sizeof(double2);       // N x N double2
    static constexpr size_t log_off = v_off + (size_t)N * N * sizeof(double2);     // kLogRounds x N/2 Rot
    static constexpr size_t diag_off = log_off + kLogRounds * (N / 2) * sizeof(Rot);
    static constexpr size_t rounds_off = diag_off + (size_t)N * sizeof(double);
    static constexpr size_t bytes = rounds_off + 256;
};

template <int NT, int N, class CT>
__global__ void __launch_bounds__(NT) k_jacobi_global(const CT* cov, char* ws) {
    using L = JacobiWs<N>;
    __shared__ Rot rot[N / 2];
    __shared__ double red[NT / 64];
    __shared__ int order[N];
    double2* A = reinterpret_cast<double2*>(ws + L::a_off);
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        const double2 c = i >= j ? cov_elem(cov, i + (size_t)j * N) : cov_elem(cov, j + (size_t)i * N);
        A[e] = make_double2(c.x, i == j ? 0.0 : (i > j ? c.y : -c.y));
    }
    __syncthreads();
    const JacobiLog<Rot> log{reinterpret_cast<Rot*>(ws + L::log_off), reinterpret_cast<int*>(ws + L::rounds_off),
                             reinterpret_cast<double*>(ws + L::diag_off)};
    jacobi_run<NT, N, true, double2>(A, nullptr, N, nullptr, nullptr, rot, red, order, nullptr, log);
}

// T = 128: the lower triangle of A packed in LDS (132 KB; the full matrix
// does not fit).  A round's 2 x 2 blocks (k, kk) and (kk, k) hold conjugate
// elements, so only blocks with k >= kk are updated: every stored element
// once per round.  Rotations go to the log as in k_jacobi_global.
template <int NT, int N, class CT, class E = double2>
__global__ void __launch_bounds__(NT) k_jacobi_packed(const CT* cov, char* ws) {
    using L = JacobiWs<N>;
    using R = RotOf<E>;
    constexpr int n = N, half = N / 2, nblocks = half * (half + 1) / 2;
    extern __shared__ double2 Lraw[];  // element (i, j), i >= j, at i (i + 1) / 2 + j
    E* Lt = reinterpret_cast<E*>(Lraw);
    __shared__ R rot[half];
    __shared__ double red[NT / 64];
    const int tid = threadIdx.x;
    auto at = [&](int i, int j) -> E& { return Lt[i * (i + 1) / 2 + j]; };
    auto get = [&](int i, int j) -> E {
        if (i >= j) return at(i, j);
        return el_conj(at(j, i));
    };
    auto set = [&](int i, int j, E v) {
        if (i >= j) at(i, j) = v;
        else at(j, i) = el_conj(v);
    };
    // lower triangle of cov (CUBLAS_FILL_MODE_LOWER), real diagonal
    for (int i = 0; i < n; ++i)
        for (int j = tid; j <= i; j += NT) set_cov_el(at(i, j), cov_elem(cov, i + (size_t)j * n), i, j);
    __syncthreads();
    auto tri_sum = [&](bool with_diag) {
        double acc = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = tid; j <= i; j += NT) {
                const double m2 = el_norm2(at(i, j));
                acc += i == j ? (with_diag ? m2 : 0.0) : 2.0 * m2;
            }
        return block_sum<NT>(acc, red);
    };
    const double fro = tri_sum(true);
    const JacobiLog<R> log{reinterpret_cast<R*>(ws + L::log_off), reinterpret_cast<int*>(ws + L::rounds_off),
                           reinterpret_cast<double*>(ws + L::diag_off)};
    // this thread's blocks (k, kk), k >= kk, with task = k (k + 1) / 2 + kk (the same every round)
    constexpr int kTasks = (nblocks + NT - 1) / NT;
    int tk[kTasks], tkk[kTasks];
#pragma unroll
    for (int u = 0; u < kTasks; ++u) {
        const int task = tid + u * NT;
        tk[u] = tkk[u] = -1;
        if (task < nblocks) {
            int k = (int)((sqrtf(8.0f * task + 1.0f) - 1.0f) * 0.5f);
            while (k * (k + 1) / 2 > task) --k;
            while ((k + 1) * (k + 2) / 2 <= task) ++k;
            tk[u] = k;
            tkk[u] = task - k * (k + 1) / 2;
        }
    }
    int g = 0;
    for (int sweep = 0; sweep < HV_JACOBI_MAX_SWEEPS; ++sweep) {
        const double off = tri_sum(false);
        if (off <= 1e-26 * fro || off == 0.0) break;
        for (int r = 0; r < n - 1; ++r) {
            for (int k = tid; k < half; k += NT) {
                int p, q;
                rr_pair(n, r, k, p, q);
                const R rk = make_rot(el_re(at(p, p)), el_re(at(q, q)), get(p, q));
                rot[k] = rk;
                log.rot[(size_t)g * half + k] = rk;
            }
            ++g;
            __syncthreads();
            // a round's blocks touch disjoint elements: all loads, then all
            // rotations, then all stores, so that the LDS latencies overlap
            int p[kTasks], q[kTasks], pp[kTasks], qq[kTasks];
            E b[kTasks][4];
#pragma unroll
            for (int u = 0; u < kTasks; ++u) {
                if (tk[u] < 0) continue;
                rr_pair(n, r, tk[u], p[u], q[u]);
                rr_pair(n, r, tkk[u], pp[u], qq[u]);
                b[u][0] = get(p[u], pp[u]);
                b[u][1] = get(p[u], qq[u]);
                b[u][2] = get(q[u], pp[u]);
                b[u][3] = get(q[u], qq[u]);
            }
#pragma unroll
            for (int u = 0; u < kTasks; ++u) {
                if (tk[u] < 0) continue;
                const R Rr = rot[tk[u]], Rc = rot[tkk[u]];
                rot_right(Rc, b[u][0], b[u][1]);
                rot_right(Rc, b[u][2], b[u][3]);
                rot_left(Rr, b[u][0], b[u][2]);
                rot_left(Rr, b[u][1], b[u][3]);
                if (tk[u] == tkk[u]) {  // the annihilated pair: exact zeros, real diagonal
                    el_zero(b[u][1]);
                    el_zero(b[u][2]);
                    el_real(b[u][0]);
                    el_real(b[u][3]);
                }
            }
#pragma unroll
            for (int u = 0; u < kTasks; ++u) {
                if (tk[u] < 0) continue;
                set(p[u], pp[u], b[u][0]);
                set(p[u], qq[u], b[u][1]);
                set(q[u], pp[u], b[u][2]);
                set(q[u], qq[u], b[u][3]);
            }
            __syncthreads();
        }
    }
    for (int j = tid; j < n; j += NT) log.diag[j] = el_re(at(j, j));
    if (tid == 0) *log.rounds = g;
}

template <class E = double2>
constexpr size_t jacobi_packed_lds(int n) { return (size_t)n * (n + 1) / 2 * sizeof(E); }

// V = J_1 J_2 ... J_G row by row: row m of the identity times the logged
// rotations, one wavefront per row in LDS (a round's rotations touch
// disjoint pairs; the wavefront's LDS accesses complete in order, so rounds
// need only a fence).
template <int N, class E = double2>
__global__ void __launch_bounds__(256) k_jacobi_apply(const char* __restrict__ ws_in, char* ws) {
    using L = JacobiWs<N>;
    using R = RotOf<E>;
    constexpr int half = N / 2;
    __shared__ E rows[4][N];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = blockIdx.x * 4 + w;
    E* row = rows[w];
    const R* log = reinterpret_cast<const R*>(ws_in + L::log_off);
    const int G = *reinterpret_cast<const int*>(ws_in + L::rounds_off);
    for (int j = lane; j < N; j += 64) {
        el_zero(row[j]);
        if (j == m) el_one(row[j]);
    }
    fft_sync<true>();
    for (int g = 0; g < G; ++g) {
        const int r = g % (N - 1);
#pragma unroll
        for (int k = lane; k < half; k += 64) {
            int p, q;
            rr_pair(N, r, k, p, q);
            const R rk = log[(size_t)g * half + k];
            E x = row[p], y = row[q];
            rot_right(rk, x, y);
            row[p] = x;
            row[q] = y;
        }
        fft_sync<true>();
    }
    E* V = reinterpret_cast<E*>(ws + L::v_off);
    for (int j = lane; j < N; j += 64) V[m + (size_t)j * N] = row[j];
}

template <int NT, int N, class E = double2>
__global__ void __launch_bounds__(NT) k_jacobi_finish(double* evals, double2* evecs, const char* __restrict__ ws,
                                                      float2* evecs32) {
    using L = JacobiWs<N>;
    __shared__ int order[N];
    const double* diag = reinterpret_cast<const double*>(ws + L::diag_off);
    const E* V = reinterpret_cast<const E*>(ws + L::v_off);
    jacobi_finish<NT, N>([&](int j) { return diag[j]; }, [&](int k, int j) { return V[k + (size_t)j * N]; }, evals,
                         evecs, order, evecs32);
}

size_t eig_workspace_bytes(int T) {
    return T <= 64 ? 0 : (T == 128 ? JacobiWs<128>::bytes : JacobiWs<256>::bytes);
}

#ifndef HV_JACOBI_PACKED
#define HV_JACOBI_PACKED 1      // T = 128: A as a packed triangle in LDS
#endif
#ifndef HV_JACOBI_NT128
#define HV_JACOBI_NT128 1024    // threads of the packed T = 128 eigensolver
#endif

// E = double: a real symmetric matrix (the real part of cov), real rotations.
template <int N, class CT, class E = double2>
static hv_status eig_deferred(double* evals, double2* evecs, const CT* cov, char* ws, hipStream_t st,
                              float2* evecs32) {
    if constexpr (N == 128 && HV_JACOBI_PACKED) {
        constexpr size_t lds = jacobi_packed_lds<E>(N);
        static const hipError_t attr = hipFuncSetAttribute(
            reinterpret_cast<const void*>(&k_jacobi_packed<HV_JACOBI_NT128, N, CT, E>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (attr != hipSuccess) return hip_fail(attr, "hipFuncSetAttribute(k_jacobi_packed)");
        k_jacobi_packed<HV_JACOBI_NT128, N, CT, E><<<1, HV_JACOBI_NT128, lds, st>>>(cov, ws);
        HV_CHECK_LAUNCH("k_jacobi_packed");
    } else {
        if constexpr (!std::is_same<E, double2>::value) return fail(HV_ERR_UNSUPPORTED, "real Jacobi: T = 256");
        k_jacobi_global<1024, N, CT><<<1, 1024, 0, st>>>(cov, ws);
        HV_CHECK_LAUNCH("k_jacobi_global");
    }
    k_jacobi_apply<N, E><<<N / 4, 256, 0, st>>>(ws, ws);
    HV_CHECK_LAUNCH("k_jacobi_apply");
    k_jacobi_finish<256, N, E><<<1, 256, 0, st>>>(evals, evecs, ws, evecs32);
    HV_CHECK_LAUNCH("k_jacobi_finish");
    return HV_OK;
}

template <class CT, class E = double2>
static hv_status eig_jacobi(double* evals, double2* evecs, const CT* cov, int T, void* work, hipStream_t st,
                            float2* e32) {
    // windows that are not powers of two (T < 32) run padded to the next power of two
    const int N = T <= 1 ? T : 1 << ilog2(T);
    if (N != T && T > 32) return fail(HV_ERR_UNSUPPORTED, "eigensolver: T must be a power of two or below 32");
    switch (N) {
        case 1: k_jacobi_lds<64, 1, CT, E><<<1, 64, 0, st>>>(evals, evecs, cov, e32, T); break;
        case 2: k_jacobi_lds<64, 2, CT, E><<<1, 64, 0, st>>>(evals, evecs, cov, e32, T); break;
        case 4: k_jacobi_lds<64, 4, CT, E><<<1, 64, 0, st>>>(evals, evecs, cov, e32, T); break;
        case 8: k_jacobi_lds<64, 8, CT, E><<<1, 64, 0, st>>>(evals, evecs, cov, e32, T); break;
        case 16: k_jacobi_lds<256, 16, CT, E><<<1, 256, 0, st>>>(evals, evecs, cov, e32, T); break;
        case 32: k_jacobi_lds<HV_JACOBI_NT32, 32, CT, E><<<1, HV_JACOBI_NT32, 0, st>>>(evals, evecs, cov, e32, T); break;
        case 64: k_jacobi_lds<HV_JACOBI_NT64, 64, CT, E><<<1, HV_JACOBI_NT64, 0, st>>>(evals, evecs, cov, e32, T); break;
        case 128:
        case 256:
            if (!work) return fail(HV_ERR_INVALID_ARG, "T > 64 needs an eigensolver workspace");
            return T == 128 ? eig_deferred<128, CT, E>(evals, evecs, cov, (char*)work, st, e32)
                            : eig_deferred<256, CT, double2>(evals, evecs, cov, (char*)work, st, e32);
        default: return fail(HV_ERR_UNSUPPORTED, "T must be a power of two in [1, 256]");
    }
    HV_CHECK_LAUNCH("k_jacobi");
    return HV_OK;
}

hv_status launch_eig_jacobi(double* evals, double2* evecs, const void* cov, int cov_double, int T, void* work,
                            hipStream_t st, float2* evecs32, int real) {
    if (real) {
        if (!cov_double) return fail(HV_ERR_INVALID_ARG, "the real eigensolver takes a double2 covariance");
        return eig_jacobi<double2, double>(evals, evecs, (const double2*)cov, T, work, st, evecs32);
    }
    return cov_double ? eig_jacobi(evals, evecs, (const double2*)cov, T, work, st, evecs32)
                      : eig_jacobi(evals, evecs, (const float2*)cov, T, work, st, evecs32);
}

// ---------------------------------------------------------------------------
// projection C[:, t] = H v_t, t in [pmin, pmax], fused with the image type or
// a moment
// ---------------------------------------------------------------------------
// What a pixel's image accumulates from component t of its projection.
//   image types (converts.cc:22-63): modulus / squared modulus / argument,
//     then the mean over [pmin, pmax] and x131071 (converts.cc:50-61);
//   moments (pipe.cc:429-455 over the PCA output, tools_conversion.cu:33-66,
//     tools_compute.cu:38-97): |C_t|^2 weighted by 1, f_t, f_t^2 with
//     f = fftfreq(T) * fps / T, the Nyquist component (T even) counted twice
//     in M0 and M2 and cancelled in M1; no scale (the reference's x131071
//     acts on the frame the moment copy then overwrites).
struct PcaOut {
    int kind;   // HV_IMG_*
    int T;
    float fps;
    __device__ __forceinline__ bool moment() const { return kind >= HV_IMG_MOMENTS_0; }
    __device__ __forceinline__ float weight(int t) const {
        const float d = fps / T;
        const float f = (T % 2 == 0) ? (t < T / 2 ? t * d : -((float)T - t) * d)
                                     : (t < (T + 1) / 2 ? t * d : (t - (float)T) * d);
        const bool nyq = (T % 2 == 0) && t == T / 2;
        if (kind == HV_IMG_MOMENTS_0) return nyq ? 2.f : 1.f;
        if (kind == HV_IMG_MOMENTS_1) return nyq ? 0.f : f;
        return nyq ? 2.f * (f * f) : f * f;
    }
    __device__ __forceinline__ float term(float2 z, int t) const {
        if (moment()) return weight(t) * (z.x * z.x + z.y * z.y);
        if (kind == HV_IMG_MODULUS) return hypotf(z.x, z.y);
        if (kind == HV_IMG_SQUARED_MODULUS) {
            const float h = hypotf(z.x, z.y);
            return h * h;
        }
        return atan2f(z.y, z.x);
    }
    __device__ __forceinline__ float finish(float acc, int nsel) const {
        if (moment()) return acc;
        return acc / nsel * (float)((2 << 16) - 1);
    }
};

// T <= 64: the pixel's window in registers, every selected eigenvector in LDS.
template <int T>
__global__ void __launch_bounds__(256)
    k_pca_image(float* __restrict__ out, const float2* __restrict__ ring, const double2* __restrict__ evecs, size_t P,
                PcaOut po, int pmin, int pmax, uint32_t* mm, int mm_slots) {
    extern __shared__ float2 vs[];  // T x nsel, s fastest
    const int nsel = pmax - pmin + 1;
    for (int e = threadIdx.x; e < T * nsel; e += blockDim.x) {
        const int s = e % T, t = pmin + e / T;
        const double2 d = evecs[s + (size_t)t * T];
        vs[e] = make_float2((float)d.x, (float)d.y);
    }
    __syncthreads();
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < P;
    float2 x[T];
#pragma unroll
    for (int s = 0; s < T; ++s) x[s] = valid ? ring[(size_t)s * P + p] : make_float2(0.f, 0.f);
    float acc = 0.f;
    for (int t = 0; t < nsel; ++t) {
        float re = 0.f, im = 0.f;
#pragma unroll
        for (int s = 0; s < T; ++s) {
            const float2 w = vs[t * T + s];
            re += x[s].x * w.x - x[s].y * w.y;
            im += x[s].x * w.y + x[s].y * w.x;
        }
        acc += po.term(make_float2(re, im), pmin + t);
    }
    const float val = po.finish(acc, nsel);
    float vmin = INFINITY, vmax = -INFINITY;
    if (valid) {
        out[p] = val;
        vmin = vmax = val;
    }
    if (mm) block_minmax_atomic(vmin, vmax, mm, mm_slots);
}

// T > 64: the window does not fit in registers.  The selected components go
// in chunks of kPcaChunk; per chunk the block stages those eigenvector
// columns in LDS and each thread streams its pixel's window once, keeping
// the chunk's accumulators in registers (the ring is read once per chunk,
// not once per component).
template <int kPcaChunk>
__global__ void __launch_bounds__(256)
    k_pca_image_big(float* __restrict__ out, const float2* __restrict__ ring, const double2* __restrict__ evecs, int T,
                    size_t P, PcaOut po, int pmin, int pmax, uint32_t* mm, int mm_slots) {
    extern __shared__ float2 vs[];  // T x kPcaChunk, component fastest
    const int nsel = pmax - pmin + 1;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < P;
    const float2* xp = ring + (valid ? p : 0);
    float acc = 0.f;
    for (int c0 = 0; c0 < nsel; c0 += kPcaChunk) {
        const int nc = min(kPcaChunk, nsel - c0);
        __syncthreads();  // the previous chunk's columns have been read
        for (int e = threadIdx.x; e < T * kPcaChunk; e += blockDim.x) {
            const int j = e % kPcaChunk, sidx = e / kPcaChunk;
            const double2 d = j < nc ? evecs[sidx + (size_t)(pmin + c0 + j) * T] : make_double2(0.0, 0.0);
            vs[e] = make_float2((float)d.x, (float)d.y);
        }
        __syncthreads();
        float re[kPcaChunk], im[kPcaChunk];
#pragma unroll
        for (int j = 0; j < kPcaChunk; ++j) re[j] = im[j] = 0.f;
#pragma unroll 2
        for (int sidx = 0; sidx < T; ++sidx) {
            const float2 xs = xp[(size_t)sidx * P];
            const float2* w = vs + sidx * kPcaChunk;
#pragma unroll
            for (int j = 0; j < kPcaChunk; ++j) {
                const float2 v = w[j];
                re[j] += xs.x * v.x - xs.y * v.y;
                im[j] += xs.x * v.y + xs.y * v.x;
            }
        }
#pragma unroll
        for (int j = 0; j < kPcaChunk; ++j)
            if (j < nc) acc += po.term(make_float2(re[j], im[j]), pmin + c0 + j);
    }
    const float val = po.finish(acc, nsel);
    float vmin = INFINITY, vmax = -INFINITY;
    if (valid) {
        out[p] = val;
        vmin = vmax = val;
    }
    if (mm) block_minmax_atomic(vmin, vmax, mm, mm_slots);
}

// T >= 32: the projection on MFMA (v_mfma_f32_32x32x2_f32, FP32 products and
// sums as on the VALU).  A wavefront owns 32 pixels; per chunk of 32
// components C (32 px x 32 components) = X (32 x T) V (T x 32) in steps of 2
// frames: Re C = Xr Vr - Xi Vi, Im C = Xr Vi + Xi Vr.  Lane l feeds
// X[pixel l % 32][frame 2 s + l / 32] as the A operand and V[frame 2 s + l / 32]
// [component l % 32] (the chunk's columns staged in LDS) as the B operand; its
// 16 outputs are C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  Each output's term
// is summed over the components (the 32 lanes of a half-wavefront) by
// shuffles.  NC chunks are held at once so that the window is read once for up
// to 32 NC components.
// M3: the three moments at once (out, out1, out2), each weighted |projection|^2
// summed exactly as the single-moment kernel sums it.
template <int T, int NC, bool M3 = false>
__global__ void __launch_bounds__(256)
    k_pca_image_mfma(float* __restrict__ out, const float2* __restrict__ ring, const double2* __restrict__ evecs,
                     size_t P, PcaOut po, int pmin, int pmax, uint32_t* mm, int mm_slots, float* __restrict__ out1 = nullptr,
                     float* __restrict__ out2 = nullptr) {
    constexpr int W = 32 * NC;  // components staged per pass
    __shared__ float2 vs[T * W];   // [frame][component]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 31, half = lane >> 5;
    const size_t px0 = ((size_t)blockIdx.x * 4 + wave) * 32;
    const bool tile_ok = px0 + 32 <= P;
    const float2* xp = ring + (size_t)half * P + (tile_ok ? px0 : 0) + col;
    const int nsel = pmax - pmin + 1;
    float rowacc[16];
    float rowacc1[M3 ? 16 : 1], rowacc2[M3 ? 16 : 1];
#pragma unroll
    for (int r = 0; r < 16; ++r) rowacc[r] = 0.f;
    if constexpr (M3) {
#pragma unroll
        for (int r = 0; r < 16; ++r) rowacc1[r] = rowacc2[r] = 0.f;
    }
    const PcaOut po1{HV_IMG_MOMENTS_1, po.T, po.fps}, po2{HV_IMG_MOMENTS_2, po.T, po.fps};
    for (int c0 = 0; c0 < nsel; c0 += W) {
        __syncthreads();  // the previous pass's columns have been read
        for (int e = threadIdx.x; e < T * W; e += blockDim.x) {
            const int j = e % W, fr = e / W, comp = pmin + c0 + j;
            const double2 d = comp <= pmax ? evecs[fr + (size_t)comp * T] : make_double2(0.0, 0.0);
            vs[e] = make_float2((float)d.x, (float)d.y);
        }
        __syncthreads();
        v16f cr[NC], ci[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) cr[c] = ci[c] = v16f{};
#pragma unroll 4
        for (int st = 0; st < T / 2; ++st) {
            const float2 x = xp[(size_t)(2 * st) * P];
            const float2* vrow = vs + (2 * st + half) * W + col;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const float2 v = vrow[32 * c];
                cr[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.x, v.x, cr[c], 0, 0, 0);
                cr[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(-x.y, v.y, cr[c], 0, 0, 0);
                ci[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.x, v.y, ci[c], 0, 0, 0);
                ci[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(x.y, v.x, ci[c], 0, 0, 0);
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int j = c0 + 32 * c + col;  // selected component of this lane's outputs
            const bool ok = j < nsel;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float2 z = make_float2(cr[c][r], ci[c][r]);
                float tv = ok ? po.term(z, pmin + j) : 0.f;
#pragma unroll
                for (int o = 16; o >= 1; o >>= 1) tv += __shfl_xor(tv, o, 64);
                rowacc[r] += tv;
                if constexpr (M3) {
                    float t1 = ok ? po1.term(z, pmin + j) : 0.f;
                    float t2 = ok ? po2.term(z, pmin + j) : 0.f;
#pragma unroll
                    for (int o = 16; o >= 1; o >>= 1) {
                        t1 += __shfl_xor(t1, o, 64);
                        t2 += __shfl_xor(t2, o, 64);
                    }
                    rowacc1[r] += t1;
                    rowacc2[r] += t2;
                }
            }
        }
    }
    // lane (half, col < 16) writes pixel row 8 (col / 4) + 4 half + col % 4
    float v = 0.f, v1 = 0.f, v2 = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r)
        if (r == col) {
            v = rowacc[r];
            if constexpr (M3) {
                v1 = rowacc1[r];
                v2 = rowacc2[r];
            }
        }
    float vmin = INFINITY, vmax = -INFINITY;
    if (tile_ok && col < 16) {
        const float val = po.finish(v, nsel);
        const size_t o = px0 + 8 * (col / 4) + 4 * half + (col % 4);
        out[o] = val;
        if constexpr (M3) {
            out1[o] = v1;
            out2[o] = v2;
        }
        vmin = vmax = val;
    }
    if (mm) block_minmax_atomic(vmin, vmax, mm, mm_slots);
}

#ifndef HV_PCA_PROJ_MFMA
#define HV_PCA_PROJ_MFMA 1      // T >= 32: projection on MFMA
#endif

template <int T>
static hv_status pca_image_mfma(float* out, const float2* ring, const double2* evecs, size_t P, const PcaOut& po,
                                int pmin, int pmax, uint32_t* mm, hipStream_t st, int mm_slots) {
    const unsigned g = (unsigned)((P + 127) / 128);
    // two chunks at once (64 components, T x 64 float2 of LDS) up to T = 128
    if (T <= 128 && pmax - pmin + 1 > 32)
        k_pca_image_mfma<T, (T <= 128 ? 2 : 1)><<<g, 256, 0, st>>>(out, ring, evecs, P, po, pmin, pmax, mm, mm_slots);
    else
        k_pca_image_mfma<T, 1><<<g, 256, 0, st>>>(out, ring, evecs, P, po, pmin, pmax, mm, mm_slots);
    HV_CHECK_LAUNCH("k_pca_image_mfma");
    return HV_OK;
}

template <int T>
static hv_status pca_moments3_mfma(float* m0, float* m1, float* m2, const float2* ring, const double2* evecs, size_t P,
                                   const PcaOut& po, int fs, int fe, hipStream_t st) {
    const unsigned g = (unsigned)((P + 127) / 128);
    if (T <= 128 && fe - fs + 1 > 32)
        k_pca_image_mfma<T, (T <= 128 ? 2 : 1), true>
            <<<g, 256, 0, st>>>(m0, ring, evecs, P, po, fs, fe, nullptr, 1, m1, m2);
    else
        k_pca_image_mfma<T, 1, true><<<g, 256, 0, st>>>(m0, ring, evecs, P, po, fs, fe, nullptr, 1, m1, m2);
    HV_CHECK_LAUNCH("k_pca_image_mfma<moments>");
    return HV_OK;
}

hv_status launch_pca_moments3(float* m0, float* m1, float* m2, const float2* ring, const double2* evecs, int T,
                              size_t P, int fs, int fe, float fps, hipStream_t st) {
    const PcaOut po{HV_IMG_MOMENTS_0, T, fps};
    if (HV_PCA_PROJ_MFMA && T >= 32 && P % 128 == 0) {
        switch (T) {
            case 32: return pca_moments3_mfma<32>(m0, m1, m2, ring, evecs, P, po, fs, fe, st);
            case 64: return pca_moments3_mfma<64>(m0, m1, m2, ring, evecs, P, po, fs, fe, st);
            case 128: return pca_moments3_mfma<128>(m0, m1, m2, ring, evecs, P, po, fs, fe, st);
            default: break;
        }
    }
    // one projection per moment elsewhere
    for (int k = 0; k < 3; ++k)
        HV_TRY(launch_pca_image(k == 0 ? m0 : (k == 1 ? m1 : m2), ring, evecs, T, P, HV_IMG_MOMENTS_0 + k, fs, fe,
                                nullptr, st, 1, fps));
    return HV_OK;
}

hv_status launch_pca_image(float* out, const float2* ring, const double2* evecs, int T, size_t P, int img_type,
                           int pmin, int pmax, uint32_t* mm, hipStream_t st, int mm_slots, float fps) {
    const unsigned g = (unsigned)((P + 255) / 256);
    const PcaOut po{img_type, T, fps};
    if (HV_PCA_PROJ_MFMA && T >= 32 && P % 128 == 0) {
        switch (T) {
            case 32: return pca_image_mfma<32>(out, ring, evecs, P, po, pmin, pmax, mm, st, mm_slots);
            case 64: return pca_image_mfma<64>(out, ring, evecs, P, po, pmin, pmax, mm, st, mm_slots);
            case 128: return pca_image_mfma<128>(out, ring, evecs, P, po, pmin, pmax, mm, st, mm_slots);
            default: return pca_image_mfma<256>(out, ring, evecs, P, po, pmin, pmax, mm, st, mm_slots);
        }
    }
    if (T <= 64 && is_pow2(T)) {
        const size_t shm = (size_t)T * (pmax - pmin + 1) * sizeof(float2);
        if (shm > 64 * 1024) return fail(HV_ERR_UNSUPPORTED, "PCA eigenvector range too large");
        switch (T) {
            case 1: k_pca_image<1><<<g, 256, shm, st>>>(out, ring, evecs, P, po, pmin, pmax, mm, mm_slots); break;
            case 2: k_pca_image<2><<<g, 256, shm, st>>>(out, ring, evecs, P, po, pmin, pmax, mm, mm_slots); break;
            case 4: k_pca_image<4><<<g, 256, shm, st>>>(out, ring, evecs, P, po, pmin, pmax, mm, mm_slots); break;
            case 8: k_pca_image<8><<<g, 256, shm, st>>>(out, ring, evecs, P, po, pmin, pmax, mm, mm_slots); break;
            case 16: k_pca_image<16><<<g, 256, shm, st>>>(out, ring, evecs, P, po, pmin, pmax, mm, mm_slots); break;
            case 32: k_pca_image<32><<<g, 256, shm, st>>>(out, ring, evecs, P, po, pmin, pmax, mm, mm_slots); break;
            default: k_pca_image<64><<<g, 256, shm, st>>>(out, ring, evecs, P, po, pmin, pmax, mm, mm_slots); break;
        }
    } else {
        // 32 components per chunk (64 held 151 registers and ran slower despite one pass less)
        // T x 32 float2 = 64 KiB at T = 256
        const size_t shm = (size_t)T * 32 * sizeof(float2);
        k_pca_image_big<32><<<g, 256, shm, st>>>(out, ring, evecs, T, P, po, pmin, pmax, mm, mm_slots);
    }
    HV_CHECK_LAUNCH("k_pca_image");
    return HV_OK;
}

}  // namespace hv
