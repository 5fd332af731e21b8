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
// Eigendecomposition.  cusolverDnCheevd (lower, ascending) is replaced by a
// cyclic Jacobi in FP64 on one workgroup with the round-robin (circle)
// ordering: T/2 disjoint rotations per round, T - 1 rounds per sweep.
//
// Projection.  C = H V restricted to the columns [pmin, pmax] the image uses,
// fused with the image type, the mean, the x131071 PCA scale
// (converts.cc:50-61) and the image min/max.
#include <cfloat>

#include "hv_kernels.hpp"

namespace hv {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// covariance
// ---------------------------------------------------------------------------
static int pca_groups(int T) { return T < 32 ? 1 : T / 32; }
static int pca_pairs(int T) {
    const int g = pca_groups(T);
    return g * (g + 1) / 2;
}
static constexpr int kCovWaves = 4;
static constexpr int kTile = 2 * 32 * 32;  // floats of one wave's two accumulators

// Pixel slices per frame-group pair: about 512 workgroups in all, each
// wave covering at least 256 pixels.
static int pca_slices(int T, size_t P) {
    const int np = pca_pairs(T);
    int ns = 1;
    while (np * ns * 2 <= 512 && (size_t)(ns * 2) * kCovWaves * 256 <= P) ns *= 2;
    return ns;
}

struct CovLayout {
    size_t partial_floats;  // slices * pairs * kTile
    size_t offset_scale;    // byte offset of the FP16 scale word
    size_t bytes;
};
static CovLayout cov_layout(int T, size_t P) {
    CovLayout l;
    l.partial_floats = (size_t)pca_slices(T, P) * pca_pairs(T) * kTile;
    l.offset_scale = l.partial_floats * sizeof(float);
    l.bytes = l.offset_scale + 256;
    return l;
}

size_t pca_workspace_bytes(int T, size_t P, int precision) {
    (void)precision;
    return cov_layout(T, P).bytes;
}

__device__ __forceinline__ void pair_of(int idx, int& F, int& G) {
    // pairs enumerated (0,0), (1,0), (1,1), (2,0), ... with F >= G
    F = 0;
    while ((F + 1) * (F + 2) / 2 <= idx) ++F;
    G = idx - F * (F + 1) / 2;
}

// Max |component| of the ring, as an ordered uint (FP16 scaling).
__global__ void __launch_bounds__(256) k_ring_absmax(uint32_t* out, const float* __restrict__ ring, size_t n) {
    float m = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(ring[i]));
    block_minmax_atomic(m, m, out);
}

__global__ void k_scale_init(uint32_t* mm) {
    if (threadIdx.x == 0) {
        mm[0] = float_to_ordered(INFINITY);
        mm[1] = float_to_ordered(0.f);
    }
}

// Power of two bringing max |x| to about 2^12 (FP16 max 65504; products of
// 2^12-sized values stay far from the FP32 accumulator's limits).
__device__ __forceinline__ float fp16_scale(const uint32_t* mm) {
    const float mx = ordered_to_float(mm[1]);
    if (!(mx > 0.f) || !isfinite(mx)) return 1.0f;
    const int e = ilogbf(mx);
    return ldexpf(1.0f, 12 - e);
}

template <int PREC>
struct CovStep {
    static constexpr int K = (PREC == HV_PCA_FP32) ? 2 : 16;      // pixels per MFMA
    static constexpr int KPL = (PREC == HV_PCA_FP32) ? 1 : 8;     // pixels per lane
};

template <int PREC>
__global__ void __launch_bounds__(64 * kCovWaves)
    k_pca_cov(float* __restrict__ partial, const float2* __restrict__ ring, int T, size_t P, int slices,
              const uint32_t* __restrict__ scale_mm) {
    using S = CovStep<PREC>;
    const int pair = blockIdx.x, slice = blockIdx.y;
    const int npairs = gridDim.x;
    int F, G;
    pair_of(pair, F, G);
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

    v16f acc_re = {}, acc_im = {};
    for (size_t p0 = p_begin; p0 < p_end; p0 += S::K) {
        if constexpr (PREC == HV_PCA_FP32) {
            const float2 a = okF ? rowF[p0 + koff] : make_float2(0.f, 0.f);
            const float2 b = okG ? rowG[p0 + koff] : make_float2(0.f, 0.f);
            acc_re = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b.x, acc_re, 0, 0, 0);
            acc_re = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b.y, acc_re, 0, 0, 0);
            acc_im = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b.y, acc_im, 0, 0, 0);
            acc_im = __builtin_amdgcn_mfma_f32_32x32x2f32(-a.y, b.x, acc_im, 0, 0, 0);
        } else {
            float2 xa[8], xb[8];
            {
                const float4* pa = reinterpret_cast<const float4*>(rowF + p0 + koff);
                const float4* pb = reinterpret_cast<const float4*>(rowG + p0 + koff);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float4 va = okF ? pa[q] : make_float4(0.f, 0.f, 0.f, 0.f);
                    float4 vb = okG ? pb[q] : make_float4(0.f, 0.f, 0.f, 0.f);
                    xa[2 * q] = make_float2(va.x * scale, va.y * scale);
                    xa[2 * q + 1] = make_float2(va.z * scale, va.w * scale);
                    xb[2 * q] = make_float2(vb.x * scale, vb.y * scale);
                    xb[2 * q + 1] = make_float2(vb.z * scale, vb.w * scale);
                }
            }
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
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        red[wave][r * 64 + lane] = acc_re[r];
        red[wave][1024 + r * 64 + lane] = acc_im[r];
    }
    __syncthreads();
    float* out = partial + ((size_t)slice * npairs + pair) * kTile;
    for (int i = threadIdx.x; i < kTile; i += blockDim.x) {
        float s = 0.f;
#pragma unroll
        for (int w = 0; w < kCovWaves; ++w) s += red[w][i];
        out[i] = s;
    }
}

// Sum the slices and write the full Hermitian matrix, column-major.
// Tile element (acc, r, lane) is row 8 (r / 4) + 4 (lane / 32) + r % 4 and
// column lane % 32 of the 32 x 32 MFMA output.  Each workgroup owns 64
// elements of one pair's tile; its four waves sum every fourth slice, then
// one wave adds the four (a fixed order: deterministic).
__global__ void __launch_bounds__(256)
    k_pca_cov_reduce(float2* __restrict__ cov, const float* __restrict__ partial, int T, int slices, int npairs,
                     const uint32_t* __restrict__ scale_mm, int use_scale) {
    const int pair = blockIdx.x;
    int F, G;
    pair_of(pair, F, G);
    const int g = threadIdx.x >> 6, el = threadIdx.x & 63;
    const int e = blockIdx.y * 64 + el;  // 0 .. 1023
    float re = 0.f, im = 0.f;
    for (int sl = g; sl < slices; sl += 4) {
        const float* t = partial + ((size_t)sl * npairs + pair) * kTile;
        re += t[e];
        im += t[1024 + e];
    }
    __shared__ float sr[4][64], si[4][64];
    sr[g][el] = re;
    si[g][el] = im;
    __syncthreads();
    if (g != 0) return;
    re = (sr[0][el] + sr[1][el]) + (sr[2][el] + sr[3][el]);
    im = (si[0][el] + si[1][el]) + (si[2][el] + si[3][el]);
    if (use_scale) {
        const float sc = fp16_scale(scale_mm);
        const float inv = 1.0f / (sc * sc);
        re *= inv;
        im *= inv;
    }
    const int r = e / 64, lane = e % 64;
    const int i = 8 * (r / 4) + 4 * (lane / 32) + (r % 4), jj = lane % 32;
    const int fi = 32 * F + i, fj = 32 * G + jj;
    if (fi >= T || fj >= T) return;
    cov[fi + (size_t)fj * T] = make_float2(re, im);
    if (F != G) cov[fj + (size_t)fi * T] = make_float2(re, -im);
}

hv_status launch_pca_cov(float2* cov, const float2* ring, int T, size_t P, int precision, void* ws,
                         hipStream_t st) {
    const CovLayout l = cov_layout(T, P);
    const int slices = pca_slices(T, P), np = pca_pairs(T);
    float* partial = reinterpret_cast<float*>(ws);
    uint32_t* scale_mm = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ws) + l.offset_scale);
    const bool fp16 = precision == HV_PCA_FP16 || precision == HV_PCA_FP16_SPLIT;
    if (P % ((size_t)slices * kCovWaves * 16)) return fail(HV_ERR_UNSUPPORTED, "frame_res not a multiple of the PCA tile");
    if (fp16) {
        k_scale_init<<<1, 64, 0, st>>>(scale_mm);
        const size_t n = (size_t)T * P * 2;
        k_ring_absmax<<<(unsigned)std::min<size_t>((n + 255) / 256, 4096), 256, 0, st>>>(scale_mm, (const float*)ring, n);
        HV_CHECK_LAUNCH("k_ring_absmax");
    }
    dim3 grid(np, slices);
    switch (precision) {
        case HV_PCA_FP32: k_pca_cov<HV_PCA_FP32><<<grid, 64 * kCovWaves, 0, st>>>(partial, ring, T, P, slices, scale_mm); break;
        case HV_PCA_BF16_SPLIT: k_pca_cov<HV_PCA_BF16_SPLIT><<<grid, 64 * kCovWaves, 0, st>>>(partial, ring, T, P, slices, scale_mm); break;
        case HV_PCA_FP16_SPLIT: k_pca_cov<HV_PCA_FP16_SPLIT><<<grid, 64 * kCovWaves, 0, st>>>(partial, ring, T, P, slices, scale_mm); break;
        case HV_PCA_BF16: k_pca_cov<HV_PCA_BF16><<<grid, 64 * kCovWaves, 0, st>>>(partial, ring, T, P, slices, scale_mm); break;
        case HV_PCA_FP16: k_pca_cov<HV_PCA_FP16><<<grid, 64 * kCovWaves, 0, st>>>(partial, ring, T, P, slices, scale_mm); break;
        default: return fail(HV_ERR_INVALID_ARG, "unknown PCA precision");
    }
    HV_CHECK_LAUNCH("k_pca_cov");
    k_pca_cov_reduce<<<dim3(np, 1024 / 64), 256, 0, st>>>(cov, partial, T, slices, np, scale_mm, fp16 ? 1 : 0);
    HV_CHECK_LAUNCH("k_pca_cov_reduce");
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

#ifndef HV_JACOBI_MAX_SWEEPS
#define HV_JACOBI_MAX_SWEEPS 40
#endif
#ifndef HV_JACOBI_NT32
#define HV_JACOBI_NT32 512      // threads of the T <= 32 eigensolver
#endif
#ifndef HV_JACOBI_NT64
#define HV_JACOBI_NT64 512      // threads of the T <= 64 eigensolver
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

template <int NT, int N>
__device__ void jacobi_run(double2* A, double2* V, int ld, double* evals, double2* evecs, Rot* rot, double* red,
                           int* order) {
    constexpr int n = N, half = N / 2;
    const int tid = threadIdx.x;
    auto a = [&](int i, int j) -> double2& { return A[i + (size_t)j * ld]; };
    auto v = [&](int i, int j) -> double2& { return V[i + (size_t)j * ld]; };
    double fro = 0.0;
    for (int e = tid; e < n * n; e += NT) {
        const double2 x = a(e % n, e / n);
        fro += x.x * x.x + x.y * x.y;
    }
    fro = block_sum<NT>(fro, red);
    for (int sweep = 0; sweep < HV_JACOBI_MAX_SWEEPS && n > 1; ++sweep) {
        double off = 0.0;
        for (int e = tid; e < n * n; e += NT) {
            const int i = e % n, j = e / n;
            if (i != j) {
                const double2 x = a(i, j);
                off += x.x * x.x + x.y * x.y;
            }
        }
        off = block_sum<NT>(off, red);
        // relative off-diagonal norm below 1e-13: far under the FP32 input's own precision
        if (off <= 1e-26 * fro || off == 0.0) break;
        for (int r = 0; r < n - 1; ++r) {
            for (int k = tid; k < half; k += NT) {
                int p, q;
                rr_pair(n, r, k, p, q);
                const double2 apq = a(p, q);
                const double mag2 = apq.x * apq.x + apq.y * apq.y;
                Rot R;
                if (mag2 == 0.0) {
                    R.c = 1.0;
                    R.s = 0.0;
                    R.ec = make_double2(1.0, 0.0);
                } else {
                    // t = sgn(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c
                    const double inv = 1.0 / sqrt(mag2);
                    R.ec = make_double2(apq.x * inv, -apq.y * inv);
                    const double tau = (a(q, q).x - a(p, p).x) * (0.5 * inv);
                    const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    R.c = 1.0 / sqrt(1.0 + t * t);
                    R.s = t * R.c;
                }
                rot[k] = R;
            }
            __syncthreads();
            // A <- J^H A J, block by block; V <- V J, row pair by row pair
            for (int task = tid; task < half * half; task += NT) {
                const int k = task % half, kk = task / half;  // row pair k, column pair kk
                int p, q, pp, qq;
                rr_pair(n, r, k, p, q);
                rr_pair(n, r, kk, pp, qq);
                const Rot Rr = rot[k], Rc = rot[kk];
                double2 b00 = a(p, pp), b01 = a(p, qq), b10 = a(q, pp), b11 = a(q, qq);
                rot_right(Rc, b00, b01);
                rot_right(Rc, b10, b11);
                rot_left(Rr, b00, b10);
                rot_left(Rr, b01, b11);
                if (k == kk) {  // the annihilated pair: exact zeros, real diagonal
                    b01 = make_double2(0.0, 0.0);
                    b10 = make_double2(0.0, 0.0);
                    b00.y = 0.0;
                    b11.y = 0.0;
                }
                a(p, pp) = b00;
                a(p, qq) = b01;
                a(q, pp) = b10;
                a(q, qq) = b11;
            }
            for (int task = tid; task < half * n; task += NT) {
                const int m = task % n, kk = task / n;
                int pp, qq;
                rr_pair(n, r, kk, pp, qq);
                double2 x = v(m, pp), y = v(m, qq);
                rot_right(rot[kk], x, y);
                v(m, pp) = x;
                v(m, qq) = y;
            }
            __syncthreads();
        }
    }
    // ascending order (stable: ties keep their index order), by parallel ranks
    for (int j = tid; j < n; j += NT) {
        const double dj = a(j, j).x;
        int rank = 0;
#pragma unroll 4
        for (int i = 0; i < n; ++i) {
            const double di = a(i, i).x;
            rank += (di < dj || (di == dj && i < j)) ? 1 : 0;
        }
        order[rank] = j;
    }
    __syncthreads();
    // phase-normalised eigenvectors: the largest component real positive
    // (the first of equal magnitudes within a relative 1e-12)
    for (int jj = tid; jj < n; jj += NT) {
        const int j = order[jj];
        evals[jj] = a(j, j).x;
        int kmax = 0;
        double best = -1.0;
#pragma unroll 4
        for (int k = 0; k < n; ++k) {
            const double2 x = v(k, j);
            const double m2 = x.x * x.x + x.y * x.y;
            if (m2 > best * (1 + 2e-12)) {
                best = m2;
                kmax = k;
            }
        }
        const double2 ph = v(kmax, j);
        const double inv = 1.0 / sqrt(ph.x * ph.x + ph.y * ph.y);
        const double2 rotp = make_double2(ph.x * inv, -ph.y * inv);
#pragma unroll 4
        for (int k = 0; k < n; ++k) evecs[k + (size_t)jj * n] = zmul(v(k, j), rotp);
    }
}

template <int NT, int N>
__global__ void __launch_bounds__(NT) k_jacobi_lds(double* evals, double2* evecs, const float2* cov) {
    constexpr int LD = N + 1;  // padded columns: row and column walks spread over the banks
    __shared__ double2 A[N * LD], V[N * LD];
    __shared__ Rot rot[N / 2 > 0 ? N / 2 : 1];
    __shared__ double red[NT / 64];
    __shared__ int order[N];
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        // Hermitian from the lower triangle (CUBLAS_FILL_MODE_LOWER)
        const float2 c = i >= j ? cov[i + (size_t)j * N] : cov[j + (size_t)i * N];
        A[i + j * LD] = make_double2(c.x, i == j ? 0.0 : (i > j ? c.y : -c.y));
        V[i + j * LD] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    jacobi_run<NT, N>(A, V, LD, evals, evecs, rot, red, order);
}

template <int NT, int N>
__global__ void __launch_bounds__(NT) k_jacobi_global(double* evals, double2* evecs, const float2* cov,
                                                     double2* work) {
    __shared__ Rot rot[N / 2];
    __shared__ double red[NT / 64];
    __shared__ int order[N];
    double2* A = work;
    double2* V = work + (size_t)N * N;
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        const float2 c = i >= j ? cov[i + (size_t)j * N] : cov[j + (size_t)i * N];
        A[e] = make_double2(c.x, i == j ? 0.0 : (i > j ? c.y : -c.y));
        V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    jacobi_run<NT, N>(A, V, N, evals, evecs, rot, red, order);
}

size_t eig_workspace_bytes(int T) { return T <= 64 ? 0 : (size_t)2 * T * T * sizeof(double2); }

hv_status launch_eig_jacobi(double* evals, double2* evecs, const float2* cov, int T, void* work, hipStream_t st) {
    switch (T) {
        case 1: k_jacobi_lds<64, 1><<<1, 64, 0, st>>>(evals, evecs, cov); break;
        case 2: k_jacobi_lds<64, 2><<<1, 64, 0, st>>>(evals, evecs, cov); break;
        case 4: k_jacobi_lds<64, 4><<<1, 64, 0, st>>>(evals, evecs, cov); break;
        case 8: k_jacobi_lds<64, 8><<<1, 64, 0, st>>>(evals, evecs, cov); break;
        case 16: k_jacobi_lds<256, 16><<<1, 256, 0, st>>>(evals, evecs, cov); break;
        case 32: k_jacobi_lds<HV_JACOBI_NT32, 32><<<1, HV_JACOBI_NT32, 0, st>>>(evals, evecs, cov); break;
        case 64: k_jacobi_lds<HV_JACOBI_NT64, 64><<<1, HV_JACOBI_NT64, 0, st>>>(evals, evecs, cov); break;
        case 128:
        case 256:
            if (!work) return fail(HV_ERR_INVALID_ARG, "T > 64 needs an eigensolver workspace");
            if (T == 128) k_jacobi_global<1024, 128><<<1, 1024, 0, st>>>(evals, evecs, cov, (double2*)work);
            else k_jacobi_global<1024, 256><<<1, 1024, 0, st>>>(evals, evecs, cov, (double2*)work);
            break;
        default: return fail(HV_ERR_UNSUPPORTED, "T must be a power of two in [1, 256]");
    }
    HV_CHECK_LAUNCH("k_jacobi");
    return HV_OK;
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
constexpr int kPcaChunk = 32;

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

hv_status launch_pca_image(float* out, const float2* ring, const double2* evecs, int T, size_t P, int img_type,
                           int pmin, int pmax, uint32_t* mm, hipStream_t st, int mm_slots, float fps) {
    const unsigned g = (unsigned)((P + 255) / 256);
    const PcaOut po{img_type, T, fps};
    if (T <= 64) {
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
        // T x kPcaChunk float2 = 64 KiB at T = 256
        const size_t shm = (size_t)T * kPcaChunk * sizeof(float2);
        k_pca_image_big<<<g, 256, shm, st>>>(out, ring, evecs, T, P, po, pmin, pmax, mm, mm_slots);
    }
    HV_CHECK_LAUNCH("k_pca_image");
    return HV_OK;
}

}  // namespace hv
