// holo_oracle.cpp -- CPU restatement of the Holovibes per-batch compute path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is compared
// against (tests/, the smoke check and the CPU-baseline leg of the benchmark).
// Nothing in the product (holovibes_amd/) links, loads or calls it.
//
// Every function restates the reference algorithm from the cited reference
// file:line (paths relative to the Holovibes 14.8.2 source tree).  The code is
// templated on the real type: the double instantiation is the accuracy oracle,
// the float instantiation is the timed CPU baseline.  Third-party routines the
// reference calls (cuFFT, cuBLAS Cgemm3m, cuSOLVER Cheevd, thrust::sort; CUDA
// Toolkit >= 12.2, closed binaries) are restated from their mathematical
// definitions: DFT (own radix-2 FFT), complex GEMM, Hermitian eigensolver
// (own cyclic Jacobi), sort (std::sort / nth_element).
//
// Pinning (tests/test_oracle.py): FFT vs numpy.fft; Jacobi vs numpy.linalg.eigh
// and the rank-1 closed form; lenses vs closed forms; STFT vs a pure tone;
// angular spectrum vs the plane-wave identity; subsample vs the reference's
// known-answer vectors (Backend/tests/test_subsample_frame.cc:18-33, 56-69);
// reductions vs test_reduce.cc patterns; ring order vs test_queue.cc:255-319;
// unwrap vs a periodic ramp; contrast/u16 vs exact integer rules.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../include/holo_hip.h"

namespace {

template <typename R>
struct C {
    R re, im;
};

template <typename R>
inline C<R> cmul(C<R> a, C<R> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <typename R>
inline C<R> cadd(C<R> a, C<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
inline C<R> csub(C<R> a, C<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
inline C<R> cconj(C<R> a) { return {a.re, -a.im}; }

bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

// ---------------------------------------------------------------------------
// FFT: iterative radix-2, twiddles evaluated in double (DFT definition,
// X[k] = sum_n x[n] exp(dir * 2 pi i k n / N), dir = -1 forward, unnormalised,
// the cuFFT C2C convention used at angular_spectrum.cu:47,59 and stft.cu:12).
// ---------------------------------------------------------------------------
// Lengths that are not powers of two (2^a 3^b 5^c camera sides) take a
// recursive mixed-radix decimation in time: x splits by its smallest prime
// factor p into p interleaved subsequences of length m = n / p, each is
// transformed, and X[k + r m] = sum_q W_n^{q (k + r m)} Y_q[k] (a direct
// p-point combination per output; twiddles from the table of the full length).
template <typename R>
struct FftPlan {
    int n = 0;
    bool pow2 = true;
    std::vector<int> rev;
    std::vector<C<R>> tw;      // exp(-2 pi i k / n), k < n/2
    std::vector<C<R>> tw_all;  // exp(-2 pi i k / n), k < n (other lengths)
    explicit FftPlan(int n_) : n(n_), pow2(is_pow2(n_)), rev(n_), tw(n_ / 2 > 0 ? n_ / 2 : 1) {
        if (!pow2) {
            tw_all.resize(n);
            for (int k = 0; k < n; ++k) {
                double a = -2.0 * M_PI * (double)k / (double)n;
                tw_all[k] = {(R)std::cos(a), (R)std::sin(a)};
            }
            return;
        }
        int lg = 0;
        while ((1 << lg) < n) ++lg;
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < lg; ++b)
                if (i & (1 << b)) r |= 1 << (lg - 1 - b);
            rev[i] = r;
        }
        for (int k = 0; k < n / 2; ++k) {
            double a = -2.0 * M_PI * (double)k / (double)n;
            tw[k] = {(R)std::cos(a), (R)std::sin(a)};
        }
    }
};

// out[0 .. len) = DFT of in[0], in[s], ..., in[(len - 1) s]; twiddle
// W_len^e = tw_all[e (n / len) mod n]; scratch holds len values.
template <typename R>
void fft_mixed(const C<R>* in, C<R>* out, int len, int s, const FftPlan<R>& p, int dir, C<R>* scratch) {
    if (len == 1) {
        out[0] = in[0];
        return;
    }
    int f = 2;
    while (len % f) ++f;
    const int m = len / f;
    for (int q = 0; q < f; ++q) fft_mixed(in + (size_t)q * s, out + (size_t)q * m, m, s * f, p, dir, scratch);
    const int step = p.n / len;
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < f; ++r) {
            C<R> acc{0, 0};
            for (int q = 0; q < f; ++q) {
                const long e = ((long)q * (k + (long)r * m)) % len;
                C<R> w = p.tw_all[(size_t)(e * step)];
                if (dir > 0) w.im = -w.im;
                acc = cadd(acc, cmul(out[(size_t)q * m + k], w));
            }
            scratch[(size_t)r * m + k] = acc;
        }
    }
    for (int i = 0; i < len; ++i) out[i] = scratch[i];
}

template <typename R>
void fft1d(C<R>* a, const FftPlan<R>& p, int dir) {
    const int n = p.n;
    if (!p.pow2) {
        std::vector<C<R>> in(a, a + n), scratch(n);
        fft_mixed(in.data(), a, n, 1, p, dir, scratch.data());
        return;
    }
    for (int i = 0; i < n; ++i)
        if (i < p.rev[i]) std::swap(a[i], a[p.rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len)
            for (int j = 0; j < half; ++j) {
                C<R> w = p.tw[j * step];
                if (dir > 0) w.im = -w.im;
                C<R> u = a[i + j], v = cmul(a[i + j + half], w);
                a[i + j] = cadd(u, v);
                a[i + j + half] = csub(u, v);
            }
    }
}

// Batched 2-D FFT, in place, unnormalised (cufftXtMakePlanMany n={H,W},
// Backend/sources/core/icompute.cc:151-175).
template <typename R>
void fft2(C<R>* buf, int batch, int W, int H, int dir) {
    FftPlan<R> pw(W), ph(H);
    const size_t P = (size_t)W * H;
    for (int b = 0; b < batch; ++b) {
        C<R>* f = buf + b * P;
#pragma omp parallel for schedule(static)
        for (int y = 0; y < H; ++y) fft1d(f + (size_t)y * W, pw, dir);
#pragma omp parallel
        {
            std::vector<C<R>> col(H);
#pragma omp for schedule(static)
            for (int x = 0; x < W; ++x) {
                for (int y = 0; y < H; ++y) col[y] = f[(size_t)y * W + x];
                fft1d(col.data(), ph, dir);
                for (int y = 0; y < H; ++y) f[(size_t)y * W + x] = col[y];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// a3: input_queue_to_input_buffer (Backend/cuda_sources/tools_conversion.cu:107-171)
// ---------------------------------------------------------------------------
template <typename R>
void convert(const void* in, int depth, size_t P, int B, C<R>* out) {
    const size_t n = P * (size_t)B;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        float v;
        if (depth == HV_DEPTH_8) v = (float)((const uint8_t*)in)[i] * 257;  // :134-139
        else if (depth == HV_DEPTH_16) v = (float)((const uint16_t*)in)[i];  // :140-141
        else v = ((const float*)in)[i];
        out[i] = {(R)v, (R)0};
    }
}

// ---------------------------------------------------------------------------
// a5: kernel_spectral_lens (Backend/cuda_sources/masks.cu:31-59) then the
// quadrant swap (shift_corners.cu:8-44, caller fourier_transform.cc:121-134).
// ---------------------------------------------------------------------------
void lens_as_centred(int Nx, int Ny, float z, float lambda, float x_step, float y_step, int precision,
                     double* out /* Nx*Ny interleaved */) {
    for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x) {
            double c, s;
            if (precision == HV_LENS_REFERENCE_FP32) {
                // float expressions in the reference's order (masks.cu:45-57)
                float u_step = 1.0f / (Nx * x_step);
                float v_step = 1.0f / (Ny * y_step);
                float u = (x - (Nx >> 1)) * u_step;
                float v = (y - (Ny >> 1)) * v_step;
                float tmp = 1.0f - (lambda * lambda * (u * u + v * v));
                if (tmp < 0.0f) tmp = 0.0f;
                float phase = 2.0f * M_PI * z / lambda * sqrtf(tmp);
                c = (double)cosf(phase);
                s = (double)sinf(phase);
            } else {
                double u = (double)(x - (Nx >> 1)) / ((double)Nx * (double)x_step);
                double v = (double)(y - (Ny >> 1)) / ((double)Ny * (double)y_step);
                double lam = (double)lambda;
                double tmp = 1.0 - lam * lam * (u * u + v * v);
                if (tmp < 0.0) tmp = 0.0;
                double phase = 2.0 * M_PI * (double)z / lam * std::sqrt(tmp);
                c = std::cos(phase);
                s = std::sin(phase);
            }
            out[2 * ((size_t)y * Nx + x)] = c;
            out[2 * ((size_t)y * Nx + x) + 1] = s;
        }
}

// kernel_shift_corners (Backend/cuda_sources/shift_corners.cu:8-44): for each
// (i, j) with j < size_y/2, swap with (i +/- size_x/2, j + size_y/2).
template <typename T>
void shift_corners(T* buf, int batch, int sx, int sy) {
    const int sx2 = sx / 2, sy2 = sy / 2;
    const size_t P = (size_t)sx * sy;
    for (int b = 0; b < batch; ++b)
        for (int j = 0; j < sy2; ++j)
            for (int i = 0; i < sx; ++i) {
                int ni = i < sx2 ? i + sx2 : i - sx2;
                int nj = j + sy2;
                std::swap(buf[b * P + (size_t)j * sx + i], buf[b * P + (size_t)nj * sx + ni]);
            }
}

// a7: kernel_quadratic_lens (masks.cu:7-29) on a square of side lens_side,
// cropped centrally (fresnel_transform.cu:14-63).
void lens_fresnel(int side, int H, int W, float lambda, float z, float pixel_size, int precision,
                  double* out /* W*H interleaved */) {
    std::vector<double> sq((size_t)side * side * 2);
    for (int j = 0; j < side; ++j)
        for (int i = 0; i < side; ++i) {
            double c, s;
            if (precision == HV_LENS_REFERENCE_FP32) {
                const float cc = M_PI / (lambda * z);
                const float dx = pixel_size * 1.0e-6f;
                const float dy = dx;
                float x = (i - static_cast<float>(side >> 1)) * dx;
                float y = (j - static_cast<float>(side >> 1)) * dy;
                float csq = cc * (x * x + y * y);
                c = cosf(csq);
                s = sinf(csq);
            } else {
                double cc = M_PI / ((double)lambda * (double)z);
                double dx = (double)pixel_size * 1.0e-6;
                double x = (i - (double)(side >> 1)) * dx, y = (j - (double)(side >> 1)) * dx;
                double ph = cc * (x * x + y * y);
                c = std::cos(ph);
                s = std::sin(ph);
            }
            sq[2 * ((size_t)j * side + i)] = c;
            sq[2 * ((size_t)j * side + i) + 1] = s;
        }
    // central crop (fresnel_transform.cu:39-60)
    const int oy = (side - H) / 2, ox = (side - W) / 2;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t src = (size_t)(y + oy) * side + (x + ox);
            out[2 * ((size_t)y * W + x)] = sq[2 * src];
            out[2 * ((size_t)y * W + x) + 1] = sq[2 * src + 1];
        }
}

// ---------------------------------------------------------------------------
// a6 / a7: angular_spectrum (angular_spectrum.cu:33-63) and fresnel_transform
// (fresnel_transform.cu:65-83).
// ---------------------------------------------------------------------------
template <typename R>
void propagate_as(C<R>* buf, int B, int W, int H, const C<R>* lens_shifted) {
    const size_t P = (size_t)W * H;
    fft2(buf, B, W, H, -1);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < P * B; ++i) buf[i] = cmul(buf[i], lens_shifted[i % P]);  // apply_mask.cu:21-35
    fft2(buf, B, W, H, +1);
    const R n = (R)(float)P;  // kernel_complex_divide (tools_compute.cu:10-25)
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < P * B; ++i) {
        buf[i].re /= n;
        buf[i].im /= n;
    }
}

template <typename R>
void propagate_fresnel(C<R>* buf, int B, int W, int H, const C<R>* chirp) {
    const size_t P = (size_t)W * H;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < P * B; ++i) buf[i] = cmul(buf[i], chirp[i % P]);
    fft2(buf, B, W, H, -1);
}

// ---------------------------------------------------------------------------
// a9: stft (stft.cu:9-15; plan icompute.cc:109-115): per pixel, length-T
// forward FFT along time with stride P, over the ring's physical slot order.
// ---------------------------------------------------------------------------
template <typename R>
void stft(const C<R>* ring, int T, size_t P, C<R>* out) {
    FftPlan<R> pt(T);
#pragma omp parallel
    {
        std::vector<C<R>> v(T);
#pragma omp for schedule(static)
        for (long p = 0; p < (long)P; ++p) {
            for (int t = 0; t < T; ++t) v[t] = ring[(size_t)t * P + p];
            fft1d(v.data(), pt, -1);
            for (int t = 0; t < T; ++t) out[(size_t)t * P + p] = v[t];
        }
    }
}

// ---------------------------------------------------------------------------
// a12: complex_to_modulus / squared_modulus / argument
// (tools_conversion.cu:11-30, 84-103, 199-232): inclusive-range mean.
// ---------------------------------------------------------------------------
template <typename R>
void image_type(const C<R>* spec, size_t P, int type, int pmin, int pmax, R* out) {
#pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)P; ++p) {
        R val = 0;
        for (int i = pmin; i <= pmax; ++i) {
            C<R> z = spec[(size_t)i * P + p];
            if (type == HV_IMG_MODULUS) val += std::hypot(z.re, z.im);
            else if (type == HV_IMG_SQUARED_MODULUS) {
                R h = std::hypot(z.re, z.im);
                val += h * h;
            } else val += std::atan2(z.im, z.re);
        }
        out[p] = val / (R)(pmax - pmin + 1);
    }
}

// ---------------------------------------------------------------------------
// a13: moments (tools_conversion.cu:33-66, tools_compute.cu:38-97,
// icompute.cc:26-69, fourier_transform.cc:244-293, pipe.cc:429-455).
// ---------------------------------------------------------------------------
void fft_freqs(int T, float camera_fps, std::vector<float>& f1) {
    f1.assign(T, 0.f);
    float d = camera_fps / T;  // icompute.cc:29
    if (T % 2 == 0) {
        for (int i = 0; i < T / 2; ++i) f1[i] = i * d;
        for (int i = T / 2; i < T; ++i) f1[i] = -((float)T - i) * d;
    } else {
        for (int i = 0; i < (T + 1) / 2; ++i) f1[i] = i * d;
        for (int i = T - 1; i > T / 2; --i) f1[i] = (i - (float)T) * d;
    }
}

template <typename R>
void moments(const C<R>* spec, int T, size_t P, int f_start, int f_end, float fps, R* m0, R* m1, R* m2) {
    std::vector<float> f1;
    fft_freqs(T, fps, f1);
    const size_t nyq = T / 2;
    const bool even = T % 2 == 0;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)P; ++p) {
        R a0 = 0, a1 = 0, a2 = 0;
        for (int i = f_start; i <= f_end; ++i) {
            C<R> z = spec[(size_t)i * P + p];
            R s = z.re * z.re + z.im * z.im;
            a0 += s;
            a1 += s * (R)f1[i];
            a2 += s * (R)(f1[i] * f1[i]);
        }
        if (even && nyq >= (size_t)f_start && nyq <= (size_t)f_end) {
            C<R> z = spec[nyq * P + p];
            R s = z.re * z.re + z.im * z.im;
            a0 += s;
            a1 -= s * (R)f1[nyq];
            a2 += s * (R)(f1[nyq] * f1[nyq]);
        }
        if (m0) m0[p] = a0;
        if (m1) m1[p] = a1;
        if (m2) m2[p] = a2;
    }
}

// ---------------------------------------------------------------------------
// a10: PCA.  cov = H^H H (cov_matrix, matrix_operations.cc:9-26), Hermitian
// eigendecomposition (eigen_values_vectors :28-65, cusolverDnCheevd: ascending
// eigenvalues, eigenvectors as columns), C = H V (matrix_multiply_complex :77-102).
// H is the ring viewed as P x T column-major: H[p, t] = ring[t*P + p].
// ---------------------------------------------------------------------------
template <typename R>
void pca_cov(const C<R>* ring, int T, size_t P, double* cov /* T*T complex col-major */) {
    // pixel tiles of kTile pixels gathered frame-contiguous; the lower
    // triangle cov(i, j) = sum_p conj(x_i[p]) x_j[p], i >= j, accumulated in
    // double, the upper one is its conjugate
    constexpr size_t kTile = 32;
    const size_t tiles = (P + kTile - 1) / kTile;
    std::vector<double> acc((size_t)T * T * 2, 0.0);
#pragma omp parallel
    {
        std::vector<double> loc((size_t)T * T * 2, 0.0);
        std::vector<double> x(kTile * (size_t)T * 2);
#pragma omp for schedule(static)
        for (long b = 0; b < (long)tiles; ++b) {
            const size_t p0 = (size_t)b * kTile, n = std::min(kTile, P - p0);
            for (size_t k = 0; k < n; ++k)
                for (int t = 0; t < T; ++t) {
                    const C<R> v = ring[(size_t)t * P + p0 + k];
                    x[2 * (k * T + t)] = (double)v.re;
                    x[2 * (k * T + t) + 1] = (double)v.im;
                }
            for (size_t k = 0; k < n; ++k) {
                const double* xk = &x[2 * k * T];
                for (int j = 0; j < T; ++j) {
                    const double jr = xk[2 * j], ji = xk[2 * j + 1];
                    double* a = &loc[2 * (size_t)j * T];
                    for (int i = j; i < T; ++i) {
                        a[2 * i] += xk[2 * i] * jr + xk[2 * i + 1] * ji;
                        a[2 * i + 1] += xk[2 * i] * ji - xk[2 * i + 1] * jr;
                    }
                }
            }
        }
#pragma omp critical
        for (size_t k = 0; k < acc.size(); ++k) acc[k] += loc[k];
    }
    for (int j = 0; j < T; ++j)
        for (int i = 0; i < j; ++i) {  // upper triangle: cov(i, j) = conj(cov(j, i))
            acc[2 * (i + (size_t)j * T)] = acc[2 * (j + (size_t)i * T)];
            acc[2 * (i + (size_t)j * T) + 1] = -acc[2 * (j + (size_t)i * T) + 1];
        }
    std::memcpy(cov, acc.data(), acc.size() * sizeof(double));
}

// Cyclic Jacobi for a Hermitian matrix (FP64).  Rotation on (p, q):
// a_pq = |a_pq| e^{i phi}; J = diag(1, e^{-i phi}) * [[c, s], [-s, c]] zeroes
// a_pq with t = sgn(tau) / (|tau| + sqrt(1 + tau^2)), tau = (a_qq - a_pp) / (2|a_pq|).
// Returns the number of sweeps.  Eigenvalues ascending; each eigenvector's
// largest-magnitude component made real positive.
int eig_jacobi(const double* cov, int n, double* evals, double* evecs) {
    std::vector<C<double>> A((size_t)n * n), V((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            // Hermitian from the lower triangle (CUBLAS_FILL_MODE_LOWER)
            const int r = std::max(i, j), c = std::min(i, j);
            C<double> v = {cov[2 * (r + (size_t)c * n)], cov[2 * (r + (size_t)c * n) + 1]};
            A[i + (size_t)j * n] = (i >= j) ? v : cconj(v);
            if (i == j) A[i + (size_t)j * n].im = 0.0;
            V[i + (size_t)j * n] = {i == j ? 1.0 : 0.0, 0.0};
        }
    auto a = [&](int i, int j) -> C<double>& { return A[i + (size_t)j * n]; };
    auto v = [&](int i, int j) -> C<double>& { return V[i + (size_t)j * n]; };
    double fro = 0;
    for (auto& e : A) fro += e.re * e.re + e.im * e.im;
    int sweep = 0;
    for (; sweep < 64; ++sweep) {
        double off = 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                if (i != j) off += a(i, j).re * a(i, j).re + a(i, j).im * a(i, j).im;
        if (off <= 1e-30 * fro || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                C<double> apq = a(p, q);
                double mag = std::hypot(apq.re, apq.im);
                if (mag == 0.0) continue;
                C<double> e = {apq.re / mag, apq.im / mag};  // e^{i phi}
                double tau = (a(q, q).re - a(p, p).re) / (2.0 * mag);
                double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                // J: Jpp = c, Jpq = s, Jqp = -s conj(e), Jqq = c conj(e)
                C<double> Jpp = {c, 0}, Jpq = {s, 0}, Jqp = cmul(C<double>{-s, 0}, cconj(e)),
                          Jqq = cmul(C<double>{c, 0}, cconj(e));
                for (int k = 0; k < n; ++k) {  // A <- A J, V <- V J
                    C<double> akp = a(k, p), akq = a(k, q);
                    a(k, p) = cadd(cmul(akp, Jpp), cmul(akq, Jqp));
                    a(k, q) = cadd(cmul(akp, Jpq), cmul(akq, Jqq));
                    C<double> vkp = v(k, p), vkq = v(k, q);
                    v(k, p) = cadd(cmul(vkp, Jpp), cmul(vkq, Jqp));
                    v(k, q) = cadd(cmul(vkp, Jpq), cmul(vkq, Jqq));
                }
                for (int k = 0; k < n; ++k) {  // A <- J^H A
                    C<double> apk = a(p, k), aqk = a(q, k);
                    a(p, k) = cadd(cmul(cconj(Jpp), apk), cmul(cconj(Jqp), aqk));
                    a(q, k) = cadd(cmul(cconj(Jpq), apk), cmul(cconj(Jqq), aqk));
                }
                a(p, q) = {0, 0};
                a(q, p) = {0, 0};
                a(p, p).im = 0;
                a(q, q).im = 0;
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return a(x, x).re < a(y, y).re; });
    for (int jj = 0; jj < n; ++jj) {
        const int j = order[jj];
        evals[jj] = a(j, j).re;
        int kmax = 0;
        double best = -1;
        for (int k = 0; k < n; ++k) {
            double m = std::hypot(v(k, j).re, v(k, j).im);
            if (m > best * (1 + 1e-12)) {
                best = m;
                kmax = k;
            }
        }
        C<double> ph = v(kmax, j);
        double m = std::hypot(ph.re, ph.im);
        C<double> rot = {ph.re / m, -ph.im / m};
        for (int k = 0; k < n; ++k) {
            C<double> r = cmul(v(k, j), rot);
            evecs[2 * (k + (size_t)jj * n)] = r.re;
            evecs[2 * (k + (size_t)jj * n) + 1] = r.im;
        }
    }
    return sweep;
}

// Components [t0, t1] of C = H V (the others are left untouched).
template <typename R>
void pca_project(const C<R>* ring, const double* evecs, int T, size_t P, C<R>* out, int t0 = 0, int t1 = -1) {
    if (t1 < 0) t1 = T - 1;
    const int nt = t1 - t0 + 1;
    // V rows s with the selected components contiguous
    std::vector<double> v((size_t)T * nt * 2);
    for (int s = 0; s < T; ++s)
        for (int t = t0; t <= t1; ++t) {
            v[2 * ((size_t)s * nt + (t - t0))] = evecs[2 * (s + (size_t)t * T)];
            v[2 * ((size_t)s * nt + (t - t0)) + 1] = evecs[2 * (s + (size_t)t * T) + 1];
        }
    constexpr size_t kTile = 32;
    const size_t tiles = (P + kTile - 1) / kTile;
#pragma omp parallel
    {
        std::vector<double> acc((size_t)nt * 2);
#pragma omp for schedule(static)
        for (long b = 0; b < (long)tiles; ++b) {
            const size_t p0 = (size_t)b * kTile, n = std::min(kTile, P - p0);
            for (size_t k = 0; k < n; ++k) {
                std::fill(acc.begin(), acc.end(), 0.0);
                for (int s = 0; s < T; ++s) {
                    const C<R> xv = ring[(size_t)s * P + p0 + k];
                    const double xr = (double)xv.re, xi = (double)xv.im;
                    const double* vs = &v[2 * (size_t)s * nt];
                    for (int t = 0; t < nt; ++t) {
                        acc[2 * t] += xr * vs[2 * t] - xi * vs[2 * t + 1];
                        acc[2 * t + 1] += xr * vs[2 * t + 1] + xi * vs[2 * t];
                    }
                }
                for (int t = 0; t < nt; ++t)
                    out[(size_t)(t0 + t) * P + p0 + k] = {(R)acc[2 * t], (R)acc[2 * t + 1]};
            }
        }
    }
}

// ---------------------------------------------------------------------------
// a14: unwrap_2d (tools_unwrap.cu:79-208, 269-369; circ_shift tools.cu:50-79).
// Output is the raw Im(IFFT2(...)) / (-2 pi) before rescale_float_unwrap2d.
// ---------------------------------------------------------------------------
template <typename R>
void unwrap2d(const R* phase, int W, int H, R* out) {
    const size_t P = (size_t)W * H;
    std::vector<C<R>> z(P), gx(P), gy(P);
    std::vector<R> kx(P), ky(P);
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            size_t idx = (size_t)j * W + i;
            // kernel_init_unwrap_2d then circ_shift by (-W/2, -H/2): the FFT frequency index
            kx[idx] = (R)(i < W / 2 ? i : i - W);
            ky[idx] = (R)(j < H / 2 ? j : j - H);
            z[idx] = {std::cos(phase[idx]), std::sin(phase[idx])};
        }
    // gradient_unwrap_2d (:298-322)
    for (size_t i = 0; i < P; ++i) gx[i] = z[i];
    fft2(gx.data(), 1, W, H, -1);
    gy = gx;
    for (size_t i = 0; i < P; ++i) {
        gx[i].re *= kx[i];
        gx[i].im *= kx[i];
        gy[i].re *= ky[i];
        gy[i].im *= ky[i];
    }
    fft2(gx.data(), 1, W, H, +1);
    fft2(gy.data(), 1, W, H, +1);
    const C<R> two_pi_i = {0, (R)(2 * M_PI)};
    for (size_t i = 0; i < P; ++i) {
        gx[i] = cmul(gx[i], two_pi_i);
        gy[i] = cmul(gy[i], two_pi_i);
    }
    // eq_unwrap_2d (:324-351): z <- conj(z * i); g *= z; FFT; g *= k / |k|^2
    for (size_t i = 0; i < P; ++i) {
        C<R> zi = cconj(cmul(z[i], C<R>{0, 1}));
        gx[i] = cmul(gx[i], zi);
        gy[i] = cmul(gy[i], zi);
    }
    fft2(gx.data(), 1, W, H, -1);
    fft2(gy.data(), 1, W, H, -1);
    for (size_t i = 0; i < P; ++i) {
        R norm = kx[i] * kx[i] + ky[i] * ky[i];
        if (norm != 0) {
            R cx_ = kx[i] / norm, cy_ = ky[i] / norm;
            gx[i].re *= cx_;
            gx[i].im *= cx_;
            gy[i].re *= cy_;
            gy[i].im *= cy_;
        } else {
            gx[i] = {0, 0};
            gy[i] = {0, 0};
        }
    }
    // phi_unwrap_2d (:353-369)
    for (size_t i = 0; i < P; ++i) gx[i] = cadd(gx[i], gy[i]);
    fft2(gx.data(), 1, W, H, +1);
    for (size_t i = 0; i < P; ++i) out[i] = gx[i].im / (R)(-2 * M_PI);
}

// rescale_float_unwrap2d (tools_conversion.cu:315-335)
void rescale_unwrap(const float* in, size_t n, float* out) {
    float mn = in[0], mx = in[0];
    for (size_t i = 1; i < n; ++i) {
        mn = std::min(mn, in[i]);
        mx = std::max(mx, in[i]);
    }
    const float m = 65535;
    for (size_t i = 0; i < n; ++i) {
        if (mn < 0.f) out[i] = (in[i] + std::fabs(mn)) / (std::fabs(mn) + mx) * m;
        else out[i] = (in[i] - mn) / (mx - mn) * m;
    }
}

// ---------------------------------------------------------------------------
// a16-a19: subsample, percentile, contrast, float -> u16
// ---------------------------------------------------------------------------
// kernel_subsample_frame (frame_reshape.cu:278-305)
template <typename T>
void subsample(const T* in, int W, int H, int s, T* out) {
    const int ow = W / s, oh = H / s;
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            const int xo = (y % 2 == 0) ? 0 : s / 2;
            const int yo = (x % 2 == 0) ? 0 : s / 2;
            out[(size_t)y * ow + x] = in[(size_t)y * s * W + (size_t)yo * W + (size_t)x * s + xo];
        }
}

// compute_percentile (percentile.cu:75-94): sort, value at uint(p/100 * N)
// (float arithmetic, truncation); indices past the end read the last element.
void percentile(const float* in, size_t n, const float* pct, int npct, float* out) {
    std::vector<float> v(in, in + n);
    std::sort(v.begin(), v.end());
    for (int i = 0; i < npct; ++i) {
        uint32_t idx = (uint32_t)(pct[i] / 100 * (float)n);
        if (idx >= n) idx = (uint32_t)n - 1;
        out[i] = v[idx];
    }
}

// device_float_to_ushort (tools_conversion.cu:350-359)
inline uint16_t f2u16(float x) {
    if (x <= 0.0f) return 0;
    if (x >= 65536.0f) return 65535;
    return (uint16_t)(uint32_t)x;
}

// apply_contrast_correction (contrast_correction.cu:5-17)
// compute_percentile_xy_view with the contrast reticle (percentile.cu:15-43,
// 150-190): the pixels with (x - W/2)^2 + (y - H/2)^2 <= r^2,
// r = uint(scale * min(W, H) / 2), in any order (they are sorted next).
std::vector<float> reticle_pixels(const float* img, int W, int H, float scale) {
    const unsigned r = (unsigned)(scale * (float)std::min(W, H) / 2);
    const long r2 = (long)r * r;
    const long cx = W / 2, cy = H / 2;
    std::vector<float> out;
    for (long y = 0; y < H; ++y)
        for (long x = 0; x < W; ++x)
            if ((x - cx) * (x - cx) + (y - cy) * (y - cy) <= r2) out.push_back(img[y * W + x]);
    return out;
}

void apply_contrast(float* img, size_t n, float mn, float mx) {
    const float factor = (unsigned short)65535 / (mx - mn + FLT_EPSILON);
    for (size_t i = 0; i < n; ++i) img[i] = factor * (img[i] - mn);
}

// ---------------------------------------------------------------------------
// a8 + a21: the Pipe cycle (pipe.cc:295-455, 400-427; queue.cc:332-376).
// ---------------------------------------------------------------------------
template <typename R>
struct Pipe {
    hv_compute_settings s;
    size_t P;
    int T;
    std::vector<C<R>> lens, buf, ring, spec;
    std::vector<R> image_r;
    std::vector<float> image, contrast_img;
    std::vector<uint16_t> out;
    std::vector<double> cov, evals, evecs;
    uint32_t batch_index = 0;
    uint32_t q_start = 0, q_size = 0;
    bool autocontrast_pending = true;
    float range[2] = {0, 0};

    explicit Pipe(const hv_compute_settings& st) : s(st) {
        P = (size_t)s.width * s.height;
        T = (int)s.time_transformation_size;
        buf.resize(P * s.batch_size);
        ring.assign(P * T, {0, 0});
        spec.resize(P * T);
        image_r.resize(P);
        image.resize(P);
        contrast_img.resize(P);
        out.resize(P);
        std::vector<double> l(P * 2);
        if (s.space_transformation == HV_SPACE_ANGULARSP) {
            float step = s.pixel_size * 1e-6f;
            lens_as_centred(s.width, s.height, s.z_distance, s.lambda, step, step, s.lens_precision, l.data());
            shift_corners(reinterpret_cast<C<double>*>(l.data()), 1, s.width, s.height);
        } else if (s.space_transformation == HV_SPACE_FRESNELTR) {
            int side = std::max(s.width, s.height);
            lens_fresnel(side, s.height, s.width, s.lambda, s.z_distance, s.pixel_size, s.lens_precision, l.data());
        }
        lens.resize(P);
        for (size_t i = 0; i < P; ++i) lens[i] = {(R)l[2 * i], (R)l[2 * i + 1]};
        range[0] = s.contrast_min;
        range[1] = s.contrast_max;
        cov.resize((size_t)T * T * 2);
        evals.resize(T);
        evecs.resize((size_t)T * T * 2);
    }

    // Queue::enqueue_multiple (queue.cc:332-376)
    void enqueue(const C<R>* elts, uint32_t nb) {
        const uint32_t max = T;
        if (nb > max) {
            elts += (size_t)(nb - max) * P;
            q_start = (q_start + nb - max) % max;
            nb = max;
        }
        const uint32_t begin = (q_start + q_size) % max;
        for (uint32_t k = 0; k < nb; ++k)
            std::memcpy(&ring[(size_t)((begin + k) % max) * P], elts + (size_t)k * P, P * sizeof(C<R>));
        q_size += nb;
        if (q_size > max) {
            q_start = (q_start + q_size - max) % max;
            q_size = max;
        }
    }

    int push(const void* frames) {
        // insert_wait_time_stride (pipe.cc:400-416)
        batch_index += s.batch_size;
        if (batch_index != s.time_stride) return 0;
        batch_index = 0;
        convert(frames, s.depth, P, s.batch_size, buf.data());
        if (s.space_transformation == HV_SPACE_ANGULARSP)
            propagate_as(buf.data(), s.batch_size, s.width, s.height, lens.data());
        else if (s.space_transformation == HV_SPACE_FRESNELTR)
            propagate_fresnel(buf.data(), s.batch_size, s.width, s.height, lens.data());
        enqueue(buf.data(), s.batch_size);
        // insert_wait_time_transformation_size (pipe.cc:418-427)
        if (q_size < (uint32_t)T) return 0;
        // time transform (fourier_transform.cc:195-229)
        if (s.time_transformation == HV_TIME_STFT) stft(ring.data(), T, P, spec.data());
        // insert_compute_p_accu (converts.cc:76-92)
        int pmin = s.p_start;
        int pmax = s.p_width ? std::max(0, std::min<int>(pmin + s.p_width, T)) : pmin;
        const int it = s.img_type;
        if (s.time_transformation == HV_TIME_PCA) {
            pca_cov(ring.data(), T, P, cov.data());
            eig_jacobi(cov.data(), T, evals.data(), evecs.data());
            // the image and the moments read only components [p_start, p_start + p_width]
            pca_project(ring.data(), evecs.data(), T, P, spec.data(), (int)s.p_start,
                        std::min<int>((int)(s.p_start + s.p_width), T - 1));
        } else if (s.time_transformation != HV_TIME_STFT) spec = ring;
        if (it == HV_IMG_MODULUS || it == HV_IMG_SQUARED_MODULUS || it == HV_IMG_ARGUMENT) {
            image_type(spec.data(), P, it, pmin, pmax, image_r.data());
            for (size_t i = 0; i < P; ++i) image[i] = (float)image_r[i];
            if (it == HV_IMG_ARGUMENT && s.unwrap_2d) {
                std::vector<R> uw(P);
                unwrap2d(image_r.data(), s.width, s.height, uw.data());
                std::vector<float> uwf(P);
                for (size_t i = 0; i < P; ++i) uwf[i] = (float)uw[i];
                rescale_unwrap(uwf.data(), P, image.data());
            }
            if (s.time_transformation == HV_TIME_PCA)  // converts.cc:50-61
                for (size_t i = 0; i < P; ++i) image[i] = image[i] * (float)((2 << 16) - 1);
        } else {
            // insert_moments (pipe.cc:429-455): f_end = min(start + width, T - 1)
            int f_start = s.p_start, f_end = std::min<int>(s.p_start + s.p_width, T - 1);
            std::vector<R> m(P);
            R* m0 = it == HV_IMG_MOMENTS_0 ? m.data() : nullptr;
            R* m1 = it == HV_IMG_MOMENTS_1 ? m.data() : nullptr;
            R* m2 = it == HV_IMG_MOMENTS_2 ? m.data() : nullptr;
            moments(spec.data(), T, P, f_start, f_end, s.camera_fps, m0, m1, m2);
            for (size_t i = 0; i < P; ++i) image[i] = (float)m[i];
        }
        if (s.fft_shift) shift_corners(image.data(), 1, s.width, s.height);  // rendering.cc:21-40
        contrast_img = image;
        if (s.contrast_enabled) {
            bool recompute = false, clamp = true;
            if (s.contrast_mode == HV_CONTRAST_REFERENCE) {
                recompute = autocontrast_pending;  // rendering.cc:239-258
                autocontrast_pending = false;
            } else if (s.contrast_mode == HV_CONTRAST_PERCENTILE) recompute = true;
            if (recompute) {
                const float pct[2] = {s.contrast_lower, s.contrast_upper};
                if (s.contrast_reticle_enabled) {
                    std::vector<float> disc = reticle_pixels(image.data(), s.width, s.height, s.contrast_reticle_scale);
                    percentile(disc.data(), disc.size(), pct, 2, range);
                } else if (s.contrast_subsample > 1) {
                    int st = s.contrast_subsample;
                    std::vector<float> sub((size_t)(s.width / st) * (s.height / st));
                    subsample(image.data(), s.width, s.height, st, sub.data());
                    percentile(sub.data(), sub.size(), pct, 2, range);
                } else percentile(image.data(), P, pct, 2, range);
                if (clamp) {  // contrast_api.cc:137-146
                    range[0] = range[0] > 1.0f ? range[0] : 1.0f;
                    range[1] = range[1] > 1.0f ? range[1] : 1.0f;
                }
            }
            if (s.contrast_mode == HV_CONTRAST_MINMAX) {
                range[0] = *std::min_element(image.begin(), image.end());
                range[1] = *std::max_element(image.begin(), image.end());
            }
            float mn = range[0], mx = range[1];
            if (s.contrast_invert) std::swap(mn, mx);
            apply_contrast(contrast_img.data(), P, mn, mx);
        }
        for (size_t i = 0; i < P; ++i) out[i] = f2u16(contrast_img[i]);
        return 1;
    }
};

template <typename R>
C<R>* cp(void* p) { return reinterpret_cast<C<R>*>(p); }
template <typename R>
const C<R>* ccp(const void* p) { return reinterpret_cast<const C<R>*>(p); }

}  // namespace

// ===========================================================================
// extern "C" surface for the Python checker (oracle/oracle.py)
// ===========================================================================
extern "C" {

int or_max_threads(void) { return omp_get_max_threads(); }
void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

void or_convert_f64(const void* in, int depth, uint64_t P, int B, double* out) { convert(in, depth, P, B, cp<double>(out)); }
void or_convert_f32(const void* in, int depth, uint64_t P, int B, float* out) { convert(in, depth, P, B, cp<float>(out)); }

void or_lens_as(int nx, int ny, float z, float lambda, float xs, float ys, int precision, double* out) {
    lens_as_centred(nx, ny, z, lambda, xs, ys, precision, out);
}
void or_shift_corners_c128(double* buf, int batch, int sx, int sy) { shift_corners(cp<double>(buf), batch, sx, sy); }
void or_shift_corners_f32(float* buf, int batch, int sx, int sy) { shift_corners(buf, batch, sx, sy); }
void or_lens_fresnel(int side, int H, int W, float lambda, float z, float px, int precision, double* out) {
    lens_fresnel(side, H, W, lambda, z, px, precision, out);
}

int or_fft2_f64(double* buf, int B, int W, int H, int dir) {
    if (W < 1 || H < 1) return HV_ERR_INVALID_ARG;
    fft2(cp<double>(buf), B, W, H, dir);
    return 0;
}
int or_fft2_f32(float* buf, int B, int W, int H, int dir) {
    if (W < 1 || H < 1) return HV_ERR_INVALID_ARG;
    fft2(cp<float>(buf), B, W, H, dir);
    return 0;
}
int or_fft1_f64(double* buf, int n, int dir) {
    if (n < 1) return HV_ERR_INVALID_ARG;
    FftPlan<double> p(n);
    fft1d(cp<double>(buf), p, dir);
    return 0;
}
void or_propagate_as_f64(double* buf, int B, int W, int H, const double* lens) {
    propagate_as(cp<double>(buf), B, W, H, ccp<double>(lens));
}
void or_propagate_fresnel_f64(double* buf, int B, int W, int H, const double* chirp) {
    propagate_fresnel(cp<double>(buf), B, W, H, ccp<double>(chirp));
}
void or_stft_f64(const double* ring, int T, uint64_t P, double* out) { stft(ccp<double>(ring), T, P, cp<double>(out)); }
void or_image_type_f64(const double* spec, uint64_t P, int type, int pmin, int pmax, double* out) {
    image_type(ccp<double>(spec), P, type, pmin, pmax, out);
}
void or_moments_f64(const double* spec, int T, uint64_t P, int fs, int fe, float fps, double* m0, double* m1,
                    double* m2) {
    moments(ccp<double>(spec), T, P, fs, fe, fps, m0, m1, m2);
}
void or_fft_freqs(int T, float fps, float* out) {
    std::vector<float> f;
    fft_freqs(T, fps, f);
    std::memcpy(out, f.data(), T * sizeof(float));
}
void or_pca_cov_f64(const double* ring, int T, uint64_t P, double* cov) { pca_cov(ccp<double>(ring), T, P, cov); }
int or_eig_jacobi_f64(const double* cov, int n, double* evals, double* evecs) {
    return eig_jacobi(cov, n, evals, evecs);
}
void or_pca_project_f64(const double* ring, const double* evecs, int T, uint64_t P, double* out) {
    pca_project(ccp<double>(ring), evecs, T, P, cp<double>(out));
}
void or_unwrap2d_f64(const double* phase, int W, int H, double* out) { unwrap2d(phase, W, H, out); }
void or_rescale_unwrap_f32(const float* in, uint64_t n, float* out) { rescale_unwrap(in, n, out); }
void or_subsample_f32(const float* in, int W, int H, int s, float* out) { subsample(in, W, H, s, out); }
void or_subsample_u8(const uint8_t* in, int W, int H, int s, uint8_t* out) { subsample(in, W, H, s, out); }
void or_percentile_f32(const float* in, uint64_t n, const float* pct, int npct, float* out) {
    percentile(in, n, pct, npct, out);
}
uint64_t or_reticle_pixels_f32(const float* img, int W, int H, float scale, float* out) {
    std::vector<float> d = reticle_pixels(img, W, H, scale);
    if (out) std::memcpy(out, d.data(), d.size() * sizeof(float));
    return d.size();
}
void or_apply_contrast_f32(float* img, uint64_t n, float mn, float mx) { apply_contrast(img, n, mn, mx); }
void or_float_to_u16(const float* in, uint64_t n, uint16_t* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = f2u16(in[i]);
}

// Pipe contexts: precision 64 -> double, 32 -> float.
void* or_pipe_create(const hv_compute_settings* s, int precision) {
    if (precision == 32) return new Pipe<float>(*s);
    return new Pipe<double>(*s);
}
int or_pipe_push(void* h, int precision, const void* frames, float* image, float* contrast_img, uint16_t* out,
                 float* range) {
    auto run = [&](auto* p) {
        int r = p->push(frames);
        if (r) {
            if (image) std::memcpy(image, p->image.data(), p->P * sizeof(float));
            if (contrast_img) std::memcpy(contrast_img, p->contrast_img.data(), p->P * sizeof(float));
            if (out) std::memcpy(out, p->out.data(), p->P * sizeof(uint16_t));
            if (range) {
                range[0] = p->range[0];
                range[1] = p->range[1];
            }
        }
        return r;
    };
    if (precision == 32) return run(static_cast<Pipe<float>*>(h));
    return run(static_cast<Pipe<double>*>(h));
}
void or_pipe_ring_state(void* h, int precision, uint32_t* start, uint32_t* size) {
    if (precision == 32) {
        *start = static_cast<Pipe<float>*>(h)->q_start;
        *size = static_cast<Pipe<float>*>(h)->q_size;
    } else {
        *start = static_cast<Pipe<double>*>(h)->q_start;
        *size = static_cast<Pipe<double>*>(h)->q_size;
    }
}
// Enqueue nb frames (host complex, precision 64) into the pipe's ring directly
// (Queue::enqueue_multiple semantics), for the ring-order known-answer tests.
void or_pipe_enqueue_raw(void* h, const double* frames, uint32_t nb) {
    static_cast<Pipe<double>*>(h)->enqueue(ccp<double>(frames), nb);
}
void or_pipe_ring_copy(void* h, double* out) {
    auto* p = static_cast<Pipe<double>*>(h);
    std::memcpy(out, p->ring.data(), p->ring.size() * sizeof(C<double>));
}
// The latest cycle's moments M0, M1, M2 over the settings' range (the
// window's STFT or PCA output), and its PCA covariance, eigenvalues and
// eigenvectors (column-major, complex double).
void or_pipe_moments(void* h, int precision, double* m0, double* m1, double* m2) {
    auto run = [&](auto* p) {
        const auto& s = p->s;
        const int fe = std::min<int>((int)(s.p_start + s.p_width), p->T - 1);
        std::vector<decltype(p->spec[0].re)> a(p->P), b(p->P), c(p->P);
        moments(p->spec.data(), p->T, p->P, (int)s.p_start, fe, s.camera_fps, a.data(), b.data(), c.data());
        for (size_t i = 0; i < p->P; ++i) {
            m0[i] = a[i];
            m1[i] = b[i];
            m2[i] = c[i];
        }
    };
    if (precision == 64) run(static_cast<Pipe<double>*>(h));
    else run(static_cast<Pipe<float>*>(h));
}

void or_pipe_pca_state(void* h, int precision, double* cov, double* evals, double* evecs) {
    auto run = [&](auto* p) {
        std::memcpy(cov, p->cov.data(), p->cov.size() * sizeof(double));
        std::memcpy(evals, p->evals.data(), p->evals.size() * sizeof(double));
        std::memcpy(evecs, p->evecs.data(), p->evecs.size() * sizeof(double));
    };
    if (precision == 64) run(static_cast<Pipe<double>*>(h));
    else run(static_cast<Pipe<float>*>(h));
}

void or_pipe_destroy(void* h, int precision) {
    if (precision == 32) delete static_cast<Pipe<float>*>(h);
    else delete static_cast<Pipe<double>*>(h);
}

}  // extern "C"
