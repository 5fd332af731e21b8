// hv_fft.hpp -- LDS-staged mixed-radix Stockham FFT engine for gfx950.
//
// One transform of length N is carried by NT = N / E threads, each holding
// E complex values in registers.  Stage s applies radix R_s (16, 8, 4 or 2)
// with Ns = R_0 ... R_{s-1}: thread j reads x[j + r N/R], multiplies by
// exp(dir 2 pi i r (j mod Ns) / (Ns R)), does a radix-R DFT in registers and
// writes x'[(j / Ns) Ns R + (j mod Ns) + r Ns].  Stage 0 reads through a
// caller-supplied loader (global memory, fused conversion / chirp), the last
// stage writes through a caller-supplied storer (global memory, fused lens,
// ring slot, image epilogue), and stages exchange through LDS in between.
// Twiddles come from a 4096-entry table of exp(-2 pi i k / 4096) rounded
// from double (hv_twiddles.inc); compile-time indices fold to constants.
#pragma once

#include <type_traits>

#include "hv_common.hpp"
#include "hv_twiddles.inc"  // generated into _build/gen by build.py

namespace hv {

template <int I, int S, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < S) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, S>(f);
    }
}

// exp(dir * 2 pi i * k / n) for n | 4096
template <int DIR>
__device__ __forceinline__ float2 twiddle(int k, int log2n) {
    const int idx = (k << (12 - log2n)) & (kTwN - 1);
    const float c = kTw[idx][0], s = kTw[idx][1];
    return DIR < 0 ? make_float2(c, s) : make_float2(c, -s);
}

// d * exp(dir * 2 pi i * k / m), m <= 16, k compile-time after unrolling.
template <int DIR>
__device__ __forceinline__ float2 tw_small(float2 d, int k, int m) {
    const int k16 = (k * (16 / m)) & 15;
    if (k16 == 0) return d;
    if (k16 == 8) return make_float2(-d.x, -d.y);
    if (k16 == 4)  // exp(dir i pi/2) = dir * i
        return DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
    if (k16 == 12) return DIR < 0 ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    const float c = kTw[k16 * 256][0];
    const float s = DIR < 0 ? kTw[k16 * 256][1] : -kTw[k16 * 256][1];
    return make_float2(d.x * c - d.y * s, d.x * s + d.y * c);
}

__host__ __device__ constexpr int bitrev(int x, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b)
        if (x & (1 << b)) r |= 1 << (bits - 1 - b);
    return r;
}
__host__ __device__ constexpr int clog2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

// In-register radix-R DFT (R <= 16), natural order in and out.
template <int R, int DIR>
__device__ __forceinline__ void dft_reg(float2* v) {
    if constexpr (R == 1) return;
    // radix-2 decimation in frequency: natural in, bit-reversed out
#pragma unroll
    for (int half = R / 2; half >= 1; half >>= 1) {
#pragma unroll
        for (int blk = 0; blk < R; blk += 2 * half) {
#pragma unroll
            for (int k = 0; k < half; ++k) {
                const float2 a = v[blk + k], b = v[blk + k + half];
                v[blk + k] = cadd(a, b);
                v[blk + k + half] = tw_small<DIR>(csub(a, b), k, 2 * half);
            }
        }
    }
    float2 t[R];
#pragma unroll
    for (int i = 0; i < R; ++i) t[i] = v[bitrev(i, clog2(R))];
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = t[i];
}

// Stage plan for length N with E values per thread: radix R_s = min(E, 16,
// what is left of N), each thread running E / R_s butterflies per stage.
template <int N, int E>
struct FftShape {
    static constexpr int kE = E < N ? E : N;
    static constexpr int NT = N / kE;
    static constexpr int RMAX = kE < 16 ? kE : 16;
    static constexpr int stages() {
        int s = 0, rem = N;
        while (rem > 1) {
            rem /= (rem >= RMAX ? RMAX : rem);
            ++s;
        }
        return s;
    }
    static constexpr int radix(int s) {
        int rem = N;
        for (int i = 0; i < s; ++i) rem /= (rem >= RMAX ? RMAX : rem);
        return rem >= RMAX ? RMAX : rem;
    }
    static constexpr int ns(int s) {
        int n = 1, rem = N;
        for (int i = 0; i < s; ++i) {
            const int r = rem >= RMAX ? RMAX : rem;
            n *= r;
            rem /= r;
        }
        return n;
    }
};

// f(i, n) for every value slot i of thread t and the input index n that the
// first stage expects there: slot k R + r holds x[t + k NT + r N / R].
template <int N, int E, class F>
__device__ __forceinline__ void fft_load_slots(int t, F&& f) {
    using S = FftShape<N, E>;
    constexpr int kE = S::kE, NT = S::NT, R = S::radix(0), NB = kE / R;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int j = t + k * NT;
#pragma unroll
        for (int r = 0; r < R; ++r) f(k * R + r, j + r * (N / R));
    }
}

// Stage-0 inputs of thread t.  Split from fft_compute so a kernel can issue
// the loads of its next transform before computing the current one.
template <int N, int E, class Load>
__device__ __forceinline__ void fft_load(int t, Load&& load, float2* v) {
    fft_load_slots<N, E>(t, [&](int i, int n) { v[i] = load(n); });
}

template <int N, int E, int DIR, bool STORE_SYNC = false, class Store, class Lds>
__device__ __forceinline__ void fft_compute(int t, float2* v, Store&& store, Lds&& lds);

// Runs one length-N transform on thread t (0 <= t < NT) of its group.
//   load(n)      -> float2 : element n of the input (stage 0)
//   store(n, v)            : element n of the output (last stage)
//   lds(n)       -> float2&: this transform's LDS staging slot n
// Every thread of the workgroup must call fft_run the same number of times
// (it contains __syncthreads()).  STORE_SYNC puts a barrier before the last
// stage's stores, for storers that write into the same LDS staging area.
template <int N, int E, int DIR, bool STORE_SYNC = false, class Load, class Store, class Lds>
__device__ __forceinline__ void fft_run(int t, Load&& load, Store&& store, Lds&& lds) {
    float2 v[FftShape<N, E>::kE];
    fft_load<N, E>(t, load, v);
    fft_compute<N, E, DIR, STORE_SYNC>(t, v, store, lds);
}

// Per-thread stage twiddles.  Thread t's twiddles depend only on t, so a
// kernel that runs many transforms (frames, rows, forward + inverse) builds
// them once: tw holds, for every stage s >= 1 and butterfly k of the thread,
// the forward twiddles w^r, r = 1 .. R_s - 1 (the inverse uses conjugates).
template <int N, int E>
struct FftTw {
    using S = FftShape<N, E>;
    static constexpr int base(int s) {  // offset of stage s
        int b = 0;
        for (int i = 1; i < s; ++i) b += (S::kE / S::radix(i)) * (S::radix(i) - 1);
        return b;
    }
    static constexpr int COUNT = base(S::stages()) > 0 ? base(S::stages()) : 1;
};

// Twiddles w^r, r = 1 .. R - 1, of butterfly k of thread t at stage S
// (w = exp(-2 pi i (j mod Ns) / (Ns R)), j = t + k NT): one table load, the
// powers by a product tree of depth <= 4.
template <int N, int E, int S>
__device__ __forceinline__ void stage_twiddles_gen(int t, int k, float2* w) {
    using FS = FftShape<N, E>;
    constexpr int R = FS::radix(S), Ns = FS::ns(S), NT = FS::NT, LOGN = clog2(N);
    const int kk = (t + k * NT) & (Ns - 1);
    w[1] = twiddle<-1>(kk * (N / (Ns * R)), LOGN);
    if constexpr (R > 2) {
        w[2] = cmul(w[1], w[1]);
        w[3] = cmul(w[2], w[1]);
    }
    if constexpr (R > 4) {
        w[4] = cmul(w[2], w[2]);
        w[5] = cmul(w[4], w[1]);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
    }
    if constexpr (R > 8) {
        w[8] = cmul(w[4], w[4]);
#pragma unroll
        for (int r = 9; r < R; ++r) w[r] = cmul(w[8], w[r - 8]);
    }
}

template <int N, int E>
__device__ __forceinline__ void fft_twiddles(int t, float2* tw) {
    using S = FftShape<N, E>;
    constexpr int kE = S::kE, NS = S::stages();
    static_for<1, NS>([&](auto s_) {
        constexpr int s = decltype(s_)::value;
        constexpr int R = S::radix(s), NB = kE / R;
#pragma unroll
        for (int k = 0; k < NB; ++k) stage_twiddles_gen<N, E, s>(t, k, tw + FftTw<N, E>::base(s) + k * (R - 1) - 1);
    });
}

// Twiddle source of a run: a thread's table from fft_twiddles (const
// float2*), or TwOnTheFly -- each stage's twiddles generated where they are
// used (fewer registers held across the run, for long transforms).
struct TwOnTheFly {};

template <int N, int E, int S>
__device__ __forceinline__ void stage_twiddles(const float2* tw, int, int k, float2* w) {
    constexpr int R = FftShape<N, E>::radix(S);
    const float2* src = tw + FftTw<N, E>::base(S) + k * (R - 1) - 1;
#pragma unroll
    for (int r = 1; r < R; ++r) w[r] = src[r];
}
template <int N, int E, int S>
__device__ __forceinline__ void stage_twiddles(TwOnTheFly, int t, int k, float2* w) {
    stage_twiddles_gen<N, E, S>(t, k, w);
}

// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

#ifndef HV_FFT_WAVE_SYNC
#define HV_FFT_WAVE_SYNC 1
#endif

// Exchange barrier of one transform.  WAVE: the transform's threads are one
// wavefront (or part of one); LDS operations of a wavefront are performed in
// order, so the exchange needs no workgroup barrier, only a fence that keeps
// the compiler from moving LDS accesses across it.
template <bool WAVE>
__device__ __forceinline__ void fft_sync() {
    if constexpr (WAVE && HV_FFT_WAVE_SYNC) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Stages [FIRST, LAST) only (LAST = 0: to the end).  A run that starts past
// stage 0 reads its input from LDS; a run that stops early leaves its output
// in LDS (exchange order, no barrier) for a later run, possibly with another
// thread mapping after a workgroup barrier.
template <int N, int E, int DIR, bool STORE_SYNC = false, bool WAVE = false, int FIRST = 0, int LAST = 0,
          class TW, class Store, class Lds>
__device__ __forceinline__ void fft_compute_tw(int t, float2* v, const TW& tw, Store&& store, Lds&& lds);

template <int N, int E, int DIR, bool STORE_SYNC, class Store, class Lds>
__device__ __forceinline__ void fft_compute(int t, float2* v, Store&& store, Lds&& lds) {
    float2 tw[FftTw<N, E>::COUNT];
    fft_twiddles<N, E>(t, tw);
    fft_compute_tw<N, E, DIR, STORE_SYNC>(t, v, tw, store, lds);
}

template <int N, int E, int DIR, bool STORE_SYNC, bool WAVE, int FIRST, int LAST, class TW, class Store, class Lds>
__device__ __forceinline__ void fft_compute_tw(int t, float2* v, const TW& tw, Store&& store, Lds&& lds) {
    static_assert(!WAVE || FftShape<N, E>::NT <= 64, "a wave-local transform has at most 64 threads");
    using S = FftShape<N, E>;
    constexpr int kE = S::kE, NT = S::NT, NS = S::stages();
    constexpr int END = LAST > 0 ? LAST : NS;
    static_for<FIRST, END>([&](auto s_) {
        constexpr int s = decltype(s_)::value;
        constexpr int R = S::radix(s), Ns = S::ns(s), NB = kE / R;
        if constexpr (s > 0) {
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int j = t + k * NT;
#pragma unroll
                for (int r = 0; r < R; ++r) v[k * R + r] = lds(j + r * (N / R));
            }
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int j = t + k * NT;
            if constexpr (Ns > 1) {
                // w^r, w = exp(dir 2 pi i (j mod Ns) / (Ns R))
                float2 w[R];
                stage_twiddles<N, E, s>(tw, t, k, w);
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[k * R + r] = DIR < 0 ? cmul(v[k * R + r], w[r]) : cmulc(v[k * R + r], w[r]);
            }
            dft_reg<R, DIR>(&v[k * R]);
        }
        if constexpr (s == END - 1 && END < NS) {
            // partial run: leave the exchange in LDS for the next run
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int j = t + k * NT;
                const int d = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
                for (int r = 0; r < R; ++r) lds(d + r * Ns) = v[k * R + r];
            }
        } else if constexpr (s == NS - 1) {
            if constexpr (STORE_SYNC) fft_sync<WAVE>();
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int j = t + k * NT;
                const int d = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    // a storer may also take the register slot (k R + r) of the value
                    if constexpr (std::is_invocable_v<Store, int, float2, int>) store(d + r * Ns, v[k * R + r], k * R + r);
                    else store(d + r * Ns, v[k * R + r]);
                }
            }
        } else {
            fft_sync<WAVE>();
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int j = t + k * NT;
                const int d = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
                for (int r = 0; r < R; ++r) lds(d + r * Ns) = v[k * R + r];
            }
            fft_sync<WAVE>();
        }
    });
}

// MT independent transforms of one length on the same thread mapping,
// interleaved stage by stage: the instruction-level parallelism hides LDS
// and ALU latency in kernels with few wavefronts per SIMD, and the twiddles
// are shared.  v[m kE + i] holds transform m; lds(m, n) is transform m's
// staging slot n; store(m, n, value, slot) receives the outputs.
template <int N, int E, int DIR, int MT, bool WAVE, class TW, class Store, class Lds>
__device__ __forceinline__ void fft_compute_multi(int t, float2* v, const TW& tw, Store&& store, Lds&& lds) {
    static_assert(!WAVE || FftShape<N, E>::NT <= 64, "a wave-local transform has at most 64 threads");
    using S = FftShape<N, E>;
    constexpr int kE = S::kE, NT = S::NT, NS = S::stages();
    static_for<0, NS>([&](auto s_) {
        constexpr int s = decltype(s_)::value;
        constexpr int R = S::radix(s), Ns = S::ns(s), NB = kE / R;
        if constexpr (s > 0) {
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const int j = t + k * NT;
#pragma unroll
                    for (int r = 0; r < R; ++r) v[m * kE + k * R + r] = lds(m, j + r * (N / R));
                }
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            float2 w[R];
            if constexpr (Ns > 1) stage_twiddles<N, E, s>(tw, t, k, w);
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                float2* x = v + m * kE + k * R;
                if constexpr (Ns > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r) x[r] = DIR < 0 ? cmul(x[r], w[r]) : cmulc(x[r], w[r]);
                }
                dft_reg<R, DIR>(x);
            }
        }
        if constexpr (s == NS - 1) {
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const int j = t + k * NT;
                    const int d = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
                    for (int r = 0; r < R; ++r) store(m, d + r * Ns, v[m * kE + k * R + r], k * R + r);
                }
        } else {
            fft_sync<WAVE>();
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const int j = t + k * NT;
                    const int d = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
                    for (int r = 0; r < R; ++r) lds(m, d + r * Ns) = v[m * kE + k * R + r];
                }
            fft_sync<WAVE>();
        }
    });
}

// Adjoint run: the stages of the DIR_FWD transform applied backwards, each as
// its adjoint (DFT of the opposite direction, then conjugate twiddles,
// exchanges with read and write positions swapped).  Input: v in the slot
// order a forward run leaves it (slot k R + r of thread t holds element
// fft_out_index(t, k R + r)); output: the opposite-direction transform
// (unnormalised), written through store(n, value) at the forward's load
// positions.  A forward run whose results stay in registers, a pointwise
// product and this adjoint run make a filter with one exchange fewer than a
// forward and an inverse transform in natural order.
template <int N, int E, int DIR_FWD, bool WAVE = false, class TW, class Store, class Lds>
__device__ __forceinline__ void fft_compute_adj_tw(int t, float2* v, const TW& tw, Store&& store, Lds&& lds) {
    using S = FftShape<N, E>;
    constexpr int kE = S::kE, NT = S::NT, NS = S::stages();
    constexpr int DIR = -DIR_FWD;
    static_for<0, NS>([&](auto i_) {
        constexpr int s = NS - 1 - decltype(i_)::value;
        constexpr int R = S::radix(s), Ns = S::ns(s), NB = kE / R;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            dft_reg<R, DIR>(&v[k * R]);
            if constexpr (Ns > 1) {
                float2 w[R];
                stage_twiddles<N, E, s>(tw, t, k, w);
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[k * R + r] = DIR_FWD < 0 ? cmulc(v[k * R + r], w[r]) : cmul(v[k * R + r], w[r]);
            }
        }
        if constexpr (s == 0) {
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int j = t + k * NT;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    // a storer may also take the register slot, the one fft_load filled with element n
                    if constexpr (std::is_invocable_v<Store, int, float2, int>) store(j + r * (N / R), v[k * R + r], k * R + r);
                    else store(j + r * (N / R), v[k * R + r]);
                }
            }
        } else {
            // the forward stage s read positions j + r N/R; stage s - 1 wrote d + r' Ns'
            fft_sync<WAVE>();
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int j = t + k * NT;
#pragma unroll
                for (int r = 0; r < R; ++r) lds(j + r * (N / R)) = v[k * R + r];
            }
            fft_sync<WAVE>();
            constexpr int Rp = S::radix(s - 1), Nsp = S::ns(s - 1), NBp = kE / Rp;
#pragma unroll
            for (int k = 0; k < NBp; ++k) {
                const int j = t + k * NT;
                const int d = (j / Nsp) * Nsp * Rp + (j & (Nsp - 1));
#pragma unroll
                for (int r = 0; r < Rp; ++r) v[k * Rp + r] = lds(d + r * Nsp);
            }
        }
    });
}

// Output index of register slot (k R + r) of thread t in the last stage
// (the n a storer receives for that value).
template <int N, int E>
__device__ __forceinline__ int fft_out_index(int t, int slot) {
    using S = FftShape<N, E>;
    constexpr int NS = S::stages(), R = S::radix(NS - 1), Ns = S::ns(NS - 1), NT = S::NT;
    const int k = slot / R, r = slot % R;
    const int j = t + k * NT;
    return (j / Ns) * Ns * R + (j & (Ns - 1)) + r * Ns;
}

// Whole transform in one thread's registers (N <= 64): radix-2 DIF with
// table twiddles, natural order in and out.
template <int N, int DIR>
__device__ __forceinline__ void fft_reg(float2* v) {
    if constexpr (N <= 16) {
        dft_reg<N, DIR>(v);
    } else {
        constexpr int LOGN = clog2(N);
#pragma unroll
        for (int half = N / 2; half >= 1; half >>= 1) {
#pragma unroll
            for (int blk = 0; blk < N; blk += 2 * half) {
#pragma unroll
                for (int k = 0; k < half; ++k) {
                    const float2 a = v[blk + k], b = v[blk + k + half];
                    v[blk + k] = cadd(a, b);
                    const float2 d = csub(a, b);
                    // exp(dir 2 pi i k / (2 half))
                    const int kk = k * (N / (2 * half));
                    v[blk + k + half] = (kk == 0) ? d : cmul(d, twiddle<DIR>(kk, LOGN));
                }
            }
        }
        float2 t[N];
#pragma unroll
        for (int i = 0; i < N; ++i) t[i] = v[bitrev(i, LOGN)];
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = t[i];
    }
}

}  // namespace hv
