// hv_spatial.hip -- input conversion, lenses and the space transforms.
//
// A batched FFT2 is two HBM passes: a row pass (transforms along x, each
// workgroup owns M whole rows) and a column pass (transforms along y, each
// workgroup owns a CW-column tile; lanes run along x so every load/store is a
// run of CW consecutive complex values).  The angular spectrum
// (angular_spectrum.cu:33-63) is three passes:
//   1. row FFT fused with the u8/u16/f32 -> complex conversion
//      (tools_conversion.cu:107-171), written straight into the frame's
//      time-window ring slot;
//   2. column FFT -> x lens/(W H) -> inverse column FFT, in LDS;
//   3. inverse row FFT, in place in the ring slot.
// The Fresnel transform (fresnel_transform.cu:65-83) is two: the chirp is
// applied in the row pass's loader.  No pass needs a separate buffer.
#include <type_traits>

#include "hv_fft.hpp"
#include "hv_kernels.hpp"

namespace hv {

// ---------------------------------------------------------------------------
// conversion (tools_conversion.cu:107-171)
// ---------------------------------------------------------------------------
template <typename TIn>
__global__ void __launch_bounds__(256) k_convert(float2* __restrict__ out, const TIn* __restrict__ in, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = make_float2(convert_px<TIn>(in[i]), 0.0f);
}

static int grid_for(size_t n, int threads, int max_blocks = 8192) {
    size_t b = (n + threads - 1) / threads;
    return (int)(b < (size_t)max_blocks ? (b > 0 ? b : 1) : max_blocks);
}

hv_status launch_convert(float2* out, const void* in, size_t n, int depth, hipStream_t st) {
    const int g = grid_for(n, 256);
    if (depth == HV_DEPTH_8) k_convert<uint8_t><<<g, 256, 0, st>>>(out, (const uint8_t*)in, n);
    else if (depth == HV_DEPTH_16) k_convert<uint16_t><<<g, 256, 0, st>>>(out, (const uint16_t*)in, n);
    else if (depth == HV_DEPTH_32F) k_convert<float><<<g, 256, 0, st>>>(out, (const float*)in, n);
    else return fail(HV_ERR_INVALID_ARG, "depth must be 1, 2 or 4");
    HV_CHECK_LAUNCH("k_convert");
    return HV_OK;
}

// ---------------------------------------------------------------------------
// lenses (masks.cu:7-59, fresnel_transform.cu:14-63)
// ---------------------------------------------------------------------------
// Angular-spectrum transfer function.  `shift` writes element (x, y) at its
// corner-shifted position; `scale` multiplies the result (1/(W H) folds the
// inverse-FFT normalisation of angular_spectrum.cu:61 into the lens; for
// power-of-two sides the scale is a power of two and the product exact).
__global__ void k_lens_as(float2* __restrict__ out, int nx, int ny, float z, float lambda, float x_step,
                          float y_step, int precision, int shift, float scale) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= nx || y >= ny) return;
    float c, s;
    if (precision == HV_LENS_REFERENCE_FP32) {
        // the reference's float expression order (masks.cu:45-57)
        float u_step = 1.0f / (nx * x_step);
        float v_step = 1.0f / (ny * y_step);
        float u = (x - (nx >> 1)) * u_step;
        float v = (y - (ny >> 1)) * v_step;
        float tmp = 1.0f - (lambda * lambda * (u * u + v * v));
        if (tmp < 0.0f) tmp = 0.0f;
        float phase = 2.0f * M_PI * z / lambda * sqrtf(tmp);
        c = cosf(phase);
        s = sinf(phase);
    } else {
        double u = (double)(x - (nx >> 1)) / ((double)nx * (double)x_step);
        double v = (double)(y - (ny >> 1)) / ((double)ny * (double)y_step);
        double lam = (double)lambda;
        double tmp = 1.0 - lam * lam * (u * u + v * v);
        if (tmp < 0.0) tmp = 0.0;
        double phase = 2.0 * M_PI * (double)z / lam * sqrt(tmp);
        c = (float)cos(phase);
        s = (float)sin(phase);
    }
    int ox = x, oy = y;
    if (shift) {  // kernel_shift_corners for even sizes: (x, y) <-> (x +- nx/2, y +- ny/2)
        ox = (x + nx / 2) % nx;
        oy = (y + ny / 2) % ny;
    }
    out[(size_t)oy * nx + ox] = make_float2(c * scale, s * scale);
}

// Fresnel chirp on a square of side `side`, written cropped to width x height.
__global__ void k_lens_fresnel(float2* __restrict__ out, int side, int width, int height, float lambda, float z,
                               float pixel_size, int precision) {
#pragma clang fp contract(off)
    const int xo = blockIdx.x * blockDim.x + threadIdx.x;
    const int yo = blockIdx.y;
    if (xo >= width || yo >= height) return;
    const int i = xo + (side - width) / 2, j = yo + (side - height) / 2;
    float c, s;
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
        c = (float)cos(ph);
        s = (float)sin(ph);
    }
    out[(size_t)yo * width + xo] = make_float2(c, s);
}

hv_status launch_lens_as(float2* out, int nx, int ny, float z, float lambda, float xs, float ys, int precision,
                         int shift, float scale, hipStream_t st) {
    dim3 grid((nx + 255) / 256, ny);
    k_lens_as<<<grid, 256, 0, st>>>(out, nx, ny, z, lambda, xs, ys, precision, shift, scale);
    HV_CHECK_LAUNCH("k_lens_as");
    return HV_OK;
}

hv_status launch_lens_fresnel(float2* out, int side, int width, int height, float lambda, float z, float px,
                              int precision, hipStream_t st) {
    dim3 grid((width + 255) / 256, height);
    k_lens_fresnel<<<grid, 256, 0, st>>>(out, side, width, height, lambda, z, px, precision);
    HV_CHECK_LAUNCH("k_lens_fresnel");
    return HV_OK;
}

// kernel_shift_corners (shift_corners.cu:8-44): in-place swap of (i, j),
// j < sy/2, with (i +- sx/2, j + sy/2).
template <typename T>
__global__ void k_shift_corners(T* buf, int batch, int sx, int sy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= sx || j >= sy / 2) return;
    const int sx2 = sx / 2, sy2 = sy / 2;
    const int ni = i < sx2 ? i + sx2 : i - sx2, nj = j + sy2;
    const size_t P = (size_t)sx * sy;
    for (int b = 0; b < batch; ++b) {
        T* f = buf + b * P;
        T a = f[(size_t)j * sx + i];
        f[(size_t)j * sx + i] = f[(size_t)nj * sx + ni];
        f[(size_t)nj * sx + ni] = a;
    }
}

hv_status launch_shift_corners_c64(float2* buf, int batch, int sx, int sy, hipStream_t st) {
    dim3 grid((sx + 255) / 256, sy / 2);
    k_shift_corners<float2><<<grid, 256, 0, st>>>(buf, batch, sx, sy);
    HV_CHECK_LAUNCH("k_shift_corners");
    return HV_OK;
}
hv_status launch_shift_corners_f32(float* buf, int batch, int sx, int sy, hipStream_t st) {
    dim3 grid((sx + 255) / 256, sy / 2);
    k_shift_corners<float><<<grid, 256, 0, st>>>(buf, batch, sx, sy);
    HV_CHECK_LAUNCH("k_shift_corners");
    return HV_OK;
}

// ---------------------------------------------------------------------------
// row pass
// ---------------------------------------------------------------------------
// Input kinds of the row pass.
enum { IN_U8 = 0, IN_U16 = 1, IN_F32 = 2, IN_C64 = 3 };

// Tile shapes (compile-time; the defaults were chosen by measurement on MI355X)
#ifndef HV_ROW_E_SMALL
#define HV_ROW_E_SMALL 8        // values per thread, row lengths <= 512
#endif
#ifndef HV_ROW_THREADS
#define HV_ROW_THREADS 256      // threads per row workgroup
#endif
#ifndef HV_ROW_MIN_WG
#define HV_ROW_MIN_WG 1024      // keep at least this many row workgroups when pipelining
#endif
#ifndef HV_COL_E_SMALL
#define HV_COL_E_SMALL 8        // values per thread, column lengths <= 512
#endif
#ifndef HV_COL_CW_SMALL
#define HV_COL_CW_SMALL 8       // columns per workgroup, column lengths <= 512
#endif
#ifndef HV_COL_E_1024
#define HV_COL_E_1024 16        // values per thread, column length 1024
#endif
#ifndef HV_COL_E_2048
#define HV_COL_E_2048 16        // values per thread, column length 2048
#endif
#ifndef HV_COL_CW_1024
#define HV_COL_CW_1024 8        // columns per workgroup, column length 1024
#endif
#ifndef HV_COL_CW_BIG
#define HV_COL_CW_BIG 4         // columns per workgroup, column length 2048 (4096: 2)
#endif
#ifndef HV_COL_MIN_WG
#define HV_COL_MIN_WG 400       // keep at least this many column workgroups when pipelining (~ two per CU)
#endif

template <int N>
struct RowCfg {
    static constexpr int E = N >= 1024 ? 16 : HV_ROW_E_SMALL;
    static constexpr int NT = FftShape<N, E>::NT;
    static constexpr int M0 = (HV_ROW_THREADS / NT) > 0 ? (HV_ROW_THREADS / NT) : 1;  // rows per workgroup
    static constexpr int THREADS0 = M0 * NT;
    static constexpr int PITCH = N + N / 8;                      // padded LDS row
};

#ifndef HV_ROW_TILE_MUL
#define HV_ROW_TILE_MUL 2       // rows per workgroup x this when the row pass writes column tiles
#endif

// Rows per workgroup of a row pass with layout LAY (1: writes column tiles,
// see LAY_TO_TILES below): tiles hold 2^tl columns x H rows, so 4-row groups
// write whole 128-byte lines of 4-column tiles (57 planes of 2048^2:
// 1009 -> 910 us; doubling the rows of the pass that reads the tiles made it
// 45 us slower).
template <int N, int LAY>
struct RowCfgL : RowCfg<N> {
    static constexpr int M = RowCfg<N>::M0 * (LAY == 1 ? HV_ROW_TILE_MUL : 1);
    static constexpr int THREADS = M * RowCfg<N>::NT;
};

// Frames of the batch map to output frame (slot0 + b) mod ring (ring > 0), or
// to frame b of `out` (ring == 0).  Input frame b is frame b of `in`.
struct FrameMap {
    int ring;
    int slot0;
    __device__ __forceinline__ size_t out_frame(int b) const {
        return ring > 0 ? (size_t)((slot0 + b) % ring) : (size_t)b;
    }
};

// Input pixel as loaded: integer pixels stay unconverted until the row is
// transformed, so that the loads of the next row stay in flight meanwhile.
template <int KIN>
using RawPx = std::conditional_t<KIN == IN_C64, float2, std::conditional_t<KIN == IN_F32, float, uint32_t>>;

template <int KIN>
__device__ __forceinline__ RawPx<KIN> load_px(const void* in, size_t i) {
    if constexpr (KIN == IN_U8) return static_cast<const uint8_t*>(in)[i];
    else if constexpr (KIN == IN_U16) return static_cast<const uint16_t*>(in)[i];
    else if constexpr (KIN == IN_F32) return static_cast<const float*>(in)[i];
    else return static_cast<const float2*>(in)[i];
}

template <int KIN>
__device__ __forceinline__ float2 px_complex(RawPx<KIN> r) {
    if constexpr (KIN == IN_U8) return make_float2(convert_px<uint8_t>(static_cast<uint8_t>(r)), 0.f);
    else if constexpr (KIN == IN_U16) return make_float2(convert_px<uint16_t>(static_cast<uint16_t>(r)), 0.f);
    else if constexpr (KIN == IN_F32) return make_float2(r, 0.f);
    else return r;
}

// Column-tiled plane layout (long columns, see ColTiled below): element
// (y, x) of a frame at ((x >> tl) H + y) 2^tl + (x mod 2^tl), i.e. tiles of
// 2^tl columns x H rows stored contiguously.
__device__ __forceinline__ size_t tiled_off(int y, int x, int height, int tl) {
    return ((((size_t)(x >> tl) * height) + y) << tl) | (size_t)(x & ((1 << tl) - 1));
}

// Row-pass layouts: LAY_ROWS row-major in and out; LAY_TO_TILES row-major in,
// column-tiled out (the tiles are read back by k_row_ifft_epi_tiled).
enum { LAY_ROWS = 0, LAY_TO_TILES = 1 };

// Each workgroup transforms `iters` groups of M rows of one frame; the loads
// of group i + 1 (pixels, and chirp factors with CHIRP) are issued before
// group i is computed, so a workgroup keeps two groups of input in flight.
template <int N, int DIR, int KIN, bool CHIRP, int LAY = LAY_ROWS>
__global__ void __launch_bounds__((RowCfgL<N, LAY>::THREADS))
    k_row_fft(const void* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ chirp, int height,
              FrameMap fin, FrameMap fm, float scale, uint32_t* mm_reset, int mm_slots, int iters, int tl = 0) {
    using C = RowCfgL<N, LAY>;
    constexpr int kE = FftShape<N, C::E>::kE;
    __shared__ float2 lds[C::M * C::PITCH];
    // the cycle's first kernel re-arms the image min/max accumulators of its last kernels
    if (mm_reset && blockIdx.x == 0 && blockIdx.y == 0) mm_reset_slots(mm_reset, mm_slots);
    const int m = threadIdx.x / C::NT, t = threadIdx.x % C::NT;
    const int b = blockIdx.y;
    const size_t P = (size_t)N * height;
    const size_t in_frame = fin.out_frame(b) * P, out_frame = fm.out_frame(b) * P;
    float2* my = lds + m * C::PITCH;
    auto lds_at = [&](int n) -> float2& { return my[n + (n >> 3)]; };
    auto row_of = [&](int it) { return (blockIdx.x * iters + it) * C::M + m; };
    RawPx<KIN> px[kE];
    float2 cf[CHIRP ? kE : 1];
    auto load_row = [&](int yr) {
        // idle rows of the last workgroup read row 0; their results are not stored
        const int yc = yr < height ? yr : 0;
        const size_t row = (size_t)yc * N;
        fft_load_slots<N, C::E>(t, [&](int i, int n) {
            px[i] = load_px<KIN>(in, in_frame + row + n);
            if constexpr (CHIRP) cf[i] = chirp[row + n];
        });
    };
    load_row(row_of(0));
    float2 tw[FftTw<N, C::E>::COUNT];
    fft_twiddles<N, C::E>(t, tw);
    for (int it = 0; it < iters; ++it) {
        const int yr = row_of(it);
        float2 v[kE];
#pragma unroll
        for (int i = 0; i < kE; ++i) {
            v[i] = px_complex<KIN>(px[i]);
            if constexpr (CHIRP) v[i] = cmul(v[i], cf[i]);
        }
        if (it + 1 < iters) load_row(row_of(it + 1));
        const int yc = yr < height ? yr : 0;
        float2* dst = out + out_frame + (size_t)yc * N;
        const bool valid = yr < height;
        fft_compute_tw<N, C::E, DIR, false, (C::NT <= 64)>(t, v, tw, [&](int n, float2 z) {
            if (!valid) return;
            const float2 r = scale == 1.0f ? z : cscale(z, scale);
            if constexpr (LAY == LAY_TO_TILES) out[out_frame + tiled_off(yc, n, height, tl)] = r;
            else dst[n] = r;
        }, lds_at);
    }
}

// ---------------------------------------------------------------------------
// column pass
// ---------------------------------------------------------------------------
enum { COL_FWD = 0, COL_INV = 1, COL_LENS = 2 };

template <int N>
struct ColCfg {
    static constexpr int E = N >= 4096 ? 16 : (N == 2048 ? HV_COL_E_2048 : (N == 1024 ? HV_COL_E_1024 : HV_COL_E_SMALL));
    static constexpr int NT = FftShape<N, E>::NT;
    // columns per workgroup: 64-byte rows up to N = 512 (two 512-thread
    // workgroups per CU; xcd_tile gives the two tiles of a 128-byte line to
    // one XCD -- 16-column tiles of 1024 threads left one workgroup per CU
    // and were 5 % slower alone, 10 % with the lanes' kernels beside them);
    // longer columns narrow the tile (the N x CW tile and the registers bound it)
    static constexpr int CW = N <= 512 ? HV_COL_CW_SMALL : (N == 1024 ? HV_COL_CW_1024 : (N == 2048 ? HV_COL_CW_BIG : 2));
    static constexpr int THREADS = NT * CW;
};

// Column tile of workgroup bx of nx.  Workgroups go round-robin to the 8
// XCDs (linear id mod 8), each with its own L2; tiles narrower than a
// 128-byte line share lines with their neighbours, so the LG tiles of one
// line group are given to consecutive workgroups of one XCD and that XCD's L2
// fetches each line once (instead of once per XCD touching it).
template <int LG>
__device__ __forceinline__ int xcd_tile(int bx, int nx) {
    if constexpr (LG <= 1) {
        return bx;
    } else {
        if (nx % (8 * LG)) return bx;
        const int q = bx >> 3;
        return ((q / LG) * 8 + (bx & 7)) * LG + q % LG;
    }
}

// Each workgroup owns a CW-column tile and walks `fpb` frames of the batch,
// loading the next frame's tile while the current one is transformed; the
// lens values of its output positions stay in registers for all frames.
// Successive frames need no barrier between them: a transform's first LDS
// write follows a barrier inside the engine.
// ONE (columns <= 512): one frame per workgroup (fpb = 1), its tile's loads
// issued before the thread builds its twiddles, no next-frame registers: more
// workgroups per CU and the twiddle work under the load latency (512-long
// columns x 14 planes: 17.6 -> 15.5 us on MI355X).
// TILED: planes (and the lens) in the column-tiled layout with tiles of CW
// columns: each workgroup's tile is one contiguous block.  2048-long columns
// in 4-column row-major tiles read 32-byte runs, and a plain copy with that
// access pattern runs at 3.2 TB/s (5.5 TB/s contiguous); the tiled pass moves
// the transposition cost into the row passes, which write / read the tiles
// from whole rows (57 planes of 2048^2 on MI355X: column pass 1.77 -> 1.15 ms,
// row passes +0.15 and +0.19 ms).
template <int N, int MODE, bool ONE, bool TILED = false>
__global__ void __launch_bounds__(ColCfg<N>::THREADS)
    k_col_fft(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ lens, int width,
              FrameMap fin, FrameMap fm, int batch, int fpb) {
    using C = ColCfg<N>;
    constexpr int kE = FftShape<N, C::E>::kE;
    __shared__ float2 lds[N * C::CW];
    const int c = threadIdx.x % C::CW, t = threadIdx.x / C::CW;
    const int tile = TILED ? (int)blockIdx.x : xcd_tile<(16 / C::CW)>(blockIdx.x, gridDim.x);
    const int x = tile * C::CW + c;
    const int b0 = blockIdx.y * fpb;
    const size_t P = (size_t)N * width;
    // element n of this thread's column: x + n width, or the tile's block
    const size_t col0 = TILED ? (size_t)tile * N * C::CW + c : (size_t)x;
    const size_t step = TILED ? (size_t)C::CW : (size_t)width;
    auto lds_at = [&](int n) -> float2& { return lds[n * C::CW + c]; };
    auto load_frame = [&](int b, float2* v) {
        const float2* src = in + fin.out_frame(b < batch ? b : 0) * P + col0;
        fft_load<N, C::E>(t, [&](int n) -> float2 { return src[(size_t)n * step]; }, v);
    };
    if constexpr (ONE) {
        float2 v[kE], lr[MODE == COL_LENS ? kE : 1];
        load_frame(b0, v);
        if constexpr (MODE == COL_LENS) {
#pragma unroll
            for (int i = 0; i < kE; ++i) lr[i] = lens[col0 + (size_t)fft_out_index<N, C::E>(t, i) * step];
        }
        float2 tw[FftTw<N, C::E>::COUNT];
        fft_twiddles<N, C::E>(t, tw);
        float2* dst = out + fm.out_frame(b0) * P + col0;
        auto store = [&](int n, float2 z) { dst[(size_t)n * step] = z; };
        if constexpr (MODE == COL_LENS) {
            fft_compute_tw<N, C::E, -1>(t, v, static_cast<const float2*>(tw), [](int, float2) {}, lds_at);
#pragma unroll
            for (int k = 0; k < kE; ++k) v[k] = cmul(v[k], lr[k]);
            fft_compute_adj_tw<N, C::E, -1>(t, v, static_cast<const float2*>(tw), store, lds_at);
        } else {
            fft_compute_tw<N, C::E, MODE == COL_FWD ? -1 : +1>(t, v, static_cast<const float2*>(tw), store, lds_at);
        }
        return;
    }
    float2 lr[kE];
    if constexpr (MODE == COL_LENS) {
#pragma unroll
        for (int i = 0; i < kE; ++i) lr[i] = lens[col0 + (size_t)fft_out_index<N, C::E>(t, i) * step];
    }
    float2 cur[kE], nxt[kE];
    load_frame(b0, cur);
    float2 tw[FftTw<N, C::E>::COUNT];
    fft_twiddles<N, C::E>(t, tw);
    for (int i = 0; i < fpb; ++i) {
        const int b = b0 + i;
        if (b >= batch) break;  // uniform across the workgroup
        if (i + 1 < fpb) load_frame(b + 1, nxt);
        float2* dst = out + fm.out_frame(b) * P + col0;
        auto store = [&](int n, float2 v) { dst[(size_t)n * step] = v; };
        if constexpr (MODE == COL_LENS) {
            // forward column FFT with its results left in registers, x lens
            // (already / (W H)), inverse column FFT as the forward's adjoint
            fft_compute_tw<N, C::E, -1>(t, cur, static_cast<const float2*>(tw), [](int, float2) {}, lds_at);
#pragma unroll
            for (int k = 0; k < kE; ++k) cur[k] = cmul(cur[k], lr[k]);
            fft_compute_adj_tw<N, C::E, -1>(t, cur, static_cast<const float2*>(tw), store, lds_at);
        } else {
            fft_compute_tw<N, C::E, MODE == COL_FWD ? -1 : +1>(t, cur, static_cast<const float2*>(tw), store,
                                                                lds_at);
        }
#pragma unroll
        for (int k = 0; k < kE; ++k) cur[k] = nxt[k];
    }
}

// ---------------------------------------------------------------------------
// Frame sides 2^a 3^b 5^c that are not powers of two (camera geometries such
// as 1536 x 2048 or 1920 x 1080; cufftPlanMany accepts any n,
// icompute.cc:151-175).  Runtime-length Stockham FFT: the plan's radices
// (8, 4, 2, then 3s, 5s) are walked at run time, each stage in place in LDS
// (a transform of CW interleaved columns): every thread first reads the
// inputs of all its butterflies into registers (n / nt values), the
// workgroup synchronises, then each writes its outputs.  Twiddles come from
// an LDS table of exp(-2 pi i k / n), k < n, built once per workgroup with
// sincospif: exp(dir 2 pi i e / M) = table[e n / M] for M | n.
// ---------------------------------------------------------------------------
struct AnyPlan {
    int n, nst;
    int r[16];
};

constexpr int kAnyVals = 40;  // values per thread a stage may hold (n / nt <= 32 with nt = 64 or 128)

static bool any_plan(int n, AnyPlan& p) {
    if (n < 16 || n > 4096 || (n & 1)) return false;
    p.n = n;
    p.nst = 0;
    int m = n;
    while (m % 8 == 0) { p.r[p.nst++] = 8; m /= 8; }
    if (m % 4 == 0) { p.r[p.nst++] = 4; m /= 4; }
    if (m % 2 == 0) { p.r[p.nst++] = 2; m /= 2; }
    while (m % 3 == 0) { p.r[p.nst++] = 3; m /= 3; }
    while (m % 5 == 0) { p.r[p.nst++] = 5; m /= 5; }
    return m == 1 && p.nst <= 16;
}

// every stage's butterflies fit the threads' registers: ceil(n / (R nt)) <= kAnyVals / R
static bool any_plan_fits(const AnyPlan& p, int nt) {
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.r[s], m = p.n / R;
        if ((m + nt - 1) / nt > kAnyVals / R) return false;
    }
    return true;
}

bool fft_len_any_ok(long n) {
    AnyPlan p;
    return n <= 4096 && any_plan((int)n, p);
}

// radix-3 and radix-5 DFTs, natural order in and out
template <int DIR>
__device__ __forceinline__ void dft3(float2* v) {
    constexpr float s3 = 0.866025403784438647f;  // sin(2 pi / 3)
    const float2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    // (-DIR i) s3 t2: forward y1 = m - i s3 t2
    const float2 q = DIR < 0 ? make_float2(s3 * t2.y, -s3 * t2.x) : make_float2(-s3 * t2.y, s3 * t2.x);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, q);
    v[2] = csub(m, q);
}
template <int DIR>
__device__ __forceinline__ void dft5(float2* v) {
    constexpr float c1 = 0.309016994374947424f, c2 = -0.809016994374947424f;  // cos(2 pi / 5), cos(4 pi / 5)
    constexpr float s1 = 0.951056516295153572f, s2 = 0.587785252292473129f;   // sin(2 pi / 5), sin(4 pi / 5)
    const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    const float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    const float2 b1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    const float2 b2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    // forward: y1 = a1 - i b1, y4 = a1 + i b1, y2 = a2 - i b2, y3 = a2 + i b2
    const float2 q1 = DIR < 0 ? make_float2(b1.y, -b1.x) : make_float2(-b1.y, b1.x);
    const float2 q2 = DIR < 0 ? make_float2(b2.y, -b2.x) : make_float2(-b2.y, b2.x);
    v[0] = make_float2(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
    v[1] = cadd(a1, q1);
    v[4] = csub(a1, q1);
    v[2] = cadd(a2, q2);
    v[3] = csub(a2, q2);
}

// table[k] = exp(-2 pi i k / n), k < n, by the whole workgroup
__device__ __forceinline__ void any_twiddle_table(float2* table, int n) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        float s, c;
        sincospif(2.0f * (float)k / (float)n, &s, &c);
        table[k] = make_float2(c, -s);
    }
}

// One Stockham stage in place: butterfly j reads a[j + r m], m = n / R,
// multiplies by w^r (w = exp(dir 2 pi i (j mod ns) / (ns R))), runs the
// radix-R DFT and writes a[(j / ns) ns R + (j mod ns) + r ns].
template <int R, int DIR, int CW>
__device__ __forceinline__ void any_stage(float2* a, const float2* table, int n, int ns, int c, int t, int nt) {
    constexpr int MAXB = kAnyVals / R;
    const int m = n / R, step = n / (ns * R);
    float2 v[MAXB * R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = t + b * nt;
        if (j < m) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b * R + r] = a[(j + r * m) * CW + c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = t + b * nt;
        if (j < m) {
            float2* x = v + b * R;
            const int k = j % ns;
            if (ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 w = table[k * r * step];  // exp(-2 pi i k r / (ns R))
                    x[r] = DIR < 0 ? cmul(x[r], w) : cmulc(x[r], w);
                }
            }
            if constexpr (R == 3) dft3<DIR>(x);
            else if constexpr (R == 5) dft5<DIR>(x);
            else dft_reg<R, DIR>(x);
            const int d = (j / ns) * ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) a[(d + r * ns) * CW + c] = x[r];
        }
    }
    __syncthreads();
}

// The whole transform, in place in a (natural order in and out).  Every
// thread of the workgroup calls it (it contains barriers); the twiddle table
// is ready.
template <int DIR, int CW>
__device__ void any_fft(float2* a, const float2* table, const AnyPlan& pl, int c, int t, int nt) {
    const int n = pl.n;
    int ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.r[s];
        switch (R) {
            case 8: any_stage<8, DIR, CW>(a, table, n, ns, c, t, nt); break;
            case 4: any_stage<4, DIR, CW>(a, table, n, ns, c, t, nt); break;
            case 2: any_stage<2, DIR, CW>(a, table, n, ns, c, t, nt); break;
            case 3: any_stage<3, DIR, CW>(a, table, n, ns, c, t, nt); break;
            default: any_stage<5, DIR, CW>(a, table, n, ns, c, t, nt); break;
        }
        ns *= R;
    }
}

constexpr int kAnyRowThreads = 256;

// Row pass of any supported length: RPW rows of one frame per workgroup
// (256 / RPW threads per row), sharing one twiddle table; rows past the
// frame's height are computed on row 0 and not stored.
template <int KIN, bool CHIRP, int DIR, int RPW>
__global__ void __launch_bounds__(kAnyRowThreads)
    k_any_row(const void* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ chirp, int height,
              FrameMap fin, FrameMap fm, float scale, uint32_t* mm_reset, int mm_slots, AnyPlan pl) {
    constexpr int NT = kAnyRowThreads / RPW;
    extern __shared__ float2 any_sh[];
    if (mm_reset && blockIdx.x == 0 && blockIdx.y == 0) mm_reset_slots(mm_reset, mm_slots);
    const int n = pl.n, b = blockIdx.y;
    const int g = threadIdx.x / NT, t = threadIdx.x % NT;
    const int y = blockIdx.x * RPW + g;
    const bool valid = y < height;
    const size_t P = (size_t)n * height, row = (size_t)(valid ? y : 0) * n;
    const size_t fi = fin.out_frame(b) * P;
    float2* a = any_sh + (size_t)g * n;
    float2* table = any_sh + (size_t)RPW * n;
    for (int i = t; i < n; i += NT) {
        float2 v = px_complex<KIN>(load_px<KIN>(in, fi + row + i));
        if constexpr (CHIRP) v = cmul(v, chirp[row + i]);
        a[i] = v;
    }
    any_twiddle_table(table, n);
    __syncthreads();
    any_fft<DIR, 1>(a, table, pl, 0, t, NT);
    float2* dst = out + fm.out_frame(b) * P + row;
    if (valid)
        for (int i = t; i < n; i += NT) dst[i] = scale == 1.0f ? a[i] : cscale(a[i], scale);
}

// Column pass of any supported length: CW columns of one frame per workgroup,
// NT threads per column (columns past the frame's width are computed on zeros
// and not stored).
template <int MODE, int CW, int NT>
__global__ void __launch_bounds__(NT * CW)
    k_any_col(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ lens, int width,
              FrameMap fin, FrameMap fm, AnyPlan pl) {
    extern __shared__ float2 any_sh[];
    const int n = pl.n, b = blockIdx.y;
    const int c = threadIdx.x % CW, t = threadIdx.x / CW;
    const int x = blockIdx.x * CW + c;
    const bool valid = x < width;
    const size_t P = (size_t)n * width;
    const float2* src = in + fin.out_frame(b) * P + x;
    float2* a = any_sh;
    float2* table = any_sh + (size_t)n * CW;
    for (int i = t; i < n; i += NT) a[i * CW + c] = valid ? src[(size_t)i * width] : make_float2(0.f, 0.f);
    any_twiddle_table(table, n);
    __syncthreads();
    if constexpr (MODE == COL_LENS) {
        any_fft<-1, CW>(a, table, pl, c, t, NT);
        for (int i = t; i < n; i += NT)
            if (valid) a[i * CW + c] = cmul(a[i * CW + c], lens[(size_t)i * width + x]);
        __syncthreads();
        any_fft<+1, CW>(a, table, pl, c, t, NT);
    } else {
        any_fft<MODE == COL_FWD ? -1 : +1, CW>(a, table, pl, c, t, NT);
    }
    float2* dst = out + fm.out_frame(b) * P + x;
    if (valid)
        for (int i = t; i < n; i += NT) dst[(size_t)i * width] = a[i * CW + c];
}

template <class K>
static hv_status any_lds_attr(K* kernel, size_t bytes) {
    // more than the default 64 KiB of dynamic LDS for the longest transforms
    if (bytes > 64 * 1024) HV_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return HV_OK;
}

template <int DIR, int RPW>
static hv_status any_row_launch(const void* in, int kin, float2* out, const float2* chirp, int width, int height,
                                int batch, FrameMap fin, FrameMap fm, float scale, hipStream_t st, uint32_t* mm_reset,
                                int mm_slots, const AnyPlan& pl) {
    if (!any_plan_fits(pl, kAnyRowThreads / RPW)) return fail(HV_ERR_UNSUPPORTED, "row length not supported");
    const size_t lds = (size_t)(RPW + 1) * width * sizeof(float2);  // the rows and the twiddle table
    const dim3 grid((height + RPW - 1) / RPW, batch);
#define HV_ANY_ROW(KIN, CH)                                                                                       \
    do {                                                                                                          \
        HV_TRY(any_lds_attr(&k_any_row<KIN, CH, DIR, RPW>, lds));                                                 \
        k_any_row<KIN, CH, DIR, RPW><<<grid, kAnyRowThreads, lds, st>>>(in, out, chirp, height, fin, fm, scale,   \
                                                                        mm_reset, mm_slots, pl);                  \
    } while (0)
    if (chirp) {
        if constexpr (DIR < 0) {
            switch (kin) {
                case IN_U8: HV_ANY_ROW(IN_U8, true); break;
                case IN_U16: HV_ANY_ROW(IN_U16, true); break;
                case IN_F32: HV_ANY_ROW(IN_F32, true); break;
                default: HV_ANY_ROW(IN_C64, true); break;
            }
        }
    } else {
        switch (kin) {
            case IN_U8: HV_ANY_ROW(IN_U8, false); break;
            case IN_U16: HV_ANY_ROW(IN_U16, false); break;
            case IN_F32: HV_ANY_ROW(IN_F32, false); break;
            default: HV_ANY_ROW(IN_C64, false); break;
        }
    }
#undef HV_ANY_ROW
    HV_CHECK_LAUNCH("k_any_row");
    return HV_OK;
}

template <int DIR>
static hv_status any_row_pass(const void* in, int kin, float2* out, const float2* chirp, int width, int height,
                              int batch, FrameMap fin, FrameMap fm, float scale, hipStream_t st, uint32_t* mm_reset,
                              int mm_slots) {
    AnyPlan pl;
    if (!any_plan(width, pl)) return fail(HV_ERR_UNSUPPORTED, "width must be an even 2^a 3^b 5^c in [16, 4096]");
    if (chirp && DIR > 0) return fail(HV_ERR_UNSUPPORTED, "chirp factors on an inverse row pass");
    // four rows per workgroup share the twiddle table (two above 2048: 128 threads per row)
    if (width <= 2048)
        return any_row_launch<DIR, 4>(in, kin, out, chirp, width, height, batch, fin, fm, scale, st, mm_reset, mm_slots,
                                      pl);
    return any_row_launch<DIR, 2>(in, kin, out, chirp, width, height, batch, fin, fm, scale, st, mm_reset, mm_slots,
                                  pl);
}

template <int MODE, int CW, int NT>
static hv_status any_col_launch(const float2* in, float2* out, const float2* lens, int width, int height, int batch,
                                FrameMap fin, FrameMap fm, const AnyPlan& pl, hipStream_t st) {
    if (!any_plan_fits(pl, NT)) return fail(HV_ERR_UNSUPPORTED, "column length not supported");
    const size_t lds = (size_t)height * (CW + 1) * sizeof(float2);  // the tile and the twiddle table
    HV_TRY(any_lds_attr(&k_any_col<MODE, CW, NT>, lds));
    k_any_col<MODE, CW, NT><<<dim3((width + CW - 1) / CW, batch), NT * CW, lds, st>>>(in, out, lens, width, fin, fm, pl);
    HV_CHECK_LAUNCH("k_any_col");
    return HV_OK;
}

template <int MODE>
static hv_status any_col_pass(const float2* in, float2* out, const float2* lens, int width, int height, int batch,
                              FrameMap fin, FrameMap fm, hipStream_t st) {
    AnyPlan pl;
    if (!any_plan(height, pl)) return fail(HV_ERR_UNSUPPORTED, "height must be an even 2^a 3^b 5^c in [16, 4096]");
    // 4-column tiles (32-byte runs): 2048 x 4 x 8 B + the table = 80 KiB, two workgroups per CU
    if (height <= 2048) return any_col_launch<MODE, 4, 64>(in, out, lens, width, height, batch, fin, fm, pl, st);
    return any_col_launch<MODE, 2, 128>(in, out, lens, width, height, batch, fin, fm, pl, st);
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <int N, int DIR>
static hv_status row_pass_n(const void* in, int kin, float2* out, const float2* chirp, int height, int batch,
                            FrameMap fin, FrameMap fm, float scale, hipStream_t st, uint32_t* mm_reset,
                            int mm_slots) {
    using C = RowCfgL<N, LAY_ROWS>;
    // enough workgroups to fill the chip, each with a few row groups to pipeline
    const int groups = (height + C::M - 1) / C::M;
    int iters = 1;
    while (iters < 8 && groups % (iters * 2) == 0 && (size_t)(groups / (iters * 2)) * batch >= HV_ROW_MIN_WG) iters *= 2;
    dim3 grid(groups / iters, batch);
#define HV_ROW_LAUNCH(KIN, CH) \
    k_row_fft<N, DIR, KIN, CH><<<grid, C::THREADS, 0, st>>>(in, out, chirp, height, fin, fm, scale, mm_reset, mm_slots, iters, 0)
    if (chirp) {
        // chirp factors only multiply the forward pass's input (Fresnel)
        if constexpr (DIR < 0) {
            switch (kin) {
                case IN_U8: HV_ROW_LAUNCH(IN_U8, true); break;
                case IN_U16: HV_ROW_LAUNCH(IN_U16, true); break;
                case IN_F32: HV_ROW_LAUNCH(IN_F32, true); break;
                default: HV_ROW_LAUNCH(IN_C64, true); break;
            }
        } else {
            return fail(HV_ERR_UNSUPPORTED, "chirp factors on an inverse row pass");
        }
    } else {
        switch (kin) {
            case IN_U8: HV_ROW_LAUNCH(IN_U8, false); break;
            case IN_U16: HV_ROW_LAUNCH(IN_U16, false); break;
            case IN_F32: HV_ROW_LAUNCH(IN_F32, false); break;
            default: HV_ROW_LAUNCH(IN_C64, false); break;
        }
    }
#undef HV_ROW_LAUNCH
    HV_CHECK_LAUNCH("k_row_fft");
    return HV_OK;
}

template <int DIR>
static hv_status row_pass(const void* in, int kin, float2* out, const float2* chirp, int width, int height,
                          int batch, FrameMap fin, FrameMap fm, float scale, hipStream_t st,
                          uint32_t* mm_reset = nullptr, int mm_slots = 1) {
    switch (width) {
        case 16: return row_pass_n<16, DIR>(in, kin, out, chirp, height, batch, fin, fm, scale, st, mm_reset, mm_slots);
        case 32: return row_pass_n<32, DIR>(in, kin, out, chirp, height, batch, fin, fm, scale, st, mm_reset, mm_slots);
        case 64: return row_pass_n<64, DIR>(in, kin, out, chirp, height, batch, fin, fm, scale, st, mm_reset, mm_slots);
        case 128: return row_pass_n<128, DIR>(in, kin, out, chirp, height, batch, fin, fm, scale, st, mm_reset, mm_slots);
        case 256: return row_pass_n<256, DIR>(in, kin, out, chirp, height, batch, fin, fm, scale, st, mm_reset, mm_slots);
        case 512: return row_pass_n<512, DIR>(in, kin, out, chirp, height, batch, fin, fm, scale, st, mm_reset, mm_slots);
        case 1024: return row_pass_n<1024, DIR>(in, kin, out, chirp, height, batch, fin, fm, scale, st, mm_reset, mm_slots);
        case 2048: return row_pass_n<2048, DIR>(in, kin, out, chirp, height, batch, fin, fm, scale, st, mm_reset, mm_slots);
        case 4096: return row_pass_n<4096, DIR>(in, kin, out, chirp, height, batch, fin, fm, scale, st, mm_reset, mm_slots);
        default:
            return any_row_pass<DIR>(in, kin, out, chirp, width, height, batch, fin, fm, scale, st, mm_reset,
                                     mm_slots);
    }
}

template <int N, int MODE>
static hv_status col_pass_n(const float2* in, float2* out, const float2* lens, int width, int batch,
                            FrameMap fin, FrameMap fm, hipStream_t st) {
    using C = ColCfg<N>;
    // widths that are not a multiple of the column tile take the runtime-length kernel
    if (width % C::CW) return any_col_pass<MODE>(in, out, lens, width, N, batch, fin, fm, st);
    const int tiles = width / C::CW;
    if constexpr (N <= 512) {
        // one frame per workgroup
        dim3 grid(tiles, batch);
        k_col_fft<N, MODE, true><<<grid, C::THREADS, 0, st>>>(in, out, lens, width, fin, fm, batch, 1);
        HV_CHECK_LAUNCH("k_col_fft");
        return HV_OK;
    }
    // frames per workgroup: pipeline several while keeping >= 512 workgroups
    int fpb = 1;
    while (fpb < 8 && (size_t)tiles * ((batch + fpb * 2 - 1) / (fpb * 2)) >= HV_COL_MIN_WG) fpb *= 2;
    dim3 grid(tiles, (batch + fpb - 1) / fpb);
    k_col_fft<N, MODE, false><<<grid, C::THREADS, 0, st>>>(in, out, lens, width, fin, fm, batch, fpb);
    HV_CHECK_LAUNCH("k_col_fft");
    return HV_OK;
}

template <int MODE>
static hv_status col_pass(const float2* in, float2* out, const float2* lens, int width, int height, int batch,
                          FrameMap fin, FrameMap fm, hipStream_t st) {
    switch (height) {
        case 16: return col_pass_n<16, MODE>(in, out, lens, width, batch, fin, fm, st);
        case 32: return col_pass_n<32, MODE>(in, out, lens, width, batch, fin, fm, st);
        case 64: return col_pass_n<64, MODE>(in, out, lens, width, batch, fin, fm, st);
        case 128: return col_pass_n<128, MODE>(in, out, lens, width, batch, fin, fm, st);
        case 256: return col_pass_n<256, MODE>(in, out, lens, width, batch, fin, fm, st);
        case 512: return col_pass_n<512, MODE>(in, out, lens, width, batch, fin, fm, st);
        case 1024: return col_pass_n<1024, MODE>(in, out, lens, width, batch, fin, fm, st);
        case 2048: return col_pass_n<2048, MODE>(in, out, lens, width, batch, fin, fm, st);
        case 4096: return col_pass_n<4096, MODE>(in, out, lens, width, batch, fin, fm, st);
        default: return any_col_pass<MODE>(in, out, lens, width, height, batch, fin, fm, st);
    }
}

// ---------------------------------------------------------------------------
// column-tiled planes (2048- and 4096-long columns)
// ---------------------------------------------------------------------------
int col_tile_log2(int height) {
    if (height == 2048) return ilog2(ColCfg<2048>::CW);
    if (height == 4096) return ilog2(ColCfg<4096>::CW);
    return -1;
}

template <int N, int DIR, int LAY>
static hv_status row_pass_lay_n(const float2* in, float2* out, int height, int batch, int tl, hipStream_t st) {
    using C = RowCfgL<N, LAY>;
    const int groups = (height + C::M - 1) / C::M;
    int iters = 1;
    while (iters < 8 && groups % (iters * 2) == 0 && (size_t)(groups / (iters * 2)) * batch >= HV_ROW_MIN_WG) iters *= 2;
    const FrameMap fm{0, 0};
    k_row_fft<N, DIR, IN_C64, false, LAY><<<dim3(groups / iters, batch), C::THREADS, 0, st>>>(
        in, out, nullptr, height, fm, fm, 1.0f, nullptr, 1, iters, tl);
    HV_CHECK_LAUNCH("k_row_fft");
    return HV_OK;
}

template <int DIR, int LAY>
static hv_status row_pass_lay(const float2* in, float2* out, int width, int height, int batch, int tl,
                              hipStream_t st) {
    switch (width) {
        case 64: return row_pass_lay_n<64, DIR, LAY>(in, out, height, batch, tl, st);
        case 128: return row_pass_lay_n<128, DIR, LAY>(in, out, height, batch, tl, st);
        case 256: return row_pass_lay_n<256, DIR, LAY>(in, out, height, batch, tl, st);
        case 512: return row_pass_lay_n<512, DIR, LAY>(in, out, height, batch, tl, st);
        case 1024: return row_pass_lay_n<1024, DIR, LAY>(in, out, height, batch, tl, st);
        case 2048: return row_pass_lay_n<2048, DIR, LAY>(in, out, height, batch, tl, st);
        case 4096: return row_pass_lay_n<4096, DIR, LAY>(in, out, height, batch, tl, st);
        default: return fail(HV_ERR_UNSUPPORTED, "column-tiled planes: width must be a power of two in [64, 4096]");
    }
}

// The split path's last pass (column-tiled or row-major planes): inverse row
// FFT of every plane and the per-pixel epilogue in one kernel, so the
// propagated planes never return to HBM.  Each workgroup owns M rows and walks the K
// planes (the next plane's tile values loaded before the current one is
// transformed); a thread's outputs sit in the same register slots for every
// plane, where their sums accumulate:
//   EPI = kEpiMoments: moments (k_bins_moments' formulas: bins fs + k with
//            f = fftfreq(T) fps, the Nyquist bin counted twice, its M1 terms
//            cancelling -- added where the bin comes, not after the loop);
//   EPI = HV_IMG_*: image term mean over the planes x scale (k_bins_image).
// Then the outputs and the min/max partials of the image (or selected moment).
constexpr int kEpiMoments = -1;  // (image types are HV_IMG_MODULUS = 0, ...)
struct EpiArgs {
    float* m0;
    float* m1;
    float* m2;
    float* img;
    int T, fs, K, sel;
    float fps, count, scale;
    uint32_t* mm;
    int mm_slots;
};

// 8 values per thread (the plane sums, the next plane's values and the
// transform in registers: 16 values per thread left one wavefront per SIMD),
// rows per workgroup so that the tile runs of consecutive rows are read
// together (two rows: 64-byte runs of 4-column tiles).
#ifndef HV_EPI_ROWS
#define HV_EPI_ROWS 2           // rows per workgroup of the tiled epilogue pass for 2048-long rows
#endif
#ifndef HV_EPI_WAVES_PER_EU
#define HV_EPI_WAVES_PER_EU 1
#endif
template <int N>
struct EpiCfg {
    static constexpr int E = 8;
    static constexpr int NT = FftShape<N, E>::NT;
    static constexpr int M = NT >= 512 ? 1 : (NT >= 256 ? HV_EPI_ROWS : 512 / NT);
    static constexpr int THREADS = M * NT;
    static constexpr int PITCH = N + N / 8;
};

template <int N, int EPI>
__global__ void __launch_bounds__(EpiCfg<N>::THREADS) __attribute__((amdgpu_waves_per_eu(HV_EPI_WAVES_PER_EU)))
    k_row_ifft_epi_tiled(const float2* __restrict__ tiles, int height, int tl, EpiArgs ea) {
    using C = EpiCfg<N>;
    constexpr int kE = FftShape<N, C::E>::kE;
    __shared__ float2 lds[C::M * C::PITCH];
    const int m = threadIdx.x / C::NT, t = threadIdx.x % C::NT;
    const size_t P = (size_t)N * height;
    // tiles: a 128-byte line of a tile holds 16 >> tl rows; the workgroups
    // whose rows share lines go to one XCD one after another (its L2 then
    // fetches each line once)
    int grp = blockIdx.x;
    if (tl >= 0) {
        const int lg = (16 >> tl) / C::M;
        if (lg == 2) grp = xcd_tile<2>(blockIdx.x, gridDim.x);
        else if (lg == 4) grp = xcd_tile<4>(blockIdx.x, gridDim.x);
        else if (lg >= 8) grp = xcd_tile<8>(blockIdx.x, gridDim.x);
    }
    const int y = grp * C::M + m;
    const bool valid = y < height;
    const int yc = valid ? y : 0;
    float2* my = lds + m * C::PITCH;
    auto lds_at = [&](int n) -> float2& { return my[n + (n >> 3)]; };
    // the slots' element offsets within a plane, the same for every plane
    // (tl < 0: row-major planes)
    uint32_t off[kE];
    fft_load_slots<N, C::E>(t, [&](int i, int n) {
        off[i] = tl >= 0 ? (uint32_t)tiled_off(yc, n, height, tl) : (uint32_t)(yc * N + n);
    });
    float2 px[kE];
    auto load_plane = [&](int k) {
        const float2* src = tiles + (size_t)k * P;
#pragma unroll
        for (int i = 0; i < kE; ++i) px[i] = src[off[i]];
    };
    load_plane(0);
    constexpr bool MOM = EPI == kEpiMoments;
    float a0[kE], a1[MOM ? kE : 1], a2[MOM ? kE : 1];
#pragma unroll
    for (int i = 0; i < kE; ++i) {
        a0[i] = 0.f;
        if constexpr (MOM) {
            a1[i] = 0.f;
            a2[i] = 0.f;
        }
    }
    const float d = MOM ? ea.fps / ea.T : 0.f;
    const int nyq = ea.T / 2;
    for (int k = 0; k < ea.K; ++k) {
        float2 v[kE];
#pragma unroll
        for (int i = 0; i < kE; ++i) v[i] = px[i];
        if (k + 1 < ea.K) load_plane(k + 1);
        const int bin = ea.fs + k;
        const float f = bin < nyq ? bin * d : -((float)ea.T - bin) * d;
        // twiddles generated per stage: the sums and the next plane's values hold the registers
        fft_compute_tw<N, C::E, +1, false, (C::NT <= 64)>(t, v, TwOnTheFly{}, [&](int, float2 z, int slot) {
            if constexpr (MOM) {
                const float sq = z.x * z.x + z.y * z.y;
                a0[slot] += sq * 1.0f;
                a1[slot] += sq * f;
                a2[slot] += sq * (f * f);
                if (bin == nyq) {  // the Nyquist bin counted again: its frequency's sign cancels in M1
                    a0[slot] += sq * 1.0f;
                    a1[slot] -= sq * f;
                    a2[slot] += sq * (f * f);
                }
            } else {
                a0[slot] += image_term_fast<EPI>(z);
            }
        }, lds_at);
    }
    float vmin = INFINITY, vmax = -INFINITY;
    if (valid) {
#pragma unroll
        for (int i = 0; i < kE; ++i) {
            const size_t p = (size_t)y * N + fft_out_index<N, C::E>(t, i);
            float v;
            if constexpr (MOM) {
                if (ea.m0) ea.m0[p] = a0[i];
                if (ea.m1) ea.m1[p] = a1[i];
                if (ea.m2) ea.m2[p] = a2[i];
                v = ea.sel == 0 ? a0[i] : ea.sel == 1 ? a1[i] : a2[i];
            } else {
                v = a0[i] / ea.count * ea.scale;
            }
            if (ea.img) ea.img[p] = v;
            vmin = fminf(vmin, v);
            vmax = fmaxf(vmax, v);
        }
    }
    if (ea.mm) block_minmax_atomic(vmin, vmax, ea.mm, ea.mm_slots);
}

template <int N, int EPI>
static hv_status row_ifft_epi_n(const float2* tiles, int height, int tl, const EpiArgs& ea, hipStream_t st) {
    using C = EpiCfg<N>;
    k_row_ifft_epi_tiled<N, EPI><<<(height + C::M - 1) / C::M, C::THREADS, 0, st>>>(tiles, height, tl, ea);
    HV_CHECK_LAUNCH("k_row_ifft_epi_tiled");
    return HV_OK;
}

bool row_ifft_epi_supported(int width) { return is_pow2(width) && width >= 64 && width <= 4096; }

template <int EPI>
static hv_status row_ifft_epi(const float2* tiles, int width, int height, const EpiArgs& ea, hipStream_t st,
                              bool tiled) {
    const int tl = tiled ? col_tile_log2(height) : -1;
    switch (width) {
        case 64: return row_ifft_epi_n<64, EPI>(tiles, height, tl, ea, st);
        case 128: return row_ifft_epi_n<128, EPI>(tiles, height, tl, ea, st);
        case 256: return row_ifft_epi_n<256, EPI>(tiles, height, tl, ea, st);
        case 512: return row_ifft_epi_n<512, EPI>(tiles, height, tl, ea, st);
        case 1024: return row_ifft_epi_n<1024, EPI>(tiles, height, tl, ea, st);
        case 2048: return row_ifft_epi_n<2048, EPI>(tiles, height, tl, ea, st);
        case 4096: return row_ifft_epi_n<4096, EPI>(tiles, height, tl, ea, st);
        default: return fail(HV_ERR_UNSUPPORTED, "column-tiled planes: width must be a power of two in [64, 4096]");
    }
}

hv_status launch_row_ifft_moments(float* m0, float* m1, float* m2, float* img, const float2* planes, int width,
                                  int height, int T, int fs, int K, float fps, int sel, uint32_t* mm, int mm_slots,
                                  hipStream_t st, bool tiled) {
    const EpiArgs ea{m0, m1, m2, img, T, fs, K, sel, fps, 1.f, 1.f, mm, mm_slots};
    return row_ifft_epi<kEpiMoments>(planes, width, height, ea, st, tiled);
}

hv_status launch_row_ifft_image(float* out, const float2* planes, int width, int height, int K, int img_type,
                                float count, float scale, uint32_t* mm, int mm_slots, hipStream_t st, bool tiled) {
    const EpiArgs ea{nullptr, nullptr, nullptr, out, 0, 0, K, 0, 0.f, count, scale, mm, mm_slots};
    switch (img_type) {
        case HV_IMG_MODULUS: return row_ifft_epi<HV_IMG_MODULUS>(planes, width, height, ea, st, tiled);
        case HV_IMG_SQUARED_MODULUS: return row_ifft_epi<HV_IMG_SQUARED_MODULUS>(planes, width, height, ea, st, tiled);
        case HV_IMG_ARGUMENT: return row_ifft_epi<HV_IMG_ARGUMENT>(planes, width, height, ea, st, tiled);
        default: return fail(HV_ERR_UNSUPPORTED, "bins-first path: image type must be modulus, squared or argument");
    }
}

bool col_tiled_supported(int width, int height) {
    const int tl = col_tile_log2(height);
    return tl >= 0 && is_pow2(width) && width >= 64 && width <= 4096 && width % (1 << tl) == 0;
}

hv_status launch_row_pass_tiled(const float2* in, float2* out, int width, int height, int planes, int direction,
                                hipStream_t st) {
    if (!col_tiled_supported(width, height)) return fail(HV_ERR_UNSUPPORTED, "column-tiled planes: unsupported shape");
    if (in == out) return fail(HV_ERR_INVALID_ARG, "a layout-changing row pass runs out of place");
    if (direction >= 0) return fail(HV_ERR_UNSUPPORTED, "column tiles are written by the forward row pass only");
    return row_pass_lay<-1, LAY_TO_TILES>(in, out, width, height, planes, col_tile_log2(height), st);
}

template <int N>
static hv_status col_lens_tiled_n(float2* buf, const float2* lens_t, int width, int batch, hipStream_t st) {
    using C = ColCfg<N>;
    const int tiles = width / C::CW;
    int fpb = 1;
    while (fpb < 8 && (size_t)tiles * ((batch + fpb * 2 - 1) / (fpb * 2)) >= HV_COL_MIN_WG) fpb *= 2;
    const FrameMap fm{0, 0};
    k_col_fft<N, COL_LENS, false, true><<<dim3(tiles, (batch + fpb - 1) / fpb), C::THREADS, 0, st>>>(
        buf, buf, lens_t, width, fm, fm, batch, fpb);
    HV_CHECK_LAUNCH("k_col_fft");
    return HV_OK;
}

hv_status launch_col_lens_tiled(float2* buf, const float2* lens_t, int width, int height, int planes,
                                hipStream_t st) {
    if (!col_tiled_supported(width, height)) return fail(HV_ERR_UNSUPPORTED, "column-tiled planes: unsupported shape");
    return height == 2048 ? col_lens_tiled_n<2048>(buf, lens_t, width, planes, st)
                          : col_lens_tiled_n<4096>(buf, lens_t, width, planes, st);
}

__global__ void k_tile_frame(float2* __restrict__ out, const float2* __restrict__ in, int width, int height, int tl) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x < width) out[tiled_off(y, x, height, tl)] = in[(size_t)y * width + x];
}

hv_status launch_tile_frame(float2* out, const float2* in, int width, int height, hipStream_t st) {
    if (!col_tiled_supported(width, height)) return fail(HV_ERR_UNSUPPORTED, "column-tiled planes: unsupported shape");
    k_tile_frame<<<dim3((width + 255) / 256, height), 256, 0, st>>>(out, in, width, height, col_tile_log2(height));
    HV_CHECK_LAUNCH("k_tile_frame");
    return HV_OK;
}

static int depth_to_kin(int depth) {
    return depth == HV_DEPTH_8 ? IN_U8 : depth == HV_DEPTH_16 ? IN_U16 : depth == HV_DEPTH_32F ? IN_F32 : -1;
}

hv_status launch_col_pass(float2* buf, const float2* lens, int width, int height, int planes, int mode,
                          hipStream_t st) {
    const FrameMap fm{0, 0};
    switch (mode) {
        case COLPASS_FWD: return col_pass<COL_FWD>(buf, buf, nullptr, width, height, planes, fm, fm, st);
        case COLPASS_INV: return col_pass<COL_INV>(buf, buf, nullptr, width, height, planes, fm, fm, st);
        default: return col_pass<COL_LENS>(buf, buf, lens, width, height, planes, fm, fm, st);
    }
}

hv_status launch_row_pass(float2* buf, int width, int height, int planes, int direction, hipStream_t st) {
    const FrameMap fm{0, 0};
    if (direction < 0) return row_pass<-1>(buf, IN_C64, buf, nullptr, width, height, planes, fm, fm, 1.0f, st);
    return row_pass<+1>(buf, IN_C64, buf, nullptr, width, height, planes, fm, fm, 1.0f, st);
}

hv_status launch_fft2(float2* buf, int batch, int width, int height, int direction, hipStream_t st) {
    FrameMap fm{0, 0};
    if (direction < 0) {
        HV_TRY(row_pass<-1>(buf, IN_C64, buf, nullptr, width, height, batch, fm, fm, 1.0f, st));
        return col_pass<COL_FWD>(buf, buf, nullptr, width, height, batch, fm, fm, st);
    }
    HV_TRY(row_pass<+1>(buf, IN_C64, buf, nullptr, width, height, batch, fm, fm, 1.0f, st));
    return col_pass<COL_INV>(buf, buf, nullptr, width, height, batch, fm, fm, st);
}

hv_status launch_angular_spectrum(float2* buf, int batch, const float2* lens, int width, int height,
                                  float post_scale, hipStream_t st) {
    FrameMap fm{0, 0};
    HV_TRY(row_pass<-1>(buf, IN_C64, buf, nullptr, width, height, batch, fm, fm, 1.0f, st));
    HV_TRY(col_pass<COL_LENS>(buf, buf, lens, width, height, batch, fm, fm, st));
    return row_pass<+1>(buf, IN_C64, buf, nullptr, width, height, batch, fm, fm, post_scale, st);
}

hv_status launch_fresnel(float2* buf, int batch, const float2* chirp, int width, int height, hipStream_t st) {
    FrameMap fm{0, 0};
    HV_TRY(row_pass<-1>(buf, IN_C64, buf, chirp, width, height, batch, fm, fm, 1.0f, st));
    return col_pass<COL_FWD>(buf, buf, nullptr, width, height, batch, fm, fm, st);
}

hv_status launch_propagate_to_ring(const void* raw, int depth, int batch, int width, int height, int space,
                                   const float2* lens, float2* ring, int ring_size, int slot0, hipStream_t st,
                                   hipEvent_t* marks, bool spectral_rows, uint32_t* mm_reset, int mm_slots) {
    // marks (optional): events recorded after pass 1 and after pass 2.
    // spectral_rows: stop the angular spectrum before its inverse row pass
    // (the fused row-IFFT + STFT kernel finishes it).
    const int kin = depth_to_kin(depth);
    if (kin < 0) return fail(HV_ERR_INVALID_ARG, "depth must be 1, 2 or 4");
    const size_t P = (size_t)width * height;
    // Only the last ring_size frames of an oversized batch survive the enqueue
    // (queue.cc:337-343); the earlier ones are skipped.
    const int skip = batch > ring_size ? batch - ring_size : 0;
    const char* src = (const char*)raw + (size_t)skip * P * depth;
    const int nb = batch - skip;
    const FrameMap plain{0, 0};
    const FrameMap slots{ring_size, (slot0 + skip) % ring_size};
    if (space == HV_SPACE_ANGULARSP) {
        // pass 1 writes the frames into their ring slots; passes 2-3 run in place there
        HV_TRY(row_pass<-1>(src, kin, ring, nullptr, width, height, nb, plain, slots, 1.0f, st, mm_reset, mm_slots));
        if (marks) HV_HIP(hipEventRecord(marks[0], st));
        HV_TRY(col_pass<COL_LENS>(ring, ring, lens, width, height, nb, slots, slots, st));
        if (marks) HV_HIP(hipEventRecord(marks[1], st));
        if (spectral_rows) return HV_OK;
        return row_pass<+1>(ring, IN_C64, ring, nullptr, width, height, nb, slots, slots, 1.0f, st);
    }
    if (space == HV_SPACE_FRESNELTR) {
        HV_TRY(row_pass<-1>(src, kin, ring, lens, width, height, nb, plain, slots, 1.0f, st, mm_reset, mm_slots));
        if (marks) {
            HV_HIP(hipEventRecord(marks[0], st));
            HV_HIP(hipEventRecord(marks[1], st));
        }
        return col_pass<COL_FWD>(ring, ring, nullptr, width, height, nb, slots, slots, st);
    }
    // no space transform: conversion only, frame by frame into the slots
    if (mm_reset) HV_TRY(launch_minmax_init(mm_reset, st, mm_slots));
    for (int b = 0; b < nb; ++b) {
        float2* dst = ring + (size_t)((slots.slot0 + b) % ring_size) * P;
        HV_TRY(launch_convert(dst, src + (size_t)b * P * depth, P, depth, st));
    }
    if (marks) {
        HV_HIP(hipEventRecord(marks[0], st));
        HV_HIP(hipEventRecord(marks[1], st));
    }
    return HV_OK;
}

// ---------------------------------------------------------------------------
// 2-D unwrap (unwrap_2d, tools_unwrap.cu:269-369) in four passes, power-of-two
// sides.  The reference's steps: z = exp(i phase); (gx, gy) = IFFT2(FFT2(z)
// (fx, fy)); g' = g 2 pi i conj(i z); G = FFT2(gx') fx / |f|^2 + FFT2(gy')
// fy / |f|^2 (0 at the origin); out = Im IFFT2(G) / -2 pi -- seven 2-D
// transforms of one or two frames and five element-wise kernels.  fx depends
// on the row frequency only, so the column transforms around its product
// cancel to a factor H: gx = H IFFT_row(fx FFT_row(z)), a 1-D filter along
// rows; likewise gy = W IFFT_col(fy FFT_col(z)).  The passes:
//   1. columns: phase -> z -> gy' -> column FFT                   (Gy_c)
//   2. rows, two jobs: phase -> z -> gx' -> row FFT (Gx_r), and Gy_c -> row FFT (Gy)
//   3. columns: Gx_r -> column FFT (Gx), G from Gx and Gy, inverse column FFT
//   4. rows: inverse row FFT, Im / -2 pi, the output's min/max
// Every product and ratio is the reference's element formula (kfreq,
// unwrap_eq_term, the |f|^2 ratio); transforms are unnormalised as cuFFT's.
// ---------------------------------------------------------------------------
// One frame only: 1024-long columns in 4-column tiles (256 workgroups
// instead of 128 of 8 columns), the other lengths as the column pass.
template <int N>
struct UnwrapColCfg : ColCfg<N> {
    static constexpr int CW = N == 1024 ? 4 : ColCfg<N>::CW;
    static constexpr int THREADS = ColCfg<N>::NT * CW;
};

template <int N>
__global__ void __launch_bounds__(UnwrapColCfg<N>::THREADS)
    k_unwrap_cols_a(const float* __restrict__ phase, float2* __restrict__ gyc, int width, uint32_t* mm_reset,
                    int mm_slots) {
    using C = UnwrapColCfg<N>;
    constexpr int kE = FftShape<N, C::E>::kE;
    __shared__ float2 lds[N * C::CW];
    // the unwrap's last pass accumulates the output's min/max
    if (mm_reset && blockIdx.x == 0) mm_reset_slots(mm_reset, mm_slots);
    const int c = threadIdx.x % C::CW, t = threadIdx.x / C::CW;
    const int x = xcd_tile<(16 / C::CW)>(blockIdx.x, gridDim.x) * C::CW + c;
    auto lds_at = [&](int n) -> float2& { return lds[n * C::CW + c]; };
    float ph[kE];
    fft_load_slots<N, C::E>(t, [&](int i, int n) { ph[i] = phase[(size_t)n * width + x]; });
    float2 tw[FftTw<N, C::E>::COUNT];
    fft_twiddles<N, C::E>(t, tw);
    float2 z[kE], v[kE];
#pragma unroll
    for (int i = 0; i < kE; ++i) {
        sincosf(ph[i], &z[i].y, &z[i].x);
        v[i] = z[i];
    }
    fft_compute_tw<N, C::E, -1>(t, v, static_cast<const float2*>(tw), [](int, float2) {}, lds_at);
#pragma unroll
    for (int i = 0; i < kE; ++i) {
        const float fy = kfreq(fft_out_index<N, C::E>(t, i), N);
        v[i] = make_float2(v[i].x * fy, v[i].y * fy);
    }
    const float wf = (float)width;
    fft_compute_adj_tw<N, C::E, -1>(t, v, static_cast<const float2*>(tw), [&](int, float2 g, int slot) {
        v[slot] = unwrap_eq_term(make_float2(g.x * wf, g.y * wf), z[slot]);
    }, lds_at);
    float2* dst = gyc + x;
    fft_compute_tw<N, C::E, -1>(t, v, static_cast<const float2*>(tw),
                                [&](int n, float2 g) { dst[(size_t)n * width] = g; }, lds_at);
}

// blockIdx.y 0: phase rows -> gx' -> row FFT into gxr; 1: gy rows (column
// transformed) -> row FFT in place.
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::THREADS0)
    k_unwrap_rows_a(const float* __restrict__ phase, float2* __restrict__ gxr, float2* __restrict__ gy, int height) {
    using C = RowCfg<N>;
    constexpr int kE = FftShape<N, C::E>::kE;
    __shared__ float2 lds[C::M0 * C::PITCH];
    const int m = threadIdx.x / C::NT, t = threadIdx.x % C::NT;
    const int y = blockIdx.x * C::M0 + m;
    const bool valid = y < height;
    const size_t row = (size_t)(valid ? y : 0) * N;
    float2* my = lds + m * C::PITCH;
    auto lds_at = [&](int n) -> float2& { return my[n + (n >> 3)]; };
    float2 tw[FftTw<N, C::E>::COUNT];
    float2 v[kE];
    if (blockIdx.y == 0) {
        float2 z[kE];
        fft_load_slots<N, C::E>(t, [&](int i, int n) {
            sincosf(phase[row + n], &z[i].y, &z[i].x);
        });
        fft_twiddles<N, C::E>(t, tw);
#pragma unroll
        for (int i = 0; i < kE; ++i) v[i] = z[i];
        fft_compute_tw<N, C::E, -1, false, (C::NT <= 64)>(t, v, static_cast<const float2*>(tw), [](int, float2) {}, lds_at);
#pragma unroll
        for (int i = 0; i < kE; ++i) {
            const float fx = kfreq(fft_out_index<N, C::E>(t, i), N);
            v[i] = make_float2(v[i].x * fx, v[i].y * fx);
        }
        const float hf = (float)height;
        fft_compute_adj_tw<N, C::E, -1, (C::NT <= 64)>(t, v, static_cast<const float2*>(tw), [&](int, float2 g, int slot) {
            v[slot] = unwrap_eq_term(make_float2(g.x * hf, g.y * hf), z[slot]);
        }, lds_at);
        float2* dst = gxr + row;
        fft_compute_tw<N, C::E, -1, false, (C::NT <= 64)>(t, v, static_cast<const float2*>(tw), [&](int n, float2 g) {
            if (valid) dst[n] = g;
        }, lds_at);
    } else {
        fft_load<N, C::E>(t, [&](int n) -> float2 { return gy[row + n]; }, v);
        fft_twiddles<N, C::E>(t, tw);
        float2* dst = gy + row;
        // a row is loaded whole before its first exchange (one wavefront, or a barrier): in place is safe
        fft_compute_tw<N, C::E, -1, false, (C::NT <= 64)>(t, v, static_cast<const float2*>(tw), [&](int n, float2 g) {
            if (valid) dst[n] = g;
        }, lds_at);
    }
}

// Gx = column FFT of gxr; G = Gx fx / |f|^2 + Gy fy / |f|^2 (kernel_norm_ratio
// + kernel_add_complex_frames, 0 where |f| = 0); inverse column FFT, in place.
template <int N>
__global__ void __launch_bounds__(UnwrapColCfg<N>::THREADS)
    k_unwrap_cols_b(float2* __restrict__ g, const float2* __restrict__ gy, int width) {
    using C = UnwrapColCfg<N>;
    constexpr int kE = FftShape<N, C::E>::kE;
    __shared__ float2 lds[N * C::CW];
    const int c = threadIdx.x % C::CW, t = threadIdx.x / C::CW;
    const int x = xcd_tile<(16 / C::CW)>(blockIdx.x, gridDim.x) * C::CW + c;
    auto lds_at = [&](int n) -> float2& { return lds[n * C::CW + c]; };
    float2 v[kE], b[kE];
    fft_load<N, C::E>(t, [&](int n) -> float2 { return g[(size_t)n * width + x]; }, v);
#pragma unroll
    for (int i = 0; i < kE; ++i) b[i] = gy[(size_t)fft_out_index<N, C::E>(t, i) * width + x];
    float2 tw[FftTw<N, C::E>::COUNT];
    fft_twiddles<N, C::E>(t, tw);
    fft_compute_tw<N, C::E, -1>(t, v, static_cast<const float2*>(tw), [](int, float2) {}, lds_at);
    const float fx = kfreq(x, width);
#pragma unroll
    for (int i = 0; i < kE; ++i) {
        const float fy = kfreq(fft_out_index<N, C::E>(t, i), N);
        const float norm = fx * fx + fy * fy;
        float2 a = make_float2(0.f, 0.f), e = make_float2(0.f, 0.f);
        if (norm != 0) {
            const float cx = fx / norm, cy = fy / norm;
            a = make_float2(v[i].x * cx, v[i].y * cx);
            e = make_float2(b[i].x * cy, b[i].y * cy);
        }
        v[i] = cadd(a, e);
    }
    float2* dst = g + x;
    fft_compute_adj_tw<N, C::E, -1>(t, v, static_cast<const float2*>(tw),
                                    [&](int n, float2 z) { dst[(size_t)n * width] = z; }, lds_at);
}

// Inverse row FFT, out = Im / -2 pi (kernel_unwrap2d_last_step), the
// output's min/max into mm (nullptr: none).
template <int N>
__global__ void __launch_bounds__(RowCfg<N>::THREADS0)
    k_unwrap_rows_b(const float2* __restrict__ g, float* __restrict__ out, int height, uint32_t* mm, int mm_slots) {
    using C = RowCfg<N>;
    constexpr int kE = FftShape<N, C::E>::kE;
    __shared__ float2 lds[C::M0 * C::PITCH];
    const int m = threadIdx.x / C::NT, t = threadIdx.x % C::NT;
    const int y = blockIdx.x * C::M0 + m;
    const bool valid = y < height;
    const size_t row = (size_t)(valid ? y : 0) * N;
    float2* my = lds + m * C::PITCH;
    auto lds_at = [&](int n) -> float2& { return my[n + (n >> 3)]; };
    float2 v[kE];
    fft_load<N, C::E>(t, [&](int n) -> float2 { return g[row + n]; }, v);
    float2 tw[FftTw<N, C::E>::COUNT];
    fft_twiddles<N, C::E>(t, tw);
    float vmin = INFINITY, vmax = -INFINITY;
    float* dst = out + row;
    fft_compute_tw<N, C::E, +1, false, (C::NT <= 64)>(t, v, static_cast<const float2*>(tw), [&](int n, float2 z) {
        if (!valid) return;
        const float o = (float)((double)z.y / -(2.0 * M_PI));
        dst[n] = o;
        vmin = fminf(vmin, o);
        vmax = fmaxf(vmax, o);
    }, lds_at);
    if (mm) block_minmax_atomic(vmin, vmax, mm, mm_slots);
}

template <int NX, int NY>
static hv_status unwrap_fused_n(float* out, const float* phase, void* scratch, hipStream_t st, uint32_t* mm,
                                int slots) {
    const size_t P = (size_t)NX * NY;
    float2* gx = reinterpret_cast<float2*>(scratch);
    float2* gy = gx + P;
    using CC = UnwrapColCfg<NY>;
    using RC = RowCfg<NX>;
    const unsigned col_groups = NX / CC::CW, row_groups = (NY + RC::M0 - 1) / RC::M0;
    k_unwrap_cols_a<NY><<<col_groups, CC::THREADS, 0, st>>>(phase, gy, NX, mm, slots);
    HV_CHECK_LAUNCH("k_unwrap_cols_a");
    k_unwrap_rows_a<NX><<<dim3(row_groups, 2), RC::THREADS0, 0, st>>>(phase, gx, gy, NY);
    HV_CHECK_LAUNCH("k_unwrap_rows_a");
    k_unwrap_cols_b<NY><<<col_groups, CC::THREADS, 0, st>>>(gx, gy, NX);
    HV_CHECK_LAUNCH("k_unwrap_cols_b");
    k_unwrap_rows_b<NX><<<row_groups, RC::THREADS0, 0, st>>>(gx, out, NY, mm, slots);
    HV_CHECK_LAUNCH("k_unwrap_rows_b");
    return HV_OK;
}

bool unwrap_fused_supported(int width, int height) {
    return is_pow2(width) && is_pow2(height) && width >= 16 && height >= 16 && width <= 4096 && height <= 4096;
}

template <int NX>
static hv_status unwrap_fused_x(float* out, const float* phase, int height, void* scratch, hipStream_t st,
                                uint32_t* mm, int slots) {
    switch (height) {
        case 16: return unwrap_fused_n<NX, 16>(out, phase, scratch, st, mm, slots);
        case 32: return unwrap_fused_n<NX, 32>(out, phase, scratch, st, mm, slots);
        case 64: return unwrap_fused_n<NX, 64>(out, phase, scratch, st, mm, slots);
        case 128: return unwrap_fused_n<NX, 128>(out, phase, scratch, st, mm, slots);
        case 256: return unwrap_fused_n<NX, 256>(out, phase, scratch, st, mm, slots);
        case 512: return unwrap_fused_n<NX, 512>(out, phase, scratch, st, mm, slots);
        case 1024: return unwrap_fused_n<NX, 1024>(out, phase, scratch, st, mm, slots);
        case 2048: return unwrap_fused_n<NX, 2048>(out, phase, scratch, st, mm, slots);
        case 4096: return unwrap_fused_n<NX, 4096>(out, phase, scratch, st, mm, slots);
        default: return fail(HV_ERR_UNSUPPORTED, "fused unwrap: height");
    }
}

// Scratch: two complex frames (the step's unwrap_scratch_bytes holds three).
hv_status launch_unwrap_2d_fused(float* out, const float* phase, int width, int height, void* scratch,
                                 hipStream_t st, uint32_t* mm, int slots) {
    if (!unwrap_fused_supported(width, height)) return fail(HV_ERR_UNSUPPORTED, "fused unwrap: frame sides");
    switch (width) {
        case 16: return unwrap_fused_x<16>(out, phase, height, scratch, st, mm, slots);
        case 32: return unwrap_fused_x<32>(out, phase, height, scratch, st, mm, slots);
        case 64: return unwrap_fused_x<64>(out, phase, height, scratch, st, mm, slots);
        case 128: return unwrap_fused_x<128>(out, phase, height, scratch, st, mm, slots);
        case 256: return unwrap_fused_x<256>(out, phase, height, scratch, st, mm, slots);
        case 512: return unwrap_fused_x<512>(out, phase, height, scratch, st, mm, slots);
        case 1024: return unwrap_fused_x<1024>(out, phase, height, scratch, st, mm, slots);
        case 2048: return unwrap_fused_x<2048>(out, phase, height, scratch, st, mm, slots);
        default: return unwrap_fused_x<4096>(out, phase, height, scratch, st, mm, slots);
    }
}

}  // namespace hv
