// hv_pipe.cpp -- the per-batch orchestrator (Pipe::exec / refresh,
// Backend/sources/core/pipe.cc:255-387, 650-680) and the multi-source server.
//
// A pipe owns every device buffer it will ever use (allocated at create, the
// reference's refresh time): the T-slot time-window ring, the lens, the
// image / contrast / u16 outputs, reduction and PCA scratch, and an input
// staging buffer.  The control flow the reference runs as lambdas
// (stride discard pipe.cc:400-416, ring fill wait :418-427, autocontrast once
// rendering.cc:239-258) stays on the host; the device work of a steady-state
// cycle is captured once per ring phase into a hipGraph and replayed.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <memory>
#include <vector>

#include "hv_kernels.hpp"

using namespace hv;

namespace {

bool is_image_kind(int t) {
    return t == HV_IMG_MODULUS || t == HV_IMG_SQUARED_MODULUS || t == HV_IMG_ARGUMENT;
}
bool is_moment_kind(int t) { return t == HV_IMG_MOMENTS_0 || t == HV_IMG_MOMENTS_1 || t == HV_IMG_MOMENTS_2; }

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// insert_compute_p_accu (converts.cc:76-92): inclusive bin range of the image
void p_range(const hv_compute_settings& s, int& pmin, int& pmax) {
    const int T = (int)s.time_transformation_size;
    pmin = (int)s.p_start;
    pmax = s.p_width ? std::min<int>(pmin + (int)s.p_width, T) : pmin;
}

bool window_replaced(const hv_compute_settings& s) {
    return s.batch_size == s.time_transformation_size && s.time_stride == s.batch_size;
}

// The whole window is replaced by every batch (B = S = T), so the angular
// spectrum's inverse row pass can be fused with the STFT.
bool fused_eligible(const hv_compute_settings& s) {
    return s.space_transformation == HV_SPACE_ANGULARSP && s.time_transformation == HV_TIME_STFT &&
           is_image_kind(s.img_type) && window_replaced(s) &&
           row_ifft_stft_supported((int)s.width, (int)s.time_transformation_size);
}

// The STFT bins an image reads: [pmin, pmax] for the image types, [f_start,
// f_end] with f_end = min(start + width, T - 1) for the moments (pipe.cc:429-455).
void bins_range(const hv_compute_settings& s, int& lo, int& K) {
    if (is_moment_kind(s.img_type)) {
        lo = (int)s.p_start;
        K = std::min<int>((int)(s.p_start + s.p_width), (int)s.time_transformation_size - 1) - lo + 1;
    } else {
        int pmax;
        p_range(s, lo, pmax);
        K = pmax - lo + 1;
    }
}

// B = S = T (or B < T, from a raw ring) STFT or PCA image or moments with a space transform: the needed
// bins (STFT) or components (PCA) can be taken before the propagation (both
// are linear combinations of frames and commute with the linear propagator;
// the PCA covariance of the propagated window is the raw window's up to a
// constant, the propagators being unitary up to scale).
// B < T (a sliding window): the raw frames of the window are kept in a raw
// ring (slot of frame n = n mod T, as the reference's ring) and the bins are
// taken from it each cycle; the propagation stays per bin.
bool raw_window(const hv_compute_settings& s) { return s.batch_size < s.time_transformation_size; }

bool bins_eligible(const hv_compute_settings& s) {
    if (!(s.space_transformation == HV_SPACE_ANGULARSP || s.space_transformation == HV_SPACE_FRESNELTR)) return false;
    if (s.time_transformation != HV_TIME_STFT && s.time_transformation != HV_TIME_PCA) return false;
    if (!window_replaced(s) && !raw_window(s)) return false;
    if (!is_image_kind(s.img_type) && !is_moment_kind(s.img_type)) return false;
    int lo, K;
    bins_range(s, lo, K);
    const int W = (int)s.width, H = (int)s.height, T = (int)s.time_transformation_size;
    if (s.time_transformation == HV_TIME_PCA && !pca_raw_cov_supported(T, (size_t)W * H)) return false;
    return bins_first_supported(W, H, T, K, (int)s.depth) || bins_split_supported(W, H, T, K, (int)s.depth);
}

// The bins-first split path with column-tiled planes (2048- and 4096-long
// columns, angular spectrum): the tiles take the ring slots after the K
// planes, so the window must hold 2 K frames.
bool bins_tiled(const hv_compute_settings& s) {
    if (!bins_eligible(s)) return false;
    int lo, K;
    bins_range(s, lo, K);
    const int T = (int)s.time_transformation_size;
    return 2 * K <= T && bins_tiled_supported((int)s.width, (int)s.height, T, K, (int)s.depth,
                                              s.space_transformation, s.time_transformation == HV_TIME_PCA);
}

// Execution path of a steady-state cycle.
enum { PATH_THREE_PASS = 0, PATH_ROW_STFT = 1, PATH_BINS = 2 };

int choose_path(const hv_compute_settings& s, int level) {
    if (s.cuts_enabled) return PATH_THREE_PASS;  // the cuts read the whole propagated window
    if (level >= PATH_BINS && bins_eligible(s)) return PATH_BINS;
    if (level >= PATH_ROW_STFT && fused_eligible(s)) return PATH_ROW_STFT;
    return PATH_THREE_PASS;
}

struct Layout {
    size_t ring, raw, lens, lens_t, staging, image, cimage, tmp, out, small, mm, pct, unwrap, cov, evals, evecs, evecs32, pca_ws,
        eig_ws;
    size_t accu, renorm, cuts, reg, outq;
    size_t total;
};

Layout layout_of(const hv_compute_settings& s) {
    Layout l{};
    const size_t P = (size_t)s.width * s.height;
    const size_t T = s.time_transformation_size;
    l.ring = align256(T * P * sizeof(float2));
    // the raw frames of a sliding window (bins-first with B < T)
    l.raw = raw_window(s) && choose_path(s, PATH_BINS) == PATH_BINS ? align256(T * P * s.depth) : 0;
    l.lens = s.space_transformation != HV_SPACE_NONE ? align256(P * sizeof(float2)) : 0;
    l.lens_t = bins_tiled(s) ? l.lens : 0;  // the lens again, column-tiled
    l.staging = align256((size_t)s.batch_size * P * s.depth);
    l.image = align256(P * sizeof(float));
    l.cimage = align256(P * sizeof(float));
    const bool unwrap = s.img_type == HV_IMG_ARGUMENT && s.unwrap_2d;
    l.tmp = unwrap ? align256(P * sizeof(float)) : 0;
    l.out = align256(P * sizeof(uint16_t));
    l.small = 256;   // contrast range (2 floats) at 0, output-queue slot words at 64
    // per-block image min/max partials: a header pair + one float pair per
    // block of the widest image kernel (P / 64 blocks)
    l.mm = align256(sizeof(float2) * (1 + (P + 63) / 64 + 4096));
    l.pct = align256(percentile_scratch_bytes());
    l.unwrap = unwrap ? align256(unwrap_scratch_bytes(s.width, s.height)) : 0;
    if (s.time_transformation == HV_TIME_PCA) {
        l.cov = align256(T * T * sizeof(double2));  // FP64 covariance for the FP64 eigensolver
        l.evals = align256(T * sizeof(double));
        l.evecs = align256(T * T * sizeof(double2));
        // the bins-first projection's weights (+ 8 columns read past the last)
        l.evecs32 = align256((T * T + 8) * sizeof(float2));
        l.pca_ws = align256(pca_workspace_bytes((int)T, P, s.pca_precision));
        l.eig_ws = align256(eig_workspace_bytes((int)T));
    }
    // image accumulation ring (+ its position) and renormalisation sums
    l.accu = s.img_accu_level > 1 ? align256((size_t)s.img_accu_level * P * sizeof(float)) + 256 : 0;
    l.renorm = s.renorm_enabled ? align256(renorm_scratch_bytes()) : 0;
    // cuts: the window's STFT, then the XZ and YZ images
    l.cuts = s.cuts_enabled ? align256(T * P * sizeof(float2)) + 2 * align256(T * std::max(s.width, s.height) * 4) : 0;
    // registration: mask, reference spectrum, work plane, and a shifted copy
    l.reg = s.registration_enabled ? align256(registration_scratch_bytes(P)) + align256(P * sizeof(float)) : 0;
    // the output queue: OutputBufferSize u16 images
    l.outq = align256((size_t)s.output_buffer_size * P * sizeof(uint16_t));
    l.total = l.outq + l.ring + l.raw + l.lens + l.lens_t + l.staging + l.image + l.cimage + l.tmp + l.out + l.small + l.mm + l.pct + l.unwrap + l.cov +
              l.evals + l.evecs + l.evecs32 + l.pca_ws + l.eig_ws + l.accu + l.renorm + l.cuts + l.reg;
    return l;
}

}  // namespace

constexpr int kStages = 5;
constexpr size_t kMaxGraphs = 64;

// Image min/max accumulation of a cycle's producers: per-block partials, or
// (measurement builds, -DHV_PIPE_MM_MODE=n) n atomic accumulator pairs.
#ifndef HV_PIPE_MM_MODE
#define HV_PIPE_MM_MODE kMmPartials
#endif
constexpr int kPipeMm = HV_PIPE_MM_MODE;

// Buffers of one in-flight cycle.  Lane 0 lives in the pipe's main
// allocation; further lanes (bins-first path only) have their own.
struct Lane {
    hipStream_t stream = nullptr;
    float2* planes = nullptr;  // bins-first planes (lane 0: the ring's first K frames)
    float2* tiles = nullptr;   // column-tiled planes (lane 0: the ring's next K frames), when the path tiles
    // time-window ring and PCA buffers (lane 0: the pipe's; further lanes of a
    // window-replaced three-pass pipe: their own)
    float2* ring = nullptr;
    double2* cov = nullptr;
    double* evals = nullptr;
    double2* evecs = nullptr;
    float2* evecs32 = nullptr;
    void* pca_ws = nullptr;
    void* eig_ws = nullptr;
    float* image = nullptr;
    float* cimage = nullptr;
    float* tmp = nullptr;
    uint16_t* out = nullptr;
    uint32_t* mm = nullptr;
    float* range = nullptr;
    uint32_t* oq_pos = nullptr;  // output-queue slot of this lane's next image, finished-block count
    void* pct = nullptr;
    void* unwrap_ws = nullptr;
    void* mem = nullptr;  // own allocation (lanes >= 1)
    int kind = 0;         // LANE_*: what the own allocation holds besides the image buffers
    // shared-ring lanes: recorded once the cycle's time transform has read the ring
    hipEvent_t tt_done = nullptr;
    // graphs, one per (ring phase, input buffer) pair
    std::map<std::pair<int, const void*>, hipGraphExec_t> graphs;
};

constexpr int kMaxLanes = 4;

// Lane kinds.  B = S = T: a cycle shares no device state with the next, so a
// lane holds its own bins-first planes (LANE_BINS) or ring and PCA buffers
// (LANE_OWN_RING).  Otherwise (LANE_SHARED_RING) the lanes share lane 0's ring
// and PCA buffers: cycle n + 1's propagation waits until cycle n's time
// transform has read the ring, and cycle n's image post-processing (unwrap,
// percentile, contrast) runs beside cycle n + 1's propagation.
enum { LANE_BINS = 0, LANE_OWN_RING = 1, LANE_SHARED_RING = 2 };

struct hv_pipe {
    hv_compute_settings s{};
    bool own_stream = false;
    size_t P = 0;
    int T = 0, B = 0;
    Layout lay{};
    char* base = nullptr;
    float2* ring = nullptr;
    void* raw = nullptr;            // sliding-window bins-first: the window's raw frames, slot n = frame n mod T
    float2* lens = nullptr;
    float2* lens_t = nullptr;  // the lens column-tiled (bins_tiled)
    void* staging = nullptr;
    double2* cov = nullptr;
    double* evals = nullptr;
    double2* evecs = nullptr;
    float2* evecs32 = nullptr;
    void* pca_ws = nullptr;
    void* eig_ws = nullptr;
    float* accu_ring = nullptr;     // img_accu_level images
    uint32_t* accu_state = nullptr; // {next slot, images held}
    void* renorm_ws = nullptr;
    float* rec = nullptr;           // moments record: M0, M1, M2 of the latest cycle (own allocation)
    float2* cuts_stft = nullptr;    // cuts: the window's STFT, XZ and YZ images
    float* cut_xz = nullptr;
    float* cut_yz = nullptr;
    void* reg = nullptr;            // registration scratch
    float* reg_tmp = nullptr;
    bool reg_ref_set = false;       // the reference image has been taken
    // output queue (Queue semantics, queue.cc:149-189): image k in slot k mod oq_cap
    uint16_t* outq = nullptr;
    uint32_t oq_cap = 0;
    uint64_t out_count = 0;         // images produced since creation
    uint32_t oq_start = 0, oq_size = 0;
    bool oq_overwritten = false;
    std::vector<hipEvent_t> oq_ev;  // completion of the image in each slot
    std::vector<char> oq_ev_rec;
    uint32_t accu_held = 0;         // host mirror of the accumulation queue's size
    // host mirror of the reference's state
    uint32_t batch_index = 0;  // BatchEnv::batch_index
    uint32_t q_start = 0, q_size = 0;
    bool autocontrast_pending = true;
    bool have_output = false;
    bool graph_enabled = false;
    // in-flight cycles: cycle n runs on lane n mod nlanes (1 unless set by
    // hv_pipe_set_lanes); `cur` is the lane being enqueued, `last` the lane
    // of the latest output
    std::vector<Lane> lane = std::vector<Lane>(1);
    int nlanes = 1, cur = 0, last = 0, next = 0;
    // cycle path (PATH_*): three passes + STFT, inverse row pass fused with
    // the STFT, or bins-first
    int path = PATH_THREE_PASS;
    int level = PATH_BINS;  // highest path allowed (hv_pipe_set_fused)
    // stage timer: events between row pass, column pass, inverse row pass,
    // image kernel(s) and the contrast / u16 stage
    bool timing = false;
    hipEvent_t ev[kStages + 1] = {};
    double stage_ms[kStages] = {};
    int timed_cycles = 0;

    hv_status image_work();
    hv_status enqueue_cycle(const void* frames, int slot0);
    // The propagated ring of the window the raw ring holds, for path `to`
    // (the sliding-window bins-first path keeps only the raw frames); the
    // pipe's streams are idle.
    hv_status complex_ring_from_raw(int to) {
        if (!raw_ring()) return HV_OK;
        return launch_propagate_to_ring(raw, (int)s.depth, T, (int)s.width, (int)s.height, s.space_transformation,
                                        lens, ring, T, 0, lane[0].stream, nullptr, to == PATH_ROW_STFT);
    }
    // where lane L's cycle writes its u16 image
    OutSlot out_slot(const Lane& L) const {
        OutSlot o;
        if (!oq_cap) return o;
        o.ring = outq;
        o.cap = oq_cap;
        o.step = timing ? 1u : (uint32_t)nlanes;
        o.pos = L.oq_pos;
        return o;
    }
    // After a change of lanes, timing or cycle order (next = 0): lane i
    // writes images out_count + i, out_count + i + nlanes, ...
    hv_status reset_output_slots() {
        if (!oq_cap) return HV_OK;
        for (auto& l : lane)
            if (l.stream) HV_HIP(hipStreamSynchronize(l.stream));
        const int n = timing ? 1 : nlanes;
        for (int i = 0; i < n && i < (int)lane.size(); ++i) {
            const uint32_t w[2] = {(uint32_t)((out_count + (uint64_t)i) % oq_cap), 0u};
            HV_HIP(hipMemcpy(lane[i].oq_pos, w, sizeof(w), hipMemcpyHostToDevice));
        }
        return HV_OK;
    }
    // Queue::enqueue bookkeeping of the image the current lane just queued
    hv_status output_enqueued() {
        if (!oq_cap) return HV_OK;
        const uint32_t slot = (uint32_t)(out_count % oq_cap);
        HV_HIP(hipEventRecord(oq_ev[slot], lane[cur].stream));
        oq_ev_rec[slot] = 1;
        ++out_count;
        if (oq_size < oq_cap) {
            ++oq_size;
        } else {
            oq_start = (oq_start + 1) % oq_cap;
            oq_overwritten = true;
        }
        return HV_OK;
    }
    // the bins-first path reads a sliding window (B < T) from the raw ring
    bool raw_ring() const { return raw && path == PATH_BINS; }
    // the raw ring follows the window while the bins-first path runs or may
    // resume (a moments record holds the three-pass path meanwhile)
    bool raw_keep() const { return raw && (path == PATH_BINS || (rec && level >= PATH_BINS)); }
    // cycles on several lanes write and read one window: the three-pass ring or the raw ring
    bool shared_ring() const { return nlanes > 1 && (lane[1].kind == LANE_SHARED_RING || raw_ring()); }
    // Queue::enqueue_multiple of raw frames (queue.cc:332-376): batch frame b
    // to slot (slot0 + b) mod T, B < T, so at most two contiguous runs
    hv_status raw_store(const void* frames, int slot0, hipStream_t st) {
        const size_t fb = P * s.depth;
        const int n1 = std::min(B, T - slot0);
        HV_HIP(hipMemcpyAsync((char*)raw + (size_t)slot0 * fb, frames, (size_t)n1 * fb, hipMemcpyDeviceToDevice, st));
        if (B > n1)
            HV_HIP(hipMemcpyAsync(raw, (const char*)frames + (size_t)n1 * fb, (size_t)(B - n1) * fb,
                                  hipMemcpyDeviceToDevice, st));
        return HV_OK;
    }
    // min/max folded into the image kernel, accumulators re-armed by the
    // cycle's first kernel (the row pass of a space transform)
    bool minmax_fused() const {
        return s.contrast_enabled && s.contrast_mode == HV_CONTRAST_MINMAX &&
               !(s.img_type == HV_IMG_ARGUMENT && s.unwrap_2d) && !s.fft_shift && !post_steps() &&
               s.space_transformation != HV_SPACE_NONE;
    }
    // registration's launch sequence is fixed once its reference is taken and
    // the accumulation queue is full: until then its cycles run eagerly
    bool registration_steady() const {
        const uint32_t level = std::max<uint32_t>(1, s.img_accu_level);
        return !s.registration_enabled || (reg_ref_set && (level <= 1 || accu_held == level));
    }
    // contrast reticle scale for the percentiles (< 0: the whole image)
    float reticle() const { return s.contrast_reticle_enabled ? s.contrast_reticle_scale : -1.f; }
    // renormalisation, accumulation or log10 change the image after its kernel
    bool post_steps() const {
        return s.renorm_enabled || s.img_accu_level > 1 || s.log_enabled || s.registration_enabled;
    }
    hv_status mark(int i) {
        if (timing) HV_HIP(hipEventRecord(ev[i], lane[cur].stream));
        return HV_OK;
    }
    void drop_graphs() {
        for (auto& l : lane) {
            for (auto& kv : l.graphs) (void)hipGraphExecDestroy(kv.second);
            l.graphs.clear();
        }
    }
    hv_status collect_times() {
        HV_HIP(hipEventSynchronize(ev[kStages]));
        for (int i = 0; i < kStages; ++i) {
            float t = 0.f;
            HV_HIP(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
            stage_ms[i] += t;
        }
        ++timed_cycles;
        return HV_OK;
    }
};

// Device work of one cycle after the ring holds T frames (pipe.cc:333-380).
hv_status hv_pipe::image_work() {
    Lane& L = lane[cur];
    const int it = s.img_type;
    const int Tn = T;
    const bool minmax_mode = s.contrast_enabled && s.contrast_mode == HV_CONTRAST_MINMAX;
    const bool unwrap = it == HV_IMG_ARGUMENT && s.unwrap_2d;
    // the image's min/max is folded in by the producing kernel when it is the
    // final image; the cycle's first kernel has re-armed the accumulators
    uint32_t* mm_img = minmax_fused() ? L.mm : nullptr;
    if (is_image_kind(it)) {
        int pmin, pmax;
        p_range(s, pmin, pmax);
        if (path == PATH_BINS) {
            // PCA images are scaled by 2^17 - 1 after the mean (converts.cc:50-61)
            const float scale = s.time_transformation == HV_TIME_PCA ? (float)((2 << 16) - 1) : 1.0f;
            HV_TRY(launch_bins_image(L.image, L.planes, (int)s.width, (int)s.height, pmax - pmin + 1,
                                     s.space_transformation, it, pmin, pmax, mm_img, L.stream, kPipeMm, scale,
                                     L.tiles));
        } else if (s.time_transformation == HV_TIME_STFT && path == PATH_ROW_STFT) {
            HV_TRY(launch_row_ifft_stft_image(L.image, L.ring, (int)s.width, (int)s.height, Tn, it, pmin, pmax, mm_img,
                                              L.stream, kPipeMm));
        } else if (s.time_transformation == HV_TIME_STFT) {
            HV_TRY(launch_stft_image(L.image, L.ring, Tn, P, it, pmin, pmax, mm_img, L.stream, kPipeMm));
        } else if (s.time_transformation == HV_TIME_PCA) {
            HV_TRY(launch_pca_cov(L.cov, 1, L.ring, Tn, P, s.pca_precision, L.pca_ws, L.stream));
            HV_TRY(launch_eig_jacobi(L.evals, L.evecs, L.cov, 1, Tn, L.eig_ws, L.stream));
            HV_TRY(launch_pca_image(L.image, L.ring, L.evecs, Tn, P, it, pmin, pmax, mm_img, L.stream, kPipeMm));
        } else {
            HV_TRY(launch_complex_to_image(L.image, L.ring, it, pmin, pmax, P, L.stream));
            if (mm_img) HV_TRY(launch_minmax(L.mm, L.image, P, L.stream, kPipeMm));
        }
        // the ring has been read (the bins-first path records it after its time DFT)
        if (shared_ring() && path != PATH_BINS) HV_HIP(hipEventRecord(L.tt_done, L.stream));
        if (unwrap) {
            // the unwrap's last pass leaves its output's min/max in mm for the rescale
            HV_TRY(launch_unwrap_2d(L.tmp, L.image, s.width, s.height, L.unwrap_ws, L.stream, L.mm, kPipeMm));
            HV_TRY(launch_rescale_unwrap(L.image, L.tmp, P, L.mm, L.stream, kPipeMm, true));
        }
    } else if (path == PATH_BINS) {
        // moments of the propagated bins [f_start, f_end] (insert_moments, pipe.cc:429-455)
        int lo, K;
        bins_range(s, lo, K);
        const int k = it - HV_IMG_MOMENTS_0;
        HV_TRY(launch_bins_moments(k == 0 ? L.image : nullptr, k == 1 ? L.image : nullptr, k == 2 ? L.image : nullptr,
                                   L.planes, (int)s.width, (int)s.height, Tn, lo, K, s.space_transformation,
                                   s.camera_fps, k, mm_img, L.stream, kPipeMm, L.tiles));
    } else {
        // insert_moments (pipe.cc:429-455): f_end = min(start + width, T - 1)
        const int fs = (int)s.p_start;
        const int fe = std::min<int>((int)(s.p_start + s.p_width), Tn - 1);
        if (s.time_transformation == HV_TIME_PCA) {
            // moments of the PCA output (gpu_p_acc_buffer = H V, fourier_transform.cc:340-377)
            HV_TRY(launch_pca_cov(L.cov, 1, L.ring, Tn, P, s.pca_precision, L.pca_ws, L.stream));
            HV_TRY(launch_eig_jacobi(L.evals, L.evecs, L.cov, 1, Tn, L.eig_ws, L.stream));
            if (rec) {
                // the moments record's projection gives the image too
                HV_TRY(launch_pca_moments3(rec, rec + P, rec + 2 * P, L.ring, L.evecs, Tn, P, fs, fe, s.camera_fps,
                                           L.stream));
                HV_HIP(hipMemcpyAsync(L.image, rec + (size_t)(it - HV_IMG_MOMENTS_0) * P, P * sizeof(float),
                                      hipMemcpyDeviceToDevice, L.stream));
                if (mm_img) HV_TRY(launch_minmax(L.mm, L.image, P, L.stream, kPipeMm));
            } else {
                HV_TRY(launch_pca_image(L.image, L.ring, L.evecs, Tn, P, it, fs, fe, mm_img, L.stream, kPipeMm,
                                        s.camera_fps));
            }
        } else if (rec) {
            // the moments record's pass gives the image too
            HV_TRY(launch_stft_moments(rec, rec + P, rec + 2 * P, L.ring, Tn, P, fs, fe, s.camera_fps, mm_img,
                                       it - HV_IMG_MOMENTS_0, L.stream, kPipeMm));
            HV_HIP(hipMemcpyAsync(L.image, rec + (size_t)(it - HV_IMG_MOMENTS_0) * P, P * sizeof(float),
                                  hipMemcpyDeviceToDevice, L.stream));
        } else {
            float* m0 = it == HV_IMG_MOMENTS_0 ? L.image : nullptr;
            float* m1 = it == HV_IMG_MOMENTS_1 ? L.image : nullptr;
            float* m2 = it == HV_IMG_MOMENTS_2 ? L.image : nullptr;
            HV_TRY(launch_stft_moments(m0, m1, m2, L.ring, Tn, P, fs, fe, s.camera_fps, mm_img,
                                       it - HV_IMG_MOMENTS_0, L.stream, kPipeMm));
        }
        if (shared_ring()) HV_HIP(hipEventRecord(L.tt_done, L.stream));  // the ring has been read
    }
    // XZ / YZ cuts of the STFT (insert_time_transformation_cuts_view, fourier_transform.cc:466-511)
    if (s.cuts_enabled) {
        HV_TRY(launch_stft(cuts_stft, L.ring, Tn, P, L.stream));
        HV_TRY(launch_cuts(cut_xz, cut_yz, cuts_stft, (int)s.width, (int)s.height, Tn, (int)s.x_start, (int)s.y_start,
                           (int)(s.x_start + s.x_width), (int)(s.y_start + s.y_width), it, L.stream));
    }
    // moments record (insert_moments + insert_moments_record, pipe.cc:429-455,
    // 583-598): the three moment frames of the window, whatever the image type
    if (rec && !is_moment_kind(it)) {  // a moment image's pass has written the record already
        const int fs = (int)s.p_start;
        const int fe = std::min<int>((int)(s.p_start + s.p_width), Tn - 1);
        if (s.time_transformation == HV_TIME_PCA) {
            HV_TRY(launch_pca_moments3(rec, rec + P, rec + 2 * P, L.ring, L.evecs, Tn, P, fs, fe, s.camera_fps,
                                       L.stream));
        } else {
            HV_TRY(launch_stft_moments(rec, rec + P, rec + 2 * P, L.ring, Tn, P, fs, fe, s.camera_fps, nullptr, 0,
                                       L.stream, kPipeMm));
        }
    }
    // post-processing in the reference's order (pipe.cc:353-369): renormalise,
    // fft shift, accumulate, log10
    if (s.renorm_enabled) HV_TRY(launch_renormalize(L.image, P, s.renorm_constant, renorm_ws, L.stream));
    if (s.fft_shift) HV_TRY(launch_shift_corners_f32(L.image, 1, s.width, s.height, L.stream));
    // registration runs once the accumulation queue is full (or without one)
    // against the reference taken after the accumulation of the first such
    // cycle (registration.cc:11-20, 71-90)
    const uint32_t level = std::max<uint32_t>(1, s.img_accu_level);
    const bool queue_full_before = level <= 1 || accu_held == level;
    if (s.registration_enabled && queue_full_before && reg_ref_set)
        HV_TRY(launch_registration(reg, L.image, reg_tmp, (int)s.width, (int)s.height, L.stream));
    if (s.img_accu_level > 1)
        HV_TRY(launch_accumulate(L.image, accu_ring, accu_state, (int)s.img_accu_level, P, L.stream));
    if (level > 1) accu_held = std::min(accu_held + 1, level);
    if (s.registration_enabled && !reg_ref_set && (level <= 1 || accu_held == level)) {
        HV_TRY(launch_registration_set_reference(reg, L.image, (int)s.width, (int)s.height, L.stream));
        reg_ref_set = true;
    }
    if (s.log_enabled) HV_TRY(launch_log10(L.image, P, L.stream));
    HV_TRY(mark(4));
    if (!s.contrast_enabled) return launch_float_to_ushort(L.out, L.image, P, L.stream, out_slot(L));
    switch (s.contrast_mode) {
        case HV_CONTRAST_MINMAX:
            if (!minmax_fused()) {
                // min/max of the final image (unwrap rescales it; no first-kernel reset without a space transform)
                HV_TRY(launch_minmax_init(L.mm, L.stream, kPipeMm));
                HV_TRY(launch_minmax(L.mm, L.image, P, L.stream, kPipeMm));
            }
            // the contrast kernel decodes the accumulators and records the range it applied
            return launch_contrast_to_ushort(L.out, L.cimage, L.image, P, L.mm, kPipeMm, s.contrast_invert, L.stream,
                                             L.range, out_slot(L));
        case HV_CONTRAST_PERCENTILE:
            HV_TRY(launch_percentile(L.range, L.image, s.width, s.height, (int)s.contrast_subsample, s.contrast_lower,
                                     s.contrast_upper, 1, L.pct, L.stream, reticle()));
            break;
        case HV_CONTRAST_REFERENCE:
            if (autocontrast_pending)
                HV_TRY(launch_percentile(L.range, L.image, s.width, s.height, (int)s.contrast_subsample,
                                         s.contrast_lower, s.contrast_upper, 1, L.pct, L.stream, reticle()));
            break;
        default:  // MANUAL: range uploaded at create
            break;
    }
    return launch_contrast_to_ushort(L.out, L.cimage, L.image, P, L.range, 0, s.contrast_invert, L.stream, nullptr,
                                     out_slot(L));
}

hv_status hv_pipe::enqueue_cycle(const void* frames, int slot0) {
    Lane& L = lane[cur];
    // shared ring: this cycle overwrites slots the previous cycle's time
    // transform reads (and reads the slots it keeps)
    if (shared_ring() && last != cur && have_output) HV_HIP(hipStreamWaitEvent(L.stream, lane[last].tt_done, 0));
    HV_TRY(mark(0));
    // STFT bins of a sliding window: the batch enters the raw ring within the
    // bins launch (the few-bin transform writes it as it reads the window)
    const bool insert = raw_ring() && s.time_transformation == HV_TIME_STFT;
    if (raw_keep() && !insert) HV_TRY(raw_store(frames, slot0, L.stream));
    if (path == PATH_BINS) {
        int lo, K;
        bins_range(s, lo, K);
        // B = T: the batch is the window (frame j in slot (slot0 + j) mod T);
        // B < T: the raw ring, in slot order
        const void* win = raw_ring() ? raw : frames;
        const int w0 = raw_ring() ? 0 : slot0;
        const float2* wts = nullptr;
        if (s.time_transformation == HV_TIME_PCA) {
            // covariance of the raw window, FP64 eigensolver, then components lo .. lo + K - 1 projected
            HV_TRY(launch_pca_cov_raw(L.cov, win, (int)s.depth, T, P, w0, s.pca_precision, L.pca_ws, L.stream));
            // the raw window's covariance is real symmetric: real rotations (T <= 128)
            HV_TRY(launch_eig_jacobi(L.evals, L.evecs, L.cov, 1, T, L.eig_ws, L.stream, L.evecs32, T <= 128));
            wts = L.evecs32 + lo;
        }
        // the event after the time DFT: the stage timer's, or the one the
        // next lane's raw-ring write waits for
        RawInsert ins;
        if (insert) {
            ins.ring = raw;
            ins.batch = frames;
            ins.n = B;
            ins.slot0 = slot0;
        }
        HV_TRY(launch_bins_propagate(win, (int)s.depth, (int)s.width, (int)s.height, T, w0, lo, K,
                                     s.space_transformation, lens, L.planes, L.stream,
                                     timing ? ev[1] : shared_ring() ? L.tt_done : nullptr,
                                     minmax_fused() ? L.mm : nullptr, kPipeMm, wts, L.tiles, lens_t, ins));
        HV_TRY(mark(2));
    } else {
        HV_TRY(launch_propagate_to_ring(frames, (int)s.depth, B, (int)s.width, (int)s.height,
                                        s.space_transformation, lens, L.ring, T, slot0, L.stream,
                                        timing ? &ev[1] : nullptr, path == PATH_ROW_STFT,
                                        minmax_fused() ? L.mm : nullptr, kPipeMm));
    }
    HV_TRY(mark(3));
    HV_TRY(image_work());
    return mark(5);
}

extern "C" {

void hv_default_settings(hv_compute_settings* s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    // Backend/includes/core/holovibes.hh:359-447
    s->width = 512;
    s->height = 512;
    s->depth = HV_DEPTH_16;
    s->batch_size = 1;
    s->time_stride = 1;
    s->time_transformation_size = 1;
    s->space_transformation = HV_SPACE_NONE;
    s->time_transformation = HV_TIME_NONE;
    s->img_type = HV_IMG_MODULUS;
    s->lambda = 852e-9f;
    s->z_distance = 0.0f;
    s->pixel_size = 12.0f;
    s->p_start = 0;
    s->p_width = 0;
    s->unwrap_2d = 0;
    s->fft_shift = 0;
    s->contrast_enabled = 1;
    s->contrast_invert = 0;
    s->contrast_mode = HV_CONTRAST_REFERENCE;
    s->contrast_min = 1.0f;
    s->contrast_max = 65535.0f;
    s->contrast_lower = 0.02f;
    s->contrast_upper = 99.98f;
    s->contrast_subsample = 1;
    s->camera_fps = 0.0f;
    s->lens_precision = HV_LENS_FP64;
    s->pca_precision = HV_PCA_BF16_SPLIT;
    s->renorm_enabled = 1;  // RenormEnabled{true} (holovibes.hh:399)
    s->renorm_constant = 5;
    s->img_accu_level = 1;
    s->log_enabled = 0;
    s->registration_enabled = 0;
    s->registration_zone = 0.7f;  // Registration (view_struct.hh)
    s->contrast_reticle_enabled = 0;   // ContrastReticleDisplayEnabled{false} (holovibes.hh:403)
    s->contrast_reticle_scale = 0.5f;  // ContrastReticleScale{0.5f} (holovibes.hh:400)
    s->output_buffer_size = 1024;      // OutputBufferSize{1024} (holovibes.hh:375)
}

hv_status hv_validate_settings(const hv_compute_settings* s) {
    if (!s) return fail(HV_ERR_INVALID_ARG, "settings is null");
    if (s->width == 0 || s->height == 0) return fail(HV_ERR_INVALID_ARG, "frame size must be positive");
    if (s->space_transformation != HV_SPACE_NONE || s->time_transformation == HV_TIME_STFT ||
        s->time_transformation == HV_TIME_PCA || (s->img_type == HV_IMG_ARGUMENT && s->unwrap_2d)) {
        if (!fft_len_ok(s->width) || !fft_len_ok(s->height))
            return fail(HV_ERR_UNSUPPORTED, "frame sides must be powers of two or even 2^a 3^b 5^c in [16, 4096]");
    }
    if (((size_t)s->width * s->height) % 256) return fail(HV_ERR_UNSUPPORTED, "frame must hold a multiple of 256 pixels");
    if (s->depth != HV_DEPTH_8 && s->depth != HV_DEPTH_16 && s->depth != HV_DEPTH_32F)
        return fail(HV_ERR_INVALID_ARG, "depth must be 1, 2 or 4");
    // transform_api.cc:36-79: batch <= stride, stride a multiple of batch
    if (s->batch_size == 0) return fail(HV_ERR_INVALID_ARG, "batch_size must be >= 1");
    if (s->time_stride < s->batch_size || s->time_stride % s->batch_size)
        return fail(HV_ERR_INVALID_ARG, "time_stride must be a multiple of batch_size");
    // transform_api.cc:113-116
    if (s->time_transformation_size == 0) return fail(HV_ERR_INVALID_ARG, "time_transformation_size must be >= 1");
    if (!time_len_ok(s->time_transformation_size))
        return fail(HV_ERR_UNSUPPORTED, "time_transformation_size must be <= 256");
    // PCA windows that are not powers of two: one MFMA frame group (T < 32)
    if (s->time_transformation == HV_TIME_PCA && !is_pow2(s->time_transformation_size) &&
        s->time_transformation_size > 32)
        return fail(HV_ERR_UNSUPPORTED, "PCA windows that are not powers of two must be shorter than 32 frames");
    if (s->space_transformation < HV_SPACE_NONE || s->space_transformation > HV_SPACE_ANGULARSP)
        return fail(HV_ERR_INVALID_ARG, "unknown space transformation");
    if (s->time_transformation < HV_TIME_NONE || s->time_transformation > HV_TIME_PCA)
        return fail(HV_ERR_UNSUPPORTED, "time transformation not supported (SSA_STFT is not on this path)");
    if (s->space_transformation != HV_SPACE_NONE && !(s->lambda > 0.f && s->pixel_size > 0.f))
        return fail(HV_ERR_INVALID_ARG, "lambda and pixel_size must be positive");
    if (s->space_transformation == HV_SPACE_FRESNELTR && s->z_distance == 0.f)
        return fail(HV_ERR_INVALID_ARG, "Fresnel transform needs z_distance != 0");
    const uint32_t T = s->time_transformation_size;
    if (is_image_kind(s->img_type)) {
        // pmax = start + width must stay inside the T frames the kernels read
        if (s->p_start >= T || (s->p_width && s->p_start + s->p_width >= T))
            return fail(HV_ERR_INVALID_ARG, "P range must satisfy p_start + p_width < T");
    } else if (is_moment_kind(s->img_type)) {
        if (s->p_start >= T) return fail(HV_ERR_INVALID_ARG, "p_start must be < T");
        if (s->time_transformation != HV_TIME_STFT && s->time_transformation != HV_TIME_PCA)
            return fail(HV_ERR_UNSUPPORTED, "moments are computed from the STFT or PCA output");
    } else {
        return fail(HV_ERR_UNSUPPORTED, "image type not on this path");
    }
    if (s->time_transformation == HV_TIME_PCA && (s->pca_precision < HV_PCA_FP32 || s->pca_precision > HV_PCA_FP16))
        return fail(HV_ERR_INVALID_ARG, "unknown PCA precision");
    if (s->contrast_mode < HV_CONTRAST_REFERENCE || s->contrast_mode > HV_CONTRAST_MANUAL)
        return fail(HV_ERR_INVALID_ARG, "unknown contrast mode");
    if (s->contrast_subsample > 1 && (s->width % s->contrast_subsample || s->height % s->contrast_subsample))
        return fail(HV_ERR_INVALID_ARG, "contrast_subsample must divide the frame sides");
    if (s->fft_shift && (s->width % 2 || s->height % 2)) return fail(HV_ERR_UNSUPPORTED, "fft shift needs even sides");
    if (s->contrast_reticle_enabled) {
        // ContrastApi::set_contrast_reticle_scale (contrast_api.cc:278-297)
        if (!(s->contrast_reticle_scale >= 0.f && s->contrast_reticle_scale <= 1.f))
            return fail(HV_ERR_INVALID_ARG, "Reticle scale must be in range [0., 1.]");
        if (s->contrast_subsample > 1)
            return fail(HV_ERR_INVALID_ARG, "the contrast reticle and contrast_subsample exclude each other");
    }
    if (s->renorm_enabled && s->renorm_constant > 31) return fail(HV_ERR_INVALID_ARG, "renorm_constant must be <= 31");
    if (s->img_accu_level > 4096) return fail(HV_ERR_INVALID_ARG, "img_accu_level must be <= 4096");
    if (s->output_buffer_size > 65536) return fail(HV_ERR_INVALID_ARG, "output_buffer_size must be <= 65536");
    if (s->registration_enabled && !(s->registration_zone > 0.f && s->registration_zone <= 1.f))
        return fail(HV_ERR_INVALID_ARG, "registration_zone must be in (0, 1]");
    if (s->cuts_enabled) {
        if (s->time_transformation != HV_TIME_STFT) return fail(HV_ERR_UNSUPPORTED, "cuts are taken of the STFT");
        // the reference reads past the frame when a view runs over its edge
        if ((uint64_t)s->x_start + s->x_width >= s->width || (uint64_t)s->y_start + s->y_width >= s->height)
            return fail(HV_ERR_INVALID_ARG, "cuts: x_start + x_width < width and y_start + y_width < height");
    }
    return HV_OK;
}

hv_status hv_pipe_memory_bytes(const hv_compute_settings* s, size_t* bytes) {
    if (!bytes) return fail(HV_ERR_INVALID_ARG, "bytes is null");
    HV_TRY(hv_validate_settings(s));
    *bytes = layout_of(*s).total;
    return HV_OK;
}

hv_status hv_pipe_create(const hv_compute_settings* s, void* stream, hv_pipe** out) {
    if (!out) return fail(HV_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    HV_TRY(hv_validate_settings(s));
    std::unique_ptr<hv_pipe> p(new hv_pipe);
    p->s = *s;
    p->P = (size_t)s->width * s->height;
    p->T = (int)s->time_transformation_size;
    p->B = (int)s->batch_size;
    p->lay = layout_of(*s);
    p->path = choose_path(*s, PATH_BINS);
    Lane& L0 = p->lane[0];
    if (stream) {
        L0.stream = as_stream(stream);
    } else {
        HV_HIP(hipStreamCreateWithFlags(&L0.stream, hipStreamNonBlocking));
        p->own_stream = true;
    }
    void* mem = nullptr;
    hipError_t e = hipMalloc(&mem, p->lay.total);
    if (e != hipSuccess) {
        if (p->own_stream) (void)hipStreamDestroy(L0.stream);
        return hip_fail(e, "hipMalloc(pipe)");
    }
    p->base = (char*)mem;
    char* c = p->base;
    auto take = [&](size_t n) {
        char* r = n ? c : nullptr;
        c += n;
        return r;
    };
    const Layout& l = p->lay;
    p->ring = (float2*)take(l.ring);
    p->raw = take(l.raw);
    p->lens = (float2*)take(l.lens);
    p->lens_t = (float2*)take(l.lens_t);
    p->staging = take(l.staging);
    L0.planes = p->ring;
    if (l.lens_t) {
        int lo, K;
        bins_range(*s, lo, K);
        L0.tiles = p->ring + (size_t)K * p->P;
    }
    L0.image = (float*)take(l.image);
    L0.cimage = (float*)take(l.cimage);
    L0.tmp = (float*)take(l.tmp);
    L0.out = (uint16_t*)take(l.out);
    L0.range = (float*)take(l.small);
    L0.oq_pos = (uint32_t*)((char*)L0.range + 64);
    L0.mm = (uint32_t*)take(l.mm);
    L0.pct = take(l.pct);
    L0.unwrap_ws = take(l.unwrap);
    p->cov = (double2*)take(l.cov);
    p->evals = (double*)take(l.evals);
    p->evecs = (double2*)take(l.evecs);
    p->evecs32 = (float2*)take(l.evecs32);
    p->pca_ws = take(l.pca_ws);
    p->eig_ws = take(l.eig_ws);
    if (l.accu) {
        p->accu_ring = (float*)take(l.accu - 256);
        p->accu_state = (uint32_t*)take(256);
    }
    p->renorm_ws = take(l.renorm);
    if (l.reg) {
        p->reg = take(l.reg - align256(p->P * sizeof(float)));
        p->reg_tmp = (float*)take(align256(p->P * sizeof(float)));
    }
    if (l.cuts) {
        const size_t cut = (l.cuts - align256((size_t)p->T * p->P * sizeof(float2))) / 2;
        p->cuts_stft = (float2*)take(l.cuts - 2 * cut);
        p->cut_xz = (float*)take(cut);
        p->cut_yz = (float*)take(cut);
    }
    p->outq = (uint16_t*)take(l.outq);
    p->oq_cap = s->output_buffer_size;
    L0.ring = p->ring;
    L0.cov = p->cov;
    L0.evals = p->evals;
    L0.evecs = p->evecs;
    L0.evecs32 = p->evecs32;
    L0.pca_ws = p->pca_ws;
    L0.eig_ws = p->eig_ws;
    // refresh-time work: lens, manual contrast range, zeroed ring
    hv_status st = hv_prepare_lens(p->lens, s->space_transformation, s->width, s->height, s->lambda, s->z_distance,
                                   s->pixel_size, s->lens_precision, L0.stream);
    if (st == HV_OK && p->lens_t) st = launch_tile_frame(p->lens_t, p->lens, (int)s->width, (int)s->height, L0.stream);
    if (st == HV_OK) {
        const float r[2] = {s->contrast_min, s->contrast_max};
        e = hipMemcpyAsync(L0.range, r, sizeof(r), hipMemcpyHostToDevice, L0.stream);
        if (e == hipSuccess) e = hipMemsetAsync(p->ring, 0, l.ring, L0.stream);
        if (e == hipSuccess && p->accu_state) e = hipMemsetAsync(p->accu_state, 0, 256, L0.stream);
        if (e == hipSuccess && p->evecs32) e = hipMemsetAsync(p->evecs32, 0, l.evecs32, L0.stream);
        if (e == hipSuccess && p->reg)
            e = hipMemsetAsync(p->reg, 0, l.reg - align256(p->P * sizeof(float)), L0.stream);
        // the disc is quadrant-swapped when exactly one of Fresnel and the fft shift holds (registration.cc:100-110)
        if (e == hipSuccess && p->reg &&
            launch_registration_mask(p->reg, (int)s->width, (int)s->height, s->registration_zone,
                                     (s->space_transformation == HV_SPACE_FRESNELTR) != (s->fft_shift != 0),
                                     L0.stream) != HV_OK)
            e = hipErrorLaunchFailure;
        if (e == hipSuccess) e = hipStreamSynchronize(L0.stream);
        if (e != hipSuccess) st = hip_fail(e, "pipe init");
    }
    if (st == HV_OK && p->oq_cap) {
        p->oq_ev.assign(p->oq_cap, nullptr);
        p->oq_ev_rec.assign(p->oq_cap, 0);
        for (auto& ev : p->oq_ev) {
            e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (e != hipSuccess) {
                st = hip_fail(e, "hipEventCreate(output queue)");
                break;
            }
        }
        if (st == HV_OK) st = p->reset_output_slots();
    }
    if (st != HV_OK) {
        hv_pipe_destroy(p.release());
        return st;
    }
    *out = p.release();
    return HV_OK;
}

hv_status hv_pipe_input_buffer(hv_pipe* p, void** frames) {
    if (!p || !frames) return fail(HV_ERR_INVALID_ARG, "null argument");
    *frames = p->staging;
    return HV_OK;
}

hv_status hv_pipe_push_batch(hv_pipe* p, const void* frames, int* produced) {
    if (!p || !frames) return fail(HV_ERR_INVALID_ARG, "null argument");
    if (produced) *produced = 0;
    // insert_wait_time_stride (pipe.cc:400-416): other batches are dropped
    p->batch_index += p->B;
    if (p->batch_index != p->s.time_stride) return HV_OK;
    p->batch_index = 0;
    // Queue::enqueue_multiple bookkeeping (queue.cc:332-376)
    const uint32_t maxq = (uint32_t)p->T;
    const int slot0 = (int)((p->q_start + p->q_size) % maxq);
    uint32_t nb = (uint32_t)p->B;
    if (nb > maxq) {
        p->q_start = (p->q_start + nb - maxq) % maxq;
        nb = maxq;
    }
    p->q_size += nb;
    if (p->q_size > maxq) {
        p->q_start = (p->q_start + p->q_size - maxq) % maxq;
        p->q_size = maxq;
    }
    // insert_wait_time_transformation_size (pipe.cc:418-427)
    const bool full = p->q_size >= maxq;
    if (!full) {
        if (p->raw_keep()) HV_TRY(p->raw_store(frames, slot0, p->lane[0].stream));
        if (p->raw_ring()) return HV_OK;
        return launch_propagate_to_ring(frames, (int)p->s.depth, p->B, (int)p->s.width, (int)p->s.height,
                                        p->s.space_transformation, p->lens, p->ring, p->T, slot0,
                                        p->lane[0].stream);
    }
    const bool pending = p->s.contrast_enabled && p->s.contrast_mode == HV_CONTRAST_REFERENCE &&
                         p->autocontrast_pending;
    // the cycle's lane: round robin over the in-flight lanes (lane 0 while timing)
    p->cur = p->timing ? 0 : p->next;
    p->next = (p->cur + 1) % p->nlanes;
    Lane& L = p->lane[p->cur];
    if (p->timing) {
        HV_TRY(p->enqueue_cycle(frames, slot0));
        HV_TRY(p->collect_times());
    } else if (p->graph_enabled && !pending && !p->shared_ring() && p->registration_steady()) {
        // one graph per (ring phase, input buffer); past the cache bound the
        // batch is staged into the pipe's own input buffer first (one lane)
        // or run without a graph (several lanes share no staging buffer)
        const void* src = frames;
        auto key = std::make_pair(slot0, src);
        auto it = L.graphs.find(key);
        bool eager = false;
        if (it == L.graphs.end() && L.graphs.size() >= kMaxGraphs) {
            if (p->nlanes > 1) {
                eager = true;
            } else {
                HV_HIP(hipMemcpyAsync(p->staging, frames, (size_t)p->B * p->P * p->s.depth,
                                      hipMemcpyDeviceToDevice, L.stream));
                src = p->staging;
                key = std::make_pair(slot0, src);
                it = L.graphs.find(key);
            }
        }
        if (eager) {
            HV_TRY(p->enqueue_cycle(frames, slot0));
        } else {
            if (it == L.graphs.end()) {
                hipGraph_t g = nullptr;
                // relaxed: the cycle enqueues only launches, async copies and
                // event records, and blocking HIP calls of other host threads
                // (a record writer, a read worker) must not void the capture
                HV_HIP(hipStreamBeginCapture(L.stream, hipStreamCaptureModeRelaxed));
                hv_status st = p->enqueue_cycle(src, slot0);
                hipError_t e = hipStreamEndCapture(L.stream, &g);
                if (st != HV_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return st;
                }
                if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
                hipGraphExec_t ex = nullptr;
                e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
                e = hipGraphUpload(ex, L.stream);  // device-side resources before the first launch
                if (e != hipSuccess) {
                    (void)hipGraphExecDestroy(ex);
                    return hip_fail(e, "hipGraphUpload");
                }
                it = L.graphs.emplace(key, ex).first;
            }
            HV_HIP(hipGraphLaunch(it->second, L.stream));
        }
    } else {
        HV_TRY(p->enqueue_cycle(frames, slot0));
    }
    HV_TRY(p->output_enqueued());
    if (pending) {
        p->autocontrast_pending = false;  // rendering.cc:251-258
        // the frozen range reaches the other lanes in their stream order
        if (p->nlanes > 1) {
            hipEvent_t e = nullptr;
            HV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            HV_HIP(hipEventRecord(e, L.stream));
            for (int i = 0; i < p->nlanes; ++i) {
                if (i == p->cur) continue;
                HV_HIP(hipStreamWaitEvent(p->lane[i].stream, e, 0));
                HV_HIP(hipMemcpyAsync(p->lane[i].range, L.range, 2 * sizeof(float), hipMemcpyDeviceToDevice,
                                      p->lane[i].stream));
            }
            HV_HIP(hipEventDestroy(e));
        }
    }
    p->last = p->cur;
    p->have_output = true;
    if (produced) *produced = 1;
    return HV_OK;
}

hv_status hv_pipe_push_batches(hv_pipe* p, const void* const* frames, int n, int* produced) {
    if (produced) *produced = 0;
    if (!p || (!frames && n > 0) || n < 0) return fail(HV_ERR_INVALID_ARG, "null argument or negative count");
    for (int i = 0; i < n; ++i) {
        int one = 0;
        HV_TRY(hv_pipe_push_batch(p, frames[i], &one));
        if (produced) *produced += one;
    }
    return HV_OK;
}

hv_status hv_pipe_set_timing(hv_pipe* p, int enabled) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    if (enabled && !p->ev[0])
        for (auto& e : p->ev) HV_HIP(hipEventCreate(&e));
    if (p->timing != (enabled != 0)) {
        p->timing = enabled != 0;
        p->next = 0;
        HV_TRY(p->reset_output_slots());  // timed cycles all run on lane 0
    }
    for (auto& t : p->stage_ms) t = 0.0;
    p->timed_cycles = 0;
    return HV_OK;
}

hv_status hv_pipe_stage_times(hv_pipe* p, float* ms, int* cycles) {
    if (!p || !ms) return fail(HV_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < kStages; ++i) ms[i] = p->timed_cycles ? (float)(p->stage_ms[i] / p->timed_cycles) : 0.f;
    if (cycles) *cycles = p->timed_cycles;
    return HV_OK;
}

hv_status hv_pipe_output(hv_pipe* p, const uint16_t** out_u16, const float** out_f32_contrast,
                         const float** out_f32_image) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    const Lane& L = p->lane[p->last];
    if (out_u16)
        *out_u16 = p->oq_cap && p->out_count ? p->outq + (size_t)((p->out_count - 1) % p->oq_cap) * p->P : L.out;
    if (out_f32_contrast) *out_f32_contrast = p->s.contrast_enabled ? L.cimage : L.image;
    if (out_f32_image) *out_f32_image = L.image;
    return p->have_output ? HV_OK : HV_ERR_NOT_READY;
}

hv_status hv_pipe_output_queue(hv_pipe* p, const uint16_t** base, uint32_t* max_size, uint32_t* start_index,
                               uint32_t* size, int* has_overwritten) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    if (base) *base = p->outq;
    if (max_size) *max_size = p->oq_cap;
    if (start_index) *start_index = p->oq_start;
    if (size) *size = p->oq_size;
    if (has_overwritten) *has_overwritten = p->oq_overwritten ? 1 : 0;
    return HV_OK;
}

hv_status hv_pipe_output_dequeue(hv_pipe* p, void* dst, int memcpy_kind, uint32_t nb, void* stream) {
    if (!p || !dst) return fail(HV_ERR_INVALID_ARG, "null argument");
    if (memcpy_kind != (int)hipMemcpyDeviceToHost && memcpy_kind != (int)hipMemcpyDeviceToDevice)
        return fail(HV_ERR_INVALID_ARG, "output dequeue copies to the host (2) or the device (3)");
    if (!p->oq_cap) return fail(HV_ERR_INVALID_ARG, "the pipe has no output queue (output_buffer_size 0)");
    if (nb == 0 || nb > p->oq_size) return fail(HV_ERR_NOT_READY, "fewer images queued than asked for");
    hipStream_t st = as_stream(stream);
    const size_t fb = p->P * sizeof(uint16_t);
    // the newest of the nb images completes last on its lane; wait for each
    // lane's latest of them (images of one lane complete in order)
    for (uint32_t k = nb > (uint32_t)p->nlanes ? nb - (uint32_t)p->nlanes : 0; k < nb; ++k) {
        const uint32_t slot = (p->oq_start + k) % p->oq_cap;
        if (p->oq_ev_rec[slot]) HV_HIP(hipStreamWaitEvent(st, p->oq_ev[slot], 0));
    }
    const uint32_t first = std::min(nb, p->oq_cap - p->oq_start);
    HV_HIP(hipMemcpyAsync(dst, p->outq + (size_t)p->oq_start * p->P, first * fb, (hipMemcpyKind)memcpy_kind, st));
    if (nb > first)
        HV_HIP(hipMemcpyAsync((char*)dst + first * fb, p->outq, (nb - first) * fb, (hipMemcpyKind)memcpy_kind, st));
    HV_HIP(hipStreamSynchronize(st));
    p->oq_start = (p->oq_start + nb) % p->oq_cap;
    p->oq_size -= nb;
    return HV_OK;
}

hv_status hv_pipe_output_frame(hv_pipe* p, uint64_t index, void* stream, const uint16_t** frame) {
    if (!p || !frame) return fail(HV_ERR_INVALID_ARG, "null argument");
    if (!p->oq_cap) return fail(HV_ERR_INVALID_ARG, "the pipe has no output queue (output_buffer_size 0)");
    if (index >= p->out_count || p->out_count - index > p->oq_cap) {
        set_error("image " + std::to_string(index) + " is not in the output queue");
        return HV_ERR_NOT_READY;
    }
    const uint32_t slot = (uint32_t)(index % p->oq_cap);
    if (stream && p->oq_ev_rec[slot]) HV_HIP(hipStreamWaitEvent(as_stream(stream), p->oq_ev[slot], 0));
    *frame = p->outq + (size_t)slot * p->P;
    return HV_OK;
}

hv_status hv_pipe_output_count(hv_pipe* p, uint64_t* count) {
    if (!p || !count) return fail(HV_ERR_INVALID_ARG, "null argument");
    *count = p->out_count;
    return HV_OK;
}

hv_status hv_pipe_registration_shift(hv_pipe* p, int* shift_x, int* shift_y) {
    if (!p || !shift_x || !shift_y) return fail(HV_ERR_INVALID_ARG, "null argument");
    if (!p->reg) return fail(HV_ERR_INVALID_ARG, "registration not enabled");
    HV_TRY(hv_pipe_sync(p));
    int sh[2] = {0, 0};
    if (p->reg_ref_set) HV_TRY(registration_shift(p->reg, p->P, sh));
    *shift_x = sh[0];
    *shift_y = sh[1];
    return HV_OK;
}

hv_status hv_pipe_cuts(hv_pipe* p, const float** xz, const float** yz) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    if (!p->cut_xz) return fail(HV_ERR_INVALID_ARG, "cuts not enabled");
    if (xz) *xz = p->cut_xz;
    if (yz) *yz = p->cut_yz;
    return p->have_output ? HV_OK : HV_ERR_NOT_READY;
}

hv_status hv_pipe_set_record_moments(hv_pipe* p, int enabled) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    HV_TRY(hv_pipe_sync(p));
    if (!enabled) {
        if (!p->rec) return HV_OK;
        // graphs captured with the record write into it: drop them before
        // the memory goes, and return to the path the record displaced
        p->drop_graphs();
        p->next = 0;
        (void)hipFree(p->rec);
        p->rec = nullptr;
        p->path = choose_path(p->s, p->level);
        return p->reset_output_slots();
    }
    if (p->s.time_transformation != HV_TIME_STFT && p->s.time_transformation != HV_TIME_PCA)
        return fail(HV_ERR_UNSUPPORTED, "moments are recorded from the STFT or PCA output");
    if (p->nlanes > 1) return fail(HV_ERR_UNSUPPORTED, "the moments record runs one cycle at a time");
    if (!p->rec) HV_HIP(hipMalloc((void**)&p->rec, 3 * p->P * sizeof(float)));
    // the moments read the whole time window, which only the three-pass path
    // keeps; a sliding window's ring is rebuilt from its raw frames
    HV_TRY(p->complex_ring_from_raw(PATH_THREE_PASS));
    p->path = PATH_THREE_PASS;
    p->drop_graphs();
    p->next = 0;
    return p->reset_output_slots();
}

hv_status hv_pipe_record_frames(hv_pipe* p, const float** frames) {
    if (!p || !frames) return fail(HV_ERR_INVALID_ARG, "null argument");
    if (!p->rec) return fail(HV_ERR_INVALID_ARG, "moments record not enabled");
    *frames = p->rec;
    return p->have_output ? HV_OK : HV_ERR_NOT_READY;
}

hv_status hv_pipe_contrast_range(hv_pipe* p, const float** range_dev) {
    if (!p || !range_dev) return fail(HV_ERR_INVALID_ARG, "null argument");
    *range_dev = p->lane[p->last].range;
    return HV_OK;
}

hv_status hv_pipe_ring(hv_pipe* p, const void** ring, uint32_t* start, uint32_t* size) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    if (ring) *ring = p->ring;
    if (start) *start = p->q_start;
    if (size) *size = p->q_size;
    return HV_OK;
}

hv_status hv_pipe_set_fused(hv_pipe* p, int level) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    if (level < 0) return fail(HV_ERR_INVALID_ARG, "fusion level must be >= 0");
    HV_TRY(hv_pipe_sync(p));
    const bool was_raw = p->raw_ring();
    const int path = p->rec ? PATH_THREE_PASS : choose_path(p->s, level);
    // leaving the sliding-window bins-first path: its raw window gives the
    // ring; entering it: the raw ring did not follow the window, which refills
    if (path != PATH_BINS) HV_TRY(p->complex_ring_from_raw(path));
    p->level = level;
    p->path = path;
    if (!was_raw && p->raw_ring()) p->q_start = p->q_size = 0;
    p->drop_graphs();
    p->nlanes = 1;  // extra lanes hold one path's buffers: hv_pipe_set_lanes again after a path change
    p->next = 0;
    return p->reset_output_slots();
}

hv_status hv_pipe_set_lanes(hv_pipe* p, int lanes) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    if (lanes < 1 || lanes > kMaxLanes) return fail(HV_ERR_INVALID_ARG, "lanes must be in [1, 4]");
    if (lanes > 1 && !p->own_stream)
        return fail(HV_ERR_UNSUPPORTED, "several lanes need the pipe's own streams (create it without a stream)");
    if (lanes > 1 && (p->s.renorm_enabled || p->s.img_accu_level > 1))
        return fail(HV_ERR_UNSUPPORTED, "renormalisation and image accumulation run one cycle at a time");
    if (lanes > 1 && p->rec) return fail(HV_ERR_UNSUPPORTED, "the moments record runs one cycle at a time");
    if (lanes > 1 && p->s.cuts_enabled) return fail(HV_ERR_UNSUPPORTED, "the cuts run one cycle at a time");
    if (lanes > 1 && p->s.registration_enabled)
        return fail(HV_ERR_UNSUPPORTED, "registration runs one cycle at a time");
    if (p->oq_cap && p->oq_cap < (uint32_t)lanes)
        return fail(HV_ERR_INVALID_ARG, "the output queue must hold at least one image per lane");
    HV_TRY(hv_pipe_sync(p));
    p->drop_graphs();
    int lo, Ki;
    bins_range(p->s, lo, Ki);
    const size_t K = (size_t)Ki;
    const Layout& l = p->lay;
    const int kind = p->path == PATH_BINS ? LANE_BINS : window_replaced(p->s) ? LANE_OWN_RING : LANE_SHARED_RING;
    const bool bins = kind == LANE_BINS, own_ring = kind == LANE_OWN_RING;
    // extra lanes allocated for another kind (all of them: one call allocates
    // one kind) are released first
    while (p->lane.size() > 1 && p->lane.back().kind != kind) {
        Lane& x = p->lane.back();
        if (x.mem) (void)hipFree(x.mem);
        if (x.stream) (void)hipStreamDestroy(x.stream);
        if (x.tt_done) (void)hipEventDestroy(x.tt_done);
        p->lane.pop_back();
    }
    if (!p->lane[0].tt_done)
        HV_HIP(hipEventCreateWithFlags(&p->lane[0].tt_done, hipEventDisableTiming));
    for (int i = (int)p->lane.size(); i < lanes; ++i) {
        Lane n;
        // column-tiled bins: the tiles after the planes
        const size_t planes = bins ? align256((p->lens_t ? 2 : 1) * K * p->P * sizeof(float2)) : 0;
        const size_t ring = own_ring ? l.ring : 0;
        // own PCA buffers: window-replaced three-pass lanes and bins-first PCA lanes
        const bool own_pca = own_ring || (bins && p->s.time_transformation == HV_TIME_PCA);
        const size_t pca = own_pca ? l.cov + l.evals + l.evecs + l.evecs32 + l.pca_ws + l.eig_ws : 0;
        const size_t total = planes + ring + pca + l.image + l.cimage + l.tmp + l.out + l.small + l.mm + l.pct +
                             l.unwrap;
        hipError_t e = hipMalloc(&n.mem, total);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(lane)");
        e = hipStreamCreateWithFlags(&n.stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            (void)hipFree(n.mem);
            return hip_fail(e, "hipStreamCreateWithFlags(lane)");
        }
        e = hipEventCreateWithFlags(&n.tt_done, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipFree(n.mem);
            (void)hipStreamDestroy(n.stream);
            return hip_fail(e, "hipEventCreateWithFlags(lane)");
        }
        char* c = (char*)n.mem;
        auto take = [&](size_t b) {
            char* r = b ? c : nullptr;
            c += b;
            return r;
        };
        n.kind = kind;
        n.planes = (float2*)take(planes);
        if (bins && p->lens_t) n.tiles = n.planes + K * p->P;
        if (own_ring) n.ring = (float2*)take(ring);
        if (own_pca) {
            n.cov = (double2*)take(l.cov);
            n.evals = (double*)take(l.evals);
            n.evecs = (double2*)take(l.evecs);
            n.evecs32 = (float2*)take(l.evecs32);
            n.pca_ws = take(l.pca_ws);
            n.eig_ws = take(l.eig_ws);
        } else if (!bins) {
            const Lane& z = p->lane[0];
            n.ring = z.ring;
            n.cov = z.cov;
            n.evals = z.evals;
            n.evecs = z.evecs;
            n.evecs32 = z.evecs32;
            n.pca_ws = z.pca_ws;
            n.eig_ws = z.eig_ws;
        }
        n.image = (float*)take(l.image);
        n.cimage = (float*)take(l.cimage);
        n.tmp = (float*)take(l.tmp);
        n.out = (uint16_t*)take(l.out);
        n.range = (float*)take(l.small);
        n.oq_pos = (uint32_t*)((char*)n.range + 64);
        n.mm = (uint32_t*)take(l.mm);
        n.pct = take(l.pct);
        n.unwrap_ws = take(l.unwrap);
        p->lane.push_back(n);
    }
    // every lane starts from lane 0's contrast range (manual or already frozen)
    for (int i = 1; i < lanes; ++i)
        HV_HIP(hipMemcpyAsync(p->lane[i].range, p->lane[0].range, 2 * sizeof(float), hipMemcpyDeviceToDevice,
                              p->lane[0].stream));
    HV_HIP(hipStreamSynchronize(p->lane[0].stream));
    p->nlanes = lanes;
    p->next = 0;
    return p->reset_output_slots();
}

hv_status hv_pipe_time_stage(hv_pipe* p, int stage, const void* const* frames, int n_frames, int iters, float* ms) {
    if (!p || !ms || iters < 1) return fail(HV_ERR_INVALID_ARG, "null argument or iters < 1");
    if (p->path != PATH_BINS) return fail(HV_ERR_UNSUPPORTED, "stage timing covers the bins-first path");
    // a sliding window (B < T) is read from the raw ring, a replaced one from the input batches
    const bool raw = p->raw_ring();
    if (stage == 0 && !raw && (!frames || n_frames < 1)) return fail(HV_ERR_INVALID_ARG, "stage 0 needs input frames");
    const bool pca = p->s.time_transformation == HV_TIME_PCA;
    if (stage < 0 || stage > 4 || (stage == 2 && !pca))
        return fail(HV_ERR_INVALID_ARG, "stage must be 0, 1, 3 or 4 (2: PCA covariance + eigensolver)");
    if (stage == 2 && !raw && (!frames || n_frames < 1)) return fail(HV_ERR_INVALID_ARG, "stage 2 needs input frames");
    HV_TRY(hv_pipe_sync(p));
    Lane& L = p->lane[0];
    const hv_compute_settings& s = p->s;
    int pmin, K;
    bins_range(s, pmin, K);
    const int pmax = pmin + K - 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HV_HIP(hipEventCreate(&e0));
    HV_HIP(hipEventCreate(&e1));
    hv_status st = HV_OK;
    hipError_t e = hipEventRecord(e0, L.stream);
    for (int i = 0; i < iters && st == HV_OK && e == hipSuccess; ++i) {
        switch (stage) {
            case 0:
                st = launch_bins_rows(raw ? p->raw : frames[i % n_frames], (int)s.depth, (int)s.width, (int)s.height, p->T, 0, pmin,
                                      K, s.space_transformation, p->lens, L.planes, L.stream, L.mm, kPipeMm,
                                      pca ? L.evecs32 + pmin : nullptr, L.tiles);
                break;
            case 2:
                st = launch_pca_cov_raw(L.cov, raw ? p->raw : frames[i % n_frames], (int)s.depth, p->T, p->P, 0, s.pca_precision,
                                        L.pca_ws, L.stream);
                if (st == HV_OK)
                    st = launch_eig_jacobi(L.evals, L.evecs, L.cov, 1, p->T, L.eig_ws, L.stream, L.evecs32,
                                           p->T <= 128);
                break;
            case 1:
                st = launch_bins_columns(L.planes, (int)s.width, (int)s.height, K, s.space_transformation, p->lens,
                                         L.stream, L.tiles, p->lens_t);
                break;
            case 3:
                if (is_moment_kind(s.img_type))
                    st = launch_bins_moments(L.image, nullptr, nullptr, L.planes, (int)s.width, (int)s.height, p->T,
                                             pmin, K, s.space_transformation, s.camera_fps, 0, L.mm, L.stream,
                                             kPipeMm, L.tiles);
                else
                    st = launch_bins_image(L.image, L.planes, (int)s.width, (int)s.height, K, s.space_transformation,
                                           s.img_type, pmin, pmax, L.mm, L.stream, kPipeMm,
                                           pca ? (float)((2 << 16) - 1) : 1.0f, L.tiles);
                break;
            default:
                st = launch_contrast_to_ushort(L.out, L.cimage, L.image, p->P, L.mm, kPipeMm, s.contrast_invert, L.stream,
                                               L.range);
                break;
        }
    }
    if (e == hipSuccess) e = hipEventRecord(e1, L.stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != HV_OK) return st;
    if (e != hipSuccess) return hip_fail(e, "stage timing");
    *ms = t / iters;
    return HV_OK;
}

hv_status hv_pipe_measure_latency(hv_pipe* p, const void* const* frames, int n_frames, int n_batches,
                                  double period_ns, int64_t* latency_ns, int* n_out) {
    if (!p || !frames || n_frames < 1 || n_batches < 1 || !latency_ns || !n_out || !(period_ns >= 0.0))
        return fail(HV_ERR_INVALID_ARG, "measure_latency: bad argument");
    using clock = std::chrono::steady_clock;
    const auto ns = [](clock::time_point t) {
        return std::chrono::duration_cast<std::chrono::nanoseconds>(t.time_since_epoch()).count();
    };
    constexpr int kEvents = 64;
    hipEvent_t ev[kEvents] = {};
    for (auto& e : ev) HV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    struct Wait {
        int ev;
        int64_t t_arrival;
    };
    std::deque<Wait> waits;
    std::vector<int> free_ev;
    for (int i = kEvents - 1; i >= 0; --i) free_ev.push_back(i);
    int n = 0;
    hv_status st = HV_OK;
    // every outstanding image (lanes may finish out of order); `block` waits
    // for the oldest first
    auto poll = [&](bool block) -> hv_status {
        if (block && !waits.empty()) HV_HIP(hipEventSynchronize(ev[waits.front().ev]));
        for (auto it = waits.begin(); it != waits.end();) {
            const hipError_t q = hipEventQuery(ev[it->ev]);
            if (q == hipErrorNotReady) {
                ++it;
                continue;
            }
            if (q != hipSuccess) return hip_fail(q, "latency event");
            latency_ns[n++] = ns(clock::now()) - it->t_arrival;
            free_ev.push_back(it->ev);
            it = waits.erase(it);
        }
        return HV_OK;
    };
    const auto t0 = clock::now();
    for (int i = 0; i < n_batches && st == HV_OK; ++i) {
        const auto target = t0 + std::chrono::nanoseconds((int64_t)(period_ns * i));
        while (clock::now() < target && st == HV_OK) st = poll(false);
        if (st != HV_OK) break;
        if (free_ev.empty()) st = poll(true);  // more than kEvents images outstanding
        if (st != HV_OK) break;
        int produced = 0;
        st = hv_pipe_push_batch(p, frames[i % n_frames], &produced);
        if (st != HV_OK || !produced) continue;
        const int e = free_ev.back();
        free_ev.pop_back();
        const hipError_t r = hipEventRecord(ev[e], p->lane[p->last].stream);
        if (r != hipSuccess) {
            st = hip_fail(r, "hipEventRecord(latency)");
            break;
        }
        waits.push_back({e, ns(target)});
    }
    while (st == HV_OK && !waits.empty()) st = poll(true);
    for (auto& e : ev) (void)hipEventDestroy(e);
    *n_out = n;
    return st;
}

hv_status hv_pipe_path(const hv_pipe* p, int* path) {
    if (!p || !path) return fail(HV_ERR_INVALID_ARG, "null argument");
    *path = p->path;
    return HV_OK;
}

hv_status hv_pipe_set_graph(hv_pipe* p, int enabled) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    p->graph_enabled = enabled != 0;
    return HV_OK;
}

hv_status hv_pipe_stream(hv_pipe* p, void** stream) {
    if (!p || !stream) return fail(HV_ERR_INVALID_ARG, "null argument");
    *stream = p->lane[0].stream;
    return HV_OK;
}

hv_status hv_pipe_sync(hv_pipe* p) {
    if (!p) return fail(HV_ERR_INVALID_ARG, "pipe is null");
    for (int i = 0; i < (int)p->lane.size(); ++i)
        if (p->lane[i].stream) HV_HIP(hipStreamSynchronize(p->lane[i].stream));
    return HV_OK;
}

hv_status hv_pipe_destroy(hv_pipe* p) {
    if (!p) return HV_OK;
    for (auto& l : p->lane)
        if (l.stream) (void)hipStreamSynchronize(l.stream);
    p->drop_graphs();
    for (size_t i = 1; i < p->lane.size(); ++i) {
        if (p->lane[i].mem) (void)hipFree(p->lane[i].mem);
        if (p->lane[i].stream) (void)hipStreamDestroy(p->lane[i].stream);
    }
    for (auto& l : p->lane)
        if (l.tt_done) (void)hipEventDestroy(l.tt_done);
    for (auto e : p->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : p->oq_ev)
        if (e) (void)hipEventDestroy(e);
    if (p->base) (void)hipFree(p->base);
    if (p->rec) (void)hipFree(p->rec);
    if (p->own_stream && p->lane[0].stream) (void)hipStreamDestroy(p->lane[0].stream);
    delete p;
    return HV_OK;
}

}  // extern "C"

// ===========================================================================
// server: several sources, admission by HBM budget, request slots,
// continuous batching, latency accounting
// ===========================================================================
namespace {
int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
struct Request {
    int64_t id;
    const void* frames;
    int64_t t_arrival;
};
struct Inflight {
    int64_t id;
    int source;
    hipEvent_t done;
    int64_t t_arrival;
    int64_t output_index;  // -1: the cycle produced no image
};
struct Source {
    hv_pipe* pipe = nullptr;
    size_t bytes = 0;
    int slots = 2, depth = 16;   // requests on the GPU at once, requests waiting
    int inflight = 0;
    std::deque<Request> pending;
};
}  // namespace

struct hv_server {
    size_t budget = 0, used = 0;
    int next_id = 0;
    int64_t next_request = 0;
    int slots = 2, depth = 16;
    uint64_t shed = 0;
    std::map<int, Source> sources;
    std::deque<Inflight> inflight;
    std::deque<hv_request_result> results;  // finished or shed, not yet polled
    std::vector<hipEvent_t> free_events;

    // runs one request on its source's pipe (a slot is free)
    hv_status admit(int id, Source& src, const Request& r) {
        int produced = 0;
        HV_TRY(hv_pipe_push_batch(src.pipe, r.frames, &produced));
        hipEvent_t ev;
        if (!free_events.empty()) {
            ev = free_events.back();
            free_events.pop_back();
        } else {
            HV_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        }
        HV_HIP(hipEventRecord(ev, src.pipe->lane[src.pipe->last].stream));
        inflight.push_back({r.id, id, ev, r.t_arrival, produced ? (int64_t)src.pipe->out_count - 1 : -1});
        ++src.inflight;
        return HV_OK;
    }
    // moves finished requests to `results` and admits waiting ones into the
    // slots that freed, in arrival order per source
    hv_status progress() {
        for (auto it = inflight.begin(); it != inflight.end();) {
            const hipError_t q = hipEventQuery(it->done);
            if (q == hipErrorNotReady) {
                ++it;
                continue;
            }
            if (q != hipSuccess) return hip_fail(q, "hipEventQuery");
            const int64_t t = now_ns();
            if (it->source >= 0) {
                results.push_back({it->id, it->source, it->output_index >= 0 ? HV_REQ_DONE : HV_REQ_NO_OUTPUT,
                                   t - it->t_arrival, it->output_index});
                auto sit = sources.find(it->source);
                if (sit != sources.end()) --sit->second.inflight;
            }
            free_events.push_back(it->done);
            it = inflight.erase(it);
        }
        for (auto& kv : sources) {
            Source& src = kv.second;
            while (src.inflight < src.slots && !src.pending.empty()) {
                const Request r = src.pending.front();
                src.pending.pop_front();
                HV_TRY(admit(kv.first, src, r));
            }
        }
        return HV_OK;
    }
};

extern "C" {

hv_status hv_server_create(size_t budget_bytes, hv_server** out) {
    if (!out) return fail(HV_ERR_INVALID_ARG, "out is null");
    *out = new hv_server;
    (*out)->budget = budget_bytes;
    return HV_OK;
}

hv_status hv_server_configure(hv_server* s, int slots, int pending) {
    if (!s) return fail(HV_ERR_INVALID_ARG, "server is null");
    if (slots < 1 || pending < 0) return fail(HV_ERR_INVALID_ARG, "slots must be >= 1 and pending >= 0");
    s->slots = slots;
    s->depth = pending;
    return HV_OK;
}

hv_status hv_server_add_source(hv_server* s, const hv_compute_settings* settings, int* source_id) {
    if (!s || !source_id) return fail(HV_ERR_INVALID_ARG, "null argument");
    size_t need = 0;
    HV_TRY(hv_pipe_memory_bytes(settings, &need));
    if (s->used + need > s->budget) return fail(HV_ERR_BUDGET, "source does not fit in the HBM budget");
    hv_pipe* p = nullptr;
    HV_TRY(hv_pipe_create(settings, nullptr, &p));
    hv_pipe_set_graph(p, 1);
    const int id = s->next_id++;
    Source src;
    src.pipe = p;
    src.bytes = need;
    src.slots = s->slots;
    src.depth = s->depth;
    s->sources[id] = std::move(src);
    s->used += need;
    *source_id = id;
    return HV_OK;
}

hv_status hv_server_remove_source(hv_server* s, int source_id) {
    if (!s) return fail(HV_ERR_INVALID_ARG, "server is null");
    auto it = s->sources.find(source_id);
    if (it == s->sources.end()) return fail(HV_ERR_INVALID_ARG, "unknown source");
    hv_pipe_sync(it->second.pipe);
    // retire its in-flight entries; its waiting requests are shed
    for (auto& f : s->inflight)
        if (f.source == source_id) f.source = -1;
    for (const Request& r : it->second.pending) {
        s->results.push_back({r.id, source_id, HV_REQ_SHED, 0, -1});
        ++s->shed;
    }
    hv_pipe_destroy(it->second.pipe);
    s->used -= it->second.bytes;
    s->sources.erase(it);
    return HV_OK;
}

hv_status hv_server_submit_request(hv_server* s, int source_id, const void* frames, int64_t t_arrival_ns,
                                   int64_t* request_id) {
    if (!s || !frames) return fail(HV_ERR_INVALID_ARG, "null argument");
    auto it = s->sources.find(source_id);
    if (it == s->sources.end()) return fail(HV_ERR_INVALID_ARG, "unknown source");
    Source& src = it->second;
    const Request r{s->next_request++, frames, t_arrival_ns};
    if (request_id) *request_id = r.id;
    HV_TRY(s->progress());  // free slots first, so that a request never waits behind a finished one
    if (src.inflight < src.slots && src.pending.empty()) return s->admit(source_id, src, r);
    if ((int)src.pending.size() >= src.depth) {
        // a full waiting queue sheds its oldest request (or this one without a queue)
        const Request old = src.depth > 0 ? src.pending.front() : r;
        if (src.depth > 0) src.pending.pop_front();
        s->results.push_back({old.id, source_id, HV_REQ_SHED, 0, -1});
        ++s->shed;
        if (src.depth == 0) return HV_OK;
    }
    src.pending.push_back(r);
    return HV_OK;
}

hv_status hv_server_submit(hv_server* s, int source_id, const void* frames, int64_t t_arrival_ns) {
    return hv_server_submit_request(s, source_id, frames, t_arrival_ns, nullptr);
}

hv_status hv_server_poll_requests(hv_server* s, hv_request_result* out, int max_n, int* n_out) {
    if (!s || !n_out || (max_n > 0 && !out)) return fail(HV_ERR_INVALID_ARG, "null argument");
    HV_TRY(s->progress());
    int n = 0;
    while (n < max_n && !s->results.empty()) {
        out[n++] = s->results.front();
        s->results.pop_front();
    }
    *n_out = n;
    return HV_OK;
}

hv_status hv_server_poll(hv_server* s, int* sources, int64_t* latency_ns, int max_n, int* n_out) {
    if (!s || !n_out) return fail(HV_ERR_INVALID_ARG, "null argument");
    HV_TRY(s->progress());
    int n = 0;
    while (n < max_n && !s->results.empty()) {
        const hv_request_result r = s->results.front();
        s->results.pop_front();
        if (r.status != HV_REQ_DONE) continue;
        if (sources) sources[n] = r.source;
        if (latency_ns) latency_ns[n] = r.latency_ns;
        ++n;
    }
    *n_out = n;
    return HV_OK;
}

hv_status hv_server_drain(hv_server* s) {
    if (!s) return fail(HV_ERR_INVALID_ARG, "server is null");
    for (;;) {
        HV_TRY(s->progress());
        bool waiting = false;
        for (auto& kv : s->sources) waiting = waiting || !kv.second.pending.empty();
        if (!waiting) break;
        // the oldest in-flight request frees a slot
        if (!s->inflight.empty()) HV_HIP(hipEventSynchronize(s->inflight.front().done));
    }
    for (auto& kv : s->sources) HV_TRY(hv_pipe_sync(kv.second.pipe));
    return s->progress();
}

hv_status hv_server_output(hv_server* s, int source_id, const uint16_t** out_u16, const float** out_f32_image) {
    if (!s) return fail(HV_ERR_INVALID_ARG, "server is null");
    auto it = s->sources.find(source_id);
    if (it == s->sources.end()) return fail(HV_ERR_INVALID_ARG, "unknown source");
    return hv_pipe_output(it->second.pipe, out_u16, nullptr, out_f32_image);
}

hv_status hv_server_output_frame(hv_server* s, int source_id, int64_t output_index, void* stream,
                                 const uint16_t** frame) {
    if (!s || output_index < 0) return fail(HV_ERR_INVALID_ARG, "null server or negative output index");
    auto it = s->sources.find(source_id);
    if (it == s->sources.end()) return fail(HV_ERR_INVALID_ARG, "unknown source");
    return hv_pipe_output_frame(it->second.pipe, (uint64_t)output_index, stream, frame);
}

hv_status hv_server_stats(hv_server* s, uint64_t* shed, uint32_t* pending) {
    if (!s) return fail(HV_ERR_INVALID_ARG, "server is null");
    if (shed) *shed = s->shed;
    if (pending) {
        uint32_t n = 0;
        for (auto& kv : s->sources) n += (uint32_t)kv.second.pending.size();
        *pending = n;
    }
    return HV_OK;
}

hv_status hv_server_memory_bytes(hv_server* s, size_t* bytes) {
    if (!s || !bytes) return fail(HV_ERR_INVALID_ARG, "null argument");
    *bytes = s->used;
    return HV_OK;
}

hv_status hv_server_destroy(hv_server* s) {
    if (!s) return HV_OK;
    for (auto& kv : s->sources) hv_pipe_destroy(kv.second.pipe);
    for (auto& f : s->inflight) (void)hipEventDestroy(f.done);
    for (auto e : s->free_events) (void)hipEventDestroy(e);
    delete s;
    return HV_OK;
}

}  // extern "C"
