// hv_source.cpp -- frame sources and the read worker that feeds the batch
// input queue.
//
// A source implements the camera interface of the reference
// (Camera/include/icamera.hh:91-166) as a table of C function pointers, so a
// camera plugin can be written in any language with a C ABI.  get_frames
// returns up to two regions of frames (CapturedFramesDescriptor,
// icamera.hh:8-58: the end and the start of the camera's ring buffer).  Two
// sources are built in: a synthetic camera replaying a ring of generated
// frames held in pinned memory, and a .holo file read in chunks (looping at
// the end, as the reference's file reader does).  The read worker is one
// host thread per source (CameraFrameReadWorker / FileFrameReadWorker,
// Backend/sources/thread/camera_frame_read_worker.cc:20-72,
// file_frame_read_worker.cc:35-312): get_frames, enqueue into the HBM queue
// on its own stream, optional frame-rate limit.
#include <atomic>
#include <chrono>
#include <cstring>
#include <string>
#include <thread>

#include "hv_common.hpp"

using namespace hv;

namespace {

struct SynthCamera {
    hv_synth_params p{};
    uint32_t ring_frames = 0, frames_per_call = 0, next = 0;
    char* ring = nullptr;  // pinned
    size_t frame_size = 0;
    bool started = false;
};

struct FileCamera {
    hv_holo_reader* r = nullptr;
    hv_holo_info info{};
    uint32_t frames_per_call = 0;
    uint64_t next = 0, first = 0, end = 0;
    char* buf = nullptr;  // pinned, frames_per_call frames
    bool started = false;
};

hv_status synth_init(void* self) {
    auto* c = (SynthCamera*)self;
    if (c->ring) return HV_OK;
    c->frame_size = (size_t)c->p.width * c->p.height * c->p.depth;
    HV_HIP(hipHostMalloc((void**)&c->ring, c->frame_size * c->ring_frames, hipHostMallocDefault));
    return hv_synth_frames(&c->p, 0, c->ring_frames, c->ring);
}
hv_status synth_start(void* self) {
    ((SynthCamera*)self)->started = true;
    return HV_OK;
}
hv_status synth_stop(void* self) {
    ((SynthCamera*)self)->started = false;
    return HV_OK;
}
hv_status synth_shutdown(void* self) {
    auto* c = (SynthCamera*)self;
    if (c->ring) (void)hipHostFree(c->ring);
    delete c;
    return HV_OK;
}
// frames [next, next + n) of the ring, split where the ring wraps
hv_status synth_get_frames(void* self, hv_captured_frames* out) {
    auto* c = (SynthCamera*)self;
    if (!c->started) return fail(HV_ERR_INVALID_ARG, "synthetic camera: acquisition not started");
    const uint32_t n = c->frames_per_call;
    const uint32_t c1 = std::min(n, c->ring_frames - c->next);
    out->region1 = c->ring + (size_t)c->next * c->frame_size;
    out->count1 = c1;
    out->region2 = c1 < n ? c->ring : nullptr;
    out->count2 = n - c1;
    out->on_gpu = 0;
    c->next = (c->next + n) % c->ring_frames;
    return HV_OK;
}
hv_status synth_fd(void* self, uint32_t* w, uint32_t* h, uint32_t* d) {
    auto* c = (SynthCamera*)self;
    *w = c->p.width;
    *h = c->p.height;
    *d = c->p.depth;
    return HV_OK;
}
int synth_fps(void*) { return -1; }

hv_status file_init(void* self) {
    auto* c = (FileCamera*)self;
    if (c->buf) return HV_OK;
    HV_HIP(hipHostMalloc((void**)&c->buf, c->info.frame_size * c->frames_per_call, hipHostMallocDefault));
    return HV_OK;
}
hv_status file_start(void* self) {
    auto* c = (FileCamera*)self;
    c->started = true;
    c->next = c->first;
    return HV_OK;
}
hv_status file_stop(void* self) {
    ((FileCamera*)self)->started = false;
    return HV_OK;
}
hv_status file_shutdown(void* self) {
    auto* c = (FileCamera*)self;
    if (c->buf) (void)hipHostFree(c->buf);
    (void)hv_holo_close(c->r);
    delete c;
    return HV_OK;
}
// the next frames_per_call frames of [first, end), from `first` again at the end
hv_status file_get_frames(void* self, hv_captured_frames* out) {
    auto* c = (FileCamera*)self;
    if (!c->started) return fail(HV_ERR_INVALID_ARG, "file source: acquisition not started");
    if (c->next >= c->end) c->next = c->first;
    uint64_t got = 0;
    HV_TRY(hv_holo_read_frames(c->r, c->next, std::min<uint64_t>(c->frames_per_call, c->end - c->next), c->buf,
                               &got));
    c->next += got;
    out->region1 = c->buf;
    out->count1 = (uint32_t)got;
    out->region2 = nullptr;
    out->count2 = 0;
    out->on_gpu = 0;
    return HV_OK;
}
hv_status file_fd(void* self, uint32_t* w, uint32_t* h, uint32_t* d) {
    auto* c = (FileCamera*)self;
    *w = c->info.width;
    *h = c->info.height;
    *d = c->info.depth;
    return HV_OK;
}
int file_fps(void*) { return -1; }

}  // namespace

struct hv_reader {
    hv_camera cam{};
    hv_input_queue* q = nullptr;
    hipStream_t stream = nullptr;
    float fps_limit = 0.f;
    std::thread th;
    std::atomic<bool> stop{false};
    std::atomic<uint64_t> frames{0};
    hv_status status = HV_OK;
    std::string error;
};

static void reader_loop(hv_reader* r) {
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto set = [&](hv_status s) {
        r->status = s;
        r->error = hv_last_error();
    };
    hv_status s = r->cam.start_acquisition(r->cam.self);
    if (s != HV_OK) return set(s);
    while (!r->stop.load(std::memory_order_relaxed)) {
        hv_captured_frames cf{};
        s = r->cam.get_frames(r->cam.self, &cf);
        if (s != HV_OK) break;
        const int kind = cf.on_gpu ? (int)hipMemcpyDeviceToDevice : (int)hipMemcpyHostToDevice;
        // a stop ends a wait for the slot the consumer holds (full queue)
        if (cf.count1) s = input_queue_enqueue_until(r->q, cf.region1, kind, cf.count1, r->stream, &r->stop);
        if (s == HV_OK && cf.count2)
            s = input_queue_enqueue_until(r->q, cf.region2, kind, cf.count2, r->stream, &r->stop);
        if (s == HV_ERR_NOT_READY && r->stop.load()) {
            s = HV_OK;  // stopped while waiting: not an error of the source
            break;
        }
        // the source may reuse its buffers after this call returns
        if (s == HV_OK && hipStreamSynchronize(r->stream) != hipSuccess) s = fail(HV_ERR_HIP, "reader stream");
        if (s != HV_OK) break;
        const uint64_t n = r->frames.fetch_add(cf.count1 + cf.count2) + cf.count1 + cf.count2;
        if (r->fps_limit > 0.f) {
            // hold the input rate: frame n is due at t0 + n / fps
            std::this_thread::sleep_until(t0 + std::chrono::duration_cast<clock::duration>(
                                                   std::chrono::duration<double>(n / (double)r->fps_limit)));
        }
    }
    hv_status s2 = r->cam.stop_acquisition(r->cam.self);
    set(s != HV_OK ? s : s2);
}

extern "C" {

hv_status hv_camera_synthetic(const hv_synth_params* p, uint32_t ring_frames, uint32_t frames_per_call,
                              hv_camera* out) {
    if (!p || !out) return fail(HV_ERR_INVALID_ARG, "hv_camera_synthetic: null argument");
    if (ring_frames == 0 || frames_per_call == 0 || frames_per_call > ring_frames)
        return fail(HV_ERR_INVALID_ARG, "hv_camera_synthetic: need 0 < frames_per_call <= ring_frames");
    auto* c = new SynthCamera;
    c->p = *p;
    c->ring_frames = ring_frames;
    c->frames_per_call = frames_per_call;
    *out = hv_camera{c, synth_init, synth_start, synth_stop, synth_shutdown, synth_get_frames, synth_fd, synth_fps};
    return HV_OK;
}

hv_status hv_camera_holo_file(const char* path, uint64_t first_frame, uint64_t end_frame, uint32_t frames_per_call,
                              hv_camera* out) {
    if (!path || !out || frames_per_call == 0) return fail(HV_ERR_INVALID_ARG, "hv_camera_holo_file: bad argument");
    hv_holo_reader* r = nullptr;
    HV_TRY(hv_holo_open(path, &r));
    auto* c = new FileCamera;
    c->r = r;
    (void)hv_holo_get_info(r, &c->info);
    c->frames_per_call = frames_per_call;
    // InputFileStartIndex / InputFileEndIndex (file_frame_read_worker.cc:38-40); 0 = to the end
    c->end = end_frame ? std::min<uint64_t>(end_frame, c->info.frames_in_file) : c->info.frames_in_file;
    c->first = first_frame;
    if (c->first >= c->end) {
        (void)hv_holo_close(r);
        delete c;
        return fail(HV_ERR_INVALID_ARG, "hv_camera_holo_file: empty frame range");
    }
    *out = hv_camera{c, file_init, file_start, file_stop, file_shutdown, file_get_frames, file_fd, file_fps};
    return HV_OK;
}

hv_status hv_reader_start(const hv_camera* cam, hv_input_queue* q, float fps_limit, hv_reader** out) {
    if (!cam || !q || !out) return fail(HV_ERR_INVALID_ARG, "hv_reader_start: null argument");
    if (!cam->get_frames || !cam->start_acquisition || !cam->stop_acquisition || !cam->init_camera)
        return fail(HV_ERR_INVALID_ARG, "hv_reader_start: incomplete camera table");
    HV_TRY(cam->init_camera(cam->self));
    auto* r = new hv_reader;
    r->cam = *cam;
    r->q = q;
    r->fps_limit = fps_limit;
    hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete r;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    r->th = std::thread(reader_loop, r);
    *out = r;
    return HV_OK;
}

hv_status hv_reader_frames(hv_reader* r, uint64_t* frames) {
    if (!r || !frames) return fail(HV_ERR_INVALID_ARG, "hv_reader_frames: null argument");
    *frames = r->frames.load();
    return HV_OK;
}

hv_status hv_reader_stop(hv_reader* r, uint64_t* frames) {
    if (!r) return HV_OK;
    r->stop.store(true);
    input_queue_wake(r->q);  // a worker waiting for the consumer's slot sees the stop
    if (r->th.joinable()) r->th.join();
    if (frames) *frames = r->frames.load();
    hv_status s = r->status;
    std::string err = r->error;
    (void)hipStreamDestroy(r->stream);
    delete r;
    if (s != HV_OK) return fail(s, "reader: " + err);
    return HV_OK;
}

}  // extern "C"
