// hv_input.hip -- the input side of the pipe: packed 10/12-bit unpack and the
// batch input queue in HBM.
//
// Unpack (Backend/cuda_sources/unpack.cu:4-60): a packed stream holds the
// pixel values MSB first; each 16-bit output keeps the value in its top bits
// (12-bit: value << 4, 10-bit: value << 6), exactly as the reference writes
// it.  One thread turns 8 pixels (12 bytes, 12-bit) or 16 pixels (20 bytes,
// 10-bit) into one or two 16-byte stores: HBM-bound byte work, 3.5 (12-bit)
// or 3.25 (10-bit) bytes per pixel.
//
// Queue (Backend/sources/data_struct/batch_input_queue.cc:124-444): frames
// are appended into whole batches of one HBM allocation; a completed batch
// becomes visible, the oldest is overwritten when the queue is full, and the
// consumer takes the oldest batch in place (no copy) for the pipe.  The
// reference orders producer and consumer with one stream and one mutex per
// batch; here each slot has two events (filled by the producer's stream,
// released by the consumer's), so neither side waits on the host except when
// the producer reaches the slot the consumer still holds.
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "hv_common.hpp"

namespace {

struct alignas(4) Bytes12 {
    uint32_t w0, w1, w2;
};
struct alignas(4) Bytes20 {
    uint32_t w0, w1, w2, w3, w4;
};

__device__ __forceinline__ uint32_t byte_of(uint32_t w, int k) { return (w >> (8 * k)) & 0xFFu; }

// two 12-bit pixels from 3 bytes -> two u16 in one word (low half first)
__device__ __forceinline__ uint32_t pair12(uint32_t b0, uint32_t b1, uint32_t b2) {
    uint32_t o0 = (b0 << 8) | (b1 & 0xF0u);
    uint32_t o1 = ((b1 & 0x0Fu) << 12) | (b2 << 4);
    return o0 | (o1 << 16);
}

// four 10-bit pixels from 5 bytes -> four u16 in two words
__device__ __forceinline__ uint2 quad10(uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3, uint32_t b4) {
    uint32_t o0 = (b0 << 8) | (b1 & 0xC0u);
    uint32_t o1 = ((b1 & 0x3Fu) << 10) | ((b2 & 0xF0u) << 2);
    uint32_t o2 = ((b2 & 0x0Fu) << 12) | ((b3 & 0xFCu) << 4);
    uint32_t o3 = ((b3 & 0x03u) << 14) | (b4 << 6);
    return make_uint2(o0 | (o1 << 16), o2 | (o3 << 16));
}

// groups per thread, loads issued before stores (1 measured best: 5.96 TB/s
// at 12 bits, the speed of a flat device copy; 2 the same, 4 slower)
#ifndef HV_UNPACK_GPT
#define HV_UNPACK_GPT 1
#endif
constexpr int kGpt = HV_UNPACK_GPT;

// 8 pixels per group: 12 aligned bytes in, one 16-byte store out; a block
// covers 256 * kGpt consecutive groups, thread t the groups t, t + 256, ...
__global__ void __launch_bounds__(256) k_unpack12_vec(uint4* __restrict__ out, const Bytes12* __restrict__ in,
                                                      size_t groups) {
    const size_t g0 = (size_t)blockIdx.x * (256 * kGpt) + threadIdx.x;
    Bytes12 v[kGpt];
#pragma unroll
    for (int k = 0; k < kGpt; ++k)
        if (g0 + 256 * k < groups) v[k] = in[g0 + 256 * k];
#pragma unroll
    for (int k = 0; k < kGpt; ++k) {
        if (g0 + 256 * k >= groups) break;
        uint4 o;
        o.x = pair12(byte_of(v[k].w0, 0), byte_of(v[k].w0, 1), byte_of(v[k].w0, 2));
        o.y = pair12(byte_of(v[k].w0, 3), byte_of(v[k].w1, 0), byte_of(v[k].w1, 1));
        o.z = pair12(byte_of(v[k].w1, 2), byte_of(v[k].w1, 3), byte_of(v[k].w2, 0));
        o.w = pair12(byte_of(v[k].w2, 1), byte_of(v[k].w2, 2), byte_of(v[k].w2, 3));
        out[g0 + 256 * k] = o;
    }
}

// 16 pixels per group: 20 aligned bytes in, two 16-byte stores out
__global__ void __launch_bounds__(256) k_unpack10_vec(uint4* __restrict__ out, const Bytes20* __restrict__ in,
                                                      size_t groups) {
    const size_t g0 = (size_t)blockIdx.x * (256 * kGpt) + threadIdx.x;
    Bytes20 v[kGpt];
#pragma unroll
    for (int k = 0; k < kGpt; ++k)
        if (g0 + 256 * k < groups) v[k] = in[g0 + 256 * k];
#pragma unroll
    for (int k = 0; k < kGpt; ++k) {
        const size_t g = g0 + 256 * k;
        if (g >= groups) break;
        const Bytes20& w = v[k];
        uint2 a = quad10(byte_of(w.w0, 0), byte_of(w.w0, 1), byte_of(w.w0, 2), byte_of(w.w0, 3), byte_of(w.w1, 0));
        uint2 b = quad10(byte_of(w.w1, 1), byte_of(w.w1, 2), byte_of(w.w1, 3), byte_of(w.w2, 0), byte_of(w.w2, 1));
        uint2 c = quad10(byte_of(w.w2, 2), byte_of(w.w2, 3), byte_of(w.w3, 0), byte_of(w.w3, 1), byte_of(w.w3, 2));
        uint2 d = quad10(byte_of(w.w3, 3), byte_of(w.w4, 0), byte_of(w.w4, 1), byte_of(w.w4, 2), byte_of(w.w4, 3));
        out[2 * g] = make_uint4(a.x, a.y, b.x, b.y);
        out[2 * g + 1] = make_uint4(c.x, c.y, d.x, d.y);
    }
}

// byte-wise units [first, first + n): pairs (12-bit) or quads (10-bit); for
// the tail and for buffers without the vector alignment
template <int BITS>
__global__ void __launch_bounds__(256) k_unpack_units(uint16_t* __restrict__ out, const uint8_t* __restrict__ in,
                                                      size_t first, size_t n) {
    size_t u = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    u += first;
    if (BITS == 12) {
        const uint8_t* p = in + 3 * u;
        uint32_t w = pair12(p[0], p[1], p[2]);
        out[2 * u] = (uint16_t)(w & 0xFFFFu);
        out[2 * u + 1] = (uint16_t)(w >> 16);
    } else {
        const uint8_t* p = in + 5 * u;
        uint2 w = quad10(p[0], p[1], p[2], p[3], p[4]);
        out[4 * u] = (uint16_t)(w.x & 0xFFFFu);
        out[4 * u + 1] = (uint16_t)(w.x >> 16);
        out[4 * u + 2] = (uint16_t)(w.y & 0xFFFFu);
        out[4 * u + 3] = (uint16_t)(w.y >> 16);
    }
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

namespace hv {

// pixels: output u16 count (a multiple of 2 for 12-bit, of 4 for 10-bit);
// the input holds pixels * BITS / 8 bytes
static hv_status launch_unpack(int bits, uint16_t* out, const uint8_t* in, size_t pixels, hipStream_t st) {
    const size_t px_per_group = bits == 12 ? 8 : 16, px_per_unit = bits == 12 ? 2 : 4;
    const size_t in_group = bits == 12 ? 12 : 20;
    const bool aligned = ((uintptr_t)in % 4 == 0) && ((uintptr_t)out % 16 == 0);
    size_t groups = aligned ? pixels / px_per_group : 0;
    // the grid's x dimension holds up to 2^31 - 1 blocks; split larger streams
    const size_t max_groups = (size_t)0x7FFFFFFF / 2 * 256;
    for (size_t g0 = 0; g0 < groups; g0 += max_groups) {
        const size_t g = std::min(max_groups, groups - g0);
        const unsigned nb = (unsigned)((g + 256 * kGpt - 1) / (256 * kGpt));
        if (bits == 12)
            hipLaunchKernelGGL(k_unpack12_vec, dim3(nb), dim3(256), 0, st,
                               (uint4*)(out + g0 * px_per_group), (const Bytes12*)(in + g0 * in_group), g);
        else
            hipLaunchKernelGGL(k_unpack10_vec, dim3(nb), dim3(256), 0, st,
                               (uint4*)(out + g0 * px_per_group), (const Bytes20*)(in + g0 * in_group), g);
        HV_CHECK_LAUNCH("k_unpack_vec");
    }
    const size_t units = pixels / px_per_unit, first = groups * px_per_group / px_per_unit;
    for (size_t u0 = first; u0 < units; u0 += max_groups) {
        const size_t n = std::min(max_groups, units - u0);
        if (bits == 12)
            hipLaunchKernelGGL(k_unpack_units<12>, dim3(blocks_for(n)), dim3(256), 0, st, out, in, u0, n);
        else
            hipLaunchKernelGGL(k_unpack_units<10>, dim3(blocks_for(n)), dim3(256), 0, st, out, in, u0, n);
        HV_CHECK_LAUNCH("k_unpack_units");
    }
    return HV_OK;
}

static hv_status check_unpack(int bits, const void* out, size_t output_size, const void* in, size_t input_size) {
    if (!out || !in) return fail(HV_ERR_INVALID_ARG, "unpack: null buffer");
    const size_t unit = bits == 12 ? 2 : 4;
    if (output_size == 0 || output_size % unit)
        return fail(HV_ERR_INVALID_ARG, "unpack: output_size must be a positive multiple of " + std::to_string(unit) +
                                            " pixels");
    if (input_size < output_size * (size_t)bits / 8)
        return fail(HV_ERR_INVALID_ARG, "unpack: input_size " + std::to_string(input_size) + " < " +
                                            std::to_string(output_size * (size_t)bits / 8) + " packed bytes");
    return HV_OK;
}

}  // namespace hv

// ---------------------------------------------------------------------------
// batch input queue
// ---------------------------------------------------------------------------
struct hv_input_queue {
    uint32_t frame_capacity = 0;  // frames asked for at creation
    uint32_t batch_size = 0;
    uint32_t max_size = 0;        // whole batches held
    uint32_t width = 0, height = 0, depth = 0;
    size_t frame_size = 0;
    char* data = nullptr;
    uint8_t* packed = nullptr;  // device staging of host packed frames (one batch)
    size_t packed_bytes = 0;

    // BatchInputQueue state
    uint32_t start = 0, end = 0, size = 0, curr_batch_counter = 0, curr_nb_frames = 0;
    bool has_overwritten = false;
    int held = -1;  // slot handed to the consumer by front, until pop
    std::thread::id holder;

    std::vector<hipEvent_t> filled, released;
    std::vector<char> filled_rec, released_rec;
    std::mutex m;
    std::condition_variable cv;
};

namespace {

using hv::fail;

hv_status make_slots(hv_input_queue* q, uint32_t batch_size) {
    q->batch_size = batch_size;
    const uint32_t total = q->frame_capacity - q->frame_capacity % batch_size;  // batch_input_queue.cc create_queue
    if (total == 0) return fail(HV_ERR_INVALID_ARG, "There must be more at least a frame in the queue.");
    q->max_size = total / batch_size;
    for (auto e : q->filled) (void)hipEventDestroy(e);
    for (auto e : q->released) (void)hipEventDestroy(e);
    q->filled.assign(q->max_size, nullptr);
    q->released.assign(q->max_size, nullptr);
    q->filled_rec.assign(q->max_size, 0);
    q->released_rec.assign(q->max_size, 0);
    for (uint32_t i = 0; i < q->max_size; ++i) {
        HV_HIP(hipEventCreateWithFlags(&q->filled[i], hipEventDisableTiming));
        HV_HIP(hipEventCreateWithFlags(&q->released[i], hipEventDisableTiming));
    }
    q->start = q->end = q->size = q->curr_batch_counter = q->curr_nb_frames = 0;
    q->has_overwritten = false;
    q->held = -1;
    return HV_OK;
}

hv_status sync_all(hv_input_queue* q) {
    for (uint32_t i = 0; i < q->max_size; ++i) {
        if (q->filled_rec[i]) HV_HIP(hipEventSynchronize(q->filled[i]));
        if (q->released_rec[i]) HV_HIP(hipEventSynchronize(q->released[i]));
    }
    return HV_OK;
}

// Appends n frames through `put(dst, frame_offset, count)`, which enqueues
// the copy (or unpack) of `count` source frames from `frame_offset` into dst
// on the producer stream (batch_input_queue.cc enqueue).
// `stop` (may be null) ends a wait for the consumer's slot.
template <typename Put>
hv_status enqueue_frames(hv_input_queue* q, uint32_t nb_frames, hipStream_t st, Put put,
                         const std::atomic<bool>* stop = nullptr) {
    std::unique_lock<std::mutex> lk(q->m);
    uint32_t done = 0;
    while (done < nb_frames) {
        if (q->curr_batch_counter == 0) {
            // a new batch in slot `end`: wait while the consumer holds it, and
            // order the overwrite after the consumer's last use on the GPU
            if (q->held == (int)q->end && q->holder == std::this_thread::get_id())
                return fail(HV_ERR_INVALID_ARG, "enqueue would overwrite the batch this thread holds: pop it first");
            q->cv.wait(lk, [&] { return q->held != (int)q->end || (stop && stop->load()); });
            if (q->held == (int)q->end)
                return fail(HV_ERR_NOT_READY, "enqueue stopped while the consumer holds the slot to overwrite");
            if (q->released_rec[q->end]) HV_HIP(hipStreamWaitEvent(st, q->released[q->end], 0));
        }
        const uint32_t n = std::min(nb_frames - done, q->batch_size - q->curr_batch_counter);
        char* dst = q->data + ((size_t)q->end * q->batch_size + q->curr_batch_counter) * q->frame_size;
        HV_TRY(put(dst, done, n));
        q->curr_batch_counter += n;
        if (q->curr_batch_counter == q->batch_size) {
            q->curr_batch_counter = 0;
            HV_HIP(hipEventRecord(q->filled[q->end], st));
            q->filled_rec[q->end] = 1;
            q->end = (q->end + 1) % q->max_size;
            if (q->size == q->max_size) {
                q->has_overwritten = true;
                q->start = (q->start + 1) % q->max_size;
            } else {
                q->size++;
                q->curr_nb_frames += q->batch_size;
            }
            q->cv.notify_all();
        }
        done += n;
    }
    return HV_OK;
}

// the producer is refilling the oldest slot (a full queue being overwritten)
bool producer_holds_start(const hv_input_queue* q) { return q->curr_batch_counter > 0 && q->end == q->start; }

}  // namespace

using namespace hv;

hv_status hv::input_queue_enqueue_until(hv_input_queue* q, const void* frames, int memcpy_kind, uint32_t nb_frames,
                                        void* stream, const std::atomic<bool>* stop) {
    if (!q || (!frames && nb_frames)) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_enqueue: null argument");
    if (memcpy_kind != (int)hipMemcpyHostToDevice && memcpy_kind != (int)hipMemcpyDeviceToDevice)
        return fail(HV_ERR_INVALID_ARG, "Input queue : can't cudaMemcpy to host with the queue on gpu");
    hipStream_t st = as_stream(stream);
    const char* src = (const char*)frames;
    return enqueue_frames(
        q, nb_frames, st,
        [&](char* dst, uint32_t off, uint32_t n) -> hv_status {
            HV_HIP(hipMemcpyAsync(dst, src + (size_t)off * q->frame_size, (size_t)n * q->frame_size,
                                  (hipMemcpyKind)memcpy_kind, st));
            return HV_OK;
        },
        stop);
}

void hv::input_queue_wake(hv_input_queue* q) {
    std::lock_guard<std::mutex> lk(q->m);
    q->cv.notify_all();
}

extern "C" {

hv_status hv_unpack_12_to_16bit(uint16_t* output, size_t output_size, const uint8_t* input, size_t input_size,
                                void* stream) {
    HV_TRY(check_unpack(12, output, output_size, input, input_size));
    return launch_unpack(12, output, input, output_size, as_stream(stream));
}

hv_status hv_unpack_10_to_16bit(uint16_t* output, size_t output_size, const uint8_t* input, size_t input_size,
                                void* stream) {
    HV_TRY(check_unpack(10, output, output_size, input, input_size));
    return launch_unpack(10, output, input, output_size, as_stream(stream));
}

hv_status hv_input_queue_create(uint32_t frame_capacity, uint32_t batch_size, uint32_t width, uint32_t height,
                                uint32_t depth, hv_input_queue** out) {
    if (!out) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_create: null output");
    *out = nullptr;
    if (batch_size == 0) return fail(HV_ERR_INVALID_ARG, "Batch size cannot be 0.");
    if (width == 0 || height == 0 || depth == 0) return fail(HV_ERR_INVALID_ARG, "empty frame descriptor");
    auto* q = new hv_input_queue;
    q->frame_capacity = frame_capacity;
    q->width = width;
    q->height = height;
    q->depth = depth;
    q->frame_size = (size_t)width * height * depth;
    hv_status s = make_slots(q, batch_size);
    if (s == HV_OK) {
        hipError_t e = hipMalloc((void**)&q->data, (size_t)q->max_size * batch_size * q->frame_size);
        if (e != hipSuccess) s = fail(HV_ERR_OUT_OF_MEMORY, "hv_input_queue_create: HBM allocation failed");
    }
    if (s != HV_OK) {
        for (auto e : q->filled) (void)hipEventDestroy(e);
        for (auto e : q->released) (void)hipEventDestroy(e);
        delete q;
        return s;
    }
    *out = q;
    return HV_OK;
}

hv_status hv_input_queue_enqueue(hv_input_queue* q, const void* frames, int memcpy_kind, uint32_t nb_frames,
                                 void* stream) {
    return input_queue_enqueue_until(q, frames, memcpy_kind, nb_frames, stream, nullptr);
}

hv_status hv_input_queue_enqueue_packed(hv_input_queue* q, const void* packed_frames, int memcpy_kind, int bits,
                                        uint32_t nb_frames, void* stream) {
    if (!q || (!packed_frames && nb_frames))
        return fail(HV_ERR_INVALID_ARG, "hv_input_queue_enqueue_packed: null argument");
    if (bits != 10 && bits != 12) return fail(HV_ERR_UNSUPPORTED, "Image encoding format not supported: bits");
    if (q->depth != 2) return fail(HV_ERR_INVALID_ARG, "packed frames unpack into a 16-bit queue");
    const size_t pixels = (size_t)q->width * q->height;
    if (pixels % (bits == 12 ? 2 : 4)) return fail(HV_ERR_INVALID_ARG, "frame size does not hold whole packed units");
    if (memcpy_kind != (int)hipMemcpyHostToDevice && memcpy_kind != (int)hipMemcpyDeviceToDevice)
        return fail(HV_ERR_INVALID_ARG, "packed frames come from the host or the device");
    const size_t pfs = pixels * (size_t)bits / 8;
    hipStream_t st = as_stream(stream);
    if (memcpy_kind == (int)hipMemcpyHostToDevice && q->packed_bytes < pfs * q->batch_size) {
        // one batch of packed frames; reused chunk after chunk on the producer stream
        if (q->packed) {
            // the old staging may still feed an unpack of a batch in progress
            HV_HIP(hipStreamSynchronize(st));
            HV_TRY(sync_all(q));
            HV_HIP(hipFree(q->packed));
            q->packed = nullptr;
        }
        HV_HIP(hipMalloc((void**)&q->packed, pfs * q->batch_size));
        q->packed_bytes = pfs * q->batch_size;
    }
    const uint8_t* src = (const uint8_t*)packed_frames;
    return enqueue_frames(q, nb_frames, st, [&](char* dst, uint32_t off, uint32_t n) -> hv_status {
        const uint8_t* p = src + (size_t)off * pfs;
        if (memcpy_kind == (int)hipMemcpyHostToDevice) {
            HV_HIP(hipMemcpyAsync(q->packed, p, (size_t)n * pfs, hipMemcpyHostToDevice, st));
            p = q->packed;
        }
        return launch_unpack(bits, (uint16_t*)dst, p, (size_t)n * pixels, st);
    });
}

hv_status hv_input_queue_front(hv_input_queue* q, void* stream, int timeout_ms, const void** batch) {
    if (!q || !batch) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_front: null argument");
    std::unique_lock<std::mutex> lk(q->m);
    if (q->held >= 0) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_front: the front batch is already held");
    // a whole batch that the producer is not overwriting (the reference's consumer waits for it)
    auto ready = [&] { return q->size > 0 && !producer_holds_start(q); };
    if (!ready() && (timeout_ms <= 0 || !q->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)))
        return HV_ERR_NOT_READY;
    q->held = (int)q->start;
    q->holder = std::this_thread::get_id();
    if (q->filled_rec[q->start]) HV_HIP(hipStreamWaitEvent(as_stream(stream), q->filled[q->start], 0));
    *batch = q->data + (size_t)q->start * q->batch_size * q->frame_size;
    return HV_OK;
}

hv_status hv_input_queue_pop(hv_input_queue* q, void* stream) {
    if (!q) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_pop: null queue");
    std::unique_lock<std::mutex> lk(q->m);
    if (q->held < 0) {
        // dequeue() without a copy: drops the oldest whole batch if there is one
        if (q->size == 0) return HV_OK;
        if (producer_holds_start(q)) return HV_ERR_NOT_READY;
    } else {
        HV_HIP(hipEventRecord(q->released[q->held], as_stream(stream)));
        q->released_rec[q->held] = 1;
        q->held = -1;
    }
    q->start = (q->start + 1) % q->max_size;
    q->size--;
    q->curr_nb_frames -= q->batch_size;
    q->cv.notify_all();
    return HV_OK;
}

hv_status hv_input_queue_dequeue(hv_input_queue* q, void* dst, int memcpy_kind, void* stream) {
    if (!q || !dst) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_dequeue: null argument");
    if (memcpy_kind != (int)hipMemcpyDeviceToDevice && memcpy_kind != (int)hipMemcpyDeviceToHost)
        return fail(HV_ERR_INVALID_ARG, "hv_input_queue_dequeue: copy to the device or the host");
    const void* b = nullptr;
    hv_status s = hv_input_queue_front(q, stream, 0, &b);
    if (s != HV_OK) {
        if (s == HV_ERR_NOT_READY) set_error("hv_input_queue_dequeue: no whole batch is ready");
        return s;
    }
    HV_HIP(hipMemcpyAsync(dst, b, (size_t)q->batch_size * q->frame_size, (hipMemcpyKind)memcpy_kind,
                          as_stream(stream)));
    HV_TRY(hv_input_queue_pop(q, stream));
    // the reference's dequeue synchronises the batch's stream before returning
    HV_HIP(hipStreamSynchronize(as_stream(stream)));
    return HV_OK;
}

hv_status hv_input_queue_state(hv_input_queue* q, uint32_t* size, uint32_t* max_size, uint32_t* start_index,
                               uint32_t* curr_nb_frames, int* has_overwritten) {
    if (!q) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_state: null queue");
    std::lock_guard<std::mutex> lk(q->m);
    if (size) *size = q->size;
    if (max_size) *max_size = q->max_size;
    if (start_index) *start_index = q->start;
    if (curr_nb_frames) *curr_nb_frames = q->curr_nb_frames;
    if (has_overwritten) *has_overwritten = q->has_overwritten ? 1 : 0;
    return HV_OK;
}

hv_status hv_input_queue_data(hv_input_queue* q, const void** data) {
    if (!q || !data) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_data: null argument");
    *data = q->data;
    return HV_OK;
}

hv_status hv_input_queue_resize(hv_input_queue* q, uint32_t new_batch_size) {
    if (!q) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_resize: null queue");
    if (new_batch_size == 0) return fail(HV_ERR_INVALID_ARG, "Batch size cannot be 0.");
    std::unique_lock<std::mutex> lk(q->m);
    if (q->held >= 0) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_resize: a batch is held by the consumer");
    if (q->frame_capacity - q->frame_capacity % new_batch_size == 0)
        return fail(HV_ERR_INVALID_ARG, "There must be more at least a frame in the queue.");
    HV_TRY(sync_all(q));
    // the allocation holds frame_capacity frames, enough for any batch size
    return make_slots(q, new_batch_size);
}

hv_status hv_input_queue_reset(hv_input_queue* q) {
    if (!q) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_reset: null queue");
    std::unique_lock<std::mutex> lk(q->m);
    if (q->held >= 0) return fail(HV_ERR_INVALID_ARG, "hv_input_queue_reset: a batch is held by the consumer");
    HV_TRY(sync_all(q));
    q->start = q->end = q->size = q->curr_batch_counter = q->curr_nb_frames = 0;
    q->has_overwritten = false;
    return HV_OK;
}

hv_status hv_input_queue_destroy(hv_input_queue* q) {
    if (!q) return HV_OK;
    (void)sync_all(q);
    for (auto e : q->filled) (void)hipEventDestroy(e);
    for (auto e : q->released) (void)hipEventDestroy(e);
    if (q->data) (void)hipFree(q->data);
    if (q->packed) (void)hipFree(q->packed);
    delete q;
    return HV_OK;
}

}  // extern "C"
