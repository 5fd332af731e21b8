// hv_holo.cpp -- .holo recordings: 64-byte header, raw frames, JSON footer,
// and the file <-> HBM streaming that feeds the pipe from a recording and
// records its output.
//
// File layout (Backend/includes/io/holo_file.hh:58-83, packed on 2 bytes):
//   [0, 64)                 header (magic "HOLO", version, bits/pixel, W, H,
//                           img_nb, total_data_size, endianness, data_type,
//                           34 padding bytes), little-endian
//   [64, 64 + data)         img_nb frames of W*H*(bits/8) bytes, back to back
//   [64 + data, file end)   JSON footer (compute settings + info), optional
//
// The reference reads a batch with one blocking fread into a pinned buffer,
// then copies it to the device and synchronises
// (Backend/sources/thread/file_frame_read_worker.cc:197-271).  Here a batch
// is cut into chunks that alternate between two pinned staging buffers: the
// chunk's bytes are read by several threads with pread (page-cache copies
// scale with threads) while the DMA engine copies the previous chunk to HBM.
// Recording runs the same pipeline the other way (device -> pinned -> file,
// large writes through a shared mapping of the file so that several threads
// fill its pages at once).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "hv_common.hpp"

static_assert(sizeof(hv_holo_header) == 64, "the .holo header is 64 bytes");
static_assert(offsetof(hv_holo_header, total_data_size) == 20, "pack(2) layout of the .holo header");
static_assert(offsetof(hv_holo_header, endianness) == 28, "pack(2) layout of the .holo header");

namespace {

#ifndef HV_HOLO_CHUNK_MIB
#define HV_HOLO_CHUNK_MIB 32
#endif
#ifndef HV_HOLO_IO_THREADS
#define HV_HOLO_IO_THREADS 8
#endif
// 1: frames from the device are copied into a shared mapping of the file
// (page faults of several threads proceed in parallel); 0: pwrite (a
// buffered write holds the inode lock, so parallel pwrites serialise)
#ifndef HV_HOLO_WRITE_MMAP
#define HV_HOLO_WRITE_MMAP 1
#endif
// writes below this size use pwrite: a mapping costs ~5 ms per call to set up
// and fault in, so it pays only for large writes (measured: 17 MB 2.9 ms with
// pwrite vs 7.3 ms mapped; 2 GiB 9.6 GB/s with pwrite vs 14.2 GB/s mapped)
#ifndef HV_HOLO_MAP_MIN_MIB
#define HV_HOLO_MAP_MIN_MIB 256
#endif

// chunk of one staging buffer; two are used per reader or writer
constexpr size_t kChunk = size_t(HV_HOLO_CHUNK_MIB) << 20;
// below this a chunk is read or written by one thread
constexpr size_t kParallelMin = size_t(4) << 20;
constexpr uint16_t kCurrentVersion = 7;  // holo_file.hh:98

int io_threads() {
    unsigned n = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min((unsigned)HV_HOLO_IO_THREADS, n / 2));
}

// pread until n bytes or end of file; returns the bytes read, -1 on error
long long pread_full(int fd, char* dst, size_t n, uint64_t off) {
    size_t got = 0;
    while (got < n) {
        ssize_t r = ::pread(fd, dst + got, n - got, (off_t)(off + got));
        if (r < 0) {
            if (errno == EINTR) continue;
            return -1;
        }
        if (r == 0) break;
        got += (size_t)r;
    }
    return (long long)got;
}

bool pwrite_full(int fd, const char* src, size_t n, uint64_t off) {
    size_t put = 0;
    while (put < n) {
        ssize_t r = ::pwrite(fd, src + put, n - put, (off_t)(off + put));
        if (r < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        put += (size_t)r;
    }
    return true;
}

template <typename F>
bool parallel_parts(size_t n, F f);

// writes [off, off + len) of a file whose size already covers it through a
// shared mapping, copying in parallel parts
bool mmap_write(int fd, const char* src, size_t len, uint64_t off) {
    const uint64_t base = off & ~uint64_t(4095);
    const size_t lead = (size_t)(off - base);
    void* m = ::mmap(nullptr, lead + len, PROT_READ | PROT_WRITE, MAP_SHARED, fd, (off_t)base);
    if (m == MAP_FAILED) return false;
    char* dst = (char*)m + lead;
    bool ok = parallel_parts(len, [&](size_t b, size_t l) {
        std::memcpy(dst + b, src + b, l);
        return true;
    });
    return ::munmap(m, lead + len) == 0 && ok;
}

// splits [0, n) into page-aligned parts, one per thread; f(part_off, part_len) -> ok
template <typename F>
bool parallel_parts(size_t n, F f) {
    int nt = n >= kParallelMin ? io_threads() : 1;
    if (nt <= 1) return f(size_t(0), n);
    size_t part = ((n / nt) + 4095) & ~size_t(4095);
    std::vector<std::thread> ts;
    std::vector<char> ok(nt, 1);
    for (int t = 0; t < nt; ++t) {
        size_t b = std::min(n, part * t), e = std::min(n, part * (t + 1));
        if (b >= e) {
            ok[t] = 1;
            continue;
        }
        ts.emplace_back([&, t, b, e] { ok[t] = f(b, e - b) ? 1 : 0; });
    }
    for (auto& th : ts) th.join();
    return std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; });
}

// two pinned chunks with the event of the last copy that used each
struct Staging {
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};

    hv_status ensure() {
        for (int i = 0; i < 2; ++i) {
            if (!buf[i]) HV_HIP(hipHostMalloc((void**)&buf[i], kChunk, hipHostMallocDefault));
            if (!ev[i]) HV_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
        }
        return HV_OK;
    }
    hv_status wait(int i) {
        if (pending[i]) {
            pending[i] = false;
            HV_HIP(hipEventSynchronize(ev[i]));
        }
        return HV_OK;
    }
    hv_status drain() {
        HV_TRY(wait(0));
        return wait(1);
    }
    void release() {
        for (int i = 0; i < 2; ++i) {
            if (pending[i]) (void)hipEventSynchronize(ev[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            buf[i] = nullptr;
            ev[i] = nullptr;
            pending[i] = false;
        }
    }
};

std::string errno_text() { return std::string(std::strerror(errno)); }

}  // namespace

struct hv_holo_reader {
    int fd = -1;
    std::string path;
    hv_holo_header h{};
    hv_holo_info info{};
    uint64_t file_size = 0;
    Staging st;
};

struct hv_holo_writer {
    int fd = -1;
    std::string path;
    hv_holo_header h{};
    uint64_t frame_size = 0;
    uint64_t pos = 0;  // current write position (the reference's FILE* cursor)
    Staging st;
};

using namespace hv;

// Frames of `total` bytes are about to be written at the cursor.  Large
// writes go through a shared mapping, which needs the file to cover them:
// the blocks are reserved first (fallocate fails cleanly on a full disk,
// where a write through a mapping of a sparse file would fault); file
// systems without fallocate take the pwrite path.
static hv_status reserve_frames(hv_holo_writer* w, uint64_t total, bool* use_map) {
    *use_map = false;
    if (!HV_HOLO_WRITE_MMAP || total < (size_t(HV_HOLO_MAP_MIN_MIB) << 20)) return HV_OK;
    if (::fallocate(w->fd, 0, (off_t)w->pos, (off_t)total) == 0) {
        *use_map = true;
        return HV_OK;
    }
    if (errno == ENOSPC) return fail(HV_ERR_IO, "Unable to write output holo file frame: no space left on the device");
    return HV_OK;
}

static bool write_region(int fd, const char* src, size_t len, uint64_t off, bool use_map) {
    if (use_map) return mmap_write(fd, src, len, off);
    return parallel_parts(len, [&](size_t b, size_t l) { return pwrite_full(fd, src + b, l, off + b); });
}

extern "C" {

hv_status hv_holo_make_header(uint32_t width, uint32_t height, uint32_t depth, uint64_t img_nb, int data_type,
                              hv_holo_header* out) {
    if (!out) return fail(HV_ERR_INVALID_ARG, "hv_holo_make_header: null header");
    if (depth == 0 || depth > 8) return fail(HV_ERR_INVALID_ARG, "hv_holo_make_header: depth must be 1..8 bytes");
    if (data_type < HV_HOLO_RAW || data_type > HV_HOLO_MOMENTS)
        return fail(HV_ERR_INVALID_ARG, "hv_holo_make_header: data_type must be RAW, PROCESSED or MOMENTS");
    std::memset(out, 0, sizeof(*out));
    std::memcpy(out->magic_number, "HOLO", 4);
    out->version = kCurrentVersion;
    out->bits_per_pixel = (uint16_t)(depth * 8);
    out->img_width = width;
    out->img_height = height;
    out->img_nb = (uint32_t)img_nb;
    out->total_data_size = (uint64_t)width * height * depth * img_nb;
    out->endianness = 0;  // little endian
    out->data_type = (uint8_t)data_type;
    return HV_OK;
}

hv_status hv_holo_open(const char* path, hv_holo_reader** out) {
    if (!path || !out) return fail(HV_ERR_INVALID_ARG, "hv_holo_open: null argument");
    *out = nullptr;
    int fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) return fail(HV_ERR_IO, std::string("hv_holo_open: cannot open ") + path + ": " + errno_text());
    struct stat sb;
    if (::fstat(fd, &sb) != 0) {
        ::close(fd);
        return fail(HV_ERR_IO, std::string("hv_holo_open: cannot stat ") + path);
    }
    auto* r = new hv_holo_reader;
    r->fd = fd;
    r->path = path;
    r->file_size = (uint64_t)sb.st_size;
    long long got = pread_full(fd, (char*)&r->h, sizeof(hv_holo_header), 0);
    if (got < 0) {
        delete r;
        ::close(fd);
        return fail(HV_ERR_IO, "An error was encountered while reading the file");
    }
    // input_holo_file.cc:67-71
    if (got != (long long)sizeof(hv_holo_header) || std::memcmp(r->h.magic_number, "HOLO", 4) != 0) {
        delete r;
        ::close(fd);
        return fail(HV_ERR_FORMAT, "Invalid holo file: invalid bytecode size");
    }
    hv_holo_info& in = r->info;
    in.version = r->h.version;
    in.bits_per_pixel = r->h.bits_per_pixel;
    in.width = r->h.img_width;
    in.height = r->h.img_height;
    in.depth = r->h.bits_per_pixel / 8u;  // input_holo_file.cc:87
    in.img_nb = r->h.img_nb;
    in.endianness = r->h.endianness ? 1 : 0;
    // the data type is in the header from version 7 on; RAW before (input_holo_file.cc:73-77)
    in.data_type = r->h.version == 7 ? r->h.data_type : HV_HOLO_RAW;
    in.frame_size = (uint64_t)in.width * in.height * in.depth;
    in.total_data_size = r->h.total_data_size;
    // checksum of the header (input_holo_file.cc:38-42)
    if (r->h.total_data_size != in.frame_size * in.img_nb) {
        delete r;
        ::close(fd);
        return fail(HV_ERR_FORMAT, "Invalid holo file: invalid total data size");
    }
    uint64_t data_end = sizeof(hv_holo_header) + r->h.total_data_size;
    in.footer_offset = data_end;
    in.footer_size = r->file_size > data_end ? r->file_size - data_end : 0;
    // frames actually present: a truncated recording keeps its whole frames
    uint64_t avail = r->file_size > sizeof(hv_holo_header) ? r->file_size - sizeof(hv_holo_header) : 0;
    in.frames_in_file = in.frame_size ? std::min<uint64_t>(in.img_nb, avail / in.frame_size) : 0;
    in.truncated = data_end > r->file_size ? 1 : 0;
    *out = r;
    return HV_OK;
}

hv_status hv_holo_get_info(const hv_holo_reader* r, hv_holo_info* info) {
    if (!r || !info) return fail(HV_ERR_INVALID_ARG, "hv_holo_get_info: null argument");
    *info = r->info;
    return HV_OK;
}

hv_status hv_holo_get_header(const hv_holo_reader* r, hv_holo_header* h) {
    if (!r || !h) return fail(HV_ERR_INVALID_ARG, "hv_holo_get_header: null argument");
    *h = r->h;
    return HV_OK;
}

static hv_status clamp_frames(const hv_holo_reader* r, uint64_t first, uint64_t count, uint64_t* n) {
    if (first > r->info.frames_in_file)
        return fail(HV_ERR_INVALID_ARG, "Unable to seek the frame requested: frame " + std::to_string(first) +
                                            " of " + std::to_string(r->info.frames_in_file));
    *n = std::min<uint64_t>(count, r->info.frames_in_file - first);
    return HV_OK;
}

hv_status hv_holo_read_frames(hv_holo_reader* r, uint64_t first, uint64_t count, void* host_dst,
                              uint64_t* frames_read) {
    if (!r || (!host_dst && count)) return fail(HV_ERR_INVALID_ARG, "hv_holo_read_frames: null argument");
    uint64_t n = 0;
    HV_TRY(clamp_frames(r, first, count, &n));
    const uint64_t fs = r->info.frame_size, off = sizeof(hv_holo_header) + first * fs;
    char* dst = (char*)host_dst;
    bool ok = parallel_parts(n * fs, [&](size_t b, size_t len) {
        return pread_full(r->fd, dst + b, len, off + b) == (long long)len;
    });
    if (!ok) return fail(HV_ERR_IO, "Unable to read " + std::to_string(n) + " frames");
    if (frames_read) *frames_read = n;
    return HV_OK;
}

hv_status hv_holo_load_frames(hv_holo_reader* r, uint64_t first, uint64_t count, void* dev_dst, void* stream,
                              uint64_t* frames_read) {
    if (!r || (!dev_dst && count)) return fail(HV_ERR_INVALID_ARG, "hv_holo_load_frames: null argument");
    uint64_t n = 0;
    HV_TRY(clamp_frames(r, first, count, &n));
    HV_TRY(r->st.ensure());
    hipStream_t s = (hipStream_t)stream;
    const uint64_t total = n * r->info.frame_size, off = sizeof(hv_holo_header) + first * r->info.frame_size;
    char* dst = (char*)dev_dst;
    int b = 0;
    for (uint64_t done = 0; done < total; done += kChunk, b ^= 1) {
        const size_t len = (size_t)std::min<uint64_t>(kChunk, total - done);
        HV_TRY(r->st.wait(b));  // the copy that last used this buffer has finished
        char* stage = r->st.buf[b];
        bool ok = parallel_parts(len, [&](size_t pb, size_t pl) {
            return pread_full(r->fd, stage + pb, pl, off + done + pb) == (long long)pl;
        });
        if (!ok) {
            (void)r->st.drain();
            return fail(HV_ERR_IO, "Unable to read " + std::to_string(n) + " frames");
        }
        HV_HIP(hipMemcpyAsync(dst + done, stage, len, hipMemcpyHostToDevice, s));
        HV_HIP(hipEventRecord(r->st.ev[b], s));
        r->st.pending[b] = true;
    }
    HV_TRY(r->st.drain());
    if (frames_read) *frames_read = n;
    return HV_OK;
}

hv_status hv_holo_read_footer(hv_holo_reader* r, char* dst, size_t capacity, size_t* size) {
    if (!r || !size) return fail(HV_ERR_INVALID_ARG, "hv_holo_read_footer: null argument");
    // input_holo_file.cc:100-107: the footer offset lies past the end of a truncated file
    if (r->info.truncated) return fail(HV_ERR_FORMAT, "Invalid holo file: invalid footer");
    *size = (size_t)r->info.footer_size;
    if (!dst) return HV_OK;  // size query
    if (capacity < r->info.footer_size)
        return fail(HV_ERR_INVALID_ARG, "hv_holo_read_footer: buffer of " + std::to_string(capacity) +
                                            " bytes for a footer of " + std::to_string(r->info.footer_size));
    if (pread_full(r->fd, dst, *size, r->info.footer_offset) != (long long)*size)
        return fail(HV_ERR_IO, "An error occurred while retrieving the meta data");
    return HV_OK;
}

hv_status hv_holo_close(hv_holo_reader* r) {
    if (!r) return HV_OK;
    r->st.release();
    if (r->fd >= 0) ::close(r->fd);
    delete r;
    return HV_OK;
}

hv_status hv_holo_create(const char* path, uint32_t width, uint32_t height, uint32_t depth, uint64_t img_nb,
                         int data_type, hv_holo_writer** out) {
    if (!path || !out) return fail(HV_ERR_INVALID_ARG, "hv_holo_create: null argument");
    *out = nullptr;
    hv_holo_header h;
    HV_TRY(hv_holo_make_header(width, height, depth, img_nb, data_type, &h));
    int fd = ::open(path, O_RDWR | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (fd < 0) return fail(HV_ERR_IO, std::string("hv_holo_create: cannot open ") + path + ": " + errno_text());
    auto* w = new hv_holo_writer;
    w->fd = fd;
    w->path = path;
    w->h = h;
    w->frame_size = (uint64_t)width * height * depth;
    *out = w;
    return HV_OK;
}

hv_status hv_holo_write_header(hv_holo_writer* w) {
    if (!w) return fail(HV_ERR_INVALID_ARG, "hv_holo_write_header: null writer");
    if (!pwrite_full(w->fd, (const char*)&w->h, sizeof(hv_holo_header), w->pos))
        return fail(HV_ERR_IO, "Unable to write output holo file header");
    w->pos += sizeof(hv_holo_header);
    return HV_OK;
}

hv_status hv_holo_write_frames(hv_holo_writer* w, const void* host_src, uint64_t count) {
    if (!w || (!host_src && count)) return fail(HV_ERR_INVALID_ARG, "hv_holo_write_frames: null argument");
    const uint64_t total = count * w->frame_size;
    bool use_map = false;
    HV_TRY(reserve_frames(w, total, &use_map));
    if (!write_region(w->fd, (const char*)host_src, total, w->pos, use_map))
        return fail(HV_ERR_IO, "Unable to write output holo file frame");
    w->pos += total;
    return HV_OK;
}

hv_status hv_holo_store_frames(hv_holo_writer* w, const void* dev_src, uint64_t count, void* stream) {
    if (!w || (!dev_src && count)) return fail(HV_ERR_INVALID_ARG, "hv_holo_store_frames: null argument");
    HV_TRY(w->st.ensure());
    hipStream_t s = (hipStream_t)stream;
    const uint64_t total = count * w->frame_size, pos = w->pos;
    const char* src = (const char*)dev_src;
    // chunk k is copied into buffer k&1 while chunk k-1 is written to the file
    auto issue = [&](uint64_t k) -> hv_status {
        const uint64_t at = k * kChunk;
        const size_t len = (size_t)std::min<uint64_t>(kChunk, total - at);
        const int b = (int)(k & 1);
        HV_HIP(hipMemcpyAsync(w->st.buf[b], src + at, len, hipMemcpyDeviceToHost, s));
        HV_HIP(hipEventRecord(w->st.ev[b], s));
        w->st.pending[b] = true;
        return HV_OK;
    };
    const uint64_t chunks = (total + kChunk - 1) / kChunk;
    bool use_map = false;
    HV_TRY(reserve_frames(w, total, &use_map));
    if (chunks) HV_TRY(issue(0));
    for (uint64_t k = 0; k < chunks; ++k) {
        const int b = (int)(k & 1);
        HV_TRY(w->st.wait(b));
        if (k + 1 < chunks) HV_TRY(issue(k + 1));
        const uint64_t at = k * kChunk;
        const size_t len = (size_t)std::min<uint64_t>(kChunk, total - at);
        if (!write_region(w->fd, w->st.buf[b], len, pos + at, use_map)) {
            (void)w->st.drain();
            return fail(HV_ERR_IO, "Unable to write output holo file frame");
        }
    }
    w->pos += total;
    return HV_OK;
}

hv_status hv_holo_write_footer(hv_holo_writer* w, const char* json, size_t size) {
    if (!w || (!json && size)) return fail(HV_ERR_INVALID_ARG, "hv_holo_write_footer: null argument");
    if (!pwrite_full(w->fd, json, size, w->pos)) return fail(HV_ERR_IO, "Unable to write output holo file footer");
    w->pos += size;
    return HV_OK;
}

hv_status hv_holo_correct_number_of_frames(hv_holo_writer* w, uint64_t frames_written) {
    if (!w) return fail(HV_ERR_INVALID_ARG, "hv_holo_correct_number_of_frames: null writer");
    // output_holo_file.cc:104-122: rewrite the header in place, keep the cursor
    w->h.img_nb = (uint32_t)frames_written;
    w->h.total_data_size = w->frame_size * frames_written;
    if (!pwrite_full(w->fd, (const char*)&w->h, sizeof(hv_holo_header), 0))
        return fail(HV_ERR_IO, "Unable to correct number of written frames");
    return HV_OK;
}

hv_status hv_holo_writer_close(hv_holo_writer* w) {
    if (!w) return HV_OK;
    w->st.release();
    hv_status s = HV_OK;
    if (w->fd >= 0 && ::close(w->fd) != 0) s = fail(HV_ERR_IO, "hv_holo_writer_close: close failed");
    delete w;
    return s;
}

}  // extern "C"
