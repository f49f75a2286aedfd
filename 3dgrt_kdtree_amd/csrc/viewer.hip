// Frame loop of the interactive viewer (Kd-treeConverterMain.cpp:226-251,
// 3514-3608): the reference renders each frame on a compute stream into a
// display buffer that OpenGL owns and maps it on a transfer stream.  An MI355X
// has no display, so frames go to the host through a ring of slots:
//
//  * a slot is a device framebuffer (float3 per pixel, the render's output),
//    an optional 8-bit device frame, and a pinned host frame;
//  * submit renders into a free slot on the compute stream (kdgrt_render, so
//    the scene's tile schedule orders the tiles of a close view from the
//    previous frame's durations), converts to 8 bits there if asked, and
//    copies the frame to the slot's host memory on the transfer stream after
//    an event: the copy of frame k runs while frame k+1 renders;
//  * acquire hands out finished frames in submission order; a slot returns to
//    the ring at release, so a frame the consumer holds is never overwritten.
//
// No allocation happens after create.  Streams run at the default priority.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "runtime.hpp"

using kdgrt::set_error;

namespace {

enum SlotState { kFree = 0, kInFlight = 1, kHeld = 2 };

struct Slot {
    void* d_fb = nullptr;      // float3 framebuffer
    uint8_t* d_u8 = nullptr;   // 8-bit frame (8-bit formats)
    void* h_frame = nullptr;   // pinned host frame
    hipEvent_t ev_start = nullptr, ev_end = nullptr; // render timing
    hipEvent_t ev_ready = nullptr;                   // host copy finished
    hipEvent_t ev_rendered = nullptr;                // device frame complete
    uint64_t frame_id = 0;
    int state = kFree;
    kdgrt_camera cam{};
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define VHIP(expr)                                                                                              \
    do {                                                                                                        \
        hipError_t e_ = (expr);                                                                                 \
        if (e_ != hipSuccess)                                                                                   \
            return set_error(e_ == hipErrorOutOfMemory ? KDGRT_ERR_OUT_OF_MEMORY : KDGRT_ERR_DEVICE,            \
                             std::string("viewer: ") + #expr + ": " + hipGetErrorString(e_));                    \
    } while (0)

} // namespace

struct kdgrt_viewer {
    kdgrt_context* ctx = nullptr;
    const kdgrt_scene* scene = nullptr;
    int device = 0;
    int W = 0, H = 0, format = 0, depth = 12;
    int window_start = 0, window_end = 0;
    uint64_t frame_bytes = 0;
    hipStream_t compute = nullptr, transfer = nullptr;
    std::vector<Slot> slots;
    std::deque<int> fifo;      // in-flight slots in submission order
    uint64_t next_id = 0;
    // stats
    uint64_t delivered = 0, window_frames = 0;
    double sum_ms = 0.0, window_fps_sum = 0.0, t_first_submit = -1.0, t_last_delivery = 0.0;
    std::mutex mu;             // slot states, fifo, stats
    std::condition_variable freed;
    std::mutex submit_mu;      // one submit at a time: the enqueue order is the frame order
    std::mutex acquire_mu;     // one acquire at a time: frames leave in order
};

namespace {

void viewer_free(kdgrt_viewer* v) {
    if (v->device >= 0) hipSetDevice(v->device);
    if (v->compute) hipStreamSynchronize(v->compute);
    if (v->transfer) hipStreamSynchronize(v->transfer);
    for (Slot& s : v->slots) {
        if (s.d_fb) hipFree(s.d_fb);
        if (s.d_u8) hipFree(s.d_u8);
        if (s.h_frame) hipHostFree(s.h_frame);
        for (hipEvent_t e : {s.ev_start, s.ev_end, s.ev_ready, s.ev_rendered})
            if (e) hipEventDestroy(e);
    }
    if (v->compute) hipStreamDestroy(v->compute);
    if (v->transfer) hipStreamDestroy(v->transfer);
    delete v;
}

int format_channels(int format) { return format == KDGRT_FRAME_RGBA8 ? 4 : 3; }

} // namespace

extern "C" {

kdgrt_status kdgrt_viewer_create(kdgrt_context* ctx, const kdgrt_scene* scene, int width, int height, int format,
                                 int n_slots, int stack_depth, int window_start, int window_end,
                                 kdgrt_viewer** out) {
    if (!ctx || !scene || !out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_create: null argument");
    *out = nullptr;
    if (width <= 0 || height <= 0 || (uint64_t)width * height > (1ull << 28))
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_create: bad frame size");
    if (format != KDGRT_FRAME_RGB_F32 && format != KDGRT_FRAME_RGBA8 && format != KDGRT_FRAME_RGB8)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_create: unknown frame format");
    if (n_slots < 2 || n_slots > 8) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_create: n_slots must be 2..8");
    if (stack_depth < 1 || stack_depth > 32)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_create: stack_depth must be 1..32");
    if (window_start < 0 || window_end < window_start)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_create: bad measurement window");
    auto* v = new kdgrt_viewer();
    v->ctx = ctx;
    v->scene = scene;
    v->device = kdgrt::context_device(ctx);
    v->W = width;
    v->H = height;
    v->format = format;
    v->depth = stack_depth;
    v->window_start = window_start;
    v->window_end = window_end;
    const uint64_t px = (uint64_t)width * height;
    v->frame_bytes = format == KDGRT_FRAME_RGB_F32 ? 12 * px : (uint64_t)format_channels(format) * px;
    v->slots.resize(n_slots);
    auto fail = [&](hipError_t e, const char* what) {
        viewer_free(v);
        return set_error(e == hipErrorOutOfMemory ? KDGRT_ERR_OUT_OF_MEMORY : KDGRT_ERR_DEVICE,
                         std::string("viewer_create: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipSetDevice(v->device)) != hipSuccess) return fail(e, "device");
    if ((e = hipStreamCreateWithFlags(&v->compute, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    if ((e = hipStreamCreateWithFlags(&v->transfer, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    for (Slot& s : v->slots) {
        if ((e = hipMalloc(&s.d_fb, 12 * px)) != hipSuccess) return fail(e, "framebuffer");
        if (format != KDGRT_FRAME_RGB_F32 && (e = hipMalloc(&s.d_u8, v->frame_bytes)) != hipSuccess)
            return fail(e, "8-bit frame");
        if ((e = hipHostMalloc(&s.h_frame, v->frame_bytes, hipHostMallocDefault)) != hipSuccess)
            return fail(e, "pinned host frame");
        if ((e = hipEventCreate(&s.ev_start)) != hipSuccess || (e = hipEventCreate(&s.ev_end)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&s.ev_ready, hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&s.ev_rendered, hipEventDisableTiming)) != hipSuccess)
            return fail(e, "events");
    }
    *out = v;
    return KDGRT_OK;
}

kdgrt_status kdgrt_viewer_submit(kdgrt_viewer* v, const kdgrt_camera* cam, int timeout_ms, uint64_t* frame_id) {
    if (!v || !cam) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_submit: null argument");
    kdgrt_camera_info ci;
    kdgrt_status st = kdgrt_camera_make(cam, v->W, v->H, &ci);
    if (st != KDGRT_OK) return st;
    std::lock_guard<std::mutex> sub(v->submit_mu);
    int k = -1;
    uint64_t id = 0;
    {
        std::unique_lock<std::mutex> lk(v->mu);
        auto find_free = [&] {
            for (int i = 0; i < (int)v->slots.size(); ++i)
                if (v->slots[i].state == kFree) return i;
            return -1;
        };
        k = find_free();
        if (k < 0 && timeout_ms != 0) {
            auto pred = [&] { return (k = find_free()) >= 0; };
            if (timeout_ms < 0) v->freed.wait(lk, pred);
            else v->freed.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
        }
        if (k < 0) return set_error(KDGRT_ERR_NOT_READY, "viewer_submit: every slot is held");
        id = v->next_id++;
        Slot& s = v->slots[k];
        s.state = kInFlight;
        s.frame_id = id;
        s.cam = *cam;
        if (v->t_first_submit < 0) v->t_first_submit = now_s();
    }
    Slot& s = v->slots[k];
    auto undo = [&](kdgrt_status err) {
        // the frame never reaches the fifo: its slot is free again and its id
        // is skipped (acquire follows the fifo, not the ids)
        std::lock_guard<std::mutex> lk(v->mu);
        s.state = kFree;
        v->freed.notify_all();
        return err;
    };
    if (hipSetDevice(v->device) != hipSuccess) return undo(set_error(KDGRT_ERR_DEVICE, "viewer_submit: device"));
    if (hipEventRecord(s.ev_start, v->compute) != hipSuccess)
        return undo(set_error(KDGRT_ERR_DEVICE, "viewer_submit: event record"));
    st = kdgrt_render(v->ctx, v->scene, &ci, v->W, v->H, v->depth, KDGRT_OUT_RGB_F32, s.d_fb, v->compute);
    if (st != KDGRT_OK) return undo(st);
    if (hipEventRecord(s.ev_end, v->compute) != hipSuccess)
        return undo(set_error(KDGRT_ERR_DEVICE, "viewer_submit: event record"));
    const void* src = s.d_fb;
    if (v->format != KDGRT_FRAME_RGB_F32) {
        st = kdgrt_framebuffer_to_u8(v->ctx, (const float*)s.d_fb, v->W, v->H, 3, s.d_u8, format_channels(v->format),
                                     1, v->compute);
        if (st != KDGRT_OK) return undo(st);
        src = s.d_u8;
    }
    hipError_t e = hipEventRecord(s.ev_rendered, v->compute);
    if (e == hipSuccess) e = hipStreamWaitEvent(v->transfer, s.ev_rendered, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(s.h_frame, src, v->frame_bytes, hipMemcpyDeviceToHost, v->transfer);
    if (e == hipSuccess) e = hipEventRecord(s.ev_ready, v->transfer);
    if (e != hipSuccess)
        return undo(set_error(KDGRT_ERR_DEVICE, std::string("viewer_submit: frame copy: ") + hipGetErrorString(e)));
    {
        std::lock_guard<std::mutex> lk(v->mu);
        v->fifo.push_back(k);
    }
    if (frame_id) *frame_id = id;
    return KDGRT_OK;
}

kdgrt_status kdgrt_viewer_acquire(kdgrt_viewer* v, int timeout_ms, kdgrt_viewer_frame* out) {
    if (!v || !out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_acquire: null argument");
    std::lock_guard<std::mutex> acq(v->acquire_mu);
    int k;
    {
        std::lock_guard<std::mutex> lk(v->mu);
        if (v->fifo.empty()) return set_error(KDGRT_ERR_NOT_READY, "viewer_acquire: no frame pending");
        k = v->fifo.front();
    }
    Slot& s = v->slots[k];
    if (hipSetDevice(v->device) != hipSuccess) return set_error(KDGRT_ERR_DEVICE, "viewer_acquire: device");
    if (timeout_ms < 0) {
        VHIP(hipEventSynchronize(s.ev_ready));
    } else {
        const double deadline = now_s() + timeout_ms * 1e-3;
        for (;;) {
            const hipError_t q = hipEventQuery(s.ev_ready);
            if (q == hipSuccess) break;
            if (q != hipErrorNotReady)
                return set_error(KDGRT_ERR_DEVICE, std::string("viewer_acquire: ") + hipGetErrorString(q));
            if (now_s() >= deadline) return set_error(KDGRT_ERR_NOT_READY, "viewer_acquire: frame not finished");
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
    float ms = 0.0f;
    VHIP(hipEventElapsedTime(&ms, s.ev_start, s.ev_end));
    std::lock_guard<std::mutex> lk(v->mu);
    v->fifo.pop_front();
    s.state = kHeld;
    v->delivered++;
    v->sum_ms += ms;
    v->t_last_delivery = now_s();
    const float fps = ms > 0.0f ? 1000.0f / ms : 0.0f;
    if (s.frame_id >= (uint64_t)v->window_start && s.frame_id < (uint64_t)v->window_end) {
        v->window_frames++;
        v->window_fps_sum += fps;
    }
    out->frame_id = s.frame_id;
    out->pixels = s.h_frame;
    out->bytes = v->frame_bytes;
    out->width = v->W;
    out->height = v->H;
    out->format = v->format;
    out->render_ms = ms;
    out->fps = fps;
    out->camera = s.cam;
    return KDGRT_OK;
}

kdgrt_status kdgrt_viewer_release(kdgrt_viewer* v, uint64_t frame_id) {
    if (!v) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_release: null viewer");
    std::lock_guard<std::mutex> lk(v->mu);
    for (Slot& s : v->slots)
        if (s.state == kHeld && s.frame_id == frame_id) {
            s.state = kFree;
            v->freed.notify_all();
            return KDGRT_OK;
        }
    return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_release: frame " + std::to_string(frame_id) + " is not held");
}

kdgrt_status kdgrt_viewer_stats_get(const kdgrt_viewer* cv, kdgrt_viewer_stats* out) {
    if (!cv || !out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_stats_get: null argument");
    auto* v = const_cast<kdgrt_viewer*>(cv);
    std::lock_guard<std::mutex> lk(v->mu);
    out->submitted = v->next_id;
    out->delivered = v->delivered;
    out->mean_render_ms = v->delivered ? v->sum_ms / (double)v->delivered : 0.0;
    out->window_frames = v->window_frames;
    out->window_avg_fps = v->window_frames ? v->window_fps_sum / (double)v->window_frames : 0.0;
    const double span = v->t_last_delivery - v->t_first_submit;
    out->wall_fps = (v->delivered && span > 0.0) ? (double)v->delivered / span : 0.0;
    return KDGRT_OK;
}

kdgrt_status kdgrt_viewer_destroy(kdgrt_viewer* v) {
    if (!v) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "viewer_destroy: null viewer");
    viewer_free(v);
    return KDGRT_OK;
}

} // extern "C"
