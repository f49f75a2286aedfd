// Serving loop: resident scenes, HBM-budget admission, continuous batching
// into the persistent batched kernel.
//
//  * Scenes are made resident before start; their device bytes count
//    against the budget.  The rest of the budget is one output arena,
//    allocated once at start; each admitted request takes W*H*16 bytes of it
//    (first fit) until the caller collects the result.  A request larger
//    than the whole arena is rejected at submit.
//  * One scheduler thread admits pending requests in arrival order while
//    arena memory and launch slots are free, and launches them as one
//    persistent batch on its slot's stream.  Several batches are in flight at
//    once on separate streams, so requests that arrive while others run start
//    as soon as a slot and memory free up and their work overlaps on the
//    device.  While the device is busy the scheduler waits up to
//    batch_window_ms to fill a batch; an idle device launches at once.
//  * No allocation and no implicit device synchronisation happen on the
//    launch path: slots (work table, counters, pinned staging) and streams
//    are created at start.
//  * A batch that holds a single view runs the one-view kernel (k_view with
//    its cone pre-test) on the slot's stream; larger batches run the
//    persistent kernel.
//  * persistent_kernel = 1: one long-running persistent kernel (k_serve) per
//    epoch takes every request; the scheduler publishes each admitted request
//    into the kernel's queue while the kernel runs, so requests join the
//    running kernel's tile stream as they arrive (continuous batching into
//    the persistent kernel).  An idle epoch is stopped after epoch_idle_ms
//    and the next request starts a new one; requests an epoch leaves
//    unfinished (its kernel ended early) are published again in the next.
//    With idle_plain = 1, a request that finds the device idle (no epoch, no
//    other request) runs as one plain one-view launch instead, and no epoch
//    starts until that launch has finished.
//  * plain_launches = 1 is the comparison mode: each request is one plain
//    one-view launch (k_view), run one at a time, as the reference renders.
//  * With ctx == NULL and simulate_mrays_per_s > 0 the same scheduler runs
//    against a modelled device (batches take rays / rate seconds, one at a
//    time) so admission, batching, deadlines and accounting can be tested
//    without a GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "runtime.hpp"

using kdgrt::set_error;

namespace {

// Tile-schedule buffers reserved per scene: views up to 4096 x 4096 pixels
// (3 MB per scene) follow the tile order; larger ones run in raster order.
constexpr int kServeMaxEdge = 4096;

enum ReqState { kPending = 0, kAdmitted = 1, kDone = 2, kCancelled = 3, kFailed = 4 };

struct Req {
    uint64_t id = 0;
    kdgrt_request r{};
    int state = kPending;
    int status = KDGRT_OK;
    double arrival = 0, admitted = 0, launched = 0, done = 0;
    uint64_t bytes = 0, off = 0;
    bool has_mem = false;
    bool collecting = false; // a kdgrt_server_wait is copying the result out
    uint32_t batch_id = 0, batch_views = 0;
};

struct Flight {
    int slot = -1;
    uint32_t id = 0;
    std::vector<Req*> reqs;
    double launched = 0, sim_done = 0;
    uint64_t rays = 0;
};

// First-fit allocator over [0, size) with coalescing, 256-B granules.
class Arena {
  public:
    void reset(uint64_t size) {
        free_.clear();
        size_ = size;
        if (size) free_[0] = size;
        used_ = 0;
    }
    bool alloc(uint64_t bytes, uint64_t* off) {
        bytes = (bytes + 255) & ~uint64_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const uint64_t rest = it->second - bytes, at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = (bytes + 255) & ~uint64_t(255);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }

  private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t size_ = 0, used_ = 0;
};

} // namespace

struct kdgrt_server {
    kdgrt_context* ctx = nullptr;
    kdgrt_server_config cfg{};
    bool sim = false;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<int32_t, std::pair<kdgrt_scene*, uint64_t>> scenes;
    std::map<int32_t, bool> scene_owned; // false: a caller's resident scene, not destroyed here
    int32_t next_scene = 0;
    uint64_t scene_bytes = 0;

    std::deque<Req*> pending;
    std::unordered_map<uint64_t, Req*> reqs;
    uint64_t next_id = 1;
    uint32_t next_batch = 1;

    Arena arena;
    void* d_arena = nullptr;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    std::vector<kdgrt::BatchSlot*> slots;
    std::vector<bool> slot_busy;
    std::vector<Flight> flights;
    hipStream_t copy_stream = nullptr;
    std::vector<hipStream_t> copy_pool; // result-copy streams of kdgrt_server_wait (copy_stream among them)
    double sim_gpu_free = 0;

    // persistent mode
    kdgrt::ServeQueue* sq = nullptr;
    hipStream_t serve_stream = nullptr;
    std::vector<std::pair<Req*, uint32_t>> serving; // published in the running epoch: (request, sequence)
    std::map<int32_t, int> scene_index;             // scene id -> index in the kernel's scene table
    bool epoch_running = false, epoch_stopping = false;
    double last_busy = 0;

    std::thread worker;
    bool started = false, stopping = false;
    kdgrt_server_stats st{};
    int busy_count = 0;
    double busy_since = 0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Completes flights whose work has finished.  Caller holds the lock.
void retire(kdgrt_server* s, bool* any) {
    const double t = s->now();
    for (size_t i = 0; i < s->flights.size();) {
        Flight& f = s->flights[i];
        bool finished;
        int status = KDGRT_OK;
        if (s->sim) {
            finished = t >= f.sim_done;
        } else {
            const hipError_t e = hipEventQuery(s->events[f.slot]);
            finished = e != hipErrorNotReady;
            if (finished && e != hipSuccess) status = KDGRT_ERR_DEVICE;
        }
        if (!finished) {
            ++i;
            continue;
        }
        const double done = s->sim ? f.sim_done : t;
        for (Req* r : f.reqs) {
            r->done = done;
            r->state = status == KDGRT_OK ? kDone : kFailed;
            r->status = status;
            s->st.completed += status == KDGRT_OK;
        }
        s->st.rays += f.rays;
        s->slot_busy[f.slot] = false;
        if (--s->busy_count == 0) s->st.busy_s += done - s->busy_since;
        s->flights.erase(s->flights.begin() + i);
        *any = true;
    }
}

int free_slot(kdgrt_server* s) {
    for (size_t i = 0; i < s->slot_busy.size(); ++i)
        if (!s->slot_busy[i]) return (int)i;
    return -1;
}

// Admits and launches one batch if possible.  Caller holds the lock.
bool launch_one(kdgrt_server* s) {
    const double t = s->now();
    // drop requests whose deadline passed before they started
    for (auto it = s->pending.begin(); it != s->pending.end();) {
        Req* r = *it;
        if (r->r.deadline_s > 0 && t > r->r.deadline_s) {
            r->state = kCancelled;
            r->status = KDGRT_ERR_CANCELLED;
            r->done = t;
            s->st.cancelled++;
            it = s->pending.erase(it);
        } else {
            ++it;
        }
    }
    if (s->pending.empty()) return false;
    if (s->cfg.plain_launches && s->busy_count > 0) return false; // one plain launch at a time
    const int slot = free_slot(s);
    if (slot < 0) return false;
    const int maxv = s->cfg.plain_launches ? 1 : std::max(1, s->cfg.max_batch_views);
    // While the device is busy, give a batch up to batch_window_ms to fill.
    if (s->busy_count > 0 && (int)s->pending.size() < maxv &&
        t - s->pending.front()->arrival < s->cfg.batch_window_ms * 1e-3)
        return false;
    std::vector<Req*> take;
    while (!s->pending.empty() && (int)take.size() < maxv) {
        Req* r = s->pending.front();
        if (!r->has_mem) {
            uint64_t off = 0;
            if (!s->arena.alloc(r->bytes, &off)) break; // arrival order: no overtaking
            r->off = off;
            r->has_mem = true;
        }
        take.push_back(r);
        s->pending.pop_front();
    }
    if (take.empty()) return false;
    Flight f;
    f.slot = slot;
    f.id = s->next_batch++;
    f.launched = t;
    f.reqs = take;
    for (Req* r : take) {
        r->state = kAdmitted;
        r->admitted = t;
        r->launched = t;
        r->batch_id = f.id;
        r->batch_views = (uint32_t)take.size();
        f.rays += (uint64_t)r->r.width * r->r.height;
    }
    kdgrt_status stt = KDGRT_OK;
    if (s->sim) {
        const double start = std::max(t, s->sim_gpu_free);
        f.sim_done = start + (double)f.rays / (s->cfg.simulate_mrays_per_s * 1e6);
        s->sim_gpu_free = f.sim_done;
    } else {
        std::vector<kdgrt_view> views(take.size());
        for (size_t i = 0; i < take.size(); ++i) {
            Req* r = take[i];
            kdgrt_view& v = views[i];
            v.scene = s->scenes[r->r.scene_id].first;
            v.cam = r->r.cam;
            v.width = r->r.width;
            v.height = r->r.height;
            v.out_rgbt = (float*)((char*)s->d_arena + r->off);
        }
        hipStream_t st = s->streams[slot];
        if (s->cfg.plain_launches || views.size() == 1) {
            // comparison mode, or a batch of one view: the one-view kernel
            // (its per-view cone pre-test makes it the faster path for a
            // single view); it still runs on the slot's stream and overlaps
            // with the other slots.  The tile order uses the schedule buffers
            // reserved when the scene was added, so nothing is allocated here.
            stt = kdgrt::render_served(s->ctx, views[0].scene, &views[0].cam, views[0].width, views[0].height,
                                       s->cfg.tuning.stack_depth, views[0].out_rgbt, st);
        } else {
            uint64_t rays = 0;
            stt = kdgrt::slot_launch(s->slots[slot], views.data(), (int)views.size(), s->cfg.tuning, st, &rays);
        }
        if (stt == KDGRT_OK && hipEventRecord(s->events[slot], st) != hipSuccess) stt = KDGRT_ERR_DEVICE;
    }
    if (stt != KDGRT_OK) {
        for (Req* r : take) {
            r->state = kFailed;
            r->status = stt;
            r->done = t;
        }
        return true;
    }
    s->slot_busy[slot] = true;
    if (s->busy_count++ == 0) s->busy_since = t;
    s->flights.push_back(std::move(f));
    s->st.batches++;
    s->st.bytes_peak = std::max<uint64_t>(s->st.bytes_peak, s->scene_bytes + s->arena.used());
    return true;
}

// Persistent mode: completions, epoch ends, admission into the running
// kernel, idle stop.  Caller holds the lock.  Returns whether anything changed.
bool persist_step(kdgrt_server* s) {
    bool changed = false;
    const double t = s->now();
    retire(s, &changed); // plain launches of requests that found the device idle
    for (size_t i = 0; i < s->serving.size();) {
        Req* r = s->serving[i].first;
        if (!kdgrt::serve_done(s->sq, s->serving[i].second)) {
            ++i;
            continue;
        }
        r->done = t;
        r->state = kDone;
        r->status = KDGRT_OK;
        s->st.completed++;
        s->st.rays += (uint64_t)r->r.width * r->r.height;
        s->serving.erase(s->serving.begin() + i);
        changed = true;
    }
    if (s->epoch_running) {
        kdgrt_status est = KDGRT_OK;
        if (kdgrt::serve_epoch_finished(s->sq, &est)) {
            // the kernel exited: stopped by us, or early (its waits ran out);
            // what it left unfinished runs again in the next epoch, in order
            s->epoch_running = s->epoch_stopping = false;
            for (auto it = s->serving.rbegin(); it != s->serving.rend(); ++it) {
                Req* r = it->first;
                if (est != KDGRT_OK) {
                    r->state = kFailed;
                    r->status = est;
                    r->done = t;
                } else {
                    r->state = kPending;
                    s->pending.push_front(r);
                }
            }
            s->serving.clear();
            if (--s->busy_count == 0) s->st.busy_s += t - s->busy_since;
            changed = true;
        }
    }
    // requests whose deadline passed before they were published
    for (auto it = s->pending.begin(); it != s->pending.end();) {
        Req* r = *it;
        if (r->r.deadline_s > 0 && t > r->r.deadline_s) {
            r->state = kCancelled;
            r->status = KDGRT_ERR_CANCELLED;
            r->done = t;
            s->st.cancelled++;
            it = s->pending.erase(it);
            changed = true;
        } else {
            ++it;
        }
    }
    // idle_plain: a request that finds the device idle -- no epoch running,
    // no plain launch in flight, no other request waiting -- runs as one
    // plain one-view launch (cone pre-test, tile order) on a slot's stream.
    // No epoch starts while such a launch runs: the persistent kernel's
    // waves would hold the wave slots that the launch's remaining
    // workgroups wait for.  Requests that arrive meanwhile wait and then
    // start an epoch together (or run plain again if alone).
    if (s->cfg.idle_plain && !s->epoch_running && s->flights.empty() && s->pending.size() == 1 && launch_one(s))
        changed = true;
    while (!s->pending.empty() && !s->epoch_stopping && s->flights.empty()) {
        Req* r = s->pending.front();
        if (!r->has_mem) {
            uint64_t off = 0;
            if (!s->arena.alloc(r->bytes, &off)) break; // arrival order: no overtaking
            r->off = off;
            r->has_mem = true;
        }
        if (!s->epoch_running) {
            const kdgrt_status st = kdgrt::serve_epoch_start(s->sq);
            if (st != KDGRT_OK) {
                r->state = kFailed;
                r->status = st;
                r->done = t;
                s->pending.pop_front();
                changed = true;
                continue;
            }
            s->epoch_running = true;
            s->st.batches++;
            if (s->busy_count++ == 0) s->busy_since = t;
        }
        if (!kdgrt::serve_room(s->sq, r->r.width, r->r.height)) {
            kdgrt::serve_epoch_stop(s->sq); // the epoch's tables are full: drain, then a new epoch
            s->epoch_stopping = true;
            break;
        }
        uint32_t seq = 0;
        const kdgrt_status st = kdgrt::serve_publish(s->sq, s->scene_index[r->r.scene_id], r->r.cam, r->r.width,
                                                     r->r.height, r->off, &seq);
        if (st != KDGRT_OK) break;
        s->pending.pop_front();
        r->state = kAdmitted;
        r->admitted = r->launched = t;
        r->batch_id = (uint32_t)s->st.batches;
        s->serving.push_back({r, seq});
        r->batch_views = (uint32_t)s->serving.size(); // requests in the running kernel when it joined
        s->st.bytes_peak = std::max<uint64_t>(s->st.bytes_peak, s->scene_bytes + s->arena.used());
        changed = true;
    }
    if (!s->serving.empty() || !s->pending.empty()) s->last_busy = t;
    if (s->epoch_running && !s->epoch_stopping && s->serving.empty() && s->pending.empty() &&
        (s->stopping || t - s->last_busy > std::max(0.0, s->cfg.epoch_idle_ms) * 1e-3)) {
        kdgrt::serve_epoch_stop(s->sq);
        s->epoch_stopping = true;
    }
    return changed;
}

void worker_main_persistent(kdgrt_server* s) {
    hipSetDevice(kdgrt::context_device(s->ctx));
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        if (persist_step(s)) s->cv_done.notify_all();
        if (s->stopping && s->pending.empty() && s->serving.empty() && !s->epoch_running && s->flights.empty())
            break;
        // completions are flags in host memory (epochs) or events (plain
        // launches): poll them briskly while work runs
        const double wait_s = (!s->serving.empty() || s->epoch_stopping || !s->flights.empty())
                                  ? 10e-6
                                  : (s->epoch_running ? 200e-6 : -1.0);
        if (wait_s > 0) s->cv_work.wait_for(lk, std::chrono::duration<double>(wait_s));
        else s->cv_work.wait(lk);
    }
}

void worker_main(kdgrt_server* s) {
    if (!s->sim) hipSetDevice(kdgrt::context_device(s->ctx));
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool changed = false;
        retire(s, &changed);
        while (launch_one(s)) changed = true;
        if (changed) s->cv_done.notify_all();
        if (s->stopping && s->pending.empty() && s->flights.empty()) break;
        if (!s->flights.empty() || !s->pending.empty()) {
            // poll in-flight completion / the batching window
            double wait_s = 50e-6;
            if (s->sim && !s->flights.empty()) {
                double nxt = 1e30;
                for (auto& f : s->flights) nxt = std::min(nxt, f.sim_done);
                wait_s = std::max(1e-6, std::min(2e-3, nxt - s->now()));
            }
            s->cv_work.wait_for(lk, std::chrono::duration<double>(wait_s));
        } else {
            s->cv_work.wait(lk);
        }
    }
}

} // namespace

extern "C" {

kdgrt_status kdgrt_server_create(kdgrt_context* ctx, const kdgrt_server_config* cfg, kdgrt_server** out) {
    if (!cfg || !out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_create: null");
    const bool sim = ctx == nullptr;
    if (sim && !(cfg->simulate_mrays_per_s > 0))
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_create: no context and no simulated rate");
    if (cfg->max_batch_views < 1 || cfg->max_inflight_batches < 1 || cfg->max_inflight_batches > 16)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_create: batch views >= 1, in-flight batches 1..16");
    if (cfg->hbm_budget_bytes == 0) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_create: zero budget");
    auto* s = new kdgrt_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    s->sim = sim;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return KDGRT_OK;
}

kdgrt_status kdgrt_server_add_scene(kdgrt_server* s, const kdgrt_tree* tree, const kdgrt_gaussian* g, uint64_t n_g,
                                    int scales_inverted, int32_t* scene_id) {
    if (!s || !tree || !scene_id) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_add_scene: null");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_add_scene: server already started");
    // the scene's arrays plus, on a device, the tile-schedule buffers reserved for it below
    const uint64_t bytes = kdgrt_scene_bytes_estimate((uint32_t)tree->nodes.size(), tree->offsets.size(), n_g) +
                           (s->sim ? 0 : kdgrt::tile_schedule_bytes(kServeMaxEdge, kServeMaxEdge));
    if (s->scene_bytes + bytes > s->cfg.hbm_budget_bytes)
        return set_error(KDGRT_ERR_CAPACITY, "server_add_scene: scene does not fit the HBM budget");
    kdgrt_scene* sc = nullptr;
    if (!s->sim) {
        kdgrt_status st = kdgrt_scene_create_from_tree(s->ctx, tree, g, n_g, scales_inverted, &sc);
        if (st != KDGRT_OK) return st;
        st = kdgrt::reserve_tile_schedule(sc, kServeMaxEdge, kServeMaxEdge);
        if (st != KDGRT_OK) {
            kdgrt_scene_destroy(sc);
            return st;
        }
    } else {
        kdgrt_status st = kdgrt::validate_tree(tree->nodes.data(), (uint32_t)tree->nodes.size(),
                                               tree->offsets.size(), tree->offsets.data(), n_g);
        if (st != KDGRT_OK) return st;
    }
    const int32_t id = s->next_scene++;
    s->scenes[id] = {sc, bytes};
    s->scene_owned[id] = true;
    s->scene_bytes += bytes;
    s->st.bytes_resident = s->scene_bytes;
    *scene_id = id;
    return KDGRT_OK;
}

kdgrt_status kdgrt_server_add_resident_scene(kdgrt_server* s, kdgrt_scene* scene, int32_t* scene_id) {
    if (!s || !scene || !scene_id) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_add_resident_scene: null");
    if (s->sim) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_add_resident_scene: modelled device");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_add_resident_scene: server already started");
    uint64_t bytes = 0;
    kdgrt_status st = kdgrt_scene_device_bytes(scene, &bytes);
    if (st != KDGRT_OK) return st;
    bytes += kdgrt::tile_schedule_bytes(kServeMaxEdge, kServeMaxEdge); // reserved for the scene below
    if (s->scene_bytes + bytes > s->cfg.hbm_budget_bytes)
        return set_error(KDGRT_ERR_CAPACITY, "server_add_resident_scene: scene does not fit the HBM budget");
    st = kdgrt::reserve_tile_schedule(scene, kServeMaxEdge, kServeMaxEdge);
    if (st != KDGRT_OK) return st;
    const int32_t id = s->next_scene++;
    s->scenes[id] = {scene, bytes};
    s->scene_owned[id] = false;
    s->scene_bytes += bytes;
    s->st.bytes_resident = s->scene_bytes;
    *scene_id = id;
    return KDGRT_OK;
}

kdgrt_status kdgrt_server_remove_scene(kdgrt_server* s, int32_t scene_id) {
    if (!s) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_remove_scene: null");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_remove_scene: stop the server first");
    auto it = s->scenes.find(scene_id);
    if (it == s->scenes.end()) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_remove_scene: unknown scene");
    if (it->second.first && s->scene_owned[scene_id]) kdgrt_scene_destroy(it->second.first);
    s->scene_owned.erase(scene_id);
    s->scene_bytes -= it->second.second;
    s->st.bytes_resident = s->scene_bytes;
    s->scenes.erase(it);
    return KDGRT_OK;
}

kdgrt_status kdgrt_server_start(kdgrt_server* s) {
    if (!s) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_start: null");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) return KDGRT_OK;
    if (s->scene_bytes >= s->cfg.hbm_budget_bytes)
        return set_error(KDGRT_ERR_CAPACITY, "server_start: no budget left for outputs");
    const uint64_t arena = (s->cfg.hbm_budget_bytes - s->scene_bytes) & ~uint64_t(255);
    const int n = s->cfg.max_inflight_batches;
    if (!s->sim) {
        if (hipSetDevice(kdgrt::context_device(s->ctx)) != hipSuccess)
            return set_error(KDGRT_ERR_DEVICE, "server_start: set device");
        hipError_t e = hipMalloc(&s->d_arena, arena);
        if (e != hipSuccess) return set_error(KDGRT_ERR_OUT_OF_MEMORY, "server_start: output arena allocation failed");
        if (hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess)
            return set_error(KDGRT_ERR_DEVICE, "server_start: stream");
        s->copy_pool.push_back(s->copy_stream);
        for (int i = 0; i < n; ++i) {
            hipStream_t st;
            hipEvent_t ev;
            kdgrt::BatchSlot* sl = nullptr;
            if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
                return set_error(KDGRT_ERR_DEVICE, "server_start: stream/event");
            kdgrt_status k = kdgrt::slot_create(s->ctx, std::max(1, s->cfg.max_batch_views), &sl);
            if (k != KDGRT_OK) return k;
            s->streams.push_back(st);
            s->events.push_back(ev);
            s->slots.push_back(sl);
        }
    }
    s->slot_busy.assign(n, false);
    s->arena.reset(arena);
    if (s->cfg.persistent_kernel && !s->sim) {
        std::vector<const kdgrt_scene*> table;
        for (auto& kv : s->scenes) {
            s->scene_index[kv.first] = (int)table.size();
            table.push_back(kv.second.first);
        }
        if (table.empty()) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_start: no scenes");
        if (hipStreamCreateWithFlags(&s->serve_stream, hipStreamNonBlocking) != hipSuccess)
            return set_error(KDGRT_ERR_DEVICE, "server_start: stream");
        const int max_req = s->cfg.persistent_max_requests > 0 ? s->cfg.persistent_max_requests : 4096;
        const uint32_t spins = s->cfg.persistent_spin_limit > 0 ? s->cfg.persistent_spin_limit : (1u << 22);
        kdgrt_status k = kdgrt::serve_queue_create(s->ctx, max_req, table.data(), (int)table.size(), s->d_arena,
                                                   s->serve_stream, s->cfg.tuning, spins, &s->sq);
        if (k != KDGRT_OK) return k;
    }
    s->started = true;
    s->stopping = false;
    s->worker = std::thread(s->sq ? worker_main_persistent : worker_main, s);
    return KDGRT_OK;
}

kdgrt_status kdgrt_server_submit(kdgrt_server* s, const kdgrt_request* req, uint64_t* id) {
    if (!s || !req || !id) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_submit: null");
    if (req->width <= 0 || req->height <= 0 || req->width > 32768 || req->height > 32768)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_submit: bad size");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->started || s->stopping) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_submit: not running");
    if (!s->scenes.count(req->scene_id)) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_submit: unknown scene");
    const uint64_t bytes = (uint64_t)req->width * req->height * 16;
    s->st.submitted++;
    if (((bytes + 255) & ~uint64_t(255)) > s->arena.size()) {
        s->st.rejected++;
        return set_error(KDGRT_ERR_CAPACITY, "server_submit: output larger than the arena");
    }
    Req* r = new Req();
    r->id = s->next_id++;
    r->r = *req;
    r->bytes = bytes;
    r->arrival = req->arrival_s > 0 ? req->arrival_s : s->now();
    s->reqs[r->id] = r;
    s->pending.push_back(r);
    *id = r->id;
    s->cv_work.notify_one();
    return KDGRT_OK;
}

kdgrt_status kdgrt_server_wait(kdgrt_server* s, uint64_t id, double timeout_ms, float* host_rgbt,
                               kdgrt_result_info* info) {
    if (!s) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_wait: null");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_wait: unknown request");
    Req* r = it->second;
    if (r->collecting) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_wait: request already being collected");
    auto finished = [&] { return r->state >= kDone; };
    if (timeout_ms < 0) s->cv_done.wait(lk, finished);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished))
        return set_error(KDGRT_ERR_TIMEOUT, "server_wait: not finished");
    kdgrt_status st = (kdgrt_status)r->status;
    if (r->state == kCancelled) st = KDGRT_ERR_CANCELLED;
    if (st == KDGRT_OK && host_rgbt) {
        if (s->sim) {
            std::memset(host_rgbt, 0, r->bytes);
        } else {
            // The copy runs without the lock, so the scheduler thread keeps
            // retiring and launching batches meanwhile.  The request stays in
            // s->reqs and its arena region stays allocated until it is erased
            // below (only this call erases a finished request), and each
            // waiter copies on its own stream.
            // Concurrent waiters of other requests take copy streams from a
            // pool (created on demand, destroyed with the server).
            const void* src = (const char*)s->d_arena + r->off;
            const size_t nbytes = r->bytes;
            r->collecting = true;
            hipStream_t cs = nullptr;
            if (!s->copy_pool.empty()) {
                cs = s->copy_pool.back();
                s->copy_pool.pop_back();
            } else if (hipSetDevice(kdgrt::context_device(s->ctx)) != hipSuccess ||
                       hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
                cs = nullptr;
            }
            lk.unlock();
            const bool ok = cs && hipMemcpyAsync(host_rgbt, src, nbytes, hipMemcpyDeviceToHost, cs) == hipSuccess &&
                            hipStreamSynchronize(cs) == hipSuccess;
            lk.lock();
            if (cs) s->copy_pool.push_back(cs);
            if (!ok) st = set_error(KDGRT_ERR_DEVICE, "server_wait: result copy failed");
        }
    }
    if (info) {
        info->request_id = r->id;
        info->status = st;
        info->arrival_s = r->arrival;
        info->admitted_s = r->admitted;
        info->launched_s = r->launched;
        info->done_s = r->done;
        info->batch_id = r->batch_id;
        info->batch_views = r->batch_views;
    }
    if (r->has_mem) s->arena.release(r->off, r->bytes);
    s->reqs.erase(it);
    delete r;
    s->cv_work.notify_one();
    if (st == KDGRT_ERR_CANCELLED) return set_error(st, "server_wait: request cancelled");
    if (st != KDGRT_OK) return set_error(st, "server_wait: request failed");
    return KDGRT_OK;
}

kdgrt_status kdgrt_server_cancel(kdgrt_server* s, uint64_t id) {
    if (!s) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_cancel: null");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_cancel: unknown request");
    Req* r = it->second;
    if (r->state != kPending) return set_error(KDGRT_ERR_NOT_READY, "server_cancel: already started");
    s->pending.erase(std::find(s->pending.begin(), s->pending.end(), r));
    r->state = kCancelled;
    r->status = KDGRT_ERR_CANCELLED;
    r->done = s->now();
    s->st.cancelled++;
    s->cv_done.notify_all();
    return KDGRT_OK;
}

kdgrt_status kdgrt_server_stats_get(kdgrt_server* s, kdgrt_server_stats* out) {
    if (!s || !out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_stats: null");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->wall_s = s->now();
    out->bytes_resident = s->scene_bytes;
    if (s->busy_count > 0) out->busy_s += s->now() - s->busy_since;
    return KDGRT_OK;
}

double kdgrt_server_now(kdgrt_server* s) { return s ? s->now() : 0.0; }

kdgrt_status kdgrt_server_stop(kdgrt_server* s) {
    if (!s) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "server_stop: null");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->started) return KDGRT_OK;
        s->stopping = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    std::lock_guard<std::mutex> lk(s->mu);
    s->started = false;
    return KDGRT_OK;
}

kdgrt_status kdgrt_server_destroy(kdgrt_server* s) {
    if (!s) return KDGRT_OK;
    kdgrt_server_stop(s);
    if (!s->sim) {
        hipSetDevice(kdgrt::context_device(s->ctx));
        hipDeviceSynchronize();
        for (auto sl : s->slots) kdgrt::slot_destroy(sl);
        for (auto e : s->events) hipEventDestroy(e);
        for (auto st : s->streams) hipStreamDestroy(st);
        for (auto cs : s->copy_pool) hipStreamDestroy(cs);
        if (s->sq) kdgrt::serve_queue_destroy(s->sq);
        if (s->serve_stream) hipStreamDestroy(s->serve_stream);
        if (s->d_arena) hipFree(s->d_arena);
        for (auto& kv : s->scenes)
            if (kv.second.first && s->scene_owned[kv.first]) kdgrt_scene_destroy(kv.second.first);
    }
    for (auto& kv : s->reqs) delete kv.second;
    delete s;
    return KDGRT_OK;
}

} // extern "C"
