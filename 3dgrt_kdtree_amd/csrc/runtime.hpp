// Device-runtime internals shared by the HIP translation units (not ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace kdgrt {

struct BatchSlot;
kdgrt_status slot_create(kdgrt_context* ctx, int max_items, BatchSlot** out);
kdgrt_status slot_launch(BatchSlot* sl, const kdgrt_view* views, int n, const kdgrt_tuning& t, hipStream_t s,
                         uint64_t* rays);
void slot_destroy(BatchSlot* sl);
// kdgrt_render (RGBT output) for the serving loop: the scene's tile schedule
// is followed only within the buffers reserve_tile_schedule made, so the call
// allocates nothing and never waits on the host (larger views run in raster
// order).
kdgrt_status render_served(kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W, int H,
                           int depth, void* out, hipStream_t stream);
// Tile-schedule buffers of a scene for views up to max_w x max_h pixels.
kdgrt_status reserve_tile_schedule(kdgrt_scene* sc, int max_w, int max_h);
// Device bytes that reserve_tile_schedule(sc, max_w, max_h) allocates for a scene.
uint64_t tile_schedule_bytes(int max_w, int max_h);
int context_device(const kdgrt_context* ctx);

// Persistent serving queue: one running k_serve kernel per epoch takes view
// requests published while it runs (continuous batching into the running
// kernel).  The queue's tables are allocated at create; starting an epoch,
// publishing, polling and stopping allocate nothing.
struct ServeQueue;
kdgrt_status serve_queue_create(kdgrt_context* ctx, int max_requests, const kdgrt_scene* const* scenes, int n_scenes,
                                void* arena, hipStream_t stream, const kdgrt_tuning& t, uint32_t spin_limit,
                                ServeQueue** out);
void serve_queue_destroy(ServeQueue* q);
kdgrt_status serve_epoch_start(ServeQueue* q);
// The host publishes nothing more in this epoch; the kernel drains and exits.
void serve_epoch_stop(ServeQueue* q);
// True once the epoch's kernel has exited (or no epoch runs); *status tells
// whether it exited cleanly.
bool serve_epoch_finished(ServeQueue* q, kdgrt_status* status);
bool serve_room(const ServeQueue* q, int W, int H);
kdgrt_status serve_publish(ServeQueue* q, int scene_index, const kdgrt_camera_info& cam, int W, int H,
                           uint64_t out_off, uint32_t* seq);
bool serve_done(const ServeQueue* q, uint32_t seq);

} // namespace kdgrt
