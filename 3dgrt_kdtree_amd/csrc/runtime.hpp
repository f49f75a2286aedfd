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

} // namespace kdgrt
