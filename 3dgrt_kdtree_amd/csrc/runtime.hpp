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
// kdgrt_render (RGBT output) in raster tile order, without the scene's
// frame-to-frame schedule: allocates nothing and never waits on the host.
kdgrt_status render_unscheduled(kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W, int H,
                                int depth, void* out, hipStream_t stream);
int context_device(const kdgrt_context* ctx);

} // namespace kdgrt
