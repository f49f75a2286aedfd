// Autotuner for the persistent batched launch: coordinate descent over ray
// tile size, short-stack depth, hit-buffer capacity, waves per CU and block
// size, each candidate timed with HIP events over `reps` launches of the
// caller's real batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <string>
#include <vector>

#include "internal.hpp"

using kdgrt::set_error;

namespace {

kdgrt_status time_tuning(kdgrt_context* ctx, const kdgrt_view* views, int n_views, const kdgrt_ray_set* sets,
                         int n_sets, const kdgrt_tuning& t, int reps, hipStream_t s, float* ms) {
    kdgrt_batch* b = nullptr;
    kdgrt_status st = kdgrt_batch_create(ctx, views, n_views, sets, n_sets, &t, &b);
    if (st != KDGRT_OK) return st;
    st = kdgrt_batch_run(b, s, 0); // warm-up (code object load, caches)
    if (st == KDGRT_OK && hipStreamSynchronize(s) != hipSuccess) st = set_error(KDGRT_ERR_DEVICE, "autotune: sync");
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    if (st == KDGRT_OK) {
        hipEventRecord(e0, s);
        for (int r = 0; r < reps && st == KDGRT_OK; ++r) st = kdgrt_batch_run(b, s, 0);
        hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) st = set_error(KDGRT_ERR_DEVICE, "autotune: event sync");
        float total = 0.f;
        hipEventElapsedTime(&total, e0, e1);
        *ms = total / (float)reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    kdgrt_batch_destroy(b);
    return st;
}

} // namespace

extern "C" kdgrt_status kdgrt_autotune(kdgrt_context* ctx, const kdgrt_view* views, int n_views,
                                       const kdgrt_ray_set* sets, int n_sets, int max_trials, int reps,
                                       kdgrt_tuning* best, float* best_ms) {
    if (!ctx || !best || !best_ms) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "autotune: null");
    if (reps < 1) reps = 1;
    if (max_trials < 1) max_trials = 1;
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return set_error(KDGRT_ERR_DEVICE, "autotune: stream");
    kdgrt_tuning cur;
    kdgrt_tuning_default(ctx, &cur);
    float cur_ms = FLT_MAX;
    kdgrt_status st = time_tuning(ctx, views, n_views, sets, n_sets, cur, reps, s, &cur_ms);
    int trials = 1;
    // Dimensions in order of expected impact.
    struct Dim {
        int kind;
        std::vector<int> vals;
    };
    const std::vector<Dim> dims = {
        {0, {8, 16, 24, 32, 48}}, // waves per CU
        {1, {0, 1, 2}},           // tile: 8x8, 16x8, 16x16
        {2, {4, 7, 8, 16}},       // hit capacity
        {3, {12, 8, 16}},         // stack depth
        {4, {64, 128, 256}},      // block threads
    };
    const int tiles[3][2] = {{8, 8}, {16, 8}, {16, 16}};
    for (int pass = 0; pass < 2 && st == KDGRT_OK; ++pass) {
        for (const Dim& dim : dims) {
            for (int v : dim.vals) {
                if (trials >= max_trials) break;
                kdgrt_tuning c = cur;
                switch (dim.kind) {
                case 0: c.waves_per_cu = v; break;
                case 1: c.tile_w = tiles[v][0]; c.tile_h = tiles[v][1]; break;
                case 2: c.hit_capacity = v; break;
                case 3: c.stack_depth = v; break;
                case 4: c.block_threads = v; break;
                }
                if (std::memcmp(&c, &cur, sizeof(c)) == 0) continue;
                float ms = FLT_MAX;
                kdgrt_status s2 = time_tuning(ctx, views, n_views, sets, n_sets, c, reps, s, &ms);
                ++trials;
                if (s2 != KDGRT_OK) {
                    st = s2;
                    break;
                }
                if (ms < cur_ms) {
                    cur_ms = ms;
                    cur = c;
                }
            }
            if (st != KDGRT_OK) break;
        }
    }
    hipStreamDestroy(s);
    if (st != KDGRT_OK) return st;
    *best = cur;
    *best_ms = cur_ms;
    return KDGRT_OK;
}
