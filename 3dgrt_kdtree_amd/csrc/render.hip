// HIP kernels and device runtime behind the C ABI (include/kdgrt.h).
//
// Kernels:
//  k_view      plain one-view launch (cudaRenderer.cu:2313-2390 + 3099-3484):
//              one-wave blocks of 8x8 pixels, 64x64-pixel groups of tiles
//              interleaved over the XCDs (each L2 keeps a compact
//              neighbourhood, the costly image centre is spread over all).
//  k_rays      explicit ray lists, 64 rays per one-wave block.
//  k_batch     persistent launch over a device work queue of ray tiles drawn
//              from many views / ray sets / scenes; one queue per XCD with
//              stealing, each wave pulls its own tile (no block barriers).
//  k_reset     zeroes the queue counters (graph-safe, no memset node).
//  k_to_u8     framebuffer to 8-bit pixels (saveDeviceFramebufferToPNG,
//              cudaRenderer.cu:3066-3096) on the device.
//  k_hybrid    hybrid view (SECONDARY_RAY, cudaRenderer.cu:2217-2311): per
//              pixel, segments of the Gaussian trace between mirror / glass
//              secondary primitives, tiled as k_view.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "f16_up.hpp"
#include "internal.hpp"
#include "runtime.hpp"
#include "secondary_device.hpp"
#include "trace_device.hpp"

using namespace kdgrt_dev;

namespace {

constexpr int kXcds = 8;
constexpr int kMaxDevices = 64;

// Hit-buffer capacity of the view, ray-list, hybrid and serving kernels: a
// leaf visit with more hits takes a second, exact pass.  7 rather than 8
// (1080p view -1.1 %, 4M rays -1.6 %, hybrid view -3.3 %, measured with the
// iterative instruction scheduler; 6 gains less, 10 loses 6 %): one entry
// fewer frees registers the scheduler uses better than it loses to the rare
// extra passes.  The batch kernel takes 4, 7, 8 or 16 from its tuning.
constexpr int kViewK = 7;

// ---------------------------------------------------------------------------
// Plain one-view kernel.
// ---------------------------------------------------------------------------
struct ViewArgs {
    uint64_t* prof; // 5 x u64 per ray (profile variant)
    DevScene sc;
    CamInfo cam;
    int W, H;
    int depth;
    int blocks_x, blocks_y;
    int tile4; // 4x4-pixel tiles, 16 rays per wave (views too small to fill the chip)
    const uint32_t* order; // launch order of the logical tiles (null: XCD-interleaved raster order)
    uint32_t* cost;        // per logical tile: duration of this launch in wall-clock ticks (null: not recorded)
    void* out;
    int32_t* counters;
    uint64_t* hashes;
    uint32_t* seq;
    int seq_cap;
    uint8_t* touch[6]; // DBG: touch maps (node, quad, pre-test, exact test, evaluated, blended), each may be null
};

// Hardware dispatches workgroup b to XCD b % 8.  Logical blocks are taken
// in groups of G that share one XCD (and its L2), and consecutive groups go
// round-robin over the XCDs, so each L2 sees compact neighbourhoods while the
// expensive middle of an image is spread over all XCDs.  The grid is rounded
// up to a multiple of 8G; logical indices past the end exit at once.
template <uint32_t G>
__device__ __forceinline__ uint32_t xcd_interleave(uint32_t b) {
    const uint32_t i = b / kXcds;
    return (i / G) * (G * kXcds) + (b % kXcds) * G + (i % G);
}

template <uint32_t G>
static inline uint32_t interleave_grid(uint64_t n_logical) {
    return (uint32_t)(((n_logical + G * kXcds - 1) / (G * kXcds)) * (G * kXcds));
}

// Ray packets (64 rays) per XCD group of k_rays.  Runs of a sorted ray list
// differ in cost (a run is one direction cell over part of the box), and
// long runs per XCD leave some XCDs the costly ones: 4M incoherent rays on
// the 1M scene take 8.00 ms with groups of 8, 8.14-8.19 with 64 and 9.5-9.7
// with 256 or 1024 (1, 4 and 16: 8.04-8.11).
constexpr uint32_t kRaysGroup = 8;

template <int K, bool DBG, int FMT, bool PROF = false, bool CONE = true>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(DBG || PROF ? 1 : 4))) void k_view(ViewArgs a) {
    // One wavefront per block = one 8x8-pixel tile, so a finished wave frees
    // its slot at once (per-ray cost varies widely across an image).  Groups
    // of 64 tiles (64x64 px) share an XCD, unless the launch follows a tile
    // order (longest tiles of the previous launch first).
    const uint32_t L = a.order ? a.order[blockIdx.x] : xcd_interleave<64>(blockIdx.x);
    const uint64_t t_start = wall_clock64();
    const uint32_t st = L / 64u, local = L % 64u;
    const int tx = (int)(st % (uint32_t)a.blocks_x) * 8 + (int)(local % 8u);
    const int ty = (int)(st / (uint32_t)a.blocks_x) * 8 + (int)(local / 8u);
    const int lane = threadIdx.x;
    const int sh = a.tile4 ? 2 : 3; // log2 of the tile edge
    if (st >= (uint32_t)a.blocks_x * (uint32_t)a.blocks_y || (tx << sh) >= a.W || (ty << sh) >= a.H) {
        if (a.cost && lane == 0) a.cost[L] = 0; // a logical tile outside the image
        return;
    }
    if (a.tile4 && lane >= 16) return;
    const int x = (tx << sh) + (lane & ((1 << sh) - 1));
    const int y = (ty << sh) + (lane >> sh);
    if (x >= a.W || y >= a.H) return;
    float o[3], d[3], rgb[3], A;
    make_ray(a.cam, x, y, o, d);
    const size_t rid = (size_t)y * a.W + x;
    RayDbg dbg;
    if (DBG) {
        for (int i = 0; i < 8; ++i) dbg.cnt[i] = 0;
        dbg.hash = 0xcbf29ce484222325ull;
        dbg.seq = a.seq ? a.seq + rid * (size_t)a.seq_cap : nullptr;
        dbg.seq_cap = a.seq_cap;
        dbg.nseq = 0;
        dbg.t_node = a.touch[0];
        dbg.t_quad = a.touch[1];
        dbg.t_pre = a.touch[2];
        dbg.t_exact = a.touch[3];
        dbg.t_eval = a.touch[4];
        dbg.t_blend = a.touch[5];
    }
    if (PROF) {
        for (int i = 0; i < 5; ++i) dbg.cyc[i] = 0;
        dbg.pretests = dbg.survivors = 0;
        dbg.tlast = clock64();
    }
    extern __shared__ uint2 lds_stack[];
    trace_ray<K, DBG, PROF, CONE>(a.sc, o, d, a.depth, lds_stack, lane, rgb, A, &dbg);
    if (PROF) {
        for (int i = 0; i < 5; ++i) a.prof[8 * rid + i] = dbg.cyc[i];
        a.prof[8 * rid + 5] = dbg.pretests;
        a.prof[8 * rid + 6] = dbg.survivors;
        a.prof[8 * rid + 7] = 0;
    }
    const size_t pid = (size_t)(a.H - y - 1) * a.W + x; // framebuffer row flip (cudaRenderer.cu:2367)
    if (FMT == KDGRT_OUT_RGBT_F32) {
        reinterpret_cast<float4*>(a.out)[pid] = make_float4(rgb[0], rgb[1], rgb[2], 1.0f - A);
    } else {
        float* f = reinterpret_cast<float*>(a.out) + 3 * pid;
        f[0] = rgb[0];
        f[1] = rgb[1];
        f[2] = rgb[2];
    }
    if (DBG) {
        if (a.counters)
            for (int i = 0; i < 8; ++i) a.counters[8 * rid + i] = dbg.cnt[i];
        if (a.hashes) a.hashes[rid] = dbg.hash;
    }
    // lane 0 holds the tile's first pixel, which is inside the image
    if (a.cost && lane == 0) {
        const uint64_t dt = (uint64_t)(wall_clock64() - (long long)t_start);
        a.cost[L] = dt < 0xFFFFFFFFull ? (uint32_t)dt : 0xFFFFFFFFu;
    }
}

// ---------------------------------------------------------------------------
// Tile order of a launch that follows a close view (frame-to-frame schedule).
// ---------------------------------------------------------------------------
// A launch lasts at least as long as its slowest tile after that tile
// starts, so tiles that start late and run long leave the chip idle at the
// end (a 1080p view of the 1M-Gaussian scene keeps 77 % of its wave slots
// busy in raster order).  k_view records the duration of every tile; before
// the next launch of a close view, a counting sort run by one workgroup (the
// last workgroup of k_cone_prep) turns the smoothed durations (mean of the
// previous smoothed value and the last duration; the last duration alone
// after a camera change) into a longest-first order.  Durations fall into
// 512 log-spaced classes, 16 per octave; order within a class is that of
// the LDS atomics.  Any duration values give a permutation of [0, n).
constexpr int kSchedClasses = 512;
// k_cone_prep<false> runs in workgroups of 256 threads with no LDS and few
// registers, so it fits into the gaps that concurrent launches leave on the
// CUs; k_cone_prep<true> runs in workgroups of 1024, the first of which sorts
// a tile order (the sort is written for 1024 threads).
constexpr uint32_t kCinBlock = 1024;
constexpr uint32_t kCinBlockSmall = 256;
constexpr uint32_t kSchedMaxTiles = 1u << 20;
constexpr int kPrepUnroll = 8;
constexpr uint32_t kConeCuBlocks = 16;
constexpr uint32_t kPrepCuThreads = 4096;

__device__ __forceinline__ uint32_t sched_class(uint32_t c) {
    if (c == 0) return 0;
    const uint32_t e = 31u - (uint32_t)__builtin_clz(c);
    const uint32_t m = e >= 4u ? (c >> (e - 4u)) & 15u : (c << (4u - e)) & 15u;
    return e * 16u + m;
}

// Density cells of the first-launch estimate: cells of 2^cell_shift x
// 2^cell_shift tiles, at most kDensCells of them (32 KB of LDS in the sort).
constexpr uint32_t kDensCells = 8192;
static_assert(kDensCells % kCinBlock == 0, "the sort loads its cells in whole steps");

struct TileSort {
    const uint32_t* cost; // durations of the previous launch, per logical tile
    uint32_t* smooth;     // smoothed durations (in/out)
    uint32_t* order;      // out: tile order of the coming launch
    uint32_t* dens;       // density mode: Gaussian centres per cell (zeroed after reading)
    uint32_t n;           // logical tiles
    int reset;            // smooth := cost (the previous view was another one)
    int on;               // run the sort at all
    int density;          // key by the centre density around each tile instead of durations
    int blocks_x;         // 64-tile groups per image row (logical tile -> tile position)
    int tiles_w, tiles_h; // tiles of the image
    int cell_shift, cells_w, cells_h;
};

// Density key of logical tile t: Gaussian centres counted in the 3 x 3
// cells around the tile's cell (0 outside the image).
__device__ __forceinline__ uint32_t density_key(const TileSort& ts, const uint32_t* box, uint32_t t, float inv_bx) {
    // box holds, per cell, the centre count of the 3x3 cells around it
    const uint32_t st = t / 64u, local = t % 64u, bx = (uint32_t)ts.blocks_x;
    // st / bx from the float reciprocal, exact after one correction either
    // way (st < 2^24: the product is off by far less than one)
    uint32_t gy = (uint32_t)((float)st * inv_bx);
    if (gy * bx > st) --gy;
    else if ((gy + 1u) * bx <= st) ++gy;
    const int tx = (int)(st - gy * bx) * 8 + (int)(local % 8u);
    const int ty = (int)gy * 8 + (int)(local / 8u);
    if (tx >= ts.tiles_w || ty >= ts.tiles_h) return 0;
    return box[(ty >> ts.cell_shift) * ts.cells_w + (tx >> ts.cell_shift)];
}

// The sort, run by one workgroup of kCinBlock threads.  Keys: the smoothed
// durations of the previous launch, or (ts.density) the centre density
// around each tile.
__device__ void tile_sort_block(const TileSort& ts) {
    __shared__ uint32_t cls[kSchedClasses];
    __shared__ uint32_t dl[kDensCells];
    const uint32_t* __restrict__ cost = ts.cost;
    uint32_t* __restrict__ smooth = ts.smooth;
    uint32_t* __restrict__ order = ts.order;
    const uint32_t n = ts.n;
    const int reset = ts.reset;
    const uint32_t tid = threadIdx.x;
    const float inv_bx = 1.0f / (float)max(ts.blocks_x, 1);
    for (uint32_t k = tid; k < (uint32_t)kSchedClasses; k += blockDim.x) cls[k] = 0;
    if (ts.density) {
        // the cell counts into LDS (at most kDensCells = kUc x kCinBlock:
        // each thread loads its kUc cells at once, one round trip behind the
        // cone-record copy's traffic) and zeroed for the next estimate; then,
        // in place, each cell's count becomes the sum over the 3x3 cells
        // around it (clamped at the borders): a horizontal, then a vertical
        // pass, since the clamped window is a product of two intervals
        const uint32_t nc = (uint32_t)(ts.cells_w * ts.cells_h), cw = (uint32_t)ts.cells_w;
        constexpr uint32_t kUc = kDensCells / kCinBlock;
        uint32_t v[kUc];
#pragma unroll
        for (uint32_t j = 0; j < kUc; ++j) {
            const uint32_t k = j * blockDim.x + tid;
            v[j] = k < nc ? ts.dens[k] : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < kUc; ++j) {
            const uint32_t k = j * blockDim.x + tid;
            if (k < nc) {
                dl[k] = v[j];
                ts.dens[k] = 0;
            }
        }
        for (uint32_t pass = 0; pass < 2; ++pass) {
            __syncthreads();
            const uint32_t step = pass == 0 ? 1u : cw;
#pragma unroll
            for (uint32_t j = 0; j < kUc; ++j) {
                const uint32_t k = j * blockDim.x + tid;
                if (k < nc) {
                    const uint32_t x = k % cw, y = k / cw;
                    const bool lo = pass == 0 ? x > 0u : y > 0u;
                    const bool hi = pass == 0 ? x + 1u < cw : y + 1u < (uint32_t)ts.cells_h;
                    v[j] = dl[k] + (lo ? dl[k - step] : 0u) + (hi ? dl[k + step] : 0u);
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t j = 0; j < kUc; ++j) {
                const uint32_t k = j * blockDim.x + tid;
                if (k < nc) dl[k] = v[j];
            }
        }
        __syncthreads();
        constexpr uint32_t kUd = 8;
        for (uint32_t b0 = 0; b0 < n; b0 += kUd * blockDim.x) {
#pragma unroll
            for (uint32_t j = 0; j < kUd; ++j) {
                const uint32_t t = b0 + j * blockDim.x + tid;
                if (t < n) atomicAdd(&cls[sched_class(density_key(ts, dl, t, inv_bx))], 1u);
            }
        }
    }
    __syncthreads();
    // sixteen elements per thread and step, loaded before any is used, so a
    // step waits for one memory round trip rather than sixteen, and a 1080p
    // view (32,640 logical tiles) takes two steps per pass: the sort runs
    // beside the cone-record copy, whose traffic makes every round trip long
    // (eight per thread and step: the sort outlasted the copy by about 13 us)
    constexpr uint32_t kU = 16;
    const uint32_t stride = kU * blockDim.x;
    for (uint32_t b0 = 0; !ts.density && b0 < n; b0 += stride) {
        uint32_t c[kU], e[kU];
#pragma unroll
        for (uint32_t j = 0; j < kU; ++j) {
            const uint32_t t = b0 + j * blockDim.x + tid;
            c[j] = t < n ? cost[t] : 0u;
            e[j] = t < n ? smooth[t] : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < kU; ++j) {
            const uint32_t t = b0 + j * blockDim.x + tid;
            if (t >= n) break;
            const uint32_t v = (reset || e[j] == 0) ? c[j] : (uint32_t)(((uint64_t)e[j] + c[j] + 1u) >> 1);
            smooth[t] = v;
            atomicAdd(&cls[sched_class(v)], 1u);
        }
    }
    __syncthreads();
    if (tid < 64) {
        // exclusive scan from the top class down: lane l owns the classes
        // 511 - 8l ... 504 - 8l
        uint32_t cnt[8], sum = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            cnt[i] = cls[kSchedClasses - 1 - (8 * (int)tid + i)];
            sum += cnt[i];
        }
        uint32_t incl = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t v = __shfl_up(incl, d);
            if ((int)tid >= d) incl += v;
        }
        uint32_t run = incl - sum;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            cls[kSchedClasses - 1 - (8 * (int)tid + i)] = run;
            run += cnt[i];
        }
    }
    __syncthreads();
    for (uint32_t b0 = 0; b0 < n; b0 += stride) {
        uint32_t v[kU];
#pragma unroll
        for (uint32_t j = 0; j < kU; ++j) {
            const uint32_t t = b0 + j * blockDim.x + tid;
            v[j] = t < n ? (ts.density ? density_key(ts, dl, t, inv_bx) : smooth[t]) : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < kU; ++j) {
            const uint32_t t = b0 + j * blockDim.x + tid;
            if (t >= n) break;
            order[atomicAdd(&cls[sched_class(v[j])], 1u)] = t;
        }
    }
}

// ---------------------------------------------------------------------------
// Explicit rays.
// ---------------------------------------------------------------------------
struct RaysArgs {
    DevScene sc;
    const float* ro;
    const float* rd;
    uint64_t n;
    int depth;
    float4* out;
    int32_t* counters;
    uint64_t* hashes;
    const uint32_t* perm; // processing order (ray indices), or null for list order
    uint8_t* touch[6];    // DBG: touch maps as in ViewArgs
};

// Sort key of an explicit ray: its direction cell (4 bits: the octahedral
// map of the direction cut into 4 x 4 cells), then the Morton code of its
// origin in the scene box (9 bits per axis).  Rays sorted by it share a wave
// with rays that start nearby and head the same way, so the lanes of a wave
// walk more of the same nodes and leaves (2M rays with uniform origins and
// directions: 313 -> 423 Mrays/s with the direction octant in place of the
// cell; 4M such rays on the 1M scene: 8.58 -> 8.19 ms with the 16 cells).
constexpr int kRayKeyBits = 31;
constexpr uint64_t kRaySortMin = 65536;
// Batched views of at least this many pixels (the 1920x1080 class) get cone
// records of their own, up to kBatchConeMaxBytes of tables per batch:
// preparing them takes about 70 us per view of a 1M-Gaussian scene, which
// a 1280x720 view does not win back (64-view batch: 47.88 ms with sphere
// pre-tests only, 47.33 ms with cone records for 1080p views, 47.52 ms for
// 720p views as well).
constexpr uint64_t kBatchConeMinPixels = 1500000;
constexpr uint64_t kBatchConeMaxBytes = 16ull << 30;

// Static scratch of kdgrt_render_rays_noalloc: two key buffers and two index
// buffers of kNoAllocMaxRays entries, then the radix sort's temporaries
// (36 MiB in all, held by every device the library runs on until the process
// ends).  Host code never names the array directly: kdgrt_init resolves its
// device address once per device (hipGetSymbolAddress) and the entry point
// uses that address.
constexpr uint64_t kNoAllocMaxRays = 1u << 21;
constexpr size_t kNoAllocArr = (size_t)kNoAllocMaxRays * 4;
constexpr size_t kNoAllocTemp = (size_t)4 << 20;
constexpr size_t kNoAllocBytes = 4 * kNoAllocArr + kNoAllocTemp;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

__device__ __forceinline__ uint32_t spread3(uint32_t v) { // 9 bits -> every third bit
    uint32_t r = 0;
#pragma unroll
    for (int b = 0; b < 9; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

__global__ __launch_bounds__(256) void k_ray_keys(const float* __restrict__ ro, const float* __restrict__ rd,
                                                  uint32_t n, float3 lo, float3 inv_ext, uint32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ idx) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float ox = ro[3 * (size_t)i], oy = ro[3 * (size_t)i + 1], oz = ro[3 * (size_t)i + 2];
        const float dx = rd[3 * (size_t)i], dy = rd[3 * (size_t)i + 1], dz = rd[3 * (size_t)i + 2];
        auto cell = [](float v) { // [0, 1) -> 0..511, clamped (NaN -> 0)
            const float c = v * 512.0f;
            return (uint32_t)(c > 0.0f ? (c < 511.0f ? c : 511.0f) : 0.0f);
        };
        const uint32_t m = spread3(cell((ox - lo.x) * inv_ext.x)) | (spread3(cell((oy - lo.y) * inv_ext.y)) << 1) |
                           (spread3(cell((oz - lo.z) * inv_ext.z)) << 2);
        // direction cell: the octahedral map of d (|d|_1 = 1 folded onto the
        // square [-1, 1]^2) cut into 4 x 4 cells
        const float sa = fabsf(dx) + fabsf(dy) + fabsf(dz);
        float px = sa > 0.0f ? dx / sa : 0.0f, py = sa > 0.0f ? dy / sa : 0.0f;
        if (dz < 0.0f) {
            const float fx = (1.0f - fabsf(py)) * (px < 0.0f ? -1.0f : 1.0f);
            const float fy = (1.0f - fabsf(px)) * (py < 0.0f ? -1.0f : 1.0f);
            px = fx;
            py = fy;
        }
        auto dcell = [](float v) { // [-1, 1] -> 0..3, clamped (NaN -> 0)
            const float c = (v * 0.5f + 0.5f) * 4.0f;
            return (uint32_t)(c > 0.0f ? (c < 3.0f ? c : 3.0f) : 0.0f);
        };
        keys[i] = (((dcell(py) << 2) | dcell(px)) << 27) | m;
        idx[i] = i;
    }
}

template <int K, bool DBG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(DBG ? 1 : 4))) void k_rays(RaysArgs a) {
    const uint64_t j = (uint64_t)xcd_interleave<kRaysGroup>(blockIdx.x) * 64 + threadIdx.x;
    if (j >= a.n) return;
    const uint64_t i = a.perm ? (uint64_t)a.perm[j] : j;
    const float o[3] = {a.ro[3 * i], a.ro[3 * i + 1], a.ro[3 * i + 2]};
    const float d[3] = {a.rd[3 * i], a.rd[3 * i + 1], a.rd[3 * i + 2]};
    float rgb[3], A;
    RayDbg dbg;
    if (DBG) {
        for (int k = 0; k < 8; ++k) dbg.cnt[k] = 0;
        dbg.hash = 0xcbf29ce484222325ull;
        dbg.seq = nullptr;
        dbg.seq_cap = 0;
        dbg.nseq = 0;
        dbg.t_node = a.touch[0];
        dbg.t_quad = a.touch[1];
        dbg.t_pre = a.touch[2];
        dbg.t_exact = a.touch[3];
        dbg.t_eval = a.touch[4];
        dbg.t_blend = a.touch[5];
    }
    extern __shared__ uint2 lds_stack[];
    // SH records requested early (4M incoherent rays 7.59 -> 7.45 ms); the
    // debug kernel keeps the late request
    trace_ray<K, DBG, false, false, false, true, !DBG>(a.sc, o, d, a.depth, lds_stack, (int)threadIdx.x, rgb, A, &dbg);
    a.out[i] = make_float4(rgb[0], rgb[1], rgb[2], 1.0f - A);
    if (DBG) {
        if (a.counters)
            for (int k = 0; k < 8; ++k) a.counters[8 * i + k] = dbg.cnt[k];
        if (a.hashes) a.hashes[i] = dbg.hash;
    }
}

// ---------------------------------------------------------------------------
// Hybrid view: Gaussians + secondary primitives.
// ---------------------------------------------------------------------------
struct HybridArgs {
    ViewArgs v;
    SecScene sec;
};

// renderKernelGaussian_Mesh_Hybrid (cudaRenderer.cu:2217-2311).  Each
// segment: closest secondary primitive, Gaussian trace up to it (colour and
// opacity keep accumulating over the segments, as the reference passes
// them by reference), colour += throughput * accumulated colour, throughput
// *= 1 - accumulated opacity, stop below 0.01 in every channel, then the
// material step.  The first segment starts at the eye and uses the view's
// cone records; later segments use the sphere pre-test.
// Four waves per SIMD: the unconstrained allocation is one register above
// the 128 that four waves allow (129 -> three waves, 2.36 ms for the 1080p
// hybrid view; capped at 128, no spills, 1.98 ms).
template <int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_hybrid(HybridArgs ha) {
    const ViewArgs& a = ha.v;
    const uint32_t L = a.order ? a.order[blockIdx.x] : xcd_interleave<64>(blockIdx.x);
    const uint64_t t_start = wall_clock64();
    const uint32_t st = L / 64u, local = L % 64u;
    const int tx = (int)(st % (uint32_t)a.blocks_x) * 8 + (int)(local % 8u);
    const int ty = (int)(st / (uint32_t)a.blocks_x) * 8 + (int)(local / 8u);
    const int lane = threadIdx.x;
    const int sh = a.tile4 ? 2 : 3; // log2 of the tile edge
    if (st >= (uint32_t)a.blocks_x * (uint32_t)a.blocks_y || (tx << sh) >= a.W || (ty << sh) >= a.H) {
        if (a.cost && lane == 0) a.cost[L] = 0; // a logical tile outside the image
        return;
    }
    if (a.tile4 && lane >= 16) return;
    const int x = (tx << sh) + (lane & ((1 << sh) - 1));
    const int y = (ty << sh) + (lane >> sh);
    if (x >= a.W || y >= a.H) return;
    float o[3], d[3];
    make_ray(a.cam, x, y, o, d);
    float fin[3] = {0.0f, 0.0f, 0.0f}, thr[3] = {1.0f, 1.0f, 1.0f}, acc[3] = {0.0f, 0.0f, 0.0f}, A = 0.0f;
    extern __shared__ uint2 lds_stack[];
    for (int b = 0; b < ha.sec.max_bounces; ++b) {
        float t_hit;
        const int hit = sec_closest(ha.sec, o, d, t_hit);
        trace_ray<K, false, false, false, true, true, true>(a.sc, o, d, a.depth, lds_stack, lane, acc, A, nullptr,
                                                            hit >= 0 ? t_hit : kSecFar, b == 0);
        for (int k = 0; k < 3; ++k) fin[k] = fin[k] + thr[k] * acc[k];
        for (int k = 0; k < 3; ++k) thr[k] = thr[k] * (1.0f - A);
        if (thr[0] < 0.01f && thr[1] < 0.01f && thr[2] < 0.01f) break;
        if (hit < 0) break;
        SecHit h;
        sec_record(ha.sec, hit, t_hit, o, d, h);
        if (!sec_bounce(h, b, o, d, fin, thr)) break;
    }
    float* f = reinterpret_cast<float*>(a.out) + 3 * ((size_t)(a.H - y - 1) * a.W + x);
    f[0] = fminf(fin[0], 1.0f);
    f[1] = fminf(fin[1], 1.0f);
    f[2] = fminf(fin[2], 1.0f);
    if (a.cost && lane == 0) {
        const uint64_t dt = (uint64_t)(wall_clock64() - (long long)t_start);
        a.cost[L] = dt < 0xFFFFFFFFull ? (uint32_t)dt : 0xFFFFFFFFu;
    }
}

// ---------------------------------------------------------------------------
// Persistent batched kernel.
// ---------------------------------------------------------------------------
struct WorkItem { // one view or one explicit ray set
    DevScene sc;
    CamInfo cam;
    int32_t W, H;       // view size (ray sets: W = 0)
    int32_t tiles_x;    // view tiles per row
    int32_t tiles_y;    // view tile rows
    int32_t groups_x;   // 8x8-tile groups per row (view tiles are numbered group-major)
    uint32_t tile_begin;
    uint32_t tile_count;
    float4* out;
    const float* ro;
    const float* rd;
    uint64_t n_rays;
    const uint32_t* perm; // ray sets: processing order (ray indices), or null for list order
};

// Tile queues: the tile space is cut into groups of kQueueGroup consecutive
// tiles (for a view, one 8x8-tile square), dealt round-robin to the kXcds
// queues; q_len[x] is the number of tiles queue x owns.
constexpr uint32_t kQueueGroup = 64;

__host__ __device__ __forceinline__ uint32_t queue_tile(uint32_t q, uint32_t j) {
    return ((j / kQueueGroup) * kXcds + q) * kQueueGroup + (j % kQueueGroup);
}

static inline void queue_lengths(uint64_t tiles, uint32_t q_len[kXcds]) {
    const uint64_t groups = tiles / kQueueGroup, rest = tiles % kQueueGroup;
    for (uint32_t q = 0; q < kXcds; ++q) {
        const uint64_t full = groups / kXcds + (q < groups % kXcds ? 1 : 0);
        q_len[q] = (uint32_t)(full * kQueueGroup + ((rest && groups % kXcds == q) ? rest : 0));
    }
}

struct BatchArgs {
    const WorkItem* items;
    int n_items;
    uint32_t q_len[kXcds];       // tiles owned by each queue (queue_tile maps them)
    uint32_t* counters;          // kXcds counters, kCounterStride words apart
    int depth;
    int tile_w, tile_h;          // view tile in pixels; tile_w*tile_h is a multiple of 64
};

// Plain streaming copy, the calibration kernel for HBM traffic counters.
// One output pixel per thread: clamp to [0, 1], times 255, truncated
// (cudaRenderer.cu:3083-3086); flip: read framebuffer row h-1-y for image
// row y (the reference's stbi_flip_vertically_on_write); a fourth output
// channel is 255 (opaque), a fourth input channel (T) is dropped.
__global__ __launch_bounds__(256) void k_to_u8(const float* __restrict__ fb, int w, int h, int cin,
                                               uint8_t* __restrict__ out, int cout, int flip) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint64_t)w * h) return;
    const int x = (int)(i % (uint64_t)w), y = (int)(i / (uint64_t)w);
    const int sy = flip ? h - 1 - y : y;
    const float* src = fb + (size_t)cin * ((size_t)sy * w + x);
    uint8_t* dst = out + (size_t)cout * i;
    for (int c = 0; c < 3; ++c) {
        const float v = fmaxf(0.0f, fminf(1.0f, src[c]));
        dst[c] = (uint8_t)(v * 255.0f);
    }
    if (cout == 4) dst[3] = 255;
}

// Streaming copy: each thread moves four 16-B words per chunk of 4 x 256
// words, all four loads issued before the first store, with nontemporal
// hints (the data is not reused); the remainder past the last whole chunk is
// copied word by word.  On one MI355X this moves 1 GiB each way at about
// 6.0 TB/s (read + write) against 4.5 TB/s for one word per thread per step.
typedef float kd_f4 __attribute__((ext_vector_type(4)));
constexpr int kCopyUnroll = 4;
__global__ __launch_bounds__(256) void k_copy(kd_f4* __restrict__ dst, const kd_f4* __restrict__ src, uint64_t n) {
    constexpr uint64_t chunk = 256ull * kCopyUnroll;
    const uint64_t nfull = n / chunk;
    for (uint64_t c = blockIdx.x; c < nfull; c += gridDim.x) {
        const uint64_t b = c * chunk + threadIdx.x;
        kd_f4 v[kCopyUnroll];
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) v[u] = __builtin_nontemporal_load(src + b + 256 * u);
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) __builtin_nontemporal_store(v[u], dst + b + 256 * u);
    }
    for (uint64_t i = nfull * chunk + (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        dst[i] = src[i];
}

// VALU-issue ceiling: eight independent fused multiply-add chains per lane,
// kValuChains * iters v_fma_f32 per wavefront and nothing else in the loop
// but scalar loop control, at full occupancy.  Its instruction rate
// (kdgrt_valu_probe) is the measured peak that the trace kernels' VALU
// instruction counts are compared with.
constexpr int kValuChains = 8;
__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ out, int iters, float b, float c) {
    float a[kValuChains];
#pragma unroll
    for (int k = 0; k < kValuChains; ++k) a[k] = (float)(threadIdx.x + k);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < kValuChains; ++k) a[k] = __builtin_fmaf(a[k], b, c);
    }
    float r = 0.0f;
#pragma unroll
    for (int k = 0; k < kValuChains; ++k) r += a[k];
    out[blockIdx.x * 256 + threadIdx.x] = r;
}

// First-launch estimate of the tile costs: each Gaussian centre in front of
// the camera is projected to the image and counted in its density cell.  The
// sort keys each tile by the centres in the 3 x 3 cells around it: where many
// Gaussians project, rays meet many leaves and candidates.
struct DensityArgs {
    float eye[3];
    float n[3];     // u x v (any length)
    float sn;       // (start - eye) . n
    float se[3];    // start - eye
    float pu[3];    // u / (|u|^2 step_x): pixel x of a point on the plane
    float pv[3];    // v / (|v|^2 step_y): minus pixel y
    float cell_px;  // 1 / cell edge in pixels
    int W, H, cells_w, cells_h;
};

// Counts go to a per-workgroup histogram in LDS first: the centres of a
// dense cluster fall into few cells, and one global atomic per centre on
// those cells serialises (177 us for 1M centres); from LDS each cell costs
// one global atomic per workgroup.
constexpr uint32_t kDensBlock = 1024;

__global__ __launch_bounds__(kDensBlock) void k_density(const float4* __restrict__ gsph, uint64_t n_g, DensityArgs d,
                                                        uint32_t* __restrict__ dens) {
    __shared__ uint32_t hist[kDensCells];
    const uint32_t nc = (uint32_t)(d.cells_w * d.cells_h);
    for (uint32_t k = threadIdx.x; k < nc; k += kDensBlock) hist[k] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * kDensBlock + threadIdx.x; i < n_g; i += (uint64_t)gridDim.x * kDensBlock) {
        const float4 c = gsph[i];
        const float rx = c.x - d.eye[0], ry = c.y - d.eye[1], rz = c.z - d.eye[2];
        const float t = (rx * d.n[0] + ry * d.n[1] + rz * d.n[2]) / d.sn;
        if (!(t > 0.0f)) continue;
        const float inv = 1.0f / t;
        const float qx = rx * inv - d.se[0], qy = ry * inv - d.se[1], qz = rz * inv - d.se[2];
        const float px = qx * d.pu[0] + qy * d.pu[1] + qz * d.pu[2];
        const float py = -(qx * d.pv[0] + qy * d.pv[1] + qz * d.pv[2]);
        if (!(px >= 0.0f && px < (float)d.W && py >= 0.0f && py < (float)d.H)) continue;
        const int cx = min((int)(px * d.cell_px), d.cells_w - 1), cy = min((int)(py * d.cell_px), d.cells_h - 1);
        atomicAdd(&hist[cy * d.cells_w + cx], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nc; k += kDensBlock)
        if (hist[k]) atomicAdd(&dens[k], hist[k]);
}

// Cone records for the pre-test of rays leaving one origin e (cone_pass).
// Bound-sphere bound: for the sphere test's own f32 quantities a = c - e and
// R = r + kappa*|a|_1 (same operations, same bits), sphere_pass accepting
// implies, with eps = 2^-24 and exact arithmetic on the f32 inputs:
//   |a x d| <= R' = R(1+4eps) + 10eps|a|   (rounding of the f32 cross product
//                                           and of dist2 <= R*R),
//   (a.d)^2 >= |a|^2 (1-16eps) - R'^2 =: Q^2  (|d|^2 >= 1-9eps for a
//                                           normalised f32 direction),
//   a.d > -R - 4eps|a|                      (t test with t_near >= 1e-2 > 0;
//                                           f32 a.d is within 4eps|a|).
// When Q >= R + 4eps|a| the second and third exclude a.d <= -Q, so
// a.d >= Q and the f32 a.d >= q = Q - 4eps|a|.  q is formed in double and
// rounded down; otherwise (origin near or inside the sphere, or r = inf) the
// record accepts every direction.
//
// Per-Gaussian cone record of a view from eye e: (a, q) with a = mu - e the
// f32 difference and q a lower bound of the f32 dot product a.d (cone_pass,
// three roundings) for every ray direction d from e whose exact test can
// accept the Gaussian.  q is the larger of two bounds, both formed in double
// and rounded down:
//  * the bound-sphere bound above (r, kappa of the sphere record);
//  * a bound from the projected ellipsoid.  A ray the exact test accepts
//    passes within kappa |a|_1 (the f32 rounding of that test, as for the
//    sphere) of the ellipsoid E = {x : (x - mu)^T S^-1 (x - mu) <= 1} scaled
//    by 1.001 about mu, at a point in front of e, where S = 8 (A^T A)^-1 with
//    A = diag(1/scale) M (sig holds S per Gaussian).  With v the unit vector
//    from e to mu, that dilated ellipsoid lies in the cylinder of radius
//    R = 1.001 sqrt(lambda_max(P S P)) + kappa |a|_1 about the line e + z v
//    (P = I - v v^T: the largest semi-axis of E's projection across v) and in
//    the slab |z - |mu - e|| <= h = 1.001 sqrt(v^T S v) + kappa |a|_1.  If
//    z0 = |mu - e| - h > 0, an accepted direction makes an angle theta with v
//    with tan theta <= R / z0, so (mu - e).d >= |mu - e| |d| C with
//    C = z0 / sqrt(z0^2 + R^2); the f32 a.d then exceeds
//    |mu - e| (C - 12 eps): eps for the rounding of a, 3 eps for the
//    dot product and 4.5 eps for |d| >= 1 - 4.5 eps.  lambda_max(P S P) =
//    (T + sqrt(2F - T^2)) / 2 with T = tr S - v^T S v and
//    F = tr S^2 - 2 |S v|^2 + (v^T S v)^2 (the two nonzero eigenvalues of
//    P S P have sum T and sum of squares F), plus 2e-7 tr S: sig holds S
//    rounded to f32, and S <= S_f32 + 2^-24 |S|_F I <= S_f32 + 2^-24 tr(S) I
//    (the doubled term also covers the double rounding here).
// sig holds NaN where the sphere record has r = inf (a singular or
// non-orthogonal rotation): the ellipsoid bound is skipped there.
__device__ __forceinline__ float4 gaussian_cone(const float4 sp, const float (&S)[6], float ex, float ey, float ez) {
    const double eps = 5.9604644775390625e-08;
    const float ax = sp.x - ex, ay = sp.y - ey, az = sp.z - ez;
    const uint32_t w = __float_as_uint(sp.w);
    const float r = __half2float(__ushort_as_half((unsigned short)(w & 0xFFFFu)));
    const float kap = __half2float(__ushort_as_half((unsigned short)(w >> 16)));
    const float l1 = fabsf(ax) + fabsf(ay) + fabsf(az);
    const float R = r + kap * l1;
    double qd = -INFINITY;
    if (R < 3.402823466e+38f) {
        // bound sphere
        const double a2 = (double)ax * ax + (double)ay * ay + (double)az * az;
        const double na = sqrt(a2);
        const double Rp = (double)R * (1.0 + 4.0 * eps) + 10.0 * eps * na;
        const double Q2 = a2 * (1.0 - 16.0 * eps) - Rp * Rp;
        const double lim = (double)R + 4.0 * eps * na;
        if (Q2 > 0.0 && Q2 >= lim * lim) qd = sqrt(Q2) - 4.0 * eps * na;
        // projected ellipsoid
        const double s00 = S[0], s01 = S[1], s02 = S[2], s11 = S[3], s12 = S[4], s22 = S[5];
        const double mx = (double)sp.x - (double)ex, my = (double)sp.y - (double)ey, mz = (double)sp.z - (double)ez;
        const double nm = sqrt(mx * mx + my * my + mz * mz);
        if (s00 == s00 && nm > 0.0) {
            const double inm = 1.0 / nm;
            const double vx = mx * inm, vy = my * inm, vz = mz * inm;
            const double t0 = s00 * vx + s01 * vy + s02 * vz, t1 = s01 * vx + s11 * vy + s12 * vz,
                         t2 = s02 * vx + s12 * vy + s22 * vz;
            const double vsv = fmax(vx * t0 + vy * t1 + vz * t2, 0.0);
            const double tr = s00 + s11 + s22;
            const double trs2 = s00 * s00 + s11 * s11 + s22 * s22 + 2.0 * (s01 * s01 + s02 * s02 + s12 * s12);
            const double T = tr - vsv;
            const double F = trs2 - 2.0 * (t0 * t0 + t1 * t1 + t2 * t2) + vsv * vsv;
            const double lmax = 0.5 * (fmax(T, 0.0) + sqrt(fmax(2.0 * F - T * T, 0.0))) + 2e-7 * tr;
            const double dil = (double)kap * (double)l1;
            const double Rc = 1.001 * sqrt(lmax) + dil;
            const double h = 1.001 * sqrt(vsv + 2e-7 * tr) + dil;
            const double z0 = nm - h;
            if (z0 > 0.0) {
                const double C = z0 / sqrt(z0 * z0 + Rc * Rc) - 12.0 * eps;
                if (C > 0.0) qd = fmax(qd, nm * C * (1.0 - 1e-12));
            }
        }
    }
    return make_float4(ax, ay, az, qd == -INFINITY ? -3.402823466e+38f : __double2float_rd(qd));
}

// Grid-stride over the Gaussians (about one per thread for a 1M-Gaussian
// scene, gaussian_cone_grid), the records of a thread's next Gaussian
// requested before the current one's double-precision arithmetic.
__global__ __launch_bounds__(256) void k_gaussian_cone(const float4* __restrict__ gsph, const float* __restrict__ sig,
                                                       uint64_t n, float ex, float ey, float ez,
                                                       float4* __restrict__ gcone) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float4 sp = gsph[i];
    float S[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) S[k] = sig[6 * i + k];
    for (;;) {
        const uint64_t j = i + stride;
        float4 sp_n = make_float4(0.f, 0.f, 0.f, 0.f);
        float S_n[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (j < n) {
            sp_n = gsph[j];
#pragma unroll
            for (int k = 0; k < 6; ++k) S_n[k] = sig[6 * j + k];
        }
        gcone[i] = gaussian_cone(sp, S, ex, ey, ez);
        if (j >= n) break;
        i = j;
        sp = sp_n;
#pragma unroll
        for (int k = 0; k < 6; ++k) S[k] = S_n[k];
    }
}

// Workgroups of a k_gaussian_cone pass over n Gaussians: kConeCuBlocks
// workgroups of 256 threads per CU, fewer for a small scene.
static uint32_t gaussian_cone_grid(uint64_t n, int n_cu) {
    const uint64_t want = (n + 255) / 256, cap = (uint64_t)std::max(1, n_cu) * kConeCuBlocks;
    return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

// Per-view cone records in leaf-slot order: the record of each slot's
// Gaussian (Gaussians are numbered by first appearance in the leaf lists, so
// consecutive slots read nearby records), zero records in the padding slots
// (their pre-test bits are masked off).
// SORT: the first workgroup (dispatched first, so it starts at once) sorts
// the tile order of the coming view launch instead (tile_sort_block) while
// the others copy the cone records, so the sort adds no time of its own.
template <bool SORT>
__global__ __launch_bounds__(SORT ? kCinBlock : kCinBlockSmall) void k_cone_prep(const uint32_t* __restrict__ offs,
                                                                                 const float4* __restrict__ gcone,
                                                                                 uint64_t n_slots, uint64_t n,
                                                                                 float4* __restrict__ vcone,
                                                                                 TileSort ts) {
    const uint32_t first = SORT ? 1u : 0u, prep_blocks = gridDim.x - first;
    if (SORT && blockIdx.x < first) {
        tile_sort_block(ts);
        return;
    }
    // Chunks of 64 x kPrepUnroll consecutive slots are dealt to the
    // wavefronts in turn (whatever the workgroup size): a wave issues the ID
    // loads of its chunk together, then its gathers, then its stores, so it
    // keeps several dependent gathers in flight; the rest (and the padding
    // slots, which get zero records) one record per lane at a time.
    const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x / 64u;
    const uint64_t wid = (uint64_t)(blockIdx.x - first) * wpb + threadIdx.x / 64u, nw = (uint64_t)prep_blocks * wpb;
    constexpr uint64_t chunk = 64u * kPrepUnroll;
    const uint64_t nfull = n_slots / chunk;
    for (uint64_t c = wid; c < nfull; c += nw) {
        const uint64_t b = c * chunk + lane;
        uint32_t g[kPrepUnroll];
#pragma unroll
        for (int u = 0; u < kPrepUnroll; ++u) g[u] = __builtin_nontemporal_load(offs + b + 64u * u);
        float4 v[kPrepUnroll];
#pragma unroll
        for (int u = 0; u < kPrepUnroll; ++u) v[u] = gcone[g[u]];
#pragma unroll
        for (int u = 0; u < kPrepUnroll; ++u) vcone[b + 64u * u] = v[u];
    }
    for (uint64_t i = nfull * chunk + wid * 64u + lane; i < n; i += nw * 64u)
        vcone[i] = i < n_slots ? gcone[offs[i]] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Workgroups of a cone-table pass over n slot records: kPrepCuThreads
// threads per CU, fewer when the table is short.
static uint32_t cone_prep_grid(uint64_t n, uint32_t blk, int n_cu) {
    const uint64_t want = (n + (uint64_t)blk * kPrepUnroll - 1) / ((uint64_t)blk * kPrepUnroll);
    const uint64_t cap = (uint64_t)std::max(1, n_cu) * kPrepCuThreads / blk;
    return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

// Batch queue counters, one per 128-B line: every wave of the chip claims
// its tiles through them, and atomics from all XCDs on one line serialise at
// the memory side.
constexpr int kCounterStride = 32;
constexpr size_t kCounterBytes = sizeof(uint32_t) * kCounterStride * kXcds;

__global__ void k_reset(uint32_t* c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x * kCounterStride] = 0;
}

template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(K <= 8 ? 4 : 1))) void k_batch(BatchArgs a) {
    const int lane = threadIdx.x & 63;
    extern __shared__ uint2 lds_stack[];
    uint2* const wave_lds = lds_stack + (threadIdx.x >> 6) * (wave_lds_bytes(a.depth) / sizeof(uint2));
    const int home = blockIdx.x % kXcds;
    int probe = 0;
    while (probe < kXcds) {
        const int q = (home + probe) % kXcds;
        const uint32_t qlen = a.q_len[q];
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&a.counters[q * kCounterStride], 1u);
        // readfirstlane makes the tile (and everything derived from it)
        // provably wave-uniform, so the work item lives in scalar registers
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= qlen) {
            ++probe; // this queue is drained; move on for good
            continue;
        }
        const uint32_t tile = queue_tile((uint32_t)q, t);
        // locate the work item (items are sorted by tile_begin)
        int lo = 0, hi = a.n_items - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.items[mid].tile_begin <= tile) lo = mid;
            else hi = mid - 1;
        }
        // copy the work item word by word through readfirstlane so that every
        // field (scene pointers, camera) is held in scalar registers
        WorkItem it;
        {
            const uint32_t* src = reinterpret_cast<const uint32_t*>(a.items + __builtin_amdgcn_readfirstlane(lo));
            uint32_t* dst = reinterpret_cast<uint32_t*>(&it);
#pragma unroll
            for (int w = 0; w < (int)(sizeof(WorkItem) / 4); ++w) dst[w] = __builtin_amdgcn_readfirstlane(src[w]);
        }
        const uint32_t local = tile - it.tile_begin;
        const bool is_view = it.W > 0;
        // view tiles are numbered in 8x8-tile groups, row-major within a group
        const uint32_t grp = local / 64u, gl = local % 64u;
        const int tx = is_view ? (int)((grp % (uint32_t)it.groups_x) * 8u + gl % 8u) : 0;
        const int ty = is_view ? (int)((grp / (uint32_t)it.groups_x) * 8u + gl / 8u) : 0;
        if (is_view && (tx >= it.tiles_x || ty >= it.tiles_y)) continue; // group padding past the view
        const int npx = a.tile_w * a.tile_h;
        for (int p0 = 0; p0 < npx; p0 += 64) {
            const int p = p0 + lane;
            float o[3], d[3];
            bool live;
            size_t dst;
            if (is_view) {
                // 8x8 sub-squares inside the tile keep each 64-lane step compact
                const int sub = p >> 6, l = p & 63;
                const int subs_x = a.tile_w >> 3;
                const int x = tx * a.tile_w + (sub % subs_x) * 8 + (l & 7);
                const int y = ty * a.tile_h + (sub / subs_x) * 8 + (l >> 3);
                live = x < it.W && y < it.H;
                make_ray(it.cam, x, y, o, d);
                dst = (size_t)(it.H - y - 1) * it.W + x; // framebuffer row flip
            } else {
                const uint64_t j = (uint64_t)local * npx + p;
                live = j < it.n_rays;
                const uint64_t i = (live && it.perm) ? (uint64_t)it.perm[j] : j;
                o[0] = o[1] = o[2] = 0.f;
                d[0] = d[1] = 0.f;
                d[2] = 1.f;
                if (live) {
                    o[0] = it.ro[3 * i];
                    o[1] = it.ro[3 * i + 1];
                    o[2] = it.ro[3 * i + 2];
                    d[0] = it.rd[3 * i];
                    d[1] = it.rd[3 * i + 1];
                    d[2] = it.rd[3 * i + 2];
                }
                dst = i;
            }
            if (live) {
                float rgb[3], A;
                // a view with cone records of its own takes the view
                // pre-test (the item, hence the branch, is wave-uniform); both
                // branches request SH records early (16 B of per-tile state
                // spill, and the 64-view batch still gains 45.12 -> 44.33 ms)
                if (it.sc.vcone)
                    trace_ray<K, false, false, true, false, true, true>(it.sc, o, d, a.depth, wave_lds, lane, rgb, A,
                                                                         nullptr);
                else
                    trace_ray<K, false, false, false, false, true, true>(it.sc, o, d, a.depth, wave_lds, lane, rgb, A,
                                                                          nullptr);
                it.out[dst] = make_float4(rgb[0], rgb[1], rgb[2], 1.0f - A);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Persistent serving kernel: one launch per epoch takes view requests while
// it runs.
// ---------------------------------------------------------------------------
// The host appends request records to a table in pinned host memory and
// publishes them with a release store of the request count.  The tile space
// of the epoch (requests in order, each a whole number of 64-tile groups of
// 8x8-pixel tiles) is dealt to kXcds queues as the batch kernel deals it:
// group g belongs to queue g mod kXcds, so the neighbouring tiles of a group
// run on one XCD and share its L2.  A wave claims the next tile of its own
// XCD's queue with one atomic on that queue's counter.  A claimed tile may
// not be published yet: the wave that holds the first tile of its queue's
// first unpublished group (at most one per queue) polls the host table,
// copies new records into device memory and then raises the device's
// published-tile mark (atomic maximum, so concurrent pollers never lower
// it); every other waiting wave takes published tiles that other queues have
// not handed out (compare-and-swap on their counters, never an unpublished
// one), or waits on the mark in device memory.  The wave that finishes a
// request's last tile stores the request's sequence number into the host's
// done table.  The host stops an epoch by storing 1 to `stop` after its last
// publication; a polling wave turns it into a device stop flag and every
// wave exits once its own claim is unpublished (claims of a queue are
// consecutive, so its published tiles have all been handed out by then and
// the waves holding them finish them).  Every wait is bounded
// (spin_limit polls): a wave that runs out of polls exits too, and the host
// publishes whatever an early-ended epoch left unfinished again.  Records are
// written once per epoch (no reuse), so a record below the published mark
// never changes while a wave reads it.
//
// Memory ordering without cache maintenance on the tile path: acquire and
// release operations at agent or system scope write back or invalidate a
// whole L2 on this GPU, which would evict every wave's scene data at each
// tile.  Instead every shared word goes through scope-coherent accesses --
// relaxed atomic loads and stores at agent scope for the device tables, at
// system scope for the host tables and the output pixels, which write
// through to memory -- and a wave waits for its own accesses to complete
// (workgroup-scope fence, a counter wait only) before the access that
// publishes them: the record copies before the device mark, a tile's pixels
// before its count.  A wave issues its loads in program order and a record
// is read only after the mark that covers it was seen, so it reads the
// completed record.
struct ServeRec { // one view request, 112 B
    uint32_t tile_begin, tile_count, scene, seq;
    int32_t W, H, tiles_x, tiles_y;
    int32_t groups_x, pad0;
    uint64_t out_off; // byte offset into the output arena
    CamInfo cam;
    uint32_t pad1[2];
};
static_assert(sizeof(ServeRec) == 112, "ServeRec layout");
constexpr uint32_t kRecWords = sizeof(ServeRec) / 4;

struct ServeHostCtl { // pinned host memory, written by the host
    uint32_t published_req;
    uint32_t stop;
    uint32_t pad[14];
};

// Device control words.  Every wave claims a tile through its queue's
// counter, and agent-scope accesses from all XCDs to one line serialise at
// the memory side, so each counter has a 128-B line of its own and the
// marks another.
struct ServeQueueWord {
    uint32_t next; // claims of the queue (queue_tile maps claim j of queue q to its tile)
    uint32_t pad[31];
};
struct ServeDevCtl { // device memory
    uint32_t mirror_tiles; // tiles of the records copied to the device (a multiple of 64)
    uint32_t mirror_req;   // records copied to the device
    uint32_t stop;
    uint32_t pad[29];
    ServeQueueWord queue[kXcds];
};
static_assert(sizeof(ServeDevCtl) == 128 * (kXcds + 1), "ServeDevCtl layout");

struct ServeArgs {
    ServeHostCtl* hctl;
    ServeRec* hrec;
    uint32_t* hdone;        // host: sequence + 1 of each finished request
    ServeDevCtl* dctl;
    ServeRec* drec;
    uint32_t* done_tiles;   // device: finished tiles per request
    const DevScene* scenes;
    int n_scenes;
    char* arena;
    int depth;
    int max_req;
    uint32_t spin_limit;
};

__device__ __forceinline__ uint32_t wave_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_system(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_system(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// This wave's earlier memory accesses have completed (a counter wait; no
// cache write-back or invalidation at this scope).
__device__ __forceinline__ void wave_accesses_done() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup"); }

// The wave that holds the first unpublished tile: copy newly published
// records into device memory and advance the device marks, or turn the
// host's stop into the device stop.  The host stores stop only after its
// last publication, so stop is read first.
__device__ void serve_poll_host(const ServeArgs& a, int lane) {
    uint32_t st = 0, pr = 0;
    if (lane == 0) {
        st = ld_system(&a.hctl->stop);
        pr = ld_system(&a.hctl->published_req);
    }
    st = wave_u32(st);
    pr = wave_u32(pr);
    if (pr > (uint32_t)a.max_req) pr = (uint32_t)a.max_req; // malformed count: never read past the table
    const uint32_t mr = wave_u32(ld_agent(&a.dctl->mirror_req));
    if (pr > mr) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(a.hrec + mr);
        uint32_t* dst = reinterpret_cast<uint32_t*>(a.drec + mr);
        const uint32_t n = (pr - mr) * kRecWords;
        for (uint32_t i = (uint32_t)lane; i < n; i += 64) st_agent(dst + i, ld_system(src + i));
        uint32_t end = 0;
        if (lane == 0) {
            const ServeRec* last = a.hrec + pr - 1;
            end = ld_system(&last->tile_begin) + ld_system(&last->tile_count);
        }
        wave_accesses_done();
        // pollers may overlap: each copied every record below its marks, and
        // the maxima keep the marks from going back
        if (lane == 0) {
            __hip_atomic_fetch_max(&a.dctl->mirror_req, pr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wave_accesses_done();
            __hip_atomic_fetch_max(&a.dctl->mirror_tiles, end, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (st) {
        if (lane == 0) st_agent(&a.dctl->stop, 1u);
    }
}

// The request whose tiles hold tile t (records below the device mark are
// complete and sorted by tile_begin), starting from the request k of the
// wave's previous tile: forward while t lies past it, by bisection when a
// stolen tile lies before it.  max_req when t lies in no record.
__device__ __forceinline__ uint32_t serve_find(const ServeArgs& a, uint32_t t, uint32_t k) {
    auto begin = [&](uint32_t i) { return wave_u32(ld_agent(reinterpret_cast<const uint32_t*>(a.drec + i))); };
    if (k < (uint32_t)a.max_req && t < begin(k)) {
        uint32_t lo = 0, hi = k; // the answer is below k
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (begin(mid) <= t) lo = mid;
            else hi = mid;
        }
        k = lo;
    }
    while (k < (uint32_t)a.max_req) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(a.drec + k);
        const uint32_t b = wave_u32(ld_agent(w)), c = wave_u32(ld_agent(w + 1));
        if (t < b + c) break;
        ++k;
    }
    return k;
}

// One tile of a request: trace its rays, write the pixels through to memory,
// count the tile.  r is the record of the request k that holds tile t.
template <int K>
__device__ __forceinline__ void serve_tile(const ServeArgs& a, const ServeRec& r, uint32_t k, uint32_t t, int lane,
                                           uint2* lds_stack) {
    const uint32_t local = t - r.tile_begin;
    const uint32_t grp = local / 64u, gl = local % 64u;
    const int tx = (int)((grp % (uint32_t)r.groups_x) * 8u + gl % 8u);
    const int ty = (int)((grp / (uint32_t)r.groups_x) * 8u + gl / 8u);
    if (tx < r.tiles_x && ty < r.tiles_y && r.scene < (uint32_t)a.n_scenes) {
        const DevScene sc = a.scenes[r.scene];
        const int x = tx * 8 + (lane & 7), y = ty * 8 + (lane >> 3);
        if (x < r.W && y < r.H) {
            float o[3], d[3], rgb[3], A;
            make_ray(r.cam, x, y, o, d);
            // without the look-ahead (with it the 64-view burst takes 54.3
            // instead of 48.6 ms: this kernel holds more state), SH records
            // requested early (47.9 instead of 48.7 ms)
            trace_ray<K, false, false, false, false, false, true>(sc, o, d, a.depth, lds_stack, lane, rgb, A, nullptr);
            uint32_t* px = reinterpret_cast<uint32_t*>(a.arena + r.out_off) + 4 * ((size_t)(r.H - y - 1) * r.W + x);
            st_system(px + 0, __float_as_uint(rgb[0]));
            st_system(px + 1, __float_as_uint(rgb[1]));
            st_system(px + 2, __float_as_uint(rgb[2]));
            st_system(px + 3, __float_as_uint(1.0f - A));
        }
    }
    // the tile's pixels are written through before it counts as finished
    wave_accesses_done();
    if (lane == 0) {
        const uint32_t c = __hip_atomic_fetch_add(&a.done_tiles[k], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c + 1 == r.tile_count) __hip_atomic_store(&a.hdone[k], r.seq + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// The record of the request that holds published tile t, kept across tiles:
// most tiles of a wave belong to the request of its previous tile.
__device__ __forceinline__ bool serve_record(const ServeArgs& a, uint32_t t, uint32_t& k, ServeRec& r, bool& have,
                                             int lane) {
    if (have && t >= r.tile_begin && t - r.tile_begin < r.tile_count) return true;
    k = serve_find(a, t, have ? k : 0u);
    if (k >= (uint32_t)a.max_req) return false; // malformed table: never read past it
    const uint32_t* src = reinterpret_cast<const uint32_t*>(a.drec + k);
    const uint32_t mine = lane < (int)kRecWords ? ld_agent(src + lane) : 0u;
    uint32_t* dst = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
    for (int w = 0; w < (int)kRecWords; ++w) dst[w] = __builtin_amdgcn_readlane(mine, w);
    have = true;
    return true;
}

// A published, unclaimed tile of another queue, taken only if it is
// published (compare-and-swap on the queue's counter), or ~0u.
__device__ __forceinline__ uint32_t serve_steal(const ServeArgs& a, uint32_t home, uint32_t mt, int lane) {
    uint32_t got = ~0u;
    if (lane == 0) {
        for (uint32_t j = 1; j < (uint32_t)kXcds && got == ~0u; ++j) {
            const uint32_t q = (home + j) % kXcds;
            uint32_t n = ld_agent(&a.dctl->queue[q].next);
            while (queue_tile(q, n) < mt) {
                if (__hip_atomic_compare_exchange_strong(&a.dctl->queue[q].next, &n, n + 1u, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    got = queue_tile(q, n);
                    break;
                }
            }
        }
    }
    return wave_u32(got);
}

template <int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_serve(ServeArgs a) {
    const int lane = threadIdx.x;
    extern __shared__ uint2 lds_stack[];
    const uint32_t home = blockIdx.x % kXcds;
    uint32_t k = 0, mt_seen = 0; // the request of the wave's last tile; the last published mark read
    ServeRec r;
    bool have = false;
    uint32_t own = 0;    // the wave's claim on its home queue
    bool holding = false;
    uint32_t spins = 0;
    while (true) {
        if (!holding) {
            // claim the home queue's next tile (one atomic on the queue's line)
            uint32_t c = 0;
            if (lane == 0)
                c = __hip_atomic_fetch_add(&a.dctl->queue[home].next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            own = queue_tile(home, wave_u32(c));
            holding = true;
            spins = 0;
        }
        uint32_t t = ~0u; // the tile to run now
        // marks only grow: a claim below the last mark read is published
        if (own < mt_seen) {
            t = own;
            holding = false;
        } else {
            uint32_t stop = 0;
            if (lane == 0) {
                mt_seen = ld_agent(&a.dctl->mirror_tiles);
                stop = ld_agent(&a.dctl->stop);
            }
            mt_seen = wave_u32(mt_seen);
            stop = wave_u32(stop);
            if (own < mt_seen) {
                t = own;
                holding = false;
            } else if (stop || spins >= a.spin_limit) {
                break;
            } else if ((own % kQueueGroup) == 0 && own / kQueueGroup < mt_seen / kQueueGroup + kXcds) {
                // the first tile of its queue's first unpublished group (at
                // most one per queue, within kXcds groups of the mark)
                // polls the host
                serve_poll_host(a, lane);
                __builtin_amdgcn_s_sleep(2);
                ++spins;
                continue;
            } else {
                // while its own claim waits, a wave runs published tiles that
                // other queues have not handed out yet (the end of a burst)
                t = serve_steal(a, home, mt_seen, lane);
                if (t == ~0u) {
                    // nothing to take: back off up to ~30 us
                    const uint32_t naps = spins < 4 ? 1u : (spins < 64 ? 4u : 16u);
                    for (uint32_t z = 0; z < naps; ++z) __builtin_amdgcn_s_sleep(127);
                    ++spins;
                    continue;
                }
                spins = 0;
            }
        }
        if (!serve_record(a, t, k, r, have, lane)) break;
        serve_tile<K>(a, r, k, t, lane, lds_stack);
    }
}

__global__ void k_serve_reset(ServeDevCtl* c, uint32_t* done_tiles, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        for (int q = 0; q < kXcds; ++q) c->queue[q].next = 0;
        c->mirror_tiles = 0;
        c->mirror_req = 0;
        c->stop = 0;
    }
    for (int j = i; j < n; j += gridDim.x * blockDim.x) done_tiles[j] = 0;
}

typedef void (*BatchKernel)(BatchArgs);

// Dynamic LDS of a launch: per wavefront, one 8-B short-stack ring entry per
// lane and level plus each lane's 64-B SH basis.
static inline size_t stack_lds_bytes(int depth, int block_threads, bool dbg = false) {
    return wave_lds_bytes(depth, dbg) * (size_t)(block_threads / 64);
}

BatchKernel pick_batch_kernel(int k) {
    if (k == 4) return k_batch<4>;
    if (k == 7) return k_batch<7>;
    if (k == 16) return k_batch<16>;
    return k_batch<8>;
}

} // namespace

// ===========================================================================
// Runtime objects.
// ===========================================================================
struct kdgrt_context {
    int device = 0;
    void* noalloc_scratch = nullptr; // device address of `scratch` on this device
    hipDeviceProp_t prop{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::mutex mu;
};

struct kdgrt_scene {
    kdgrt_context* ctx = nullptr;
    DevScene dev{};
    void* d_nodes = nullptr;
    void* d_quads = nullptr;
    void* d_offs = nullptr;
    void* d_sgeom = nullptr;
    void* d_shade = nullptr;
    void* d_lsph = nullptr;
    void* d_gsph = nullptr; // bound spheres per Gaussian (centres for the tile-cost estimate)
    // Cone records of the plain view launches, one table per scene: launches
    // on one stream run in order; a launch on another stream first waits for
    // the last launch that used the table (cone_ev).
    void* d_vcone = nullptr;
    void* d_gcone = nullptr; // the view's cone records per Gaussian (gathered into d_vcone)
    void* d_gsig = nullptr;  // ellipsoid shape matrices per Gaussian, 6 floats each
    void* d_ext = nullptr;   // the caller's Gaussian ID of each device number (debug outputs)
    hipEvent_t cone_ev = nullptr;
    hipStream_t cone_stream = nullptr;
    bool cone_used = false;
    std::mutex cone_mu;
    // Frame-to-frame tile schedule of kdgrt_render and kdgrt_render_hybrid
    // (guarded by cone_mu and ordered by cone_ev like the cone table):
    // d_sched holds sched_cap tile durations of the last launch (the view in
    // sched_cam / sched_w x sched_h / sched_depth, by the kernel of
    // sched_kind), sched_cap smoothed durations and sched_cap tile indices
    // (the order of a launch that follows a close view).
    uint32_t* d_sched = nullptr;
    uint32_t sched_cap = 0;
    bool sched_on = true;
    bool sched_valid = false;
    bool sched_fresh = true; // the smoothed durations are to be restarted from the next costs
    int sched_kind = 0; // 0 plain view kernel, 1 hybrid kernel
    int sched_w = 0, sched_h = 0, sched_depth = 0;
    CamInfo sched_cam{};
    uint64_t bytes = 0;
    uint32_t n_nodes = 0;
    uint64_t n_offsets = 0, n_gaussians = 0;
};

struct kdgrt_batch {
    kdgrt_context* ctx = nullptr;
    kdgrt_tuning tuning{};
    WorkItem* d_items = nullptr;
    uint32_t* d_counters = nullptr;
    BatchArgs args{};
    BatchKernel kernel = nullptr;
    int grid = 0, block = 0;
    uint32_t n_tiles = 0;
    uint64_t n_rays = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    void* graph_stream = nullptr;
    std::vector<void*> d_perms; // ray-key order of the long ray sets
    // the ray-key sort of each long ray set, rerun (eagerly, before the
    // persistent launch) by every kdgrt_batch_run: workspace allocated here
    struct SortJob {
        DevScene sc;
        const float* ro;
        const float* rd;
        uint32_t n;
        uint32_t* perm;
        void* ws;
        size_t temp;
    };
    std::vector<SortJob> sorts;
    // cone records of the batch's large views (the view pre-test), prepared
    // by every run ahead of the persistent launch: one table per view, one
    // per-Gaussian scratch table shared by the preparations in stream order
    struct ConeJob {
        const kdgrt_scene* scene;
        float eye[3];
        float4* table;
    };
    std::vector<ConeJob> cones;
    void* d_gcone = nullptr;
};

using kdgrt::set_error;

#define HIP_TRY(expr)                                                                                           \
    do {                                                                                                        \
        hipError_t e_ = (expr);                                                                                 \
        if (e_ != hipSuccess)                                                                                   \
            return set_error(e_ == hipErrorOutOfMemory ? KDGRT_ERR_OUT_OF_MEMORY : KDGRT_ERR_DEVICE,            \
                             std::string(#expr) + ": " + hipGetErrorString(e_));                                 \
    } while (0)

namespace kdgrt {

// Launch helpers shared with the serving loop.
kdgrt_status batch_build(kdgrt_context* ctx, const kdgrt_view* views, int n_views, const kdgrt_ray_set* sets,
                         int n_sets, const kdgrt_tuning* tuning, kdgrt_batch** out);
kdgrt_status batch_launch(kdgrt_batch* b, hipStream_t s);

// A reusable launch slot for the serving loop: device work table and queue
// counters plus a pinned host staging table, all allocated once, so a launch
// is an async table upload + reset + persistent kernel on the caller's
// stream with no allocation or implicit device synchronisation.
struct BatchSlot {
    kdgrt_context* ctx = nullptr;
    int max_items = 0;
    WorkItem* h_items = nullptr; // pinned
    WorkItem* d_items = nullptr;
    uint32_t* d_counters = nullptr;
};

} // namespace kdgrt


// ---------------------------------------------------------------------------
// Bound-sphere records for the conservative pre-test (trace_device.hpp,
// sphere_pass).  For a Gaussian with inverse scales g and rotation matrix M
// (the device's float computation from the quaternion), normalised
// coordinates are x' = A (x - mu) with A = diag(g) M.  A ray the exact test
// accepts passes within sqrt(8) normalised units, i.e. within
// sqrt(8) / sigma_min(A) world units of mu.  The t test uses world distances
// only when M is orthogonal, so a non-orthogonal or singular M gets r = inf.
// ---------------------------------------------------------------------------
static void sym3_eigen(double a[3][3], double ev[3]) {
    // cyclic Jacobi on a symmetric 3x3 matrix
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        if (off < 1e-30 * (a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2]) || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
            }
    }
    for (int k = 0; k < 3; ++k) ev[k] = a[k][k];
}

// quaternionWXYZToMatrixTranspose (cudaRenderer.cu:808-828), f32, q = (w, x, y, z).
static void quat_to_mt(const float4& q, float m[3][3]) {
    const float qr = q.x, qx = q.y, qy = q.z, qz = q.w;
    const float xx = qx * qx, yy = qy * qy, zz = qz * qz;
    const float xy = qx * qy, xz = qx * qz, yz = qy * qz;
    const float rx = qr * qx, ry = qr * qy, rz = qr * qz;
    m[0][0] = (1.f - 2.f * (yy + zz));
    m[0][1] = 2.f * (xy + rz);
    m[0][2] = 2.f * (xz - ry);
    m[1][0] = 2.f * (xy - rz);
    m[1][1] = (1.f - 2.f * (xx + zz));
    m[1][2] = 2.f * (yz + rx);
    m[2][0] = 2.f * (xz + ry);
    m[2][1] = 2.f * (yz - rx);
    m[2][2] = (1.f - 2.f * (xx + yy));
}

// Also writes S = 8 (A^T A)^-1, the shape matrix of the ellipsoid the exact
// test accepts (x - mu)^T S^-1 (x - mu) <= 1, as (S00, S01, S02, S11, S12,
// S22) to sig6, or NaN where r = inf.
static float4 bound_sphere(const float4& a, const float4& b, const float4& q, double* sig6) {
    for (int k = 0; k < 6; ++k) sig6[k] = NAN;
    float mf[3][3];
    quat_to_mt(q, mf);
    const double g[3] = {b.x, b.y, b.z};
    double r = INFINITY, kap = 0.0;
    bool ok = std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z);
    double ortho = 0.0, A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double mtm = 0.0;
            for (int k = 0; k < 3; ++k) mtm += (double)mf[k][i] * (double)mf[k][j];
            ortho = std::max(ortho, std::fabs(mtm - (i == j ? 1.0 : 0.0)));
            A[i][j] = g[i] * (double)mf[i][j];
            ok = ok && std::isfinite(A[i][j]);
        }
    if (ok && ortho <= 4e-6) {
        double AtA[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s2 = 0.0;
                for (int k = 0; k < 3; ++k) s2 += A[k][i] * A[k][j];
                AtA[i][j] = s2;
            }
        double ev[3];
        sym3_eigen(AtA, ev);
        const double lmin = std::min(ev[0], std::min(ev[1], ev[2]));
        const double lmax = std::max(ev[0], std::max(ev[1], ev[2]));
        if (lmin > 0.0 && std::isfinite(lmax)) {
            const double smin = std::sqrt(lmin) * (1.0 - 1e-9), smax = std::sqrt(lmax);
            r = std::sqrt(8.0) / smin * (1.0 + 1e-3);
            kap = 8e-6 * (4.0 + smax / smin);
            // 8 (A^T A)^-1 by the adjugate (A^T A is recomputed: the eigen
            // solve overwrote it)
            double P[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double s2 = 0.0;
                    for (int k = 0; k < 3; ++k) s2 += A[k][i] * A[k][j];
                    P[i][j] = s2;
                }
            const double c00 = P[1][1] * P[2][2] - P[1][2] * P[2][1], c01 = P[0][2] * P[2][1] - P[0][1] * P[2][2],
                         c02 = P[0][1] * P[1][2] - P[0][2] * P[1][1], c11 = P[0][0] * P[2][2] - P[0][2] * P[2][0],
                         c12 = P[0][2] * P[1][0] - P[0][0] * P[1][2], c22 = P[0][0] * P[1][1] - P[0][1] * P[1][0];
            const double det = P[0][0] * c00 + P[0][1] * (P[1][2] * P[2][0] - P[1][0] * P[2][2]) +
                               P[0][2] * (P[1][0] * P[2][1] - P[1][1] * P[2][0]);
            if (det > 0.0 && std::isfinite(det)) {
                const double f = 8.0 / det;
                const double S6[6] = {f * c00, f * c01, f * c02, f * c11, f * c12, f * c22};
                bool fin = true;
                for (int k = 0; k < 6; ++k) fin = fin && std::isfinite(S6[k]);
                if (fin)
                    for (int k = 0; k < 6; ++k) sig6[k] = S6[k];
            }
        }
    }
    if (!std::isfinite(r)) {
        kap = 0.0;
        for (int k = 0; k < 6; ++k) sig6[k] = NAN;
    }
    // (float) of a double may round down: the f16 rounding starts from the
    // next float up in that case
    const auto up_f32 = [](double v) {
        const float f = (float)v;
        return (double)f < v ? std::nextafter(f, INFINITY) : f;
    };
    const uint32_t packed = (uint32_t)kdgrt::f32_to_f16_up(up_f32(std::min(r, 1e30))) |
                            ((uint32_t)kdgrt::f32_to_f16_up(up_f32(std::min(kap, 1e30))) << 16);
    float w;
    std::memcpy(&w, &packed, 4);
    return make_float4(a.x, a.y, a.z, w);
}

// Runs fn(begin, end) over [0, n) in contiguous chunks on up to 16 host
// threads (the per-Gaussian repack of a multi-million-Gaussian scene).
template <class F>
static void parallel_ranges(uint64_t n, F&& fn) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint64_t chunks = std::min<uint64_t>(hw, (n + 65535) / 65536);
    if (chunks <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (uint64_t c = 0; c < chunks; ++c) th.emplace_back([&, c] { fn(n * c / chunks, n * (c + 1) / chunks); });
    for (auto& t : th) t.join();
}

extern "C" {

kdgrt_status kdgrt_device_count(int* count) {
    if (!count) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "device_count: null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return KDGRT_OK;
}

kdgrt_status kdgrt_init(int device, kdgrt_context** out) {
    if (!out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "init: null");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return set_error(KDGRT_ERR_NO_DEVICE, "init: no HIP device");
    if (device < 0 || device >= n) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "init: bad device index");
    HIP_TRY(hipSetDevice(device));
    // The address of the static scratch array on this device, resolved once
    // per device with the device current; every context on the device keeps it.
    static std::mutex scratch_mu;
    static void* scratch_ptr[kMaxDevices] = {};
    if (device >= kMaxDevices) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "init: device index above 63");
    void* sp = nullptr;
    {
        std::lock_guard<std::mutex> lk(scratch_mu);
        if (!scratch_ptr[device]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return set_error(KDGRT_ERR_DEVICE, std::string("init: hipGetSymbolAddress(scratch): ") +
                                                       hipGetErrorString(e));
            scratch_ptr[device] = p;
        }
        sp = scratch_ptr[device];
    }
    auto* c = new kdgrt_context();
    c->device = device;
    c->noalloc_scratch = sp;
    if (hipGetDeviceProperties(&c->prop, device) != hipSuccess || hipEventCreate(&c->ev0) != hipSuccess ||
        hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return set_error(KDGRT_ERR_DEVICE, "init: device query failed");
    }
    *out = c;
    return KDGRT_OK;
}

kdgrt_status kdgrt_shutdown(kdgrt_context* ctx) {
    if (!ctx) return KDGRT_OK;
    hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    delete ctx;
    return KDGRT_OK;
}

kdgrt_status kdgrt_device_info(kdgrt_context* ctx, char* name, size_t name_len, int* n_cu, size_t* total_mem,
                               size_t* lds_per_block) {
    if (!ctx) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "device_info: null");
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", ctx->prop.name, ctx->prop.gcnArchName);
    }
    if (n_cu) *n_cu = ctx->prop.multiProcessorCount;
    if (total_mem) *total_mem = ctx->prop.totalGlobalMem;
    if (lds_per_block) *lds_per_block = ctx->prop.sharedMemPerBlock;
    return KDGRT_OK;
}

kdgrt_status kdgrt_device_alloc(kdgrt_context* ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "device_alloc: null");
    HIP_TRY(hipSetDevice(ctx->device));
    *dptr = nullptr;
    if (bytes == 0) return KDGRT_OK;
    HIP_TRY(hipMalloc(dptr, bytes));
    return KDGRT_OK;
}
kdgrt_status kdgrt_device_free(kdgrt_context* ctx, void* dptr) {
    if (!ctx) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "device_free: null");
    if (!dptr) return KDGRT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipFree(dptr));
    return KDGRT_OK;
}
kdgrt_status kdgrt_memcpy_h2d(kdgrt_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "memcpy_h2d: null");
    HIP_TRY(hipSetDevice(ctx->device));
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return KDGRT_OK;
}
kdgrt_status kdgrt_memcpy_d2h(kdgrt_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "memcpy_d2h: null");
    HIP_TRY(hipSetDevice(ctx->device));
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return KDGRT_OK;
}
kdgrt_status kdgrt_memset(kdgrt_context* ctx, void* dptr, int value, size_t bytes) {
    if (!ctx || (bytes && !dptr)) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "memset: null");
    HIP_TRY(hipSetDevice(ctx->device));
    if (bytes) HIP_TRY(hipMemset(dptr, value, bytes));
    return KDGRT_OK;
}
kdgrt_status kdgrt_stream_create(kdgrt_context* ctx, void** stream) {
    if (!ctx || !stream) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "stream_create: null");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return KDGRT_OK;
}
kdgrt_status kdgrt_stream_destroy(kdgrt_context* ctx, void* stream) {
    if (!ctx) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "stream_destroy: null");
    if (!stream) return KDGRT_OK;
    HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return KDGRT_OK;
}
kdgrt_status kdgrt_stream_sync(kdgrt_context* ctx, void* stream) {
    if (!ctx) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "stream_sync: null");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return KDGRT_OK;
}
kdgrt_status kdgrt_device_sync(kdgrt_context* ctx) {
    if (!ctx) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "device_sync: null");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    return KDGRT_OK;
}
kdgrt_status kdgrt_timer_begin(kdgrt_context* ctx, void* stream) {
    if (!ctx) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "timer_begin: null");
    HIP_TRY(hipEventRecord(ctx->ev0, (hipStream_t)stream));
    return KDGRT_OK;
}
kdgrt_status kdgrt_timer_end(kdgrt_context* ctx, void* stream, float* ms) {
    if (!ctx || !ms) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "timer_end: null");
    HIP_TRY(hipEventRecord(ctx->ev1, (hipStream_t)stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    HIP_TRY(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return KDGRT_OK;
}

kdgrt_status kdgrt_copy_f32x4(kdgrt_context* ctx, void* dst, const void* src, uint64_t n_float4, void* stream) {
    if (!ctx || (n_float4 && (!dst || !src))) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "copy: null");
    if (n_float4 == 0) return KDGRT_OK;
    const int n_cu = ctx->prop.multiProcessorCount > 0 ? ctx->prop.multiProcessorCount : 256;
    const uint64_t want = (n_float4 + 256 * kCopyUnroll - 1) / (256 * kCopyUnroll);
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)n_cu * 16));
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, (hipStream_t)stream, (kd_f4*)dst, (const kd_f4*)src,
                       n_float4);
    HIP_TRY(hipGetLastError());
    return KDGRT_OK;
}

kdgrt_status kdgrt_valu_probe(kdgrt_context* ctx, int iters, void* stream, uint64_t* waves, uint64_t* fma_per_wave) {
    if (!ctx || iters < 1 || iters > (1 << 24)) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "valu_probe: iters 1..2^24");
    // eight waves per SIMD on every CU: blocks of four waves, two per SIMD each
    const int blocks = std::max(1, ctx->prop.multiProcessorCount) * 8;
    float* out = nullptr;
    HIP_TRY(hipMallocAsync((void**)&out, (size_t)blocks * 256 * sizeof(float), (hipStream_t)stream));
    hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, iters, 0.999f, 1e-3f);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipFreeAsync(out, (hipStream_t)stream));
    if (waves) *waves = (uint64_t)blocks * 4;
    if (fma_per_wave) *fma_per_wave = (uint64_t)kValuChains * (uint64_t)iters;
    return KDGRT_OK;
}

kdgrt_status kdgrt_framebuffer_to_u8(kdgrt_context* ctx, const float* fb, int width, int height, int in_channels,
                                     uint8_t* out, int out_channels, int flip_rows, void* stream) {
    if (!ctx || !fb || !out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "framebuffer_to_u8: null");
    if (width <= 0 || height <= 0 || width > 32768 || height > 32768)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "framebuffer_to_u8: bad size");
    if ((in_channels != 3 && in_channels != 4) || (out_channels != 3 && out_channels != 4))
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "framebuffer_to_u8: channels must be 3 or 4");
    const uint64_t n = (uint64_t)width * height;
    hipLaunchKernelGGL(k_to_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, fb, width,
                       height, in_channels, out, out_channels, flip_rows ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return KDGRT_OK;
}

uint64_t kdgrt_scene_bytes_estimate(uint32_t n_nodes, uint64_t n_offsets, uint64_t n_gaussians) {
    auto r = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
    return r(8ull * n_nodes) + r(48ull * n_nodes) + r(4ull * n_offsets) + r(80ull * std::max<uint64_t>(1, n_offsets)) +
           r(192ull * n_gaussians) + 2 * r(16ull * (n_offsets + kLsphPad)) + r(16ull * n_gaussians) +
           r(16ull * n_gaussians) + r(24ull * n_gaussians) + r(4ull * n_gaussians);
}

kdgrt_status kdgrt_scene_create(kdgrt_context* ctx, const kdgrt_node* nodes, uint32_t n_nodes,
                                const uint32_t* offsets, uint64_t n_offsets, const float aabb[6],
                                const kdgrt_gaussian* g, uint64_t n_g, int scales_inverted, kdgrt_scene** out) {
    if (!ctx || !out || !aabb || (!g && n_g) || (!offsets && n_offsets))
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "scene_create: null argument");
    if (n_g == 0) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "scene_create: no Gaussians");
    kdgrt_status st = kdgrt::validate_tree(nodes, n_nodes, n_offsets, offsets, n_g);
    if (st != KDGRT_OK) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    // Device numbering of the Gaussians: by first appearance in the leaf
    // lists (Gaussians no leaf lists follow in their own order).  Leaves are
    // stored depth first, so the records of Gaussians that neighbouring rays
    // meet lie close together; the composite order is the leaf-slot order and
    // does not depend on the numbering, and the debug outputs map back to the
    // caller's IDs (ext).
    std::vector<uint32_t> order(n_g), rank(n_g, UINT32_MAX), ioffs(n_offsets);
    {
        uint32_t next = 0;
        for (uint64_t k = 0; k < n_offsets; ++k) {
            const uint32_t id = offsets[k];
            if (rank[id] == UINT32_MAX) {
                rank[id] = next;
                order[next++] = id;
            }
            ioffs[k] = rank[id];
        }
        for (uint64_t i = 0; i < n_g; ++i)
            if (rank[i] == UINT32_MAX) {
                rank[i] = next;
                order[next++] = (uint32_t)i;
            }
    }
    // Repack (device numbering): 80-B geometry records (built per Gaussian,
    // stored per leaf slot), 192-B SH records, 16-B bound spheres.
    std::vector<float4> geom(5 * n_g), shade(12 * n_g), sph(n_g);
    std::vector<double> sig(6 * n_g); // ellipsoid shape matrices of the view cone bound
    std::vector<float> sigf(6 * n_g);
    // leaf-slot spheres plus kLsphPad zero records: the pre-test reads whole
    // groups of four past a leaf's end and masks the extra bits off
    std::vector<float4> lsph(n_offsets + kLsphPad, make_float4(0.f, 0.f, 0.f, 0.f));
    auto repack = [&](uint64_t i0, uint64_t i1) {
    for (uint64_t i = i0; i < i1; ++i) {
        const kdgrt_gaussian& s = g[order[i]];
        float inv[3], kinv;
        if (scales_inverted) {
            inv[0] = s.scale[0];
            inv[1] = s.scale[1];
            inv[2] = s.scale[2];
            kinv = s.k_scale;
        } else {
            inv[0] = 1 / s.scale[0];
            inv[1] = 1 / s.scale[1];
            inv[2] = 1 / s.scale[2];
            kinv = 1 / s.k_scale;
        }
        // Geometry record: pos|opacity, inverse scale|pscale.x, pscale.y,
        // pscale.z, M00, M01 | M02, M10, M11, M12 | M20, M21, M22, 0 with
        // pscale = 1 / (inverse scale) and M the transposed rotation of the
        // quaternion, in the device's f32 operations (cudaRenderer.cu:808-828,
        // 1090): computed once here instead of per test.
        const float4 q = make_float4(s.rot[0], s.rot[1], s.rot[2], s.rot[3]);
        float m[3][3];
        quat_to_mt(q, m);
        const float ps[3] = {1 / inv[0], 1 / inv[1], 1 / inv[2]};
        geom[5 * i + 0] = make_float4(s.pos[0], s.pos[1], s.pos[2], s.opacity);
        geom[5 * i + 1] = make_float4(inv[0], inv[1], inv[2], ps[0]);
        geom[5 * i + 2] = make_float4(ps[1], ps[2], m[0][0], m[0][1]);
        geom[5 * i + 3] = make_float4(m[0][2], m[1][0], m[1][1], m[1][2]);
        geom[5 * i + 4] = make_float4(m[2][0], m[2][1], m[2][2], 0.0f);
        (void)kinv;
        float* sh = reinterpret_cast<float*>(&shade[12 * i]); // channel-major SH record
        for (int ch = 0; ch < 3; ++ch) {
            sh[16 * ch] = s.f_dc[ch];
            for (int k = 0; k < 15; ++k) sh[16 * ch + 1 + k] = s.f_rest[15 * ch + k];
        }
        sph[i] = bound_sphere(make_float4(s.pos[0], s.pos[1], s.pos[2], s.opacity),
                              make_float4(inv[0], inv[1], inv[2], kinv), q, &sig[6 * i]);
    }
    };
    parallel_ranges(n_g, repack);
    parallel_ranges(n_offsets, [&](uint64_t k0, uint64_t k1) {
        for (uint64_t k = k0; k < k1; ++k) lsph[k] = sph[ioffs[k]]; // leaf-slot order
    });
    auto* sc = new kdgrt_scene();
    sc->ctx = ctx;
    auto fail = [&](hipError_t e, const char* what) {
        if (sc->d_nodes) hipFree(sc->d_nodes);
        if (sc->d_quads) hipFree(sc->d_quads);
        if (sc->d_offs) hipFree(sc->d_offs);
        if (sc->d_sgeom) hipFree(sc->d_sgeom);
        if (sc->d_shade) hipFree(sc->d_shade);
        if (sc->d_lsph) hipFree(sc->d_lsph);
        if (sc->d_gsph) hipFree(sc->d_gsph);
        if (sc->d_vcone) hipFree(sc->d_vcone);
        if (sc->d_gsig) hipFree(sc->d_gsig);
        if (sc->d_gcone) hipFree(sc->d_gcone);
        if (sc->d_ext) hipFree(sc->d_ext);
        if (sc->cone_ev) hipEventDestroy(sc->cone_ev);
        delete sc;
        return set_error(e == hipErrorOutOfMemory ? KDGRT_ERR_OUT_OF_MEMORY : KDGRT_ERR_DEVICE,
                         std::string("scene_create: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc(&sc->d_nodes, 8ull * n_nodes)) != hipSuccess) return fail(e, "nodes");
    if ((e = hipMalloc(&sc->d_offs, std::max<uint64_t>(4, 4ull * n_offsets))) != hipSuccess) return fail(e, "offsets");
    if ((e = hipMalloc(&sc->d_shade, 192ull * n_g)) != hipSuccess) return fail(e, "sh");
    if ((e = hipMalloc(&sc->d_lsph, 16ull * lsph.size())) != hipSuccess) return fail(e, "leaf spheres");
    if ((e = hipMalloc(&sc->d_gsph, 16ull * std::max<uint64_t>(1, n_g))) != hipSuccess) return fail(e, "spheres");
    if (n_g && (e = hipMemcpy(sc->d_gsph, sph.data(), 16ull * n_g, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "copy");
    if ((e = hipMalloc(&sc->d_vcone, 16ull * lsph.size())) != hipSuccess) return fail(e, "view cones");
    if ((e = hipMalloc(&sc->d_gcone, 16ull * std::max<uint64_t>(1, n_g))) != hipSuccess) return fail(e, "view cones");
    for (uint64_t k = 0; k < 6 * n_g; ++k) sigf[k] = (float)sig[k];
    if ((e = hipMalloc(&sc->d_gsig, 24ull * std::max<uint64_t>(1, n_g))) != hipSuccess) return fail(e, "shapes");
    if ((e = hipMalloc(&sc->d_ext, 4ull * std::max<uint64_t>(1, n_g))) != hipSuccess) return fail(e, "IDs");
    if (n_g && (e = hipMemcpy(sc->d_ext, order.data(), 4ull * n_g, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "copy");
    if (n_g && (e = hipMemcpy(sc->d_gsig, sigf.data(), 24ull * n_g, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "copy");
    if ((e = hipEventCreateWithFlags(&sc->cone_ev, hipEventDisableTiming)) != hipSuccess) return fail(e, "event");
    if ((e = hipMemcpy(sc->d_lsph, lsph.data(), 16ull * lsph.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "copy");
    if ((e = hipMemcpy(sc->d_nodes, nodes, 8ull * n_nodes, hipMemcpyHostToDevice)) != hipSuccess) return fail(e, "copy");
    {
        // quad records: for the child pair at c (c = first child of an inner
        // node), nodes c and c+1 and the child pairs of each (zero for a leaf)
        std::vector<uint4> quads(3 * (size_t)n_nodes, make_uint4(0, 0, 0, 0));
        auto pair_at = [&](uint32_t c) {
            return make_uint4(nodes[c].x, nodes[c].y, nodes[c + 1].x, nodes[c + 1].y);
        };
        parallel_ranges(n_nodes, [&](uint64_t p0, uint64_t p1) {
            for (uint64_t p = p0; p < p1; ++p) {
                if ((nodes[p].x & 7u) == 3u) continue;
                const uint32_t c = nodes[p].x >> 3;
                quads[3 * (size_t)c] = pair_at(c);
                for (int k = 0; k < 2; ++k) {
                    const kdgrt_node& n = nodes[c + k];
                    if ((n.x & 7u) != 3u) quads[3 * (size_t)c + 1 + k] = pair_at(n.x >> 3);
                }
            }
        });
        if ((e = hipMalloc(&sc->d_quads, 16ull * quads.size())) != hipSuccess) return fail(e, "node quads");
        if ((e = hipMemcpy(sc->d_quads, quads.data(), 16ull * quads.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "copy");
    }
    if (n_offsets && (e = hipMemcpy(sc->d_offs, ioffs.data(), 4ull * n_offsets, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "copy");
    {
        // geometry records in leaf-slot order, the device number of the
        // Gaussian in the free last word
        std::vector<float4> sgeom(5 * std::max<uint64_t>(1, n_offsets));
        parallel_ranges(n_offsets, [&](uint64_t k0, uint64_t k1) {
            for (uint64_t k = k0; k < k1; ++k) {
                for (int q = 0; q < 5; ++q) sgeom[5 * k + q] = geom[5 * (uint64_t)ioffs[k] + q];
                float w;
                std::memcpy(&w, &ioffs[k], 4);
                sgeom[5 * k + 4].w = w;
            }
        });
        if ((e = hipMalloc(&sc->d_sgeom, 16ull * sgeom.size())) != hipSuccess) return fail(e, "slot geometry");
        if ((e = hipMemcpy(sc->d_sgeom, sgeom.data(), 16ull * sgeom.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "copy");
    }
    if ((e = hipMemcpy(sc->d_shade, shade.data(), 192ull * n_g, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "copy");
    sc->dev.nodes = (const uint2*)sc->d_nodes;
    sc->dev.quads = (const uint4*)sc->d_quads;
    sc->dev.offs = (const uint32_t*)sc->d_offs;
    sc->dev.ext = (const uint32_t*)sc->d_ext;
    sc->dev.sgeom = (const float4*)sc->d_sgeom;
    sc->dev.shade = (const float4*)sc->d_shade;
    sc->dev.lsph = (const float4*)sc->d_lsph;
    for (int k = 0; k < 3; ++k) {
        sc->dev.bmin[k] = aabb[k];
        sc->dev.bmax[k] = aabb[3 + k];
    }
    sc->n_nodes = n_nodes;
    sc->n_offsets = n_offsets;
    sc->n_gaussians = n_g;
    sc->bytes = kdgrt_scene_bytes_estimate(n_nodes, n_offsets, n_g);
    *out = sc;
    return KDGRT_OK;
}

kdgrt_status kdgrt_scene_create_from_tree(kdgrt_context* ctx, const kdgrt_tree* tree, const kdgrt_gaussian* g,
                                          uint64_t n_g, int scales_inverted, kdgrt_scene** out) {
    if (!tree) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "scene_create_from_tree: null tree");
    return kdgrt_scene_create(ctx, tree->nodes.data(), (uint32_t)tree->nodes.size(), tree->offsets.data(),
                              tree->offsets.size(), tree->aabb, g, n_g, scales_inverted, out);
}

kdgrt_status kdgrt_scene_destroy(kdgrt_scene* sc) {
    if (!sc) return KDGRT_OK;
    hipSetDevice(sc->ctx->device);
    hipFree(sc->d_nodes);
    hipFree(sc->d_quads);
    hipFree(sc->d_offs);
    hipFree(sc->d_sgeom);
    hipFree(sc->d_shade);
    hipFree(sc->d_lsph);
    hipFree(sc->d_gsph);
    hipFree(sc->d_vcone);
    hipFree(sc->d_gcone);
    hipFree(sc->d_gsig);
    hipFree(sc->d_ext);
    hipFree(sc->d_sched);
    hipEventDestroy(sc->cone_ev);
    delete sc;
    return KDGRT_OK;
}

kdgrt_status kdgrt_scene_set_tile_schedule(kdgrt_scene* sc, int enable) {
    if (!sc) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "scene_set_tile_schedule: null scene");
    std::lock_guard<std::mutex> lock(sc->cone_mu);
    sc->sched_on = enable != 0;
    sc->sched_valid = false;
    return KDGRT_OK;
}

kdgrt_status kdgrt_scene_device_bytes(const kdgrt_scene* sc, uint64_t* bytes) {
    if (!sc || !bytes) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "scene_device_bytes: null");
    *bytes = sc->bytes;
    return KDGRT_OK;
}

static kdgrt_status check_view(const kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W,
                               int H, int depth, const void* out) {
    if (!ctx || !sc || !cam || !out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render: null argument");
    if (sc->ctx != ctx) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render: scene belongs to another context");
    if (W <= 0 || H <= 0 || W > 32768 || H > 32768) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render: bad size");
    if (depth < 1 || depth > 32) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render: stack depth must be 1..32");
    return KDGRT_OK;
}

static void fill_view_args(ViewArgs& a, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W, int H,
                           int depth, void* out) {
    a.prof = nullptr;
    a.sc = sc->dev;
    std::memcpy(&a.cam, cam, sizeof(CamInfo));
    a.W = W;
    a.H = H;
    a.depth = depth;
    a.blocks_x = (W + 63) / 64; // 64x64-pixel groups of 64 8x8 tiles
    a.blocks_y = (H + 63) / 64;
    // A view with fewer 8x8 tiles than half the chip's wave slots leaves the
    // chip idle and its time is that of its slowest tile: 4x4 tiles (16 rays
    // per wave, four times the waves) shorten the slowest tile instead
    // (320x240: 1.09 -> 0.85 ms).  Groups of 64 tiles are then 32x32 px.
    const uint64_t slots = (uint64_t)std::max(1, sc->ctx->prop.multiProcessorCount) * 16u;
    a.tile4 = 0;
    a.order = nullptr;
    a.cost = nullptr;
    if ((uint64_t)((W + 7) / 8) * (uint64_t)((H + 7) / 8) * 2u < slots) {
        a.tile4 = 1;
        a.blocks_x = (W + 31) / 32;
        a.blocks_y = (H + 31) / 32;
    }
    a.out = out;
    a.counters = nullptr;
    a.hashes = nullptr;
    a.seq = nullptr;
    a.seq_cap = 0;
    for (int i = 0; i < 6; ++i) a.touch[i] = nullptr;
}

// Logical tiles of a view launch (the grid of k_view and k_hybrid).
static uint32_t view_tiles(const ViewArgs& a) {
    return interleave_grid<64>((uint64_t)a.blocks_x * a.blocks_y * 64);
}

// Whether the previous view of the scene is close enough to this one for its
// tile durations to order this launch: same kernel, size and stack depth,
// image axes and centre ray within 0.05 degrees, eye moved by less than
// 0.1 % of its distance to the scene box centre, pixel footprint within
// 0.1 %.  Any other launch is ordered by the projected density of Gaussian
// centres, which measured better as soon as the camera moves: orbiting the
// 1M-Gaussian scene by 0.25 degrees per frame, durations 1055 Mrays/s,
// density about 1125; for a camera that stays put, durations 1145, density
// about 1125.
static bool sched_similar(const kdgrt_scene* sc, const ViewArgs& a, int kind) {
    if (!sc->sched_valid || sc->sched_kind != kind || sc->sched_w != a.W || sc->sched_h != a.H ||
        sc->sched_depth != a.depth)
        return false;
    const CamInfo& p = sc->sched_cam;
    const CamInfo& c = a.cam;
    auto centre_ray = [&](const CamInfo& k, double r[3]) {
        for (int i = 0; i < 3; ++i)
            r[i] = (double)k.start[i] + (double)k.u[i] * k.step_x * (0.5 * a.W) -
                   (double)k.v[i] * k.step_y * (0.5 * a.H) - (double)k.eye[i];
    };
    auto cos_angle = [](const double* x, const double* y) {
        double xy = 0, xx = 0, yy = 0;
        for (int i = 0; i < 3; ++i) {
            xy += x[i] * y[i];
            xx += x[i] * x[i];
            yy += y[i] * y[i];
        }
        return (xx > 0 && yy > 0) ? xy / std::sqrt(xx * yy) : -1.0;
    };
    const double cos_lim = 0.99999961922; // cos(0.05 degrees)
    double r0[3], r1[3], u0[3], u1[3], v0[3], v1[3], de2 = 0, dist2 = 0;
    centre_ray(p, r0);
    centre_ray(c, r1);
    for (int i = 0; i < 3; ++i) {
        u0[i] = p.u[i];
        u1[i] = c.u[i];
        v0[i] = p.v[i];
        v1[i] = c.v[i];
        const double de = (double)c.eye[i] - (double)p.eye[i];
        const double dc = 0.5 * ((double)sc->dev.bmin[i] + (double)sc->dev.bmax[i]) - (double)c.eye[i];
        de2 += de * de;
        dist2 += dc * dc;
    }
    if (cos_angle(r0, r1) < cos_lim || cos_angle(u0, u1) < cos_lim || cos_angle(v0, v1) < cos_lim) return false;
    if (de2 > 1e-6 * dist2) return false;
    auto footprint = [](const float* w, float step) {
        return std::sqrt((double)w[0] * w[0] + (double)w[1] * w[1] + (double)w[2] * w[2]) * std::fabs((double)step);
    };
    const double fx0 = footprint(p.u, p.step_x), fx1 = footprint(c.u, c.step_x);
    const double fy0 = footprint(p.v, p.step_y), fy1 = footprint(c.v, c.step_y);
    return std::fabs(fx1 - fx0) <= 1e-3 * fx0 && std::fabs(fy1 - fy0) <= 1e-3 * fy0;
}

// Schedule state for a launch of n logical tiles by the kernel of `kind`
// (inside with_cone_table, before k_cone_prep on stream s).  When the
// scene's last launch was of a view close to this one, `ts` asks k_cone_prep
// to order this launch by that launch's tile durations; otherwise k_density
// counts the Gaussian centres that project into each cell of the image and
// `ts` asks for an order by that density.  Either way the launch records its
// own durations (a.cost).  `use` is false when the schedule is off or the
// launch is too large for the one-workgroup sort.
// Schedule buffers for launches of up to n logical tiles (with cone_mu held).
static kdgrt_status sched_grow(kdgrt_scene* sc, uint32_t n) {
    if (n <= sc->sched_cap) return KDGRT_OK;
    // earlier launches of this scene (ordered by cone_ev) may still use the
    // old buffer
    if (sc->cone_used) HIP_TRY(hipEventSynchronize(sc->cone_ev));
    HIP_TRY(hipFree(sc->d_sched));
    sc->d_sched = nullptr;
    sc->sched_cap = 0;
    sc->sched_valid = false;
    HIP_TRY(hipMalloc((void**)&sc->d_sched, (3ull * n + kDensCells) * sizeof(uint32_t)));
    HIP_TRY(hipMemset(sc->d_sched + 3ull * n, 0, kDensCells * sizeof(uint32_t)));
    sc->sched_cap = n;
    return KDGRT_OK;
}

static kdgrt_status sched_before(kdgrt_scene* sc, ViewArgs& a, uint32_t n, int kind, bool may_alloc, hipStream_t s,
                                 bool& use, TileSort& ts) {
    use = sc->sched_on && n <= kSchedMaxTiles && (may_alloc || n <= sc->sched_cap);
    ts = TileSort{};
    a.order = nullptr;
    a.cost = nullptr;
    if (!use) return KDGRT_OK;
    if (n > sc->sched_cap) {
        const kdgrt_status st = sched_grow(sc, n);
        if (st != KDGRT_OK) return st;
    }
    uint32_t* cost = sc->d_sched;
    uint32_t* smooth = sc->d_sched + sc->sched_cap;
    uint32_t* order = sc->d_sched + 2ull * sc->sched_cap;
    uint32_t* dens = sc->d_sched + 3ull * sc->sched_cap;
    ts.cost = cost;
    ts.smooth = smooth;
    ts.order = order;
    ts.dens = dens;
    ts.n = n;
    a.cost = cost;
    if (sched_similar(sc, a, kind)) {
        ts.reset = sc->sched_fresh ? 1 : 0;
        ts.on = 1;
        sc->sched_fresh = false;
        a.order = order;
        return KDGRT_OK;
    }
    sc->sched_fresh = true; // the smoothed durations belong to another view
    // density estimate: tiles of te pixels, cells of 2^cell_shift tiles
    const int te = a.tile4 ? 4 : 8;
    const int tw = (a.W + te - 1) / te, th = (a.H + te - 1) / te;
    int cs = 1;
    while ((uint64_t)((tw + (1 << cs) - 1) >> cs) * (uint64_t)((th + (1 << cs) - 1) >> cs) > kDensCells) ++cs;
    const CamInfo& c = a.cam;
    DensityArgs d;
    const double nx = (double)c.u[1] * c.v[2] - (double)c.u[2] * c.v[1];
    const double ny = (double)c.u[2] * c.v[0] - (double)c.u[0] * c.v[2];
    const double nz = (double)c.u[0] * c.v[1] - (double)c.u[1] * c.v[0];
    double sn = 0, uu = 0, vv = 0;
    for (int k = 0; k < 3; ++k) {
        d.eye[k] = c.eye[k];
        d.se[k] = c.start[k] - c.eye[k];
        uu += (double)c.u[k] * c.u[k];
        vv += (double)c.v[k] * c.v[k];
    }
    d.n[0] = (float)nx;
    d.n[1] = (float)ny;
    d.n[2] = (float)nz;
    sn = (double)d.se[0] * nx + (double)d.se[1] * ny + (double)d.se[2] * nz;
    if (!(std::fabs(sn) > 0.0) || !(uu > 0.0) || !(vv > 0.0) || c.step_x == 0.0f || c.step_y == 0.0f)
        return KDGRT_OK; // degenerate camera: raster order
    d.sn = (float)sn;
    for (int k = 0; k < 3; ++k) {
        d.pu[k] = (float)(c.u[k] / (uu * c.step_x));
        d.pv[k] = (float)(c.v[k] / (vv * c.step_y));
    }
    d.W = a.W;
    d.H = a.H;
    d.cells_w = (tw + (1 << cs) - 1) >> cs;
    d.cells_h = (th + (1 << cs) - 1) >> cs;
    d.cell_px = 1.0f / (float)(te << cs);
    const uint64_t per_block = 4ull * kDensBlock; // centres per workgroup, at most 512 workgroups
    const uint32_t grid = (uint32_t)std::min<uint64_t>((sc->n_gaussians + per_block - 1) / per_block, 512);
    if (sc->n_gaussians)
        hipLaunchKernelGGL(k_density, dim3(grid), dim3(kDensBlock), 0, s, (const float4*)sc->d_gsph, sc->n_gaussians,
                           d, dens);
    ts.on = 1;
    ts.density = 1;
    ts.blocks_x = a.blocks_x;
    ts.tiles_w = tw;
    ts.tiles_h = th;
    ts.cell_shift = cs;
    ts.cells_w = d.cells_w;
    ts.cells_h = d.cells_h;
    a.order = order;
    return KDGRT_OK;
}

// After the launch: the view whose durations the cost buffer now holds.
static void sched_after(kdgrt_scene* sc, const ViewArgs& a, int kind) {
    sc->sched_valid = true;
    sc->sched_kind = kind;
    sc->sched_w = a.W;
    sc->sched_h = a.H;
    sc->sched_depth = a.depth;
    sc->sched_cam = a.cam;
}

// Launches one view: the cone records of the view's origin into the scene's
// table, then `launch` (the view kernel, reading a.sc.vcone), both on `s`.
// Launches of one scene on different streams are ordered through the
// table's event.
extern "C++" {
// sched_kind >= 0: the launch records tile durations and follows the
// scene's tile schedule for that kernel (0 plain, 1 hybrid); the tile sort,
// when one is due, runs inside k_cone_prep.  Without may_alloc a launch whose
// tiles exceed the schedule buffers runs in raster order instead of growing
// them.
template <class Launch>
static kdgrt_status with_cone_table(const kdgrt_scene* csc, ViewArgs& a, hipStream_t s, int sched_kind,
                                    bool may_alloc, Launch&& launch) {
    kdgrt_scene* sc = const_cast<kdgrt_scene*>(csc);
    std::lock_guard<std::mutex> lock(sc->cone_mu);
    if (sc->cone_used && sc->cone_stream != s) HIP_TRY(hipStreamWaitEvent(s, sc->cone_ev, 0));
    TileSort ts{};
    bool use = false;
    if (sched_kind >= 0) {
        const kdgrt_status st = sched_before(sc, a, view_tiles(a), sched_kind, may_alloc, s, use, ts);
        if (st != KDGRT_OK) return st;
    }
    const uint64_t ng = sc->n_gaussians;
    const uint32_t g_grid = gaussian_cone_grid(ng, sc->ctx->prop.multiProcessorCount);
    hipLaunchKernelGGL(k_gaussian_cone, dim3(g_grid), dim3(256), 0, s, (const float4*)sc->d_gsph,
                       (const float*)sc->d_gsig, ng, a.cam.eye[0], a.cam.eye[1], a.cam.eye[2], (float4*)sc->d_gcone);
    const uint64_t n = sc->n_offsets + kLsphPad;
    const uint32_t blk = ts.on ? kCinBlock : kCinBlockSmall;
    const uint32_t prep_grid = cone_prep_grid(n, blk, sc->ctx->prop.multiProcessorCount) + (ts.on ? 1u : 0u);
    if (ts.on)
        hipLaunchKernelGGL(k_cone_prep<true>, dim3(prep_grid), dim3(blk), 0, s, (const uint32_t*)sc->d_offs,
                           (const float4*)sc->d_gcone, sc->n_offsets, n, (float4*)sc->d_vcone, ts);
    else
        hipLaunchKernelGGL(k_cone_prep<false>, dim3(prep_grid), dim3(blk), 0, s, (const uint32_t*)sc->d_offs,
                           (const float4*)sc->d_gcone, sc->n_offsets, n, (float4*)sc->d_vcone, ts);
    a.sc.vcone = (const float4*)sc->d_vcone;
    const kdgrt_status st = launch(sc);
    if (st != KDGRT_OK) return st;
    HIP_TRY(hipGetLastError());
    if (use) sched_after(sc, a, sched_kind);
    HIP_TRY(hipEventRecord(sc->cone_ev, s));
    sc->cone_stream = s;
    sc->cone_used = true;
    return KDGRT_OK;
}
} // extern "C++"

typedef void (*ViewKernel)(ViewArgs);
// Tile order of a plain launch: raster (debug, profile and warm-up
// launches), the scene's schedule, or the schedule only within the buffers
// the scene already has (no allocation, the serving loop).
enum SchedUse { kSchedNone, kSchedFull, kSchedNoAlloc };

static kdgrt_status launch_view(const kdgrt_scene* csc, ViewArgs& a, ViewKernel kern, hipStream_t s,
                                SchedUse sched = kSchedNone, bool dbg = false) {
    return with_cone_table(csc, a, s, sched == kSchedNone ? -1 : 0, sched == kSchedFull,
                           [&](kdgrt_scene*) -> kdgrt_status {
        hipLaunchKernelGGL(kern, dim3(view_tiles(a)), dim3(64), stack_lds_bytes(a.depth, 64, dbg), s, a);
        return KDGRT_OK;
    });
}

static kdgrt_status render_plain(kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W,
                                 int H, int depth, int fmt, void* out, void* stream, SchedUse sched) {
    kdgrt_status st = check_view(ctx, sc, cam, W, H, depth, out);
    if (st != KDGRT_OK) return st;
    if (fmt != KDGRT_OUT_RGBT_F32 && fmt != KDGRT_OUT_RGB_F32)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render: unknown output format");
    ViewArgs a;
    fill_view_args(a, sc, cam, W, H, depth, out);
    return launch_view(sc, a,
                       fmt == KDGRT_OUT_RGBT_F32 ? k_view<kViewK, false, KDGRT_OUT_RGBT_F32>
                                                 : k_view<kViewK, false, KDGRT_OUT_RGB_F32>,
                       (hipStream_t)stream, sched);
}

kdgrt_status kdgrt_render(kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W, int H,
                          int depth, int fmt, void* out, void* stream) {
    return render_plain(ctx, sc, cam, W, H, depth, fmt, out, stream, kSchedFull);
}

extern "C++" {
namespace kdgrt {
kdgrt_status render_served(kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W, int H,
                           int depth, void* out, hipStream_t stream) {
    return render_plain(ctx, sc, cam, W, H, depth, KDGRT_OUT_RGBT_F32, out, stream, kSchedNoAlloc);
}

// Logical tiles of the schedule buffers that views up to max_w x max_h need.
static uint32_t schedule_tiles(int max_w, int max_h) {
    const uint64_t n = interleave_grid<64>((uint64_t)((max_w + 63) / 64) * (uint64_t)((max_h + 63) / 64) * 64);
    return (uint32_t)std::min<uint64_t>(n, kSchedMaxTiles);
}

uint64_t tile_schedule_bytes(int max_w, int max_h) {
    return (3ull * schedule_tiles(max_w, max_h) + kDensCells) * sizeof(uint32_t);
}

kdgrt_status reserve_tile_schedule(kdgrt_scene* sc, int max_w, int max_h) {
    if (!sc || max_w <= 0 || max_h <= 0) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "reserve_tile_schedule");
    hipSetDevice(sc->ctx->device);
    std::lock_guard<std::mutex> lock(sc->cone_mu);
    return sched_grow(sc, schedule_tiles(max_w, max_h));
}
} // namespace kdgrt
} // extern "C++"

kdgrt_status kdgrt_render_debug(kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W,
                                int H, int depth, float* out, int32_t* counters, uint64_t* hashes, uint32_t* seq,
                                int seq_cap, void* stream) {
    kdgrt_status st = check_view(ctx, sc, cam, W, H, depth, out);
    if (st != KDGRT_OK) return st;
    if (seq && seq_cap <= 0) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_debug: seq_cap must be > 0");
    ViewArgs a;
    fill_view_args(a, sc, cam, W, H, depth, out);
    a.counters = counters;
    a.hashes = hashes;
    a.seq = seq;
    a.seq_cap = seq ? seq_cap : 0;
    return launch_view(sc, a, k_view<kViewK, true, KDGRT_OUT_RGBT_F32>, (hipStream_t)stream, kSchedNone, true);
}

static void touch_ptrs(const kdgrt_touch_maps* m, uint8_t* t[6]) {
    t[0] = m->nodes;
    t[1] = m->quads;
    t[2] = m->slots_pretested;
    t[3] = m->slots_tested;
    t[4] = m->gaussians_evaluated;
    t[5] = m->gaussians_blended;
}

kdgrt_status kdgrt_render_touch(kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W,
                                int H, int depth, float* out, const kdgrt_touch_maps* maps, void* stream) {
    kdgrt_status st = check_view(ctx, sc, cam, W, H, depth, out);
    if (st != KDGRT_OK) return st;
    if (!maps) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_touch: null maps");
    ViewArgs a;
    fill_view_args(a, sc, cam, W, H, depth, out);
    touch_ptrs(maps, a.touch);
    return launch_view(sc, a, k_view<kViewK, true, KDGRT_OUT_RGBT_F32>, (hipStream_t)stream, kSchedNone, true);
}

kdgrt_status kdgrt_render_profile(kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W,
                                  int H, float* out, uint64_t* phase_cycles, void* stream) {
    kdgrt_status st = check_view(ctx, sc, cam, W, H, 12, out);
    if (st != KDGRT_OK) return st;
    if (!phase_cycles) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_profile: null phase buffer");
    ViewArgs a;
    fill_view_args(a, sc, cam, W, H, 12, out);
    a.prof = phase_cycles; // 8 x u64 per ray
    return launch_view(sc, a, k_view<kViewK, false, KDGRT_OUT_RGBT_F32, true>, (hipStream_t)stream);
}

// The two launches with_cone_table makes before a plain view, writing the
// table into the caller's buffer instead of the scene's.
kdgrt_status kdgrt_view_cone_table(kdgrt_context* ctx, const kdgrt_scene* csc, const float eye[3], void* out,
                                   uint64_t capacity, uint64_t* n_records, void* stream) {
    if (!ctx || !csc || !eye) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "view_cone_table: null argument");
    if (csc->ctx != ctx)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "view_cone_table: scene belongs to another context");
    if (!std::isfinite(eye[0]) || !std::isfinite(eye[1]) || !std::isfinite(eye[2]))
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "view_cone_table: eye not finite");
    const uint64_t n = csc->n_offsets + kLsphPad;
    if (n_records) *n_records = n;
    if (!out) return KDGRT_OK;
    if (capacity < n) return set_error(KDGRT_ERR_CAPACITY, "view_cone_table: capacity below n_offsets + 4 records");
    kdgrt_scene* sc = const_cast<kdgrt_scene*>(csc);
    const hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(sc->cone_mu);
    if (sc->cone_used && sc->cone_stream != s) HIP_TRY(hipStreamWaitEvent(s, sc->cone_ev, 0));
    const int n_cu = sc->ctx->prop.multiProcessorCount;
    if (sc->n_gaussians)
        hipLaunchKernelGGL(k_gaussian_cone, dim3(gaussian_cone_grid(sc->n_gaussians, n_cu)), dim3(256), 0, s,
                           (const float4*)sc->d_gsph, (const float*)sc->d_gsig, sc->n_gaussians, eye[0], eye[1],
                           eye[2], (float4*)sc->d_gcone);
    hipLaunchKernelGGL(k_cone_prep<false>, dim3(cone_prep_grid(n, kCinBlockSmall, n_cu)), dim3(kCinBlockSmall), 0, s,
                       (const uint32_t*)sc->d_offs, (const float4*)sc->d_gcone, sc->n_offsets, n, (float4*)out,
                       TileSort{});
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(sc->cone_ev, s));
    sc->cone_stream = s;
    sc->cone_used = true;
    return KDGRT_OK;
}

kdgrt_status kdgrt_secondary_demo_spheres(kdgrt_secondary_sphere out[2]) {
    if (!out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "secondary_demo_spheres: null");
    // demo_spheres (cudaRenderer.cu:2162-2167)
    out[0] = kdgrt_secondary_sphere{{0.0f, 1.0f, 0.0f}, 0.3f, KDGRT_MATERIAL_GLASS, 1.0f, 1.1f, {1.0f, 1.0f, 1.0f}};
    out[1] = kdgrt_secondary_sphere{{0.8f, 1.0f, 1.2f}, 0.3f, KDGRT_MATERIAL_MIRROR, 1.0f, 1.0f, {1.0f, 1.0f, 1.0f}};
    return KDGRT_OK;
}

kdgrt_status kdgrt_render_hybrid(kdgrt_context* ctx, const kdgrt_scene* sc, const kdgrt_camera_info* cam, int W,
                                 int H, int depth, const kdgrt_secondary* sec, float* out, void* stream) {
    kdgrt_status st = check_view(ctx, sc, cam, W, H, depth, out);
    if (st != KDGRT_OK) return st;
    if (!sec) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_hybrid: null secondary scene");
    if (sec->max_bounces < 1 || sec->max_bounces > KDGRT_SECONDARY_MAX_BOUNCES)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_hybrid: max_bounces must be 1..16");
    if (sec->count < 0) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_hybrid: negative primitive count");
    HybridArgs ha;
    std::memset(&ha.sec, 0, sizeof(ha.sec));
    ha.sec.prim = sec->prim;
    ha.sec.count = sec->count;
    ha.sec.max_bounces = sec->max_bounces;
    if (sec->prim == KDGRT_SECONDARY_SPHERE) {
        if (sec->count > KDGRT_SECONDARY_MAX_SPHERES)
            return set_error(KDGRT_ERR_CAPACITY, "render_hybrid: at most 8 secondary spheres");
        if (sec->count && !sec->spheres) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_hybrid: null spheres");
        static_assert(sizeof(SecSphere) == sizeof(kdgrt_secondary_sphere), "sphere record layout");
        if (sec->count) std::memcpy(ha.sec.sph, sec->spheres, sizeof(SecSphere) * (size_t)sec->count);
    } else if (sec->prim == KDGRT_SECONDARY_TRI) {
        if (sec->count && !sec->d_vertices)
            return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_hybrid: null triangle vertices");
        ha.sec.tri = sec->d_vertices;
        ha.sec.mesh_mat = sec->mesh_material;
        ha.sec.mesh_refl = sec->mesh_reflectivity;
        ha.sec.mesh_ior = sec->mesh_ior;
        for (int k = 0; k < 3; ++k) ha.sec.mesh_col[k] = sec->mesh_color[k];
    } else {
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_hybrid: prim must be KDGRT_SECONDARY_TRI or _SPHERE");
    }
    fill_view_args(ha.v, sc, cam, W, H, depth, out);
    const hipStream_t s = (hipStream_t)stream;
    return with_cone_table(sc, ha.v, s, 1, true, [&](kdgrt_scene*) -> kdgrt_status {
        hipLaunchKernelGGL(k_hybrid<kViewK>, dim3(view_tiles(ha.v)), dim3(64), stack_lds_bytes(depth, 64), s, ha);
        return KDGRT_OK;
    });
}

// Scene box of the ray keys: lower corner and 1 / extent per axis (0 for a
// flat axis, whose cell index is then 0).
static void ray_key_box(const DevScene& sc, float3& lo, float3& inv) {
    const float ext[3] = {sc.bmax[0] - sc.bmin[0], sc.bmax[1] - sc.bmin[1], sc.bmax[2] - sc.bmin[2]};
    lo = make_float3(sc.bmin[0], sc.bmin[1], sc.bmin[2]);
    inv = make_float3(ext[0] > 0 ? 1.0f / ext[0] : 0.0f, ext[1] > 0 ? 1.0f / ext[1] : 0.0f,
                      ext[2] > 0 ? 1.0f / ext[2] : 0.0f);
}

static uint32_t ray_key_grid(uint32_t n) { return (uint32_t)std::min<uint64_t>(((uint64_t)n + 255) / 256, 8192); }

// Workspace of the ray-key sort of n rays: keys in and out and indices in
// (256-B aligned arrays of n words each), then the radix sort's temporaries
// (*temp bytes).
static kdgrt_status ray_order_ws_bytes(uint32_t n, size_t* temp, size_t* total) {
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, *temp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                               (const uint32_t*)nullptr, (uint32_t*)nullptr, n, 0, kRayKeyBits,
                                               (hipStream_t)nullptr));
    const size_t arr = ((size_t)n * 4 + 255) & ~(size_t)255;
    *total = 3 * arr + *temp;
    return KDGRT_OK;
}

// Ray-key order of a ray list into perm (n entries) on stream s, in a
// workspace of ray_order_ws_bytes(n).
static kdgrt_status ray_order_ws(const DevScene& sc, const float* ro, const float* rd, uint32_t n, uint32_t* perm,
                                 void* ws, size_t temp, hipStream_t s) {
    const size_t arr = ((size_t)n * 4 + 255) & ~(size_t)255;
    char* base = (char*)ws;
    uint32_t* k_in = (uint32_t*)base;
    uint32_t* k_out = (uint32_t*)(base + arr);
    uint32_t* v_in = (uint32_t*)(base + 2 * arr);
    void* tmp = base + 3 * arr;
    float3 lo, inv;
    ray_key_box(sc, lo, inv);
    hipLaunchKernelGGL(k_ray_keys, dim3(ray_key_grid(n)), dim3(256), 0, s, ro, rd, n, lo, inv, k_in, v_in);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, temp, k_in, k_out, v_in, perm, n, 0, kRayKeyBits, s));
    return KDGRT_OK;
}

// Ray-key order with a workspace from the stream-ordered allocator,
// released on s after the sort.
static kdgrt_status ray_order(const DevScene& sc, const float* ro, const float* rd, uint32_t n, uint32_t* perm,
                              hipStream_t s) {
    size_t temp = 0, total = 0;
    kdgrt_status st = ray_order_ws_bytes(n, &temp, &total);
    if (st != KDGRT_OK) return st;
    void* block = nullptr;
    HIP_TRY(hipMallocAsync(&block, total, s));
    st = ray_order_ws(sc, ro, rd, n, perm, block, temp, s);
    HIP_TRY(hipFreeAsync(block, s));
    return st;
}

static kdgrt_status launch_rays(kdgrt_context* ctx, const kdgrt_scene* sc, const float* ro, const float* rd,
                                uint64_t n, int depth, float* out, int32_t* counters, uint64_t* hashes, bool dbg,
                                void* stream, const kdgrt_touch_maps* maps = nullptr) {
    if (!ctx || !sc || (n && (!ro || !rd || !out))) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_rays: null");
    if (depth < 1 || depth > 32) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_rays: stack depth must be 1..32");
    if (n == 0) return KDGRT_OK;
    if (n > (uint64_t)0x7FFFF000 * 64ull) return set_error(KDGRT_ERR_CAPACITY, "render_rays: too many rays");
    RaysArgs a{sc->dev, ro, rd, n, depth, (float4*)out, counters, hashes, nullptr, {}};
    if (maps) touch_ptrs(maps, a.touch);
    const dim3 grid(interleave_grid<kRaysGroup>((n + 63) / 64)), block(64);
    hipStream_t s = (hipStream_t)stream;
    const size_t lds = stack_lds_bytes(depth, 64, dbg);
    // Lists of at least kRaySortMin rays are traced in ray-key order (the
    // sort costs less than the coherence it buys); the scratch comes from the
    // stream-ordered allocator, so the call stays asynchronous.
    void* perm = nullptr;
    if (n >= kRaySortMin && n <= 0x7FFFFFFFull) {
        HIP_TRY(hipMallocAsync(&perm, 4 * n, s));
        const kdgrt_status st = ray_order(sc->dev, ro, rd, (uint32_t)n, (uint32_t*)perm, s);
        if (st != KDGRT_OK) {
            hipFreeAsync(perm, s);
            return st;
        }
        a.perm = (const uint32_t*)perm;
    }
    if (dbg) hipLaunchKernelGGL((k_rays<kViewK, true>), grid, block, lds, s, a);
    else hipLaunchKernelGGL((k_rays<kViewK, false>), grid, block, lds, s, a);
    HIP_TRY(hipGetLastError());
    if (perm) HIP_TRY(hipFreeAsync(perm, s));
    return KDGRT_OK;
}

// Allocation-free ray lists: the ray-key sort of kdgrt_render_rays works in
// the file-scope device array `scratch` (resolved per device at init)
// instead of stream-ordered allocations.  Layout: keys (two buffers), ray
// indices (two buffers, the sort's ping-pong pair), then the radix sort's
// temporaries.  Every call on one device shares the array, so calls that
// overlap on one device race; nothing in it has to start at zero (the key
// kernel writes every key and index, the sort every output).
kdgrt_status kdgrt_render_rays_noalloc(kdgrt_context* ctx, const kdgrt_scene* sc, const float* ro, const float* rd,
                                       uint64_t n, int depth, float* out, void* stream) {
    if (!ctx || !sc || (n && (!ro || !rd || !out)))
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_rays_noalloc: null");
    if (depth < 1 || depth > 32)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_rays_noalloc: stack depth must be 1..32");
    if (n > kNoAllocMaxRays)
        return set_error(KDGRT_ERR_CAPACITY, "render_rays_noalloc: " + std::to_string(n) +
                                                 " rays exceed the static scratch's " +
                                                 std::to_string(kNoAllocMaxRays) + "; use kdgrt_render_rays");
    if (n == 0) return KDGRT_OK;
    if (!ctx->noalloc_scratch) return set_error(KDGRT_ERR_DEVICE, "render_rays_noalloc: scratch not resolved");
    RaysArgs a{sc->dev, ro, rd, n, depth, (float4*)out, nullptr, nullptr, nullptr, {}};
    hipStream_t s = (hipStream_t)stream;
    if (n >= kRaySortMin) {
        char* base = (char*)ctx->noalloc_scratch;
        hipcub::DoubleBuffer<uint32_t> keys((uint32_t*)base, (uint32_t*)(base + kNoAllocArr));
        hipcub::DoubleBuffer<uint32_t> idx((uint32_t*)(base + 2 * kNoAllocArr), (uint32_t*)(base + 3 * kNoAllocArr));
        void* tmp = base + 4 * kNoAllocArr;
        size_t temp = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys, idx, (uint32_t)n, 0, kRayKeyBits, s));
        if (temp > kNoAllocTemp)
            return set_error(KDGRT_ERR_CAPACITY, "render_rays_noalloc: the radix sort needs " + std::to_string(temp) +
                                                     " B of temporaries, the static scratch holds " +
                                                     std::to_string(kNoAllocTemp));
        float3 lo, inv;
        ray_key_box(sc->dev, lo, inv);
        hipLaunchKernelGGL(k_ray_keys, dim3(ray_key_grid((uint32_t)n)), dim3(256), 0, s, ro, rd, (uint32_t)n, lo, inv,
                           keys.Current(), idx.Current());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, temp, keys, idx, (uint32_t)n, 0, kRayKeyBits, s));
        a.perm = idx.Current();
    }
    hipLaunchKernelGGL((k_rays<kViewK, false>), dim3(interleave_grid<kRaysGroup>((n + 63) / 64)), dim3(64),
                       stack_lds_bytes(depth, 64), s, a);
    HIP_TRY(hipGetLastError());
    return KDGRT_OK;
}

kdgrt_status kdgrt_render_rays(kdgrt_context* ctx, const kdgrt_scene* sc, const float* ro, const float* rd,
                               uint64_t n, int depth, float* out, void* stream) {
    return launch_rays(ctx, sc, ro, rd, n, depth, out, nullptr, nullptr, false, stream);
}

kdgrt_status kdgrt_render_rays_debug(kdgrt_context* ctx, const kdgrt_scene* sc, const float* ro, const float* rd,
                                     uint64_t n, int depth, float* out, int32_t* counters, uint64_t* hashes,
                                     void* stream) {
    return launch_rays(ctx, sc, ro, rd, n, depth, out, counters, hashes, true, stream);
}

kdgrt_status kdgrt_render_rays_touch(kdgrt_context* ctx, const kdgrt_scene* sc, const float* ro, const float* rd,
                                     uint64_t n, int depth, float* out, const kdgrt_touch_maps* maps, void* stream) {
    if (!maps) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "render_rays_touch: null maps");
    return launch_rays(ctx, sc, ro, rd, n, depth, out, nullptr, nullptr, true, stream, maps);
}

kdgrt_status kdgrt_warm_up(kdgrt_context* ctx, const kdgrt_scene* sc, void* stream) {
    // warmUp (cudaRenderer.cu:2532-2554) renders one frame to load the code
    // object and fault in the scene; here a 64x64 view from the box corner.
    if (!ctx || !sc) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "warm_up: null");
    void* tmp = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMalloc(&tmp, 64 * 64 * 16));
    kdgrt_camera cam{};
    for (int k = 0; k < 3; ++k) cam.pos[k] = sc->dev.bmax[k] + 1.0f;
    cam.u[0] = 1;
    cam.v[1] = 1;
    cam.n[2] = 1;
    cam.fovy_deg = 60;
    cam.aspect = 1;
    cam.near_c = 0.005;
    kdgrt_camera_info ci;
    kdgrt_camera_make(&cam, 64, 64, &ci);
    // raster order: the warm-up view leaves the scene's tile schedule alone
    kdgrt_status st = render_plain(ctx, sc, &ci, 64, 64, 12, KDGRT_OUT_RGBT_F32, tmp, stream, kSchedNone);
    hipStreamSynchronize((hipStream_t)stream);
    hipFree(tmp);
    return st;
}

kdgrt_status kdgrt_tuning_default(kdgrt_context* ctx, kdgrt_tuning* t) {
    if (!t) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "tuning_default: null");
    (void)ctx;
    t->tile_w = 8;
    t->tile_h = 8;
    t->stack_depth = 12;
    t->hit_capacity = 7; // as the other kernels (64-view batch 44.29 -> 44.15 ms against 8)
    t->waves_per_cu = 32;
    t->block_threads = 64;
    return KDGRT_OK;
}

kdgrt_status kdgrt_batch_create(kdgrt_context* ctx, const kdgrt_view* views, int n_views,
                                const kdgrt_ray_set* sets, int n_sets, const kdgrt_tuning* tuning,
                                kdgrt_batch** out) {
    return kdgrt::batch_build(ctx, views, n_views, sets, n_sets, tuning, out);
}

kdgrt_status kdgrt_batch_run(kdgrt_batch* b, void* stream, int use_graph) {
    if (!b) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_run: null");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(b->ctx->device));
    // ray-key order of the long ray sets, on the run's stream ahead of the
    // launch (outside any captured graph: the sort's library calls stay out)
    for (const auto& j : b->sorts) {
        const kdgrt_status st = ray_order_ws(j.sc, j.ro, j.rd, j.n, j.perm, j.ws, j.temp, s);
        if (st != KDGRT_OK) return st;
    }
    if (!use_graph) return kdgrt::batch_launch(b, s);
    if (b->graph_exec && b->graph_stream != stream) {
        hipGraphExecDestroy(b->graph_exec);
        hipGraphDestroy(b->graph);
        b->graph_exec = nullptr;
        b->graph = nullptr;
    }
    if (!b->graph_exec) {
        if (s == nullptr) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_run: graph capture needs a stream");
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        kdgrt_status st = kdgrt::batch_launch(b, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (st != KDGRT_OK) {
            if (g) hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return set_error(KDGRT_ERR_DEVICE, std::string("batch_run: capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hipGraphDestroy(g);
            return set_error(KDGRT_ERR_DEVICE, std::string("batch_run: instantiate: ") + hipGetErrorString(e));
        }
        b->graph = g;
        b->graph_stream = stream;
    }
    HIP_TRY(hipGraphLaunch(b->graph_exec, s));
    return KDGRT_OK;
}

kdgrt_status kdgrt_batch_info(const kdgrt_batch* b, uint64_t* n_rays, uint32_t* n_tiles, uint32_t* grid_blocks) {
    if (!b) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_info: null");
    if (n_rays) *n_rays = b->n_rays;
    if (n_tiles) *n_tiles = b->n_tiles;
    if (grid_blocks) *grid_blocks = (uint32_t)b->grid;
    return KDGRT_OK;
}

kdgrt_status kdgrt_batch_destroy(kdgrt_batch* b) {
    if (!b) return KDGRT_OK;
    hipSetDevice(b->ctx->device);
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    hipFree(b->d_items);
    hipFree(b->d_counters);
    for (void* p : b->d_perms) hipFree(p);
    for (auto& j : b->sorts) hipFree(j.ws);
    for (auto& j : b->cones) hipFree(j.table);
    hipFree(b->d_gcone);
    delete b;
    return KDGRT_OK;
}

} // extern "C"

namespace kdgrt {

kdgrt_status batch_build(kdgrt_context* ctx, const kdgrt_view* views, int n_views, const kdgrt_ray_set* sets,
                         int n_sets, const kdgrt_tuning* tuning, kdgrt_batch** out) {
    if (!ctx || !out || (n_views && !views) || (n_sets && !sets) || n_views < 0 || n_sets < 0)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: null argument");
    if (n_views + n_sets == 0) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: empty batch");
    kdgrt_tuning t;
    if (tuning) t = *tuning;
    else kdgrt_tuning_default(ctx, &t);
    if (t.tile_w < 8 || t.tile_h < 8 || t.tile_w % 8 || t.tile_h % 8 || t.tile_w * t.tile_h > 4096)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: tile must be multiples of 8, at most 4096 px");
    if (t.stack_depth < 1 || t.stack_depth > 32)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: stack depth must be 1..32");
    if (t.hit_capacity != 4 && t.hit_capacity != 7 && t.hit_capacity != 8 && t.hit_capacity != 16)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT,
                         "batch_create: hit capacity must be 4, 7, 8 or 16");
    if (t.block_threads != 64 && t.block_threads != 128 && t.block_threads != 256)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: block threads must be 64, 128 or 256");
    if (t.waves_per_cu < 1 || t.waves_per_cu > 64)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: waves per CU must be 1..64");
    if (stack_lds_bytes(t.stack_depth, t.block_threads) > ctx->prop.sharedMemPerBlock)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: stack depth x block size exceeds the LDS of a block");

    std::vector<WorkItem> items;
    items.reserve(n_views + n_sets);
    uint64_t tiles = 0, rays = 0;
    const uint32_t tile_px = (uint32_t)(t.tile_w * t.tile_h);
    for (int i = 0; i < n_views; ++i) {
        const kdgrt_view& v = views[i];
        if (!v.scene || !v.out_rgbt || v.width <= 0 || v.height <= 0 || v.width > 32768 || v.height > 32768)
            return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: bad view " + std::to_string(i));
        if (v.scene->ctx != ctx) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: foreign scene");
        WorkItem w{};
        w.sc = v.scene->dev;
        std::memcpy(&w.cam, &v.cam, sizeof(CamInfo));
        w.W = v.width;
        w.H = v.height;
        w.tiles_x = (v.width + t.tile_w - 1) / t.tile_w;
        w.tiles_y = (v.height + t.tile_h - 1) / t.tile_h;
        w.groups_x = (w.tiles_x + 7) / 8;
        w.tile_begin = (uint32_t)tiles;
        w.tile_count = (uint32_t)w.groups_x * (uint32_t)((w.tiles_y + 7) / 8) * 64u;
        w.out = (float4*)v.out_rgbt;
        tiles += w.tile_count;
        rays += (uint64_t)v.width * v.height;
        items.push_back(w);
    }
    for (int i = 0; i < n_sets; ++i) {
        const kdgrt_ray_set& r = sets[i];
        if (!r.scene || (r.n_rays && (!r.origins || !r.dirs || !r.out_rgbt)))
            return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: bad ray set " + std::to_string(i));
        if (r.scene->ctx != ctx) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: foreign scene");
        if (r.n_rays == 0) continue;
        WorkItem w{};
        w.sc = r.scene->dev;
        w.W = 0;
        w.tile_begin = (uint32_t)tiles;
        const uint64_t tc = (r.n_rays + tile_px - 1) / tile_px;
        w.tile_count = (uint32_t)tc;
        w.out = (float4*)r.out_rgbt;
        w.ro = r.origins;
        w.rd = r.dirs;
        w.n_rays = r.n_rays;
        tiles += tc;
        rays += r.n_rays;
        items.push_back(w);
    }
    if (tiles >= 0xFFFFFFF0ull) return set_error(KDGRT_ERR_CAPACITY, "batch_create: too many tiles");
    if (items.empty()) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "batch_create: no rays");
    HIP_TRY(hipSetDevice(ctx->device));
    auto* b = new kdgrt_batch();
    b->ctx = ctx;
    b->tuning = t;
    b->n_tiles = (uint32_t)tiles;
    b->n_rays = rays;
    hipError_t e = hipSuccess;
    {
        // long ray sets are traced in ray-key order; every run sorts them
        // again (the order changes no result), into buffers allocated here
        for (WorkItem& w : items) {
            if (w.W != 0 || w.n_rays < kRaySortMin || w.n_rays > 0x7FFFFFFFull) continue;
            void* perm = nullptr;
            if ((e = hipMalloc(&perm, 4 * w.n_rays)) != hipSuccess) break;
            b->d_perms.push_back(perm);
            kdgrt_batch::SortJob j{w.sc, w.ro, w.rd, (uint32_t)w.n_rays, (uint32_t*)perm, nullptr, 0};
            size_t total = 0;
            if (ray_order_ws_bytes(j.n, &j.temp, &total) != KDGRT_OK) {
                e = hipErrorUnknown;
                break;
            }
            if ((e = hipMalloc(&j.ws, total)) != hipSuccess) break;
            b->sorts.push_back(j);
            w.perm = (const uint32_t*)perm;
        }
    }
    if (e == hipSuccess) {
        // large views use the view pre-test: cone records per view, within a
        // memory cap (further large views keep the sphere pre-test)
        uint64_t cone_bytes = 0, max_g = 0;
        for (int i = 0; i < n_views && e == hipSuccess; ++i) {
            const kdgrt_view& v = views[i];
            const uint64_t bytes = 16ull * (v.scene->n_offsets + kLsphPad);
            if ((uint64_t)v.width * v.height < kBatchConeMinPixels || cone_bytes + bytes > kBatchConeMaxBytes)
                continue;
            void* table = nullptr;
            if (hipMalloc(&table, bytes) != hipSuccess) {
                // the tables are optional: with too little free memory the
                // remaining views keep the sphere pre-test
                (void)hipGetLastError();
                break;
            }
            cone_bytes += bytes;
            max_g = std::max<uint64_t>(max_g, v.scene->n_gaussians);
            kdgrt_batch::ConeJob j{v.scene, {v.cam.eye[0], v.cam.eye[1], v.cam.eye[2]}, (float4*)table};
            b->cones.push_back(j);
            items[i].sc.vcone = (const float4*)table;
        }
        if (max_g && hipMalloc(&b->d_gcone, 16ull * max_g) != hipSuccess) {
            (void)hipGetLastError();
            b->d_gcone = nullptr;
            for (auto& j : b->cones) hipFree(j.table);
            b->cones.clear();
            for (WorkItem& w : items) w.sc.vcone = nullptr;
        }
    }
    if (e == hipSuccess) e = hipMalloc(&b->d_items, items.size() * sizeof(WorkItem));
    if (e == hipSuccess) e = hipMalloc(&b->d_counters, kCounterBytes);
    if (e == hipSuccess) e = hipMemcpy(b->d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b->d_counters, 0, kCounterBytes);
    if (e != hipSuccess) {
        hipFree(b->d_items);
        hipFree(b->d_counters);
        for (void* p : b->d_perms) hipFree(p);
        for (auto& j : b->sorts) hipFree(j.ws);
        for (auto& j : b->cones) hipFree(j.table);
        hipFree(b->d_gcone);
        delete b;
        return set_error(KDGRT_ERR_DEVICE, std::string("batch_create: ") + hipGetErrorString(e));
    }
    BatchArgs& a = b->args;
    a.items = b->d_items;
    a.n_items = (int)items.size();
    queue_lengths(tiles, a.q_len);
    a.counters = b->d_counters;
    a.depth = t.stack_depth;
    a.tile_w = t.tile_w;
    a.tile_h = t.tile_h;
    b->kernel = pick_batch_kernel(t.hit_capacity);
    b->block = t.block_threads;
    const int n_cu = ctx->prop.multiProcessorCount > 0 ? ctx->prop.multiProcessorCount : 256;
    const int waves = t.waves_per_cu * n_cu;
    b->grid = std::max(1, waves * 64 / t.block_threads);
    // never launch more waves than tiles
    const int max_blocks = (int)std::max<uint64_t>(1, (tiles * 64 + t.block_threads - 1) / t.block_threads);
    b->grid = std::min(b->grid, max_blocks);
    *out = b;
    return KDGRT_OK;
}

kdgrt_status batch_launch(kdgrt_batch* b, hipStream_t s) {
    for (const auto& j : b->cones) {
        const kdgrt_scene* sc = j.scene;
        const uint64_t ng = sc->n_gaussians, n = sc->n_offsets + kLsphPad;
        hipLaunchKernelGGL(k_gaussian_cone, dim3(gaussian_cone_grid(ng, sc->ctx->prop.multiProcessorCount)), dim3(256),
                           0, s, (const float4*)sc->d_gsph, (const float*)sc->d_gsig, ng, j.eye[0], j.eye[1], j.eye[2],
                           (float4*)b->d_gcone);
        hipLaunchKernelGGL(k_cone_prep<false>,
                           dim3(cone_prep_grid(n, kCinBlockSmall, sc->ctx->prop.multiProcessorCount)),
                           dim3(kCinBlockSmall), 0, s, (const uint32_t*)sc->d_offs, (const float4*)b->d_gcone,
                           sc->n_offsets, n, j.table, TileSort{});
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, b->d_counters, kXcds);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(b->kernel, dim3(b->grid), dim3(b->block), stack_lds_bytes(b->args.depth, b->block), s,
                       b->args);
    HIP_TRY(hipGetLastError());
    return KDGRT_OK;
}

} // namespace kdgrt

namespace kdgrt {

int context_device(const kdgrt_context* ctx) { return ctx ? ctx->device : 0; }

kdgrt_status slot_create(kdgrt_context* ctx, int max_items, BatchSlot** out) {
    if (!ctx || !out || max_items < 1) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "slot_create: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    auto* sl = new BatchSlot();
    sl->ctx = ctx;
    sl->max_items = max_items;
    hipError_t e = hipHostMalloc((void**)&sl->h_items, sizeof(WorkItem) * max_items, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&sl->d_items, sizeof(WorkItem) * max_items);
    if (e == hipSuccess) e = hipMalloc(&sl->d_counters, kCounterBytes);
    if (e != hipSuccess) {
        slot_destroy(sl);
        return set_error(KDGRT_ERR_OUT_OF_MEMORY, std::string("slot_create: ") + hipGetErrorString(e));
    }
    *out = sl;
    return KDGRT_OK;
}

void slot_destroy(BatchSlot* sl) {
    if (!sl) return;
    if (sl->h_items) hipHostFree(sl->h_items);
    if (sl->d_items) hipFree(sl->d_items);
    if (sl->d_counters) hipFree(sl->d_counters);
    delete sl;
}

// Persistent serving queue (k_serve): pinned host table (control block,
// request records, done flags) and device table (control block, finished-
// tile counters, record copies, scene table), allocated once.
struct ServeQueue {
    kdgrt_context* ctx = nullptr;
    int max_req = 0;
    void* hblock = nullptr;
    ServeHostCtl* hctl = nullptr;
    ServeRec* hrec = nullptr;
    uint32_t* hdone = nullptr;
    void* dblock = nullptr;
    ServeArgs args{};
    int grid = 0;
    size_t lds = 0;
    uint32_t seq = 0;
    uint64_t tiles = 0;
    bool running = false;
    hipEvent_t ev = nullptr;
    hipStream_t stream = nullptr;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

kdgrt_status serve_queue_create(kdgrt_context* ctx, int max_req, const kdgrt_scene* const* scenes, int n_scenes,
                                void* arena, hipStream_t stream, const kdgrt_tuning& t, uint32_t spin_limit,
                                ServeQueue** out) {
    if (!ctx || !out || max_req < 1 || max_req > (1 << 20) || n_scenes < 1 || !arena || spin_limit < 1 ||
        t.stack_depth < 1 || t.stack_depth > 32 || t.waves_per_cu < 1)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "serve_queue_create: bad arguments");
    auto* q = new ServeQueue();
    q->ctx = ctx;
    q->max_req = max_req;
    q->stream = stream;
    const size_t h_ctl = align256(sizeof(ServeHostCtl)), h_rec = align256(sizeof(ServeRec) * (size_t)max_req),
                 h_done = align256(4ull * max_req);
    const size_t d_ctl = align256(sizeof(ServeDevCtl)), d_done = align256(4ull * max_req), d_rec = h_rec,
                 d_sc = align256(sizeof(DevScene) * (size_t)n_scenes);
    hipError_t e = hipSetDevice(ctx->device);
    // coherent (fine-grained) pinned memory: the kernel reads the host's
    // stores while it runs, and the host reads the kernel's done flags
    if (e == hipSuccess) e = hipHostMalloc(&q->hblock, h_ctl + h_rec + h_done, hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess) e = hipMalloc(&q->dblock, d_ctl + d_done + d_rec + d_sc);
    void* hdev = nullptr;
    if (e == hipSuccess) e = hipHostGetDevicePointer(&hdev, q->hblock, 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&q->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (q->hblock) hipHostFree(q->hblock);
        if (q->dblock) hipFree(q->dblock);
        delete q;
        return set_error(e == hipErrorOutOfMemory ? KDGRT_ERR_OUT_OF_MEMORY : KDGRT_ERR_DEVICE,
                         std::string("serve_queue_create: ") + hipGetErrorString(e));
    }
    std::memset(q->hblock, 0, h_ctl + h_rec + h_done);
    q->hctl = (ServeHostCtl*)q->hblock;
    q->hrec = (ServeRec*)((char*)q->hblock + h_ctl);
    q->hdone = (uint32_t*)((char*)q->hblock + h_ctl + h_rec);
    std::vector<DevScene> sc(n_scenes);
    for (int i = 0; i < n_scenes; ++i) sc[i] = scenes[i]->dev;
    char* db = (char*)q->dblock;
    ServeArgs& a = q->args;
    a.hctl = (ServeHostCtl*)hdev;
    a.hrec = (ServeRec*)((char*)hdev + h_ctl);
    a.hdone = (uint32_t*)((char*)hdev + h_ctl + h_rec);
    a.dctl = (ServeDevCtl*)db;
    a.done_tiles = (uint32_t*)(db + d_ctl);
    a.drec = (ServeRec*)(db + d_ctl + d_done);
    a.scenes = (const DevScene*)(db + d_ctl + d_done + d_rec);
    a.n_scenes = n_scenes;
    a.arena = (char*)arena;
    a.depth = t.stack_depth;
    a.max_req = max_req;
    a.spin_limit = spin_limit;
    e = hipMemcpy((void*)a.scenes, sc.data(), sizeof(DevScene) * n_scenes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        serve_queue_destroy(q);
        return set_error(KDGRT_ERR_DEVICE, std::string("serve_queue_create: ") + hipGetErrorString(e));
    }
    const int n_cu = ctx->prop.multiProcessorCount > 0 ? ctx->prop.multiProcessorCount : 256;
    q->grid = std::max(1, t.waves_per_cu * n_cu);
    q->lds = stack_lds_bytes(t.stack_depth, 64);
    if (q->lds > ctx->prop.sharedMemPerBlock) {
        serve_queue_destroy(q);
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "serve_queue_create: stack depth exceeds the LDS");
    }
    *out = q;
    return KDGRT_OK;
}

void serve_queue_destroy(ServeQueue* q) {
    if (!q) return;
    hipSetDevice(q->ctx->device);
    if (q->running) {
        serve_epoch_stop(q);
        hipEventSynchronize(q->ev);
    }
    if (q->ev) hipEventDestroy(q->ev);
    if (q->dblock) hipFree(q->dblock);
    if (q->hblock) hipHostFree(q->hblock);
    delete q;
}

kdgrt_status serve_epoch_start(ServeQueue* q) {
    if (q->running) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "serve_epoch_start: an epoch is running");
    // no kernel reads the host table between epochs
    q->hctl->published_req = 0;
    q->hctl->stop = 0;
    std::memset(q->hdone, 0, 4ull * q->max_req);
    std::atomic_thread_fence(std::memory_order_seq_cst);
    q->seq = 0;
    q->tiles = 0;
    const int rb = (int)std::min<int64_t>(1024, (q->max_req + 255) / 256);
    hipLaunchKernelGGL(k_serve_reset, dim3(rb), dim3(256), 0, q->stream, q->args.dctl, q->args.done_tiles, q->max_req);
    hipLaunchKernelGGL(k_serve<kViewK>, dim3(q->grid), dim3(64), q->lds, q->stream, q->args);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(q->ev, q->stream));
    q->running = true;
    return KDGRT_OK;
}

void serve_epoch_stop(ServeQueue* q) { __atomic_store_n(&q->hctl->stop, 1u, __ATOMIC_RELEASE); }

bool serve_epoch_finished(ServeQueue* q, kdgrt_status* status) {
    if (!q->running) return true;
    const hipError_t e = hipEventQuery(q->ev);
    if (e == hipErrorNotReady) return false;
    if (status) *status = e == hipSuccess ? KDGRT_OK : KDGRT_ERR_DEVICE;
    q->running = false;
    return true;
}

bool serve_room(const ServeQueue* q, int W, int H) {
    const uint64_t tx = (W + 7) / 8, ty = (H + 7) / 8;
    const uint64_t tiles = ((tx + 7) / 8) * ((ty + 7) / 8) * 64;
    return q->running && q->seq < (uint32_t)q->max_req && q->tiles + tiles < (1ull << 31);
}

kdgrt_status serve_publish(ServeQueue* q, int scene_index, const kdgrt_camera_info& cam, int W, int H,
                           uint64_t out_off, uint32_t* seq) {
    if (!serve_room(q, W, H) || scene_index < 0 || scene_index >= q->args.n_scenes || W < 1 || H < 1)
        return set_error(KDGRT_ERR_CAPACITY, "serve_publish: no room in the epoch");
    ServeRec& r = q->hrec[q->seq];
    r.tiles_x = (W + 7) / 8;
    r.tiles_y = (H + 7) / 8;
    r.groups_x = (r.tiles_x + 7) / 8;
    r.tile_begin = (uint32_t)q->tiles;
    r.tile_count = (uint32_t)r.groups_x * (uint32_t)((r.tiles_y + 7) / 8) * 64u;
    r.scene = (uint32_t)scene_index;
    r.seq = q->seq;
    r.W = W;
    r.H = H;
    r.pad0 = 0;
    r.out_off = out_off;
    std::memcpy(&r.cam, &cam, sizeof(CamInfo));
    q->tiles += r.tile_count;
    *seq = q->seq++;
    __atomic_store_n(&q->hctl->published_req, q->seq, __ATOMIC_RELEASE);
    return KDGRT_OK;
}

bool serve_done(const ServeQueue* q, uint32_t seq) {
    return __atomic_load_n(&q->hdone[seq], __ATOMIC_ACQUIRE) == seq + 1u;
}

kdgrt_status slot_launch(BatchSlot* sl, const kdgrt_view* views, int n, const kdgrt_tuning& t, hipStream_t s,
                         uint64_t* rays) {
    if (!sl || n < 1 || n > sl->max_items) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "slot_launch: bad batch");
    if (t.tile_w < 8 || t.tile_h < 8 || t.tile_w % 8 || t.tile_h % 8 || t.tile_w * t.tile_h > 4096 ||
        t.stack_depth < 1 || t.stack_depth > 32 ||
        (t.hit_capacity != 4 && t.hit_capacity != 7 && t.hit_capacity != 8 && t.hit_capacity != 16) ||
        (t.block_threads != 64 && t.block_threads != 128 && t.block_threads != 256) || t.waves_per_cu < 1 ||
        stack_lds_bytes(t.stack_depth, t.block_threads) > sl->ctx->prop.sharedMemPerBlock)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "slot_launch: bad tuning");
    uint64_t tiles = 0, nr = 0;
    for (int i = 0; i < n; ++i) {
        const kdgrt_view& v = views[i];
        WorkItem w{};
        w.sc = v.scene->dev;
        std::memcpy(&w.cam, &v.cam, sizeof(CamInfo));
        w.W = v.width;
        w.H = v.height;
        w.tiles_x = (v.width + t.tile_w - 1) / t.tile_w;
        w.tiles_y = (v.height + t.tile_h - 1) / t.tile_h;
        w.groups_x = (w.tiles_x + 7) / 8;
        w.tile_begin = (uint32_t)tiles;
        w.tile_count = (uint32_t)w.groups_x * (uint32_t)((w.tiles_y + 7) / 8) * 64u;
        w.out = (float4*)v.out_rgbt;
        tiles += w.tile_count;
        nr += (uint64_t)v.width * v.height;
        sl->h_items[i] = w;
    }
    HIP_TRY(hipMemcpyAsync(sl->d_items, sl->h_items, sizeof(WorkItem) * n, hipMemcpyHostToDevice, s));
    BatchArgs a{};
    a.items = sl->d_items;
    a.n_items = n;
    queue_lengths(tiles, a.q_len);
    a.counters = sl->d_counters;
    a.depth = t.stack_depth;
    a.tile_w = t.tile_w;
    a.tile_h = t.tile_h;
    const int n_cu = sl->ctx->prop.multiProcessorCount > 0 ? sl->ctx->prop.multiProcessorCount : 256;
    int grid = std::max(1, t.waves_per_cu * n_cu * 64 / t.block_threads);
    grid = std::min<int>(grid, (int)std::max<uint64_t>(1, (tiles * 64 + t.block_threads - 1) / t.block_threads));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, sl->d_counters, kXcds);
    hipLaunchKernelGGL(pick_batch_kernel(t.hit_capacity), dim3(grid), dim3(t.block_threads),
                       stack_lds_bytes(t.stack_depth, t.block_threads), s, a);
    HIP_TRY(hipGetLastError());
    if (rays) *rays = nr;
    return KDGRT_OK;
}

} // namespace kdgrt
