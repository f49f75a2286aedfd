/*
 * kdgrt.h — C ABI of the MI355X (gfx950) kd-tree 3D-Gaussian ray-render path.
 *
 * Every entry point returns a kdgrt_status; kdgrt_last_error() gives the
 * message of the most recent failure on the calling thread.  Nothing in the
 * library calls exit().  Plain C types only: no STL, no torch, no HIP types
 * in the signatures (streams are passed as opaque void*, i.e. hipStream_t).
 *
 * Ownership: the library owns scene device buffers (created by
 * kdgrt_scene_create, freed by kdgrt_scene_destroy); callers own every
 * output buffer (device pointers from kdgrt_device_alloc or any HIP
 * allocation).  Render calls are asynchronous on the given stream.
 *
 * Reference interface each entry replaces (paths relative to
 * Kd-treeConverter_v0.9/Kd-treeConverter/):
 *   kdgrt_init             <- initCuda()                       cudaRenderer.cu:2438-2453
 *   kdgrt_shutdown         <- cleanupCudaResources()           cudaRenderer.cu:3486-3521
 *   kdgrt_scene_create     <- renderGaussianWithCudaSetup()    cudaRenderer.cu:2557-2703
 *   kdgrt_camera_make      <- CameraInfo set-up in
 *                             renderGaussianWithCudaFrame()    cudaRenderer.cu:3108-3127
 *   kdgrt_render           <- renderGaussianWithCudaFrame()    cudaRenderer.cu:3099-3484
 *                             (kernel renderKernelGaussian_sortNode, :2313-2390)
 *   kdgrt_render_debug     <- HIT_AND_NODE_COUNT_DEBUG variant cudaRenderer.cu:1504-1606, 2345-2389
 *   kdgrt_warm_up          <- warmUp()                         cudaRenderer.cu:2532-2554
 *   kdgrt_tree_build       <- build_kd_tree_for_composite_object()  Kd-treeConverter.cpp:22-135
 *                             + create_composite_object_from_gaussians()
 *                                                              Kd-treeConverterMain.cpp:1925-2099
 *   kdgrt_kdt_write/read   <- dump_kd_tree_for_composite_object / read_kd_tree_from_file
 *                                                              Kd-treeConverter.cpp:297-343, 444-492
 *   kdgrt_ply_read         <- loadGaussiansFromPly()           Kd-treeConverterMain.cpp:1652-1769
 *   kdgrt_render_hybrid    <- renderGaussianWithCudaFrame() with SECONDARY_RAY
 *                             (kernel renderKernelGaussian_Mesh_Hybrid, :2217-2311)
 *   kdgrt_obj_read_secondary <- loadSecondaryMesh()            Kd-treeConverterMain.cpp:178-226
 *   kdgrt_framebuffer_to_u8 + kdgrt_png_write
 *                          <- saveDeviceFramebufferToPNG()     cudaRenderer.cu:3066-3096
 *   kdgrt_render_rays, kdgrt_batch_*, kdgrt_server_* have no single
 *   reference counterpart: they run the same per-ray path over explicit ray
 *   lists, many views in one persistent launch, and a serving loop.
 */
#ifndef KDGRT_H
#define KDGRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KDGRT_API __attribute__((visibility("default")))

typedef enum kdgrt_status {
    KDGRT_OK = 0,
    KDGRT_ERR_INVALID_ARGUMENT = 1,
    KDGRT_ERR_OUT_OF_MEMORY = 2,
    KDGRT_ERR_DEVICE = 3,
    KDGRT_ERR_IO = 4,
    KDGRT_ERR_FORMAT = 5,
    KDGRT_ERR_CAPACITY = 6,   /* over a fixed capacity or memory budget */
    KDGRT_ERR_NOT_READY = 7,  /* asynchronous result not yet available */
    KDGRT_ERR_NO_DEVICE = 8,
    KDGRT_ERR_CANCELLED = 9,
    KDGRT_ERR_TIMEOUT = 10
} kdgrt_status;

/* Gaussian record, 256 B: the reference layout with QUATERNION=true
 * (Kd-treeConverter.h:319-335).  f_rest is channel-major: SH coefficient
 * i+1 of channel c is f_rest[15*c + i] (cudaRenderer.cu:495-500). */
typedef struct kdgrt_gaussian {
    float pos[3];
    float opacity;
    float scale[3];
    float k_scale;
    float rot[4]; /* w, x, y, z */
    float f_dc[3];
    float f_rest[45];
    float pad[4];
} kdgrt_gaussian;

/* kd-tree node, 8 B (Kd-treeConverter.h:282-285, macros :402-410):
 * inner: x = (first_child << 3) | axis, y = bits of the f32 split position;
 * leaf : x = (count << 3) | 3,          y = offset into the GID list. */
typedef struct kdgrt_node {
    uint32_t x;
    uint32_t y;
} kdgrt_node;

/* Camera fields the render path reads (OpenGLStuffs.h:21-28). */
typedef struct kdgrt_camera {
    float pos[3];
    float u[3];
    float v[3];
    float n[3]; /* forward direction is -n */
    double fovy_deg;
    double aspect;
    double near_c;
} kdgrt_camera;

/* Per-view ray-generation constants (cudaRenderer.cu:318). */
typedef struct kdgrt_camera_info {
    float eye[3];
    float u[3];
    float v[3];
    float start[3];
    float step_x;
    float step_y;
} kdgrt_camera_info;

/* Builder parameters (Kd-treeConverterMain.cpp:46-61, Kd-treeConverter.h:43-65). */
typedef struct kdgrt_build_params {
    float traversal_cost;   /* TRAVL_COST 1.0 */
    float isect_cost;       /* ISCET_COST 5.0 */
    float empty_bonus;      /* EMTPY_BONUS 0.9 */
    uint32_t max_level;     /* MAX_LEVEL 256 */
    uint32_t min_prims;     /* MIN_TRI 16 (unique Gaussian IDs) */
    uint32_t force_split;   /* FORCE_SPLIT_THRESHOLD 256 (triangles); 0 = off */
    int32_t sah_mode;       /* SAH_MODE 1 (0 and 1 supported) */
    int32_t threads;        /* 0 = all hardware threads */
} kdgrt_build_params;

/* Statistics of a built or loaded tree. */
typedef struct kdgrt_tree_stats {
    uint32_t n_nodes;
    uint32_t n_leaves;
    uint32_t n_empty_leaves;
    uint32_t max_depth;
    uint32_t max_leaf_size;
    uint64_t n_offsets;
    uint64_t n_triangles;       /* proxy triangles fed to the builder */
    uint64_t n_kept_gaussians;  /* Gaussians that passed the opacity cull */
    double build_seconds;
    double mesh_seconds;
} kdgrt_tree_stats;

typedef struct kdgrt_context kdgrt_context;
typedef struct kdgrt_scene kdgrt_scene;
typedef struct kdgrt_tree kdgrt_tree;
typedef struct kdgrt_batch kdgrt_batch;
typedef struct kdgrt_server kdgrt_server;

/* Output pixel formats. */
#define KDGRT_OUT_RGBT_F32 0 /* float4 per pixel: r, g, b, T = 1 - accumulated opacity */
#define KDGRT_OUT_RGB_F32 1  /* float3 per pixel, the reference framebuffer layout */

/* Debug counters per ray (int32 x 8). */
#define KDGRT_CNT_NODE_VISITS 0
#define KDGRT_CNT_LEAF_VISITS 1
#define KDGRT_CNT_ISECT_TESTS 2
#define KDGRT_CNT_HITS_FOUND 3
#define KDGRT_CNT_BLEND_OPS 4
#define KDGRT_CNT_MAX_SORT 5
#define KDGRT_CNT_N_EVAL 6
#define KDGRT_CNT_RESTARTS 7
#define KDGRT_N_COUNTERS 8

/* ---------------- errors, device, memory ---------------- */
KDGRT_API const char* kdgrt_last_error(void);
KDGRT_API const char* kdgrt_version(void);
KDGRT_API kdgrt_status kdgrt_device_count(int* count);
KDGRT_API kdgrt_status kdgrt_init(int device, kdgrt_context** ctx);
KDGRT_API kdgrt_status kdgrt_shutdown(kdgrt_context* ctx);
KDGRT_API kdgrt_status kdgrt_device_alloc(kdgrt_context* ctx, size_t bytes, void** dptr);
KDGRT_API kdgrt_status kdgrt_device_free(kdgrt_context* ctx, void* dptr);
KDGRT_API kdgrt_status kdgrt_memcpy_h2d(kdgrt_context* ctx, void* dst, const void* src, size_t bytes);
KDGRT_API kdgrt_status kdgrt_memcpy_d2h(kdgrt_context* ctx, void* dst, const void* src, size_t bytes);
KDGRT_API kdgrt_status kdgrt_memset(kdgrt_context* ctx, void* dptr, int value, size_t bytes);
KDGRT_API kdgrt_status kdgrt_stream_create(kdgrt_context* ctx, void** stream);
KDGRT_API kdgrt_status kdgrt_stream_destroy(kdgrt_context* ctx, void* stream);
KDGRT_API kdgrt_status kdgrt_stream_sync(kdgrt_context* ctx, void* stream);
KDGRT_API kdgrt_status kdgrt_device_sync(kdgrt_context* ctx);
/* Timing with HIP events recorded on `stream`: begin, then end returns ms. */
KDGRT_API kdgrt_status kdgrt_timer_begin(kdgrt_context* ctx, void* stream);
KDGRT_API kdgrt_status kdgrt_timer_end(kdgrt_context* ctx, void* stream, float* ms);
KDGRT_API kdgrt_status kdgrt_device_info(kdgrt_context* ctx, char* name, size_t name_len,
                                         int* n_cu, size_t* total_mem, size_t* lds_per_block);

/* Streaming device copy of n_float4 x 16 B (256-thread blocks, four 16-B
 * nontemporal loads then four stores per thread and chunk).
 * No reference counterpart: the calibration kernel whose HBM counter
 * readings (FETCH_SIZE, WRITE_SIZE) scale the render kernel's. */
KDGRT_API kdgrt_status kdgrt_copy_f32x4(kdgrt_context* ctx, void* dst, const void* src, uint64_t n_float4,
                                        void* stream);

/* ---------------- scenes (device-resident, immutable) ---------------- */
/* nodes/offsets/aabb as in a .kdt file; aabb = xmin,ymin,zmin,xmax,ymax,zmax.
 * scales_inverted = 1 when `g` already holds 1/scale (the reference caller
 * contract, Kd-treeConverterMain.cpp:3636-3642); 0 makes the library invert
 * with the same f32 operation.  The tree is validated (child and offset
 * ranges, leaf GIDs < n_gaussians) before anything is uploaded. */
KDGRT_API kdgrt_status kdgrt_scene_create(kdgrt_context* ctx, const kdgrt_node* nodes, uint32_t n_nodes,
                                          const uint32_t* offsets, uint64_t n_offsets, const float aabb[6],
                                          const kdgrt_gaussian* g, uint64_t n_gaussians, int scales_inverted,
                                          kdgrt_scene** scene);
KDGRT_API kdgrt_status kdgrt_scene_create_from_tree(kdgrt_context* ctx, const kdgrt_tree* tree,
                                                    const kdgrt_gaussian* g, uint64_t n_gaussians,
                                                    int scales_inverted, kdgrt_scene** scene);
KDGRT_API kdgrt_status kdgrt_scene_destroy(kdgrt_scene* scene);
KDGRT_API kdgrt_status kdgrt_scene_device_bytes(const kdgrt_scene* scene, uint64_t* bytes);
/* Tile schedule of kdgrt_render and kdgrt_render_hybrid, on by default.  A
 * view launch takes its tiles longest-first instead of in raster order, which
 * shortens the launch's tail.  When the scene's previous launch was of the
 * same view (same kernel, size and stack depth; camera within 0.05 degrees,
 * eye within 0.1 % of its distance to the scene box centre), the order comes
 * from the tile durations that launch recorded; otherwise from the number of
 * Gaussian centres that project near each tile.  The order changes no result:
 * every pixel is computed by the same code.  Off (enable = 0), every launch
 * runs in raster order, as the reference's 2-D grid does
 * (cudaRenderer.cu:3169-3180).  The schedule state (12 B per tile and 32 KB)
 * is allocated on first use and held by the scene. */
KDGRT_API kdgrt_status kdgrt_scene_set_tile_schedule(kdgrt_scene* scene, int enable);
/* Rough device footprint of a scene before creating it. */
KDGRT_API uint64_t kdgrt_scene_bytes_estimate(uint32_t n_nodes, uint64_t n_offsets, uint64_t n_gaussians);

/* ---------------- camera ---------------- */
KDGRT_API kdgrt_status kdgrt_camera_make(const kdgrt_camera* cam, int width, int height,
                                         kdgrt_camera_info* out);

/* ---------------- plain one-view launch ---------------- */
/* out: width*height pixels in `out_format`, row (height-1-y) holds image
 * row y as in the reference framebuffer (cudaRenderer.cu:2367).
 * stack_depth: short-stack depth (reference 12, 1..32). */
KDGRT_API kdgrt_status kdgrt_render(kdgrt_context* ctx, const kdgrt_scene* scene, const kdgrt_camera_info* cam,
                                    int width, int height, int stack_depth, int out_format, void* out,
                                    void* stream);
/* Same path plus per-ray counters (int32 x KDGRT_N_COUNTERS, row-major by
 * image row y, not flipped), an order-sensitive 64-bit hash of the
 * composited Gaussian-ID sequence (FNV-1a over the IDs), and optionally the
 * first `seq_cap` composited IDs per ray (uint32, -1 padded). Any of
 * counters/id_hash/id_seq may be NULL. */
KDGRT_API kdgrt_status kdgrt_render_debug(kdgrt_context* ctx, const kdgrt_scene* scene,
                                          const kdgrt_camera_info* cam, int width, int height, int stack_depth,
                                          float* out_rgbt, int32_t* counters, uint64_t* id_hash,
                                          uint32_t* id_seq, int seq_cap, void* stream);
/* Measurement helper, no reference counterpart: the VALU-issue ceiling.
 * Launches k_valu_probe (eight independent fused multiply-add chains per
 * lane, `iters` iterations, eight waves per SIMD) on `stream`; the caller
 * times it.  *waves and *fma_per_wave give the instruction count: the
 * measured rate waves * fma_per_wave / time is the peak VALU instruction
 * rate that a trace kernel's SQ_INSTS_VALU count is compared with. */
KDGRT_API kdgrt_status kdgrt_valu_probe(kdgrt_context* ctx, int iters, void* stream, uint64_t* waves,
                                        uint64_t* fma_per_wave);
/* Measurement helper, no reference counterpart (the compulsory bytes of a
 * launch): the debug kernel marks, with a byte set to 1, every record a ray
 * reads.  Arrays are caller-owned device memory that the caller zeroes; any
 * may be NULL.  nodes / quads: n_nodes bytes (node records visited -- the
 * reference's 8-B node fetches -- and child-pair quad records loaded, by
 * first-child index); slots_pretested / slots_tested: n_offsets bytes (leaf
 * slots through the pre-test and through the exact test); gaussians_evaluated
 * / gaussians_blended: n_gaussians bytes (IDs entering the composite loop --
 * the reference's 256-B fetch_gaussian -- and IDs blended, whose SH record
 * is read). */
typedef struct kdgrt_touch_maps {
    uint8_t* nodes;
    uint8_t* quads;
    uint8_t* slots_pretested;
    uint8_t* slots_tested;
    uint8_t* gaussians_evaluated;
    uint8_t* gaussians_blended;
} kdgrt_touch_maps;
KDGRT_API kdgrt_status kdgrt_render_touch(kdgrt_context* ctx, const kdgrt_scene* scene,
                                          const kdgrt_camera_info* cam, int width, int height, int stack_depth,
                                          float* out_rgbt, const kdgrt_touch_maps* maps, void* stream);
KDGRT_API kdgrt_status kdgrt_render_rays_touch(kdgrt_context* ctx, const kdgrt_scene* scene,
                                               const float* origins, const float* dirs, uint64_t n_rays,
                                               int stack_depth, float* out_rgbt, const kdgrt_touch_maps* maps,
                                               void* stream);
/* Measurement helper, no reference counterpart: the plain kernel with
 * shader-clock marks, writing per ray (y*width+x) eight uint64: cycle sums
 * for descend, pre-test windows, exact tests + hit buffer, composite, stack
 * pop/restart/loop control (time a lane spends masked off is charged to its
 * next mark), then the number of pre-tested candidates, the number passed on
 * to the exact test, and 0. */
KDGRT_API kdgrt_status kdgrt_render_profile(kdgrt_context* ctx, const kdgrt_scene* scene,
                                            const kdgrt_camera_info* cam, int width, int height, float* out_rgbt,
                                            uint64_t* phase_cycles, void* stream);
/* Measurement helper, no reference counterpart (the pre-test data of a
 * view): the cone records that a plain view from `eye` pre-tests, one per
 * leaf slot in the order of the tree's offsets, then 4 zero records (the
 * padding a pre-test group may read past the last leaf).  A record is four
 * floats (a.x, a.y, a.z, q): a = mu - eye in f32 for the slot's Gaussian and
 * q a lower bound of the f32 dot product a.d over every ray direction d from
 * `eye` that the exact test can accept (-FLT_MAX: no bound).  *n_records
 * receives n_offsets + 4; out (caller-owned device memory of `capacity`
 * records) may be NULL to query it.  Asynchronous on `stream`. */
KDGRT_API kdgrt_status kdgrt_view_cone_table(kdgrt_context* ctx, const kdgrt_scene* scene, const float eye[3],
                                             void* out, uint64_t capacity, uint64_t* n_records, void* stream);
/* Explicit rays (device float3 arrays, directions used as given).  Lists of
 * 65,536 rays or more are traced in the order of a key (direction cell of
 * the octahedral map, then the Morton code of the origin in the scene box),
 * sorted on the device
 * with scratch from the stream-ordered allocator; results are written at the
 * caller's indices and do not depend on the order. */
KDGRT_API kdgrt_status kdgrt_render_rays(kdgrt_context* ctx, const kdgrt_scene* scene, const float* origins,
                                         const float* dirs, uint64_t n_rays, int stack_depth, float* out_rgbt,
                                         void* stream);
KDGRT_API kdgrt_status kdgrt_render_rays_debug(kdgrt_context* ctx, const kdgrt_scene* scene,
                                               const float* origins, const float* dirs, uint64_t n_rays,
                                               int stack_depth, float* out_rgbt, int32_t* counters,
                                               uint64_t* id_hash, void* stream);
/* No counterpart in the reference: kdgrt_render_rays for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and results as
 * kdgrt_render_rays (bit for bit); the ray-key sort works in a static device
 * array of 36 MiB per device, resolved at kdgrt_init, and the call makes no
 * hipMalloc/hipFree or other allocating call.  At most 2,097,152 rays per
 * call (KDGRT_ERR_CAPACITY above that).  Calls share the device's array:
 * calls that overlap on one device race, so serialise them (one stream, or
 * wait between streams). */
KDGRT_API kdgrt_status kdgrt_render_rays_noalloc(kdgrt_context* ctx, const kdgrt_scene* scene,
                                                 const float* origins, const float* dirs, uint64_t n_rays,
                                                 int stack_depth, float* out_rgbt, void* stream);
KDGRT_API kdgrt_status kdgrt_warm_up(kdgrt_context* ctx, const kdgrt_scene* scene, void* stream);

/* ---------------- hybrid render: Gaussians + reflective/refractive secondary primitives ---------------- */
/* Secondary primitive kinds, the values of SECONDARY_PRIM (Kd-treeConverter.h:28-30). */
#define KDGRT_SECONDARY_TRI 0
#define KDGRT_SECONDARY_SPHERE 1
/* Material types of MeshHit (cudaRenderer.cu:2050-2060). */
#define KDGRT_MATERIAL_DIFFUSE 0 /* adds its colour and ends the path */
#define KDGRT_MATERIAL_MIRROR 1  /* reflection */
#define KDGRT_MATERIAL_GLASS 2   /* refraction (total internal reflection when sinT^2 > 1) */
#define KDGRT_SECONDARY_MAX_SPHERES 8
#define KDGRT_SECONDARY_MAX_BOUNCES 16

/* DemoSphere (cudaRenderer.cu:2153-2160). */
typedef struct kdgrt_secondary_sphere {
    float center[3];
    float radius;
    int32_t material;
    float reflectivity;
    float ior;
    float color[3];
} kdgrt_secondary_sphere;

/* The secondary scene of one hybrid launch.  prim KDGRT_SECONDARY_SPHERE:
 * `count` spheres from the host array `spheres` (copied at launch; the
 * reference's constant pair is kdgrt_secondary_demo_spheres).  prim
 * KDGRT_SECONDARY_TRI: `count` triangles, 9 floats each, in device memory
 * `d_vertices`, all with the material of the mesh fields (SecondaryMesh,
 * cudaRenderer.h:11-19).  max_bounces: MAX_BOUNCES (5 in the reference). */
typedef struct kdgrt_secondary {
    int32_t prim;
    int32_t count;
    const kdgrt_secondary_sphere* spheres;
    const float* d_vertices;
    int32_t mesh_material;
    float mesh_reflectivity;
    float mesh_ior;
    float mesh_color[3];
    int32_t max_bounces;
} kdgrt_secondary;

/* The reference's two demo spheres (cudaRenderer.cu:2162-2167): a glass
 * sphere at (0, 1, 0) and a mirror sphere at (0.8, 1, 1.2), radius 0.3. */
KDGRT_API kdgrt_status kdgrt_secondary_demo_spheres(kdgrt_secondary_sphere out[2]);
/* One view through renderKernelGaussian_Mesh_Hybrid (cudaRenderer.cu:2217-2311):
 * per pixel, up to max_bounces segments; each segment intersects the
 * secondary primitives first, traces the Gaussians up to that hit
 * (t_scene_far clamped to it, cudaRenderer.cu:1538-1541, 1612-1616) and
 * follows the mirror / glass bounce.  out_rgb: width*height float3,
 * min(colour, 1) per channel, rows flipped as in kdgrt_render. */
KDGRT_API kdgrt_status kdgrt_render_hybrid(kdgrt_context* ctx, const kdgrt_scene* scene,
                                           const kdgrt_camera_info* cam, int width, int height, int stack_depth,
                                           const kdgrt_secondary* secondary, float* out_rgb, void* stream);

/* ---------------- frame output ---------------- */
/* Device conversion of a framebuffer (width*height pixels of in_channels
 * floats: 3 = KDGRT_OUT_RGB_F32, 4 = KDGRT_OUT_RGBT_F32) to 8-bit pixels of
 * out_channels (3 = RGB, 4 = RGBA with alpha 255): clamp to [0, 1], times
 * 255, truncated (cudaRenderer.cu:3083-3086).  flip_rows != 0 writes image
 * row y from framebuffer row height-1-y, undoing the framebuffer's row flip
 * as the reference's PNG dump does; out is device memory. */
KDGRT_API kdgrt_status kdgrt_framebuffer_to_u8(kdgrt_context* ctx, const float* fb, int width, int height,
                                               int in_channels, uint8_t* out, int out_channels, int flip_rows,
                                               void* stream);

/* ---------------- batched persistent launch ---------------- */
typedef struct kdgrt_view {
    const kdgrt_scene* scene;
    kdgrt_camera_info cam;
    int32_t width;
    int32_t height;
    float* out_rgbt; /* device, width*height float4, reference row flip */
} kdgrt_view;

typedef struct kdgrt_ray_set {
    const kdgrt_scene* scene;
    const float* origins; /* device float3[n] */
    const float* dirs;    /* device float3[n] */
    uint64_t n_rays;
    float* out_rgbt;      /* device float4[n] */
} kdgrt_ray_set;

typedef struct kdgrt_tuning {
    int32_t tile_w;          /* ray tile width in pixels (tile = one work-queue item) */
    int32_t tile_h;
    int32_t stack_depth;     /* 1..32 */
    int32_t hit_capacity;    /* per-ray sorted hit buffer: 4, 7, 8 or 16 */
    int32_t waves_per_cu;    /* persistent grid = waves_per_cu * CUs * 64 lanes / block size */
    int32_t block_threads;   /* 64, 128 or 256 */
} kdgrt_tuning;

KDGRT_API kdgrt_status kdgrt_tuning_default(kdgrt_context* ctx, kdgrt_tuning* t);
/* Plan a batch: uploads the work table; the views' output buffers must stay
 * valid while the batch is run.  Ray sets of 65,536 rays or more get their
 * tracing order here (as kdgrt_render_rays sorts them; waits for the
 * device); the order changes no result, also if the rays change later. */
KDGRT_API kdgrt_status kdgrt_batch_create(kdgrt_context* ctx, const kdgrt_view* views, int n_views,
                                          const kdgrt_ray_set* ray_sets, int n_ray_sets,
                                          const kdgrt_tuning* tuning, kdgrt_batch** batch);
/* One persistent launch over all tiles of the batch on `stream`.
 * use_graph = 1 captures reset+launch into a hipGraph on first use and
 * replays it afterwards. */
KDGRT_API kdgrt_status kdgrt_batch_run(kdgrt_batch* batch, void* stream, int use_graph);
KDGRT_API kdgrt_status kdgrt_batch_info(const kdgrt_batch* batch, uint64_t* n_rays, uint32_t* n_tiles,
                                        uint32_t* grid_blocks);
KDGRT_API kdgrt_status kdgrt_batch_destroy(kdgrt_batch* batch);
/* Autotune over tile size, stack depth, hit capacity, waves/CU and block
 * size on the given batch's work; writes the fastest tuning to `best`
 * and its mean time in ms to `best_ms`.  max_trials bounds the search. */
KDGRT_API kdgrt_status kdgrt_autotune(kdgrt_context* ctx, const kdgrt_view* views, int n_views,
                                      const kdgrt_ray_set* ray_sets, int n_ray_sets, int max_trials,
                                      int reps, kdgrt_tuning* best, float* best_ms);

/* ---------------- serving loop ---------------- */
typedef struct kdgrt_server_config {
    uint64_t hbm_budget_bytes;   /* resident scenes + the output arena of admitted requests */
    int32_t max_batch_views;     /* views per persistent launch */
    int32_t max_inflight_batches;/* launches overlapping on the device (one stream each) */
    int32_t plain_launches;      /* 1 = comparison mode: one plain launch per request, one at a time */
    kdgrt_tuning tuning;
    double batch_window_ms;      /* while the device is busy, wait up to this long to fill a batch */
    double simulate_mrays_per_s; /* > 0 with ctx == NULL: no device, batches take rays / rate (tests) */
    /* 1 = continuous batching into one running persistent kernel per epoch:
     * admitted requests are published into the kernel's queue while it runs
     * (coherent host memory, polled by the kernel); batch and window fields
     * then do not apply. */
    int32_t persistent_kernel;
    int32_t persistent_max_requests; /* requests per epoch (0: 4096); a full epoch drains and a new one starts */
    double epoch_idle_ms;            /* an idle epoch's kernel is stopped after this long */
    uint32_t persistent_spin_limit;  /* polls a kernel wave makes before it gives up waiting (0: 4M) */
    /* persistent mode: 1 = a request that finds the device idle (no epoch
     * running, no plain launch in flight, no other request pending) runs as
     * one plain one-view launch; no epoch starts until it has finished */
    int32_t idle_plain;
} kdgrt_server_config;

typedef struct kdgrt_request {
    int32_t scene_id;
    kdgrt_camera_info cam;
    int32_t width;
    int32_t height;
    double arrival_s;            /* caller clock; 0 = now */
    double deadline_s;           /* 0 = none; a request not started by then is cancelled */
} kdgrt_request;

typedef struct kdgrt_result_info {
    uint64_t request_id;
    int32_t status;              /* kdgrt_status */
    double arrival_s, admitted_s, launched_s, done_s; /* server clock, seconds */
    uint32_t batch_id;
    uint32_t batch_views;
} kdgrt_result_info;

typedef struct kdgrt_server_stats {
    uint64_t submitted, completed, rejected, cancelled;
    uint64_t batches, rays;
    double busy_s, wall_s;
    uint64_t bytes_resident, bytes_peak;
} kdgrt_server_stats;

KDGRT_API kdgrt_status kdgrt_server_create(kdgrt_context* ctx, const kdgrt_server_config* cfg,
                                           kdgrt_server** server);
/* Makes a scene resident (counted against the budget); before start only. */
KDGRT_API kdgrt_status kdgrt_server_add_scene(kdgrt_server* s, const kdgrt_tree* tree, const kdgrt_gaussian* g,
                                              uint64_t n_gaussians, int scales_inverted, int32_t* scene_id);
/* Serves a scene the caller created with kdgrt_scene_create*: the server
 * does not destroy it, and it must outlive the server.  Its device bytes
 * count against the budget like an added scene's.  No reference counterpart
 * (the reference holds one scene in module globals, cudaRenderer.cu:343-355). */
KDGRT_API kdgrt_status kdgrt_server_add_resident_scene(kdgrt_server* s, kdgrt_scene* scene, int32_t* scene_id);
KDGRT_API kdgrt_status kdgrt_server_remove_scene(kdgrt_server* s, int32_t scene_id);
KDGRT_API kdgrt_status kdgrt_server_start(kdgrt_server* s);
/* Thread-safe.  Returns KDGRT_ERR_CAPACITY when the request can never fit the
 * output arena (budget minus resident scenes). */
KDGRT_API kdgrt_status kdgrt_server_submit(kdgrt_server* s, const kdgrt_request* req, uint64_t* request_id);
/* Waits up to timeout_ms (< 0: forever) for the request; on completion copies
 * the RGBT image (width*height float4, reference row flip) to `host_rgbt` if
 * not NULL and releases the request's device memory.  Every completed
 * request must be waited for once: its output occupies the arena until then.
 * Returns KDGRT_ERR_TIMEOUT if not finished in time (the request stays). */
KDGRT_API kdgrt_status kdgrt_server_wait(kdgrt_server* s, uint64_t request_id, double timeout_ms,
                                         float* host_rgbt, kdgrt_result_info* info);
KDGRT_API kdgrt_status kdgrt_server_cancel(kdgrt_server* s, uint64_t request_id);
KDGRT_API kdgrt_status kdgrt_server_stats_get(kdgrt_server* s, kdgrt_server_stats* st);
KDGRT_API double kdgrt_server_now(kdgrt_server* s);
KDGRT_API kdgrt_status kdgrt_server_stop(kdgrt_server* s);
KDGRT_API kdgrt_status kdgrt_server_destroy(kdgrt_server* s);

/* ---------------- host side: data, meshing, SAH build, files ---------------- */
KDGRT_API kdgrt_status kdgrt_build_params_default(kdgrt_build_params* p);
/* Proxy-meshes the Gaussians (20-triangle icosahedra, opacity cull) and
 * builds the SAH kd-tree with Gaussian IDs in the leaves. `g` holds
 * activated, non-inverted scales. */
KDGRT_API kdgrt_status kdgrt_tree_build(const kdgrt_gaussian* g, uint64_t n_gaussians,
                                        const kdgrt_build_params* params, kdgrt_tree** tree);
KDGRT_API kdgrt_status kdgrt_tree_from_arrays(const kdgrt_node* nodes, uint32_t n_nodes, const uint32_t* offsets,
                                              uint64_t n_offsets, const float aabb[6], kdgrt_tree** tree);
KDGRT_API kdgrt_status kdgrt_tree_stats_get(const kdgrt_tree* tree, kdgrt_tree_stats* st);
KDGRT_API const kdgrt_node* kdgrt_tree_nodes(const kdgrt_tree* tree);
KDGRT_API const uint32_t* kdgrt_tree_offsets(const kdgrt_tree* tree);
KDGRT_API kdgrt_status kdgrt_tree_aabb(const kdgrt_tree* tree, float aabb[6]);
/* Checks child/offset ranges, leaf GID ranges (< n_gaussians when nonzero). */
KDGRT_API kdgrt_status kdgrt_tree_validate(const kdgrt_tree* tree, uint64_t n_gaussians);
KDGRT_API kdgrt_status kdgrt_tree_free(kdgrt_tree* tree);
KDGRT_API kdgrt_status kdgrt_kdt_write(const kdgrt_tree* tree, const char* path);
KDGRT_API kdgrt_status kdgrt_kdt_read(const char* path, kdgrt_tree** tree);

/* Proxy mesh only (for inspection/tests): writes 9 floats per triangle and the
 * Gaussian ID of each triangle; returns counts and the vertex AABB in
 * .kdt order.  Pass NULL buffers to query n_triangles first. */
KDGRT_API kdgrt_status kdgrt_proxy_mesh(const kdgrt_gaussian* g, uint64_t n_gaussians, float* tri_xyz,
                                        uint32_t* tri_gid, uint64_t capacity, uint64_t* n_triangles,
                                        float aabb[6]);

/* Synthetic scenes.  kind 0: the 10k uniform-cube scene family;
 * kind 1: the clustered family (64 anisotropic clusters + 10% background). */
KDGRT_API kdgrt_status kdgrt_generate_gaussians(int kind, uint64_t n, uint64_t seed, kdgrt_gaussian* out);
/* PLY reader (3DGS property names; activation as the reference loader). */
KDGRT_API kdgrt_status kdgrt_ply_count(const char* path, uint64_t* n);
KDGRT_API kdgrt_status kdgrt_ply_read(const char* path, kdgrt_gaussian* out, uint64_t capacity, uint64_t* n);
/* In-place 1/scale and 1/k_scale (Kd-treeConverterMain.cpp:3636-3642). */
KDGRT_API kdgrt_status kdgrt_gaussians_invert_scales(kdgrt_gaussian* g, uint64_t n);

/* NeRF transforms.json camera list (loadCameraJson, LoadCamera.hpp:25-62):
 * per entry of "frames", columns 0/1/2 of "transform_matrix" are the u, v,
 * n axes and column 3 the position; fovy 39.6 degrees and near 0.005 as in
 * the reference, aspect from the caller.  Writes min(count, capacity)
 * cameras and sets *count to the number of frames. */
KDGRT_API kdgrt_status kdgrt_cameras_from_nerf_json(const char* path, double aspect, kdgrt_camera* out, int capacity,
                                                    int* count);
/* Triangles of a Wavefront OBJ file as the secondary mesh (loadSecondaryMesh,
 * Kd-treeConverterMain.cpp:178-226): faces in file order, quads split along
 * the shorter diagonal as tinyobjloader's triangulation does, each vertex
 * stored as (x, -z, y).  Faces of more than four vertices are rejected
 * (KDGRT_ERR_FORMAT).  Writes min(n, capacity) triangles (9 floats each)
 * and sets *n_triangles; pass out_xyz = NULL to count. */
KDGRT_API kdgrt_status kdgrt_obj_read_secondary(const char* path, float* out_xyz, uint64_t capacity,
                                                uint64_t* n_triangles);
/* PNG file of 8-bit pixels in host memory (rows top to bottom, channels
 * 3 = RGB or 4 = RGBA), 8 bits per channel, zlib stored blocks. */
KDGRT_API kdgrt_status kdgrt_png_write(const char* path, int width, int height, int channels,
                                       const uint8_t* pixels);
/* Per-ray statistics files of a debug render (the HIT_AND_NODE_COUNT_DEBUG
 * products, cudaRenderer.cu:3173-3400) into directory `dir`: renderStats.txt,
 * render_stats.csv, heatmap_{node_visits,leaf_visits,intersection,hits_found,
 * blend_ops,max_sort_size}.csv, isectCnt_kd.bin, blendCnt_kd.bin.  `counters`
 * is kdgrt_render_debug's output (8 int32 per ray, ray order y*width + x). */
KDGRT_API kdgrt_status kdgrt_stats_write(const int32_t* counters, int width, int height, const char* dir);

/* ---------------- interactive viewer: camera controls ---------------- */
/* Camera motion of the reference viewer, in its f32 host arithmetic.
 *
 * kdgrt_camera_move <- the held-key moves of idle() (Kd-treeConverterMain.cpp:
 * 3463-3500): with s = speed * dt, w: pos -= n s, s: pos += n s, a: pos -= u s,
 * d: pos += u s, q: pos += v s, e: pos -= v s, applied in that order for every
 * key bit set.  *moved (may be NULL) is set when any key was held. */
#define KDGRT_KEY_W 1
#define KDGRT_KEY_S 2
#define KDGRT_KEY_A 4
#define KDGRT_KEY_D 8
#define KDGRT_KEY_Q 16
#define KDGRT_KEY_E 32
KDGRT_API kdgrt_status kdgrt_camera_move(kdgrt_camera* cam, uint32_t keys, float speed, float dt, int* moved);
/* kdgrt_camera_orbit <- updateOrbitCamera (Kd-treeConverterMain.cpp:438-480):
 * pos - centre and the u, v, n axes rotated by `angle` radians about `axis`
 * (rotateVectorAroundArbitraryAxis, :422-436: Rodrigues' formula with the
 * axis as given, not normalised), then n normalised, v = n x u, normalised,
 * u = v x n, normalised (fMyVecNormalize / fMyVecCrossProduct,
 * MyMathUtility.cpp:59-90).  The reference orbits at 0.8 rad/s about the
 * scene-specific line centre (0.393, 1.093, 0.698), axis (0.0138, 0.9038,
 * 0.4278). */
KDGRT_API kdgrt_status kdgrt_camera_orbit(kdgrt_camera* cam, const float centre[3], const float axis[3],
                                          float angle);
/* kdgrt_camera_drag <- mousemove (Kd-treeConverterMain.cpp:1193-1296), a drag
 * of (delx, dely) pixels (dely positive upwards).  zoom = 0: yaw by
 * delx * rot_speed degrees about the world y axis (applied to u, v, n), then
 * pitch by dely * rot_speed degrees about u (applied to v, n), skipped when
 * the pitch angle is positive and n.y (after the yaw) above 0.98, or
 * negative and n.y below -0.98; then the axes are re-orthonormalised as in
 * kdgrt_camera_orbit.  The rotation matrices are
 * those of glRotatef (the OpenGL specification's axis-angle matrix of the
 * normalised axis) in f32.  zoom != 0: fovy += delx * 0.25, clamped to
 * [0.5, 100] degrees.  The reference's rot_speed is 0.1 (CAM_ROT_SPEED). */
KDGRT_API kdgrt_status kdgrt_camera_drag(kdgrt_camera* cam, int delx, int dely, float rot_speed, int zoom);

/* ---------------- interactive viewer: frame loop ---------------- */
/* The reference's interactive mode renders each frame on a compute stream
 * into a display buffer that OpenGL owns (a pixel buffer object registered
 * with CUDA and mapped on a transfer stream, Kd-treeConverterMain.cpp:
 * 226-251, 3514-3608) and reports per-frame FPS averaged over frames
 * [MEASURE_START_FRAME, MEASURE_END_FRAME) = [500, 1000).  An MI355X has no
 * display, so a viewer here delivers frames to the host: a ring of slots,
 * each a device framebuffer plus a pinned host frame.  kdgrt_viewer_submit
 * renders into a free slot on the viewer's compute stream (the scene's tile
 * schedule applies) and copies the frame to the slot's host memory on a
 * transfer stream, so the copy of one frame overlaps the render of the next;
 * kdgrt_viewer_acquire hands out finished frames in submission order, and a
 * slot is reused only after kdgrt_viewer_release.  Frame formats:
 *   KDGRT_FRAME_RGB_F32: float3 per pixel, rows as the framebuffer (row
 *     height-1-y holds image row y), the reference's display-buffer layout;
 *   KDGRT_FRAME_RGBA8 / KDGRT_FRAME_RGB8: 8-bit pixels, rows top to bottom,
 *     converted on the device as kdgrt_framebuffer_to_u8 with flip_rows = 1.
 * A viewer may be driven from two threads (one submitting, one acquiring);
 * its calls are serialised internally. */
#define KDGRT_FRAME_RGB_F32 0
#define KDGRT_FRAME_RGBA8 1
#define KDGRT_FRAME_RGB8 2
typedef struct kdgrt_viewer kdgrt_viewer;
typedef struct kdgrt_viewer_frame {
    uint64_t frame_id;     /* 0, 1, 2, ... in submission order */
    const void* pixels;    /* pinned host memory of the slot, valid until release */
    uint64_t bytes;
    int width, height, format;
    float render_ms;       /* GPU time of the frame's render (HIP events) */
    float fps;             /* 1000 / render_ms, the reference's per-frame FPS */
    kdgrt_camera camera;   /* the camera the frame was rendered with */
} kdgrt_viewer_frame;
typedef struct kdgrt_viewer_stats {
    uint64_t submitted, delivered;
    double mean_render_ms;       /* over delivered frames */
    double window_avg_fps;       /* mean per-frame FPS over delivered frames [window_start, window_end) */
    uint64_t window_frames;      /* frames of the window delivered so far */
    double wall_fps;             /* delivered frames / seconds from the first submit to the last delivery */
} kdgrt_viewer_stats;
/* n_slots 2..8; stack_depth as in kdgrt_render.  window_start/window_end:
 * the FPS measurement window in frames (the reference's 500 / 1000). */
KDGRT_API kdgrt_status kdgrt_viewer_create(kdgrt_context* ctx, const kdgrt_scene* scene, int width, int height,
                                           int format, int n_slots, int stack_depth, int window_start,
                                           int window_end, kdgrt_viewer** out);
/* Renders one frame of `cam` (aspect taken from the camera as given).
 * KDGRT_ERR_NOT_READY when every slot is still held (not yet acquired and
 * released) and timeout_ms passed; timeout_ms < 0 waits without limit. */
KDGRT_API kdgrt_status kdgrt_viewer_submit(kdgrt_viewer* v, const kdgrt_camera* cam, int timeout_ms,
                                           uint64_t* frame_id);
/* The oldest submitted frame not yet acquired, once its copy has finished.
 * KDGRT_ERR_NOT_READY when none finished within timeout_ms (or none is
 * pending: then at once); timeout_ms < 0 waits without limit. */
KDGRT_API kdgrt_status kdgrt_viewer_acquire(kdgrt_viewer* v, int timeout_ms, kdgrt_viewer_frame* out);
KDGRT_API kdgrt_status kdgrt_viewer_release(kdgrt_viewer* v, uint64_t frame_id);
KDGRT_API kdgrt_status kdgrt_viewer_stats_get(const kdgrt_viewer* v, kdgrt_viewer_stats* out);
/* Waits for the viewer's device work and frees it (held frames included). */
KDGRT_API kdgrt_status kdgrt_viewer_destroy(kdgrt_viewer* v);

#ifdef __cplusplus
}
#endif
#endif /* KDGRT_H */
