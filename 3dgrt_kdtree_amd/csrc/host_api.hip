// Reference-named C++ host API (include/kdgrt_reference_api.hpp) over the C
// ABI.  Singleton state like the reference renderer: one context, one
// resident scene, per-frame constants computed on the host.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "kdgrt_reference_api.hpp"

namespace kdgrt_ref {
namespace {

struct State {
    std::mutex mu;
    kdgrt_context* ctx = nullptr;
    kdgrt_scene* scene = nullptr;
    float aabb[6] = {0, 0, 0, 0, 0, 0}; // .kdt order
    int last_w = 1920, last_h = 1080;
    // SECONDARY_RAY set-up: frames run the hybrid render
    bool hybrid = false;
    int sec_prim = KDGRT_SECONDARY_SPHERE;
    float* d_mesh = nullptr;
    int mesh_tris = 0;
    int mesh_material = 1;
    float mesh_reflectivity = 0.95f, mesh_ior = 1.1f, mesh_color[3] = {0.95f, 0.95f, 0.95f};
};

static_assert(sizeof(float3) == 12, "SecondaryMesh vertices are packed float triples");

State& state() {
    static State s;
    return s;
}

void report(const char* where) { std::fprintf(stderr, "[kdgrt] %s: %s\n", where, kdgrt_last_error()); }

} // namespace

bool initCuda() {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.ctx) return true;
    if (kdgrt_init(0, &s.ctx) != KDGRT_OK) {
        report("initCuda");
        s.ctx = nullptr;
        return false;
    }
    return true;
}

static void drop_mesh(State& s) {
    if (s.d_mesh) kdgrt_device_free(s.ctx, s.d_mesh);
    s.d_mesh = nullptr;
    s.mesh_tris = 0;
    s.hybrid = false;
}

static void setup_locked(State& s, const CompositeObject& object, const std::vector<Gaussian>& gaussians) {
    drop_mesh(s);
    const KdTree* t = object.kd_tree;
    if (!t || gaussians.empty()) { // cudaRenderer.cu:2571-2575
        std::fprintf(stderr, "[kdgrt] renderGaussianWithCudaSetup: Kd-tree or Gaussian data is empty.\n");
        return;
    }
    if (!s.ctx) {
        std::fprintf(stderr, "[kdgrt] renderGaussianWithCudaSetup: initCuda() was not called.\n");
        return;
    }
    if (s.scene) {
        kdgrt_scene_destroy(s.scene);
        s.scene = nullptr;
    }
    // CompositeObject::AABB is XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX (cudaRenderer.cu:2668-2669).
    const float aabb[6] = {object.AABB[0], object.AABB[2], object.AABB[4],
                           object.AABB[1], object.AABB[3], object.AABB[5]};
    if (kdgrt_scene_create(s.ctx, reinterpret_cast<const kdgrt_node*>(t->tree), (uint32_t)t->tree_node_count,
                           t->prim_offset_list, (uint64_t)t->prim_offset_count, aabb, gaussians.data(),
                           gaussians.size(), /*scales_inverted=*/1, &s.scene) != KDGRT_OK) {
        report("renderGaussianWithCudaSetup");
        s.scene = nullptr;
        return;
    }
    for (int k = 0; k < 6; ++k) s.aabb[k] = aabb[k];
}

void renderGaussianWithCudaSetup(const CompositeObject& object, const std::vector<Gaussian>& gaussians) {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    setup_locked(s, object, gaussians);
}

void renderGaussianWithCudaSetup(const CompositeObject& object, const std::vector<Gaussian>& gaussians,
                                 SecondaryMesh& secMesh) {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    setup_locked(s, object, gaussians);
    if (!s.scene) return;
    // initSecondaryMesh (cudaRenderer.cu:2064-2083): the vertices to device memory, the material by value
    const size_t n_tri = secMesh.vert.size() / 3;
    if (n_tri) {
        void* d = nullptr;
        if (kdgrt_device_alloc(s.ctx, n_tri * 9 * sizeof(float), &d) != KDGRT_OK ||
            kdgrt_memcpy_h2d(s.ctx, d, secMesh.vert.data(), n_tri * 9 * sizeof(float)) != KDGRT_OK) {
            report("renderGaussianWithCudaSetup (secondary mesh)");
            if (d) kdgrt_device_free(s.ctx, d);
            return;
        }
        s.d_mesh = static_cast<float*>(d);
    }
    s.mesh_tris = (int)n_tri;
    s.mesh_material = secMesh.materialType;
    s.mesh_reflectivity = secMesh.reflectivity;
    s.mesh_ior = secMesh.ior;
    s.mesh_color[0] = secMesh.color.x;
    s.mesh_color[1] = secMesh.color.y;
    s.mesh_color[2] = secMesh.color.z;
    s.hybrid = true;
}

void setSecondaryPrim(int prim) {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.sec_prim = prim == KDGRT_SECONDARY_TRI ? KDGRT_SECONDARY_TRI : KDGRT_SECONDARY_SPHERE;
}

bool loadSecondaryMesh(const std::string& filepath, SecondaryMesh& mesh) {
    uint64_t n = 0;
    if (kdgrt_obj_read_secondary(filepath.c_str(), nullptr, 0, &n) != KDGRT_OK) {
        report("loadSecondaryMesh");
        return false;
    }
    mesh.vert.assign(3 * n, float3{0.0f, 0.0f, 0.0f});
    if (n && kdgrt_obj_read_secondary(filepath.c_str(), reinterpret_cast<float*>(mesh.vert.data()), n, &n) !=
                 KDGRT_OK) {
        report("loadSecondaryMesh");
        return false;
    }
    mesh.triangleCount = (int)n;
    return true;
}

static bool frame_info(const Camera& camera, int width, int height, kdgrt_camera_info* ci) {
    kdgrt_camera c{};
    for (int k = 0; k < 3; ++k) {
        c.pos[k] = camera.pos[k];
        c.u[k] = camera.uaxis[k];
        c.v[k] = camera.vaxis[k];
        c.n[k] = camera.naxis[k];
    }
    c.fovy_deg = camera.fovy;
    c.aspect = camera.aspect;
    c.near_c = camera.near_c;
    return kdgrt_camera_make(&c, width, height, ci) == KDGRT_OK;
}

float renderGaussianWithCudaFrame(const Camera& camera, int width, int height, float* d_framebuffer,
                                  hipStream_t stream) {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.ctx || !s.scene) {
        std::fprintf(stderr, "[kdgrt] renderGaussianWithCudaFrame: no scene set up.\n");
        return -1.0f;
    }
    kdgrt_camera_info ci;
    if (!frame_info(camera, width, height, &ci)) {
        report("renderGaussianWithCudaFrame");
        return -1.0f;
    }
    float ms = 0.0f;
    kdgrt_secondary sec{};
    kdgrt_secondary_sphere demo[2];
    if (s.hybrid) {
        sec.prim = s.sec_prim;
        sec.max_bounces = 5; // MAX_BOUNCES (cudaRenderer.cu:2230)
        if (s.sec_prim == KDGRT_SECONDARY_SPHERE) {
            kdgrt_secondary_demo_spheres(demo);
            sec.count = 2;
            sec.spheres = demo;
        } else {
            sec.count = s.mesh_tris;
            sec.d_vertices = s.d_mesh;
            sec.mesh_material = s.mesh_material;
            sec.mesh_reflectivity = s.mesh_reflectivity;
            sec.mesh_ior = s.mesh_ior;
            for (int k = 0; k < 3; ++k) sec.mesh_color[k] = s.mesh_color[k];
        }
    }
    if (kdgrt_timer_begin(s.ctx, stream) != KDGRT_OK) {
        report("renderGaussianWithCudaFrame");
        return -1.0f;
    }
    const kdgrt_status launched =
        s.hybrid ? kdgrt_render_hybrid(s.ctx, s.scene, &ci, width, height, 12, &sec, d_framebuffer, stream)
                 : kdgrt_render(s.ctx, s.scene, &ci, width, height, 12, KDGRT_OUT_RGB_F32, d_framebuffer, stream);
    if (launched != KDGRT_OK || kdgrt_timer_end(s.ctx, stream, &ms) != KDGRT_OK) {
        report("renderGaussianWithCudaFrame");
        return -1.0f;
    }
    s.last_w = width;
    s.last_h = height;
    return ms > 0.0f ? 1000.0f / ms : 0.0f; // kernel-only FPS (cudaRenderer.cu:3469-3471)
}

void warmUp(float* d_framebuffer, hipStream_t stream) {
    // The reference renders one frame of the configured window size with the
    // current constants into the caller's framebuffer (cudaRenderer.cu:2532-2554);
    // here: a frame from a corner of the scene box at the last frame size.
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.ctx || !s.scene) return;
    Camera cam{};
    for (int k = 0; k < 3; ++k) cam.pos[k] = s.aabb[3 + k] + 1.0f;
    cam.uaxis[0] = 1;
    cam.vaxis[1] = 1;
    cam.naxis[2] = 1;
    cam.fovy = 60;
    cam.aspect = (double)s.last_w / s.last_h;
    cam.near_c = 0.005;
    kdgrt_camera_info ci;
    if (!frame_info(cam, s.last_w, s.last_h, &ci)) return;
    if (kdgrt_render(s.ctx, s.scene, &ci, s.last_w, s.last_h, 12, KDGRT_OUT_RGB_F32, d_framebuffer, stream) !=
            KDGRT_OK ||
        kdgrt_stream_sync(s.ctx, stream) != KDGRT_OK)
        report("warmUp");
}

void cleanupCudaResources() {
    State& s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    drop_mesh(s);
    if (s.scene) kdgrt_scene_destroy(s.scene);
    s.scene = nullptr;
    if (s.ctx) kdgrt_shutdown(s.ctx);
    s.ctx = nullptr;
}

kdgrt_context* current_context() { return state().ctx; }
const kdgrt_scene* current_scene() { return state().scene; }

} // namespace kdgrt_ref
