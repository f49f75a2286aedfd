// Drives the reference-named C++ host API (include/kdgrt_reference_api.hpp)
// the way the reference's render-test mode does (Kd-treeConverterMain.cpp:
// 3618-3677): build/load the tree, invert scales, initCuda, setup, warmUp,
// frames into a caller-owned device framebuffer; then checks the frame
// against the C ABI's RGB render of the same view.  Then the SECONDARY_RAY
// set-up: an OBJ mirror quad written to argv[1] and read back with
// loadSecondaryMesh, frames with the demo spheres and with the mesh, each
// against kdgrt_render_hybrid of the same view.  Prints one line.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kdgrt_reference_api.hpp"

static double max_diff(const float* d_a, const float* d_b, size_t n, double* sum) {
    std::vector<float> a(n), b(n);
    hipMemcpy(a.data(), d_a, n * 4, hipMemcpyDeviceToHost);
    hipMemcpy(b.data(), d_b, n * 4, hipMemcpyDeviceToHost);
    double m = 0;
    *sum = 0;
    for (size_t i = 0; i < n; ++i) {
        m = std::fmax(m, std::fabs((double)a[i] - b[i]));
        *sum += a[i];
    }
    return m;
}

int main(int argc, char** argv) {
    const int n = 20000, W = 320, H = 200;
    std::vector<kdgrt_gaussian> g(n);
    if (kdgrt_generate_gaussians(1, n, 77, g.data()) != KDGRT_OK) return 2;
    kdgrt_tree* tree = nullptr;
    if (kdgrt_tree_build(g.data(), n, nullptr, &tree) != KDGRT_OK) return 3;
    kdgrt_tree_stats st;
    kdgrt_tree_stats_get(tree, &st);
    float aabb[6];
    kdgrt_tree_aabb(tree, aabb);

    kdgrt_ref::KdTree kt{};
    kt.tree = (kdgrt_ref::KdTreeNode*)kdgrt_tree_nodes(tree);
    kt.tree_node_count = (int)st.n_nodes;
    kt.prim_offset_list = (unsigned*)kdgrt_tree_offsets(tree);
    kt.prim_offset_count = (int)st.n_offsets;
    kdgrt_ref::CompositeObject obj{};
    obj.kd_tree = &kt;
    const float ref_aabb[6] = {aabb[0], aabb[3], aabb[1], aabb[4], aabb[2], aabb[5]}; // XMIN, XMAX, ...
    std::memcpy(obj.AABB, ref_aabb, sizeof(ref_aabb));
    std::vector<kdgrt_gaussian> gi = g;
    kdgrt_gaussians_invert_scales(gi.data(), gi.size()); // caller contract, Kd-treeConverterMain.cpp:3636

    if (!kdgrt_ref::initCuda()) return 4;
    kdgrt_ref::renderGaussianWithCudaSetup(obj, gi);
    float* fb = nullptr;
    if (hipMalloc(&fb, (size_t)W * H * 3 * 4) != hipSuccess) return 5;
    hipStream_t s;
    hipStreamCreate(&s);
    kdgrt_ref::Camera cam{};
    const float c[3] = {0.5f * (aabb[0] + aabb[3]), 0.5f * (aabb[1] + aabb[4]), 0.5f * (aabb[2] + aabb[5])};
    cam.pos[0] = c[0];
    cam.pos[1] = c[1];
    cam.pos[2] = c[2] + 25.0f;
    cam.uaxis[0] = 1;
    cam.vaxis[1] = 1;
    cam.naxis[2] = 1;
    cam.fovy = 60.0;
    cam.aspect = (double)W / H;
    cam.near_c = 0.005;
    cam.far_c = 20.0;
    float fps = kdgrt_ref::renderGaussianWithCudaFrame(cam, W, H, fb, s);
    kdgrt_ref::warmUp(fb, s);
    fps = kdgrt_ref::renderGaussianWithCudaFrame(cam, W, H, fb, s);
    std::vector<float> a((size_t)W * H * 3), b((size_t)W * H * 3);
    hipMemcpy(a.data(), fb, a.size() * 4, hipMemcpyDeviceToHost);

    kdgrt_camera kc{};
    for (int k = 0; k < 3; ++k) {
        kc.pos[k] = cam.pos[k];
        kc.u[k] = cam.uaxis[k];
        kc.v[k] = cam.vaxis[k];
        kc.n[k] = cam.naxis[k];
    }
    kc.fovy_deg = cam.fovy;
    kc.aspect = cam.aspect;
    kc.near_c = cam.near_c;
    kdgrt_camera_info ci;
    kdgrt_camera_make(&kc, W, H, &ci);
    float* fb2 = nullptr;
    hipMalloc(&fb2, (size_t)W * H * 3 * 4);
    kdgrt_render(kdgrt_ref::current_context(), kdgrt_ref::current_scene(), &ci, W, H, 12, KDGRT_OUT_RGB_F32, fb2, s);
    hipStreamSynchronize(s);
    hipMemcpy(b.data(), fb2, b.size() * 4, hipMemcpyDeviceToHost);
    double maxd = 0, sum = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        maxd = std::fmax(maxd, std::fabs((double)a[i] - b[i]));
        sum += a[i];
    }
    bool ok = maxd == 0.0 && sum > 0 && fps > 0;

    // SECONDARY_RAY: a mirror quad in front of the scene centre, as an OBJ file
    // (stored as (x, -z, y), so the file holds (x, z, -y))
    const std::string obj_path = std::string(argc > 1 ? argv[1] : ".") + "/secondary.obj";
    FILE* f = std::fopen(obj_path.c_str(), "w");
    if (!f) return 6;
    const float q = 6.0f, zc = c[2] + 4.0f;
    const float quad[4][3] = {{c[0] - q, c[1] - q, zc - 3}, {c[0] + q, c[1] - q, zc + 3},
                              {c[0] + q, c[1] + q, zc + 3}, {c[0] - q, c[1] + q, zc - 3}};
    for (auto& v : quad) std::fprintf(f, "v %.6f %.6f %.6f\n", v[0], v[2], -v[1]);
    std::fprintf(f, "f 1 2 3 4\n");
    std::fclose(f);
    kdgrt_ref::SecondaryMesh mesh;
    if (!kdgrt_ref::loadSecondaryMesh(obj_path, mesh) || mesh.triangleCount != 2) return 7;
    mesh.materialType = 1; // as the reference's set-up does (Kd-treeConverterMain.cpp:3052-3055)
    kdgrt_ref::renderGaussianWithCudaSetup(obj, gi, mesh);
    double hyb_sum[2] = {0, 0}, hyb_max[2] = {1, 1};
    for (int mode = 0; mode < 2; ++mode) {
        kdgrt_ref::setSecondaryPrim(mode == 0 ? KDGRT_SECONDARY_SPHERE : KDGRT_SECONDARY_TRI);
        fps = kdgrt_ref::renderGaussianWithCudaFrame(cam, W, H, fb, s);
        kdgrt_secondary sec{};
        kdgrt_secondary_sphere demo[2];
        float* d_tri = nullptr;
        sec.prim = mode == 0 ? KDGRT_SECONDARY_SPHERE : KDGRT_SECONDARY_TRI;
        sec.max_bounces = 5;
        if (mode == 0) {
            kdgrt_secondary_demo_spheres(demo);
            sec.count = 2;
            sec.spheres = demo;
        } else {
            hipMalloc(&d_tri, mesh.vert.size() * sizeof(float3));
            hipMemcpy(d_tri, mesh.vert.data(), mesh.vert.size() * sizeof(float3), hipMemcpyHostToDevice);
            sec.count = mesh.triangleCount;
            sec.d_vertices = d_tri;
            sec.mesh_material = mesh.materialType;
            sec.mesh_reflectivity = mesh.reflectivity;
            sec.mesh_ior = mesh.ior;
            sec.mesh_color[0] = mesh.color.x;
            sec.mesh_color[1] = mesh.color.y;
            sec.mesh_color[2] = mesh.color.z;
        }
        if (kdgrt_render_hybrid(kdgrt_ref::current_context(), kdgrt_ref::current_scene(), &ci, W, H, 12, &sec, fb2,
                                s) != KDGRT_OK)
            return 8;
        hipStreamSynchronize(s);
        hyb_max[mode] = max_diff(fb, fb2, (size_t)W * H * 3, &hyb_sum[mode]);
        ok = ok && fps > 0 && hyb_max[mode] == 0.0 && hyb_sum[mode] > 0;
        if (d_tri) hipFree(d_tri);
    }
    hipFree(fb);
    hipFree(fb2);
    hipStreamDestroy(s);
    kdgrt_ref::cleanupCudaResources();
    kdgrt_tree_free(tree);
    std::printf("host_api ok=%d maxdiff=%g sum=%g fps=%g hybrid spheres maxdiff=%g sum=%g mesh maxdiff=%g sum=%g\n",
                ok, maxd, sum, fps, hyb_max[0], hyb_sum[0], hyb_max[1], hyb_sum[1]);
    return ok ? 0 : 1;
}
