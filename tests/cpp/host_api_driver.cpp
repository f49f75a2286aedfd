// Drives the reference-named C++ host API (include/kdgrt_reference_api.hpp)
// the way the reference's render-test mode does (Kd-treeConverterMain.cpp:
// 3618-3677): build/load the tree, invert scales, initCuda, setup, warmUp,
// frames into a caller-owned device framebuffer; then checks the frame
// against the C ABI's RGB render of the same view.  Prints one line.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kdgrt_reference_api.hpp"

int main() {
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
    hipFree(fb);
    hipFree(fb2);
    hipStreamDestroy(s);
    kdgrt_ref::cleanupCudaResources();
    kdgrt_tree_free(tree);
    std::printf("host_api ok=%d maxdiff=%g sum=%g fps=%g\n", maxd == 0.0 && sum > 0 && fps > 0, maxd, sum, fps);
    return (maxd == 0.0 && sum > 0 && fps > 0) ? 0 : 1;
}
