// kdgrt_reference_api.hpp — drop-in C++ host API with the reference
// renderer's entry points (Kd-treeConverter_v0.9/Kd-treeConverter/
// cudaRenderer.h:60-109), implemented on the C ABI of kdgrt.h.
//
// A caller written against cudaRenderer.h keeps its call sites:
//   initCuda()                                   cudaRenderer.cu:2438-2453
//   renderGaussianWithCudaSetup(obj, gaussians)  cudaRenderer.cu:2557-2703
//   renderGaussianWithCudaFrame(cam, w, h, d_fb, stream)  cudaRenderer.cu:3099-3484
//   warmUp(d_fb, stream)                         cudaRenderer.cu:2532-2554
//   cleanupCudaResources()                       cudaRenderer.cu:3486-3521
// with the same structs (layout-compatible: Kd-treeConverter.h:282-335,
// OpenGLStuffs.h:21-28), the same caller contract (Gaussians carry inverted
// scales, Kd-treeConverterMain.cpp:3636-3642), the same framebuffer
// (caller-owned device memory, w*h*3 float, rows flipped) and the same
// blocking frame call returning kernel-only FPS.
//
// Differences: device errors are reported on stderr and by the return value
// (false / negative FPS) instead of exit(); the scene AABB is taken from the
// CompositeObject as the reference does; the stream is a hipStream_t.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "kdgrt.h"

#define KDGRT_REF_API __attribute__((visibility("default")))

namespace kdgrt_ref {

// Kd-treeConverter.h:282-285
struct KdTreeNode {
    unsigned int x;
    unsigned int y;
};

// Kd-treeConverter.h:287-296 (PRIMITIVE_TYPE != TRI)
struct KdTree {
    KdTreeNode* tree;
    int tree_node_count;
    unsigned int* prim_offset_list;
    int prim_offset_count;
    float AABB[6];
};

// Kd-treeConverter.h:250-255
struct ExtendedVertex {
    float vertex[3];
    float normal[3];
    int material_ID;
    char pad[4];
};

// Kd-treeConverter.h:303-317; AABB order XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX.
struct CompositeObject {
    int n_triangles;
    float AABB[6];
    ExtendedVertex* extended_vertices;
    KdTree* kd_tree;
    void* leafDebug;
};

// Kd-treeConverter.h:319-335 with QUATERNION = true (256 B).
typedef kdgrt_gaussian Gaussian;

// OpenGLStuffs.h:21-28
struct Camera {
    float pos[3];
    float uaxis[3], vaxis[3], naxis[3];
    float mat[16];
    float mat_inv[16];
    int move, upanddown;
    double fovy, aspect, near_c, far_c;
};

KDGRT_REF_API bool initCuda();
KDGRT_REF_API void renderGaussianWithCudaSetup(const CompositeObject& object, const std::vector<Gaussian>& gaussians);
KDGRT_REF_API float renderGaussianWithCudaFrame(const Camera& camera, int width, int height, float* d_framebuffer,
                                  hipStream_t stream);
KDGRT_REF_API void warmUp(float* d_framebuffer, hipStream_t stream);
KDGRT_REF_API void cleanupCudaResources();

// Handles of the singleton state, for callers that mix both APIs.
KDGRT_REF_API kdgrt_context* current_context();
KDGRT_REF_API const kdgrt_scene* current_scene();

} // namespace kdgrt_ref
