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
// and, for the SECONDARY_RAY build,
//   renderGaussianWithCudaSetup(obj, gaussians, secMesh)  cudaRenderer.cu:2557-2565, 2678-2680
//   loadSecondaryMesh(path, mesh)                Kd-treeConverterMain.cpp:178-226
// with the same structs (layout-compatible: Kd-treeConverter.h:282-335,
// OpenGLStuffs.h:21-28), the same caller contract (Gaussians carry inverted
// scales, Kd-treeConverterMain.cpp:3636-3642), the same framebuffer
// (caller-owned device memory, w*h*3 float, rows flipped) and the same
// blocking frame call returning kernel-only FPS.
//
// Differences: device errors are reported on stderr and by the return value
// (false / negative FPS) instead of exit(); the scene AABB is taken from the
// CompositeObject as the reference does; the stream is a hipStream_t.  Frames
// of an unchanged or slowly moving camera start the tiles that ran longest in
// the previous frame first (kdgrt_scene_set_tile_schedule); the image is the
// same as in raster order.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
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

// cudaRenderer.h:11-19 (SECONDARY_RAY builds): the secondary mesh and its
// one material (1 reflect, 2 refract, other: opaque colour).
struct SecondaryMesh {
    std::vector<float3> vert;
    int triangleCount = 0;
    int materialType = 1;
    float reflectivity = 0.95f;
    float ior = 1.1f;
    float3 color = {0.95f, 0.95f, 0.95f};
};

KDGRT_REF_API bool initCuda();
KDGRT_REF_API void renderGaussianWithCudaSetup(const CompositeObject& object, const std::vector<Gaussian>& gaussians);
KDGRT_REF_API float renderGaussianWithCudaFrame(const Camera& camera, int width, int height, float* d_framebuffer,
                                  hipStream_t stream);
KDGRT_REF_API void warmUp(float* d_framebuffer, hipStream_t stream);

// SECONDARY_RAY setup: the scene plus the secondary mesh (initSecondaryMesh,
// cudaRenderer.cu:2064-2083); frames then run the hybrid render
// (renderKernelGaussian_Mesh_Hybrid, cudaRenderer.cu:2217-2311, MAX_BOUNCES 5)
// into the same framebuffer.
KDGRT_REF_API void renderGaussianWithCudaSetup(const CompositeObject& object, const std::vector<Gaussian>& gaussians,
                                               SecondaryMesh& secMesh);
// SECONDARY_PRIM (Kd-treeConverter.h:28-30) is a compile-time switch in the
// reference; here it is chosen at run time: KDGRT_SECONDARY_SPHERE (the
// reference's setting: its two demo spheres, cudaRenderer.cu:2162-2167) or
// KDGRT_SECONDARY_TRI (the mesh given at set-up).
KDGRT_REF_API void setSecondaryPrim(int prim);
// OBJ triangles as the secondary mesh, vertices stored as (x, -z, y);
// returns false when the file cannot be read or holds faces of more than four
// vertices.
KDGRT_REF_API bool loadSecondaryMesh(const std::string& filepath, SecondaryMesh& mesh);
KDGRT_REF_API void cleanupCudaResources();

// Handles of the singleton state, for callers that mix both APIs.
KDGRT_REF_API kdgrt_context* current_context();
KDGRT_REF_API const kdgrt_scene* current_scene();

} // namespace kdgrt_ref
