// secondary_device.hpp — the secondary primitives of the hybrid render
// (renderKernelGaussian_Mesh_Hybrid, Kd-treeConverter_v0.9/Kd-treeConverter/
// cudaRenderer.cu:2217-2311): mirror and glass spheres or a triangle mesh
// that the Gaussian trace stops at and bounces from.
//
// Arithmetic follows the reference statement by statement with the float3
// operators of SGRTx2Lib/cuda_math.h (component-wise, dot = x*x' + y*y' +
// z*z' left to right, normalize = v * (1 / sqrtf(dot(v, v)))), IEEE '/' and
// sqrtf, and no FMA contraction (built with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kdgrt_dev {

constexpr int kSecMaxSpheres = 8;
constexpr float kSecEps = 0.00001f; // EPSILON (DefineConst.h:24)
constexpr float kSecFar = 1e8f;     // "no hit" distance of MeshHit (cudaRenderer.cu:2114, 2173)

struct SecSphere { // DemoSphere (cudaRenderer.cu:2153-2160)
    float c[3];
    float r;
    int32_t mat;
    float refl, ior;
    float col[3];
};

// Secondary scene of one launch: prim 1 = spheres (in the kernel argument),
// prim 0 = triangle mesh (9 floats per triangle in device memory) with one
// material for the whole mesh (SecondaryMeshDevice, cudaRenderer.h:21-28).
struct SecScene {
    int32_t prim, count, max_bounces;
    int32_t mesh_mat;
    float mesh_refl, mesh_ior;
    float mesh_col[3];
    const float* __restrict__ tri;
    SecSphere sph[kSecMaxSpheres];
};

// MeshHit (cudaRenderer.cu:2050-2060) of the closest primitive.
struct SecHit {
    float n[3], p[3];
    int32_t mat;
    float refl, ior;
    float col[3];
    bool front;
};

__device__ __forceinline__ float sdot(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ void snormalize(float v[3]) {
    const float inv = 1.0f / sqrtf(sdot(v, v));
    v[0] = v[0] * inv;
    v[1] = v[1] * inv;
    v[2] = v[2] * inv;
}

// Sphere root of intersectSphere (cudaRenderer.cu:2176-2192); returns the
// accepted root or a value >= best when the sphere is missed.
__device__ __forceinline__ float sphere_root(const SecSphere& s, const float o[3], const float d[3]) {
    const float oc[3] = {o[0] - s.c[0], o[1] - s.c[1], o[2] - s.c[2]};
    const float a = sdot(d, d);
    const float b = 2.0f * sdot(oc, d);
    const float c = sdot(oc, oc) - s.r * s.r;
    const float disc = b * b - 4.0f * a * c;
    if (disc >= 0.0f) {
        const float sq = sqrtf(disc);
        float root = (-b - sq) / (2.0f * a);
        if (root < kSecEps) root = (-b + sq) / (2.0f * a);
        if (root > kSecEps) return root;
    }
    return 3.402823466e+38f;
}

// Moller-Trumbore of rayTriIntersect (cudaRenderer.cu:1017-1065; the
// overload at :2089-2108 computes the same values); returns t or +FLT_MAX.
__device__ __forceinline__ float tri_t(const float* __restrict__ v, const float o[3], const float d[3]) {
    const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    const float pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    const float det = sdot(e1, pv);
    if (det > -kSecEps && det < kSecEps) return 3.402823466e+38f;
    const float inv_det = 1.0f / det;
    const float tv[3] = {o[0] - v[0], o[1] - v[1], o[2] - v[2]};
    const float u = sdot(tv, pv) * inv_det;
    if (u < 0.0f || u > 1.0f) return 3.402823466e+38f;
    const float qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    const float w = sdot(d, qv) * inv_det;
    if (w < 0.0f || u + w > 1.0f) return 3.402823466e+38f;
    const float t = sdot(e2, qv) * inv_det;
    return t > kSecEps ? t : 3.402823466e+38f;
}

// Closest primitive along the ray (intersectSphere / intersectTriangleMesh,
// cudaRenderer.cu:2111-2151, 2170-2215): a primitive replaces the current
// best only when strictly closer, starting from t = 1e8.  Returns its index
// (-1: none) and sets t.  The hit record is rebuilt from (index, t) by
// sec_record, so only two values stay live across the Gaussian trace.
__device__ __forceinline__ int sec_closest(const SecScene& s, const float o[3], const float d[3], float& t) {
    int best = -1;
    t = kSecFar;
    if (s.prim == 1) {
        for (int i = 0; i < s.count; ++i) {
            const float r = sphere_root(s.sph[i], o, d);
            if (r < t) {
                t = r;
                best = i;
            }
        }
    } else {
        for (int i = 0; i < s.count; ++i) {
            const float r = tri_t(s.tri + 9 * (size_t)i, o, d);
            if (r < t) {
                t = r;
                best = i;
            }
        }
    }
    return best;
}

__device__ __forceinline__ void sec_record(const SecScene& s, int idx, float t, const float o[3], const float d[3],
                                           SecHit& h) {
    h.p[0] = o[0] + d[0] * t;
    h.p[1] = o[1] + d[1] * t;
    h.p[2] = o[2] + d[2] * t;
    if (s.prim == 1) {
        const SecSphere& q = s.sph[idx];
        h.n[0] = h.p[0] - q.c[0];
        h.n[1] = h.p[1] - q.c[1];
        h.n[2] = h.p[2] - q.c[2];
        snormalize(h.n);
        h.mat = q.mat;
        h.refl = q.refl;
        h.ior = q.ior;
        h.col[0] = q.col[0];
        h.col[1] = q.col[1];
        h.col[2] = q.col[2];
    } else {
        const float* v = s.tri + 9 * (size_t)idx;
        const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        h.n[0] = e1[1] * e2[2] - e1[2] * e2[1];
        h.n[1] = e1[2] * e2[0] - e1[0] * e2[2];
        h.n[2] = e1[0] * e2[1] - e1[1] * e2[0];
        snormalize(h.n);
        h.mat = s.mesh_mat;
        h.refl = s.mesh_refl;
        h.ior = s.mesh_ior;
        h.col[0] = s.mesh_col[0];
        h.col[1] = s.mesh_col[1];
        h.col[2] = s.mesh_col[2];
    }
    h.front = !(sdot(d, h.n) > 0.0f);
    if (!h.front) {
        h.n[0] = -h.n[0];
        h.n[1] = -h.n[1];
        h.n[2] = -h.n[2];
    }
}

// The material step after a segment (cudaRenderer.cu:2264-2302): sets the
// next segment's ray and updates colour and throughput.  Returns false when
// the path ends (an opaque material).
__device__ __forceinline__ bool sec_bounce(const SecHit& h, int bounce, float o[3], float d[3], float fin[3],
                                           float thr[3]) {
    const float I[3] = {d[0], d[1], d[2]};
    const float* N = h.n;
    if (h.mat == 1) { // mirror
        const float dn = sdot(I, N);
        float r[3];
        for (int k = 0; k < 3; ++k) r[k] = I[k] - (N[k] * 2.0f) * dn;
        for (int k = 0; k < 3; ++k) o[k] = h.p[k] + N[k] * 1e-4f;
        snormalize(r);
        for (int k = 0; k < 3; ++k) d[k] = r[k];
        const float w = 1.0f - h.refl;
        for (int k = 0; k < 3; ++k) {
            const float mc = bounce > 0 ? 0.0f : h.col[k];
            fin[k] = fin[k] + (thr[k] * w) * mc;
        }
        for (int k = 0; k < 3; ++k) thr[k] = thr[k] * h.refl;
        return true;
    }
    if (h.mat == 2) { // glass
        const float eta = h.front ? 1.0f / h.ior : h.ior;
        const float cosI = -sdot(N, I);
        const float sinT2 = eta * eta * (1.0f - cosI * cosI);
        float r[3];
        if (sinT2 > 1.0f) { // total internal reflection
            const float dn = sdot(I, N);
            for (int k = 0; k < 3; ++k) r[k] = I[k] - (N[k] * 2.0f) * dn;
            for (int k = 0; k < 3; ++k) o[k] = h.p[k] + N[k] * 1e-4f;
        } else {
            const float cosT = sqrtf(1.0f - sinT2);
            const float f = eta * cosI - cosT;
            for (int k = 0; k < 3; ++k) r[k] = I[k] * eta + N[k] * f;
            for (int k = 0; k < 3; ++k) o[k] = h.p[k] - N[k] * 1e-4f;
        }
        snormalize(r);
        for (int k = 0; k < 3; ++k) d[k] = r[k];
        for (int k = 0; k < 3; ++k) thr[k] = thr[k] * h.col[k];
        return true;
    }
    for (int k = 0; k < 3; ++k) fin[k] = fin[k] + thr[k] * h.col[k];
    return false;
}

} // namespace kdgrt_dev
