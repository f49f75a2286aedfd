// ORACLE — test infrastructure only.  Nothing in the product links or calls
// this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg load liboracle.so, to compare results against.
//
// A CPU restatement of the reference's plain render path
// (renderKernelGaussian_sortNode, SECONDARY_RAY=0, PRIMITIVE_TYPE
// ELLIPSOID_BY_TRI, QUATERNION, VOLUME_ISECT, GAUSSIAN_DEGREE 4,
// SPH_EVAL_DEGREE 3), written from the reference source, one ray per call,
// templated on the arithmetic type:
//   T = float  : the FP32 oracle.  Parity gate: the HIP kernel must match it
//                bit for bit (composited IDs, counters) and within 1e-5 (RGB, T).
//   T = double : the FP64 oracle, for error statistics only.
// Paths below are relative to Kd-treeConverter_v0.9/Kd-treeConverter/.
//
// Substitutions for CUDA intrinsics whose bits cannot be reproduced off
// NVIDIA hardware (the reference is built without fast-math, so '/', sqrtf
// are IEEE there too):
//   __fdividef(a,b) -> a * (1/b)     (cudaRenderer.cu:383-394, 1560): the
//                      intrinsic is a multiply by an approximate reciprocal
//                      (PTX div.approx.f32); here 1/b is the IEEE reciprocal
//                      of the ray direction component, formed once per ray
//   rsqrtf(x)       -> 1.0f/sqrtf(x) (cudaRenderer.cu:55)
//   expf            -> kd_expf       (own Cody-Waite exp, shared with the kernel)
//   nvcc FMA contraction -> none    (built with -ffp-contract=off)
// Parity against NVIDIA FP32 bits is therefore "parity unpinned"; this
// restatement is the definitional FP32 result.
//
// and of the hybrid render (renderKernelGaussian_Mesh_Hybrid, SECONDARY_RAY=1,
// cudaRenderer.cu:2217-2311) with sphere or triangle secondary primitives.
//
// Pins (tests/): closed forms for ray generation, slab test, intersection,
// response and SH; brute force on single-leaf scenes; hand-built trees with
// hand-counted node visits (ring-stack overwrite and restart); FP64
// agreement; golden fixtures.  Hybrid: closed-form sphere and triangle hits,
// mirror / Snell / total-internal-reflection directions computed in double,
// the running-colour sum of a mirror path by hand, segments without a
// secondary hit equal to the plain path.

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../3dgrt_kdtree_amd/csrc/kd_expf.h"

namespace {

struct Gauss256 { // Kd-treeConverter.h:319-335 (QUATERNION)
    float pos[3];
    float opacity;
    float scale[3]; // inverted scale on upload (Kd-treeConverterMain.cpp:3636-3642)
    float k_scale;
    float rot[4]; // w, x, y, z
    float f_dc[3];
    float f_rest[45];
    float pad[4];
};
static_assert(sizeof(Gauss256) == 256, "Gaussian record must be 256 B");

struct Node {
    uint32_t x, y;
};

struct CamInfo { // cudaRenderer.cu:318
    float eye[3], u[3], v[3], start[3];
    float step_x, step_y;
};

template <class T>
struct V3 {
    T x, y, z;
};
// SGRTx2Lib/cuda_math.h operator semantics (component-wise, left to right).
template <class T> inline V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> mul(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> muls(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T tsqrt(T x) { return std::sqrt(x); }
inline float tsqrt(float x) { return sqrtf(x); }
// normalize: v * (1/sqrtf(dot(v,v)))  (cuda_math.h:403-407)
template <class T> inline V3<T> normalize(V3<T> v) {
    T inv = T(1) / tsqrt(dot(v, v));
    return muls(v, inv);
}
// safe_normalize with rsqrtf -> 1/sqrtf (cudaRenderer.cu:53-56)
template <class T> inline V3<T> safe_normalize(V3<T> v) {
    const T l = v.x * v.x + v.y * v.y + v.z * v.z;
    return l > T(0) ? muls(v, T(1) / tsqrt(l)) : v;
}
template <class T> inline T texp(T x) { return std::exp(x); }
inline float texp(float x) { return kd_expf(x); }
template <class T> inline T tmin(T a, T b) { return std::fmin(a, b); }
template <class T> inline T tmax(T a, T b) { return std::fmax(a, b); }

// float33 * p: (dot(m0,p), dot(m1,p), dot(m2,p))  (cudaRenderer.cu:49-51)
template <class T> inline V3<T> rowmul(const V3<T> m[3], V3<T> p) { return {dot(m[0], p), dot(m[1], p), dot(m[2], p)}; }

// quaternionWXYZToMatrixTranspose (cudaRenderer.cu:808-828)
template <class T> inline void quat_to_mt(const float q[4], V3<T> m[3]) {
    const T r = q[0], x = q[1], y = q[2], z = q[3];
    const T xx = x * x, yy = y * y, zz = z * z;
    const T xy = x * y, xz = x * z, yz = y * z;
    const T rx = r * x, ry = r * y, rz = r * z;
    m[0] = {(T(1) - T(2) * (yy + zz)), T(2) * (xy + rz), T(2) * (xz - ry)};
    m[1] = {T(2) * (xy - rz), (T(1) - T(2) * (xx + zz)), T(2) * (yz + rx)};
    m[2] = {T(2) * (xz + ry), T(2) * (yz - rx), (T(1) - T(2) * (xx + yy))};
}

// Spherical-harmonic constants (cudaRenderer.h:33-48).  The reference
// defines every SH_C* with an f suffix, so each basis factor is a float
// product: SH_C1 * y rounds once, and the degree-3 factors
// (SH_C3_0 * y) * (3xx - yy) round twice (C++ evaluates a * b * c as
// (a * b) * c).  ShC<float> holds those float literals; ShC<double> holds the
// same decimal values as double literals for the FP64 oracle.
template <class T> struct ShC;
template <> struct ShC<float> {
    static constexpr float C0 = 0.28209479177387814f, C1 = 0.4886025119029199f, C2_0 = 1.0925484305920792f,
                           C2_1 = -1.0925484305920792f, C2_2 = 0.31539156525252005f, C2_3 = -1.0925484305920792f,
                           C2_4 = 0.5462742152960396f, C3_0 = -0.5900435899266435f, C3_1 = 2.890611442640554f,
                           C3_2 = -0.4570457994644658f, C3_3 = 0.3731763325901154f, C3_4 = -0.4570457994644658f,
                           C3_5 = 1.445305721320277f, C3_6 = -0.5900435899266435f;
};
template <> struct ShC<double> {
    static constexpr double C0 = 0.28209479177387814, C1 = 0.4886025119029199, C2_0 = 1.0925484305920792,
                            C2_1 = -1.0925484305920792, C2_2 = 0.31539156525252005, C2_3 = -1.0925484305920792,
                            C2_4 = 0.5462742152960396, C3_0 = -0.5900435899266435, C3_1 = 2.890611442640554,
                            C3_2 = -0.4570457994644658, C3_3 = 0.3731763325901154, C3_4 = -0.4570457994644658,
                            C3_5 = 1.445305721320277, C3_6 = -0.5900435899266435;
};

// The 16 basis scalars of eval_sh_final(3, dir, g) (cudaRenderer.cu:517-553)
// as the reference forms them: s[0] = SH_C0; degree 1 SH_C1 * (y, z, x) with
// the signs (-, +, -) applied in eval_sh; degree 2 SH_C2_i * poly; degree 3
// (SH_C3_i * axis) * poly, or (SH_C3_1 * xy) * z.
template <class T> inline void sh_basis(V3<T> d, T s[16]) {
    using C = ShC<T>;
    const T x = d.x, y = d.y, z = d.z;
    s[0] = C::C0;
    s[1] = C::C1 * y;
    s[2] = C::C1 * z;
    s[3] = C::C1 * x;
    const T xx = x * x, yy = y * y, zz = z * z;
    const T xy = x * y, yz = y * z, xz = x * z;
    s[4] = C::C2_0 * xy;
    s[5] = C::C2_1 * yz;
    s[6] = C::C2_2 * (T(2) * zz - xx - yy);
    s[7] = C::C2_3 * xz;
    s[8] = C::C2_4 * (xx - yy);
    s[9] = C::C3_0 * y * (T(3) * xx - yy);
    s[10] = C::C3_1 * xy * z;
    s[11] = C::C3_2 * y * (T(4) * zz - xx - yy);
    s[12] = C::C3_3 * z * (T(2) * zz - T(3) * xx - T(3) * yy);
    s[13] = C::C3_4 * x * (T(4) * zz - xx - yy);
    s[14] = C::C3_5 * z * (xx - yy);
    s[15] = C::C3_6 * x * (xx - T(3) * yy);
}

// eval_sh_final(3, dir, g) (cudaRenderer.cu:487-558): rad = SH_C0 * c0, then
// rad - s1 c1 + s2 c2 - s3 c3, then + s_i c_i for i = 4..15, left to right
// (operator*(float, float3) scales each component), + 0.5, clamped below at 0.
template <class T> inline V3<T> eval_sh(V3<T> d, const Gauss256& g) {
    V3<T> c[16];
    c[0] = {T(g.f_dc[0]), T(g.f_dc[1]), T(g.f_dc[2])};
    for (int i = 0; i < 15; ++i) c[i + 1] = {T(g.f_rest[i]), T(g.f_rest[15 + i]), T(g.f_rest[30 + i])};
    T s[16];
    sh_basis<T>(d, s);
    V3<T> rad = muls(c[0], s[0]);
    rad = sub(rad, muls(c[1], s[1]));
    rad = add(rad, muls(c[2], s[2]));
    rad = sub(rad, muls(c[3], s[3]));
    for (int i = 4; i < 16; ++i) rad = add(rad, muls(c[i], s[i]));
    rad = add(rad, V3<T>{T(0.5f), T(0.5f), T(0.5f)});
    return {tmax(rad.x, T(0)), tmax(rad.y, T(0)), tmax(rad.z, T(0))};
}

struct Hit {
    float t64_unused;
};

template <class T>
struct HitRec {
    T t;
    uint32_t id;
};

struct Scene {
    const Node* nodes;
    const uint32_t* offsets;
    const Gauss256* g;
    float bmin[3], bmax[3];
};

struct RayOut {
    int32_t cnt[8];
    uint64_t hash;
};

const uint64_t kFnvBasis = 0xcbf29ce484222325ull, kFnvPrime = 0x100000001b3ull;

// rayPrimIntersect, VOLUME_ISECT (cudaRenderer.cu:1069-1135): t of the
// closest approach in normalised space and grayDist, the squared distance
// there; accepted when t_near <= t <= t_far and grayDist < 8.
template <class T>
inline bool ray_gaussian(const V3<T>& o, const V3<T>& d, const Gauss256& g, T t_near, T t_far, T& t, T& grayDist) {
    const V3<T> ppos = {T(g.pos[0]), T(g.pos[1]), T(g.pos[2])};
    const V3<T> giscl = {T(g.scale[0]), T(g.scale[1]), T(g.scale[2])};
    const V3<T> pscale = {T(1) / giscl.x, T(1) / giscl.y, T(1) / giscl.z};
    V3<T> m[3];
    quat_to_mt<T>(g.rot, m);
    const V3<T> gposc = sub(o, ppos);
    const V3<T> gposcr = rowmul(m, gposc);
    const V3<T> gro = mul(giscl, gposcr);
    const V3<T> rayDirR = rowmul(m, d);
    const V3<T> grdu = mul(rayDirR, giscl);
    const V3<T> grd = safe_normalize(grdu);
    const T grp = -dot(grd, gro);
    const V3<T> grds = muls(mul(pscale, grd), grp);
    t = (grp < T(0) ? T(-1) : T(1)) * tsqrt(dot(grds, grds));
    const V3<T> gcrod = cross(grd, gro);
    grayDist = dot(gcrod, gcrod);
    return !(t < t_near || t > t_far) && grayDist < T(8);
}

template <class T>
inline void ray_prim(const V3<T>& o, const V3<T>& d, uint32_t id, const Gauss256& g, T t_near, T t_far,
                     std::vector<HitRec<T>>& hits) {
    T t, gd;
    if (ray_gaussian<T>(o, d, g, t_near, t_far, t, gd)) hits.push_back({t, id});
}

// singlePassIntersectGaussian_sortNode_onlyShortStack, ELLIPSOID_BY_TRI
// branch (cudaRenderer.cu:1488-1752), with the HIT_AND_NODE_COUNT_DEBUG
// counters (:1504-1606, 1679-1681).  D = short-stack depth (ring of
// cudaRenderer.cu:189-210; the reference uses SHORT_STACK_DEPTH 12).
// hybrid: one segment of the hybrid path (SECONDARY_RAY, :1500-1502,
// 1538-1541, 1612-1616): color and A are the running values passed in by
// reference and not reset, t_scene_far is clamped to max_t, and the composite
// loop stops at a hit beyond max_t.
// Records a ray reads, as byte maps (any may be null): node records visited
// (by index), leaf slots tested, leaf slots whose test accepted, Gaussians
// entering the composite loop, Gaussians blended.
struct Touch {
    uint8_t* node;
    uint8_t* slot;
    uint8_t* hit_slot;
    uint8_t* eval;
    uint8_t* blend;
};

template <class T>
void trace(const Scene& sc, V3<T> o, V3<T> d, int D, V3<T>& color, T& A, RayOut& ro, uint32_t* seq, int seq_cap,
           bool hybrid = false, T max_t = T(0), const Touch* tc = nullptr) {
    std::memset(ro.cnt, 0, sizeof(ro.cnt));
    ro.hash = kFnvBasis;
    int nseq = 0;
    if (!hybrid) {
        color = {T(0), T(0), T(0)};
        A = T(0);
    }
    // BoundsRayIntersect (cudaRenderer.cu:379-399), t in [1e-2, FLT_MAX].
    T t_scene_near = T(1e-2f), t_scene_far = T(FLT_MAX);
    {
        const T po[3] = {o.x, o.y, o.z}, pd[3] = {d.x, d.y, d.z};
        for (int k = 0; k < 3; ++k) {
            const T inv = T(1) / pd[k];
            T l1 = (T(sc.bmin[k]) - po[k]) * inv;
            T l2 = (T(sc.bmax[k]) - po[k]) * inv;
            t_scene_near = tmax(tmin(l1, l2), t_scene_near);
            t_scene_far = tmin(tmax(l1, l2), t_scene_far);
        }
        if (!((t_scene_far >= t_scene_near) & (t_scene_far >= T(0)))) return;
    }
    if (hybrid) {
        t_scene_far = tmin(t_scene_far, max_t);
        if (t_scene_near > t_scene_far) return;
    }
    // Ring stack (cudaRenderer.cu:189-210).
    std::vector<uint32_t> st_node(D);
    std::vector<T> st_t(D);
    unsigned top = (unsigned)D - 1, quant = 0;

    Node node = sc.nodes[0];
    uint32_t cur = 0; // index of node (touch maps)
    T t_near = t_scene_near, t_far = t_scene_far;
    std::vector<HitRec<T>> hits;
    const T po[3] = {o.x, o.y, o.z}, pd[3] = {d.x, d.y, d.z};
    const T pinv[3] = {T(1) / pd[0], T(1) / pd[1], T(1) / pd[2]};
    while (A < T(0.99f)) {
        while ((node.x & 7u) != 3u) {
            const unsigned child = node.x >> 3;
            const unsigned axis = node.x & 3u;
            const unsigned ax = axis > 2 ? 2 : axis; // get_dir_pos: default -> z
            float split;
            std::memcpy(&split, &node.y, 4);
            const T t_split = (T(split) - po[ax]) * pinv[ax];
            const unsigned sign = std::signbit(pd[ax]) ? 1u : 0u;
            const unsigned idx = child + (sign ^ (t_split <= t_near ? 1u : 0u));
            if (t_near < t_split && t_split < t_far) {
                if (++top == (unsigned)D) top = 0;
                quant = quant + 1 < (unsigned)D ? quant + 1 : (unsigned)D;
                st_node[top] = child + (sign ^ 1u);
                st_t[top] = t_far;
                t_far = t_split;
            }
            ro.cnt[0]++; // node_visits
            if (tc && tc->node) tc->node[cur] = 1;
            node = sc.nodes[idx];
            cur = idx;
        }
        if (tc && tc->node) tc->node[cur] = 1;
        const uint32_t base = node.y, count = node.x >> 3;
        hits.clear();
        ro.cnt[1]++; // leaf_visits
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t id = sc.offsets[base + k];
            const size_t before = hits.size();
            ray_prim<T>(o, d, id, sc.g[id], t_near, t_far, hits);
            if (tc && tc->slot) tc->slot[base + k] = 1;
            if (tc && tc->hit_slot && hits.size() > before) tc->hit_slot[base + k] = 1;
        }
        ro.cnt[2] += (int32_t)count;       // intersection_tests
        ro.cnt[3] += (int32_t)hits.size(); // hits_found
        if ((int32_t)hits.size() > ro.cnt[5]) ro.cnt[5] = (int32_t)hits.size();
        // sortHits: stable insertion sort on t (cudaRenderer.cu:415-426).
        for (size_t i = 1; i < hits.size(); ++i) {
            HitRec<T> key = hits[i];
            long j = (long)i - 1;
            while (j >= 0 && hits[j].t > key.t) {
                hits[j + 1] = hits[j];
                --j;
            }
            hits[j + 1] = key;
        }
        // Composite loop (cudaRenderer.cu:1608-1688).
        for (size_t i = 0; i < hits.size(); ++i) {
            if (hybrid && hits[i].t > max_t) break;
            ro.cnt[6]++; // N_eval
            const uint32_t gid = hits[i].id;
            if (tc && tc->eval) tc->eval[gid] = 1;
            const Gauss256& g = sc.g[gid];
            const V3<T> ppos = {T(g.pos[0]), T(g.pos[1]), T(g.pos[2])};
            const V3<T> giscl = {T(g.scale[0]), T(g.scale[1]), T(g.scale[2])};
            V3<T> m[3];
            quat_to_mt<T>(g.rot, m);
            const V3<T> gposc = sub(o, ppos);
            const V3<T> gposcr = rowmul(m, gposc);
            const V3<T> gro = mul(giscl, gposcr);
            const V3<T> rayDirR = rowmul(m, d);
            const V3<T> grdu = mul(giscl, rayDirR);
            const V3<T> grd = normalize(grdu);
            const V3<T> gcrod = cross(grd, gro);
            const T grayDist = dot(gcrod, gcrod);
            // particleResponse<4> (cudaRenderer.cu:681-685)
            const T gres = texp(T(-0.0555555555556f) * grayDist * grayDist);
            const T sample_opacity = tmin(T(0.99f), T(g.opacity) * gres);
            if (gres < T(0.0113f) || sample_opacity < T(1.0f / 255.0f)) continue;
            const V3<T> c = eval_sh<T>(d, g);
            color = add(color, muls(muls(c, sample_opacity), T(1) - A));
            A = A + sample_opacity * (T(1) - A);
            ro.cnt[4]++; // blend_ops
            if (tc && tc->blend) tc->blend[gid] = 1;
            ro.hash = (ro.hash ^ (uint64_t)gid) * kFnvPrime;
            if (seq && nseq < seq_cap) seq[nseq] = gid;
            nseq++;
            if (A > T(0.99f)) break;
        }
        if ((A > T(0.99f)) | (t_far >= t_scene_far)) break;
        if (quant == 0) {
            node = sc.nodes[0]; // restart from the root
            cur = 0;
            t_near = t_far;
            t_far = t_scene_far;
            ro.cnt[7]++;
        } else {
            const uint32_t nid = st_node[top];
            cur = nid;
            const T tm = st_t[top];
            if (top == 0) top = (unsigned)D;
            --top;
            --quant;
            node = sc.nodes[nid];
            t_near = t_far;
            t_far = tm;
        }
    }
}

// Ray generation (cudaRenderer.cu:2333-2339).
template <class T> inline void make_ray(const CamInfo& c, int x, int y, V3<T>& o, V3<T>& d) {
    const T sx = T((float)x + 0.5f), sy = T((float)y + 0.5f);
    const V3<T> start = {T(c.start[0]), T(c.start[1]), T(c.start[2])};
    const V3<T> u = {T(c.u[0]), T(c.u[1]), T(c.u[2])}, v = {T(c.v[0]), T(c.v[1]), T(c.v[2])};
    const V3<T> eye = {T(c.eye[0]), T(c.eye[1]), T(c.eye[2])};
    const V3<T> dir = sub(add(start, muls(muls(u, sx), T(c.step_x))), muls(muls(v, sy), T(c.step_y)));
    o = eye;
    d = normalize(sub(dir, eye));
}

template <class T>
void render_view(const Scene& sc, const CamInfo& cam, int W, int H, int D, T* out_rgbt, int32_t* counters,
                 uint64_t* hashes, uint32_t* seq, int seq_cap, int threads, const Touch* tc = nullptr) {
    if (threads > 0) omp_set_num_threads(threads);
    const int tiles_x = (W + 15) / 16, tiles_y = (H + 15) / 16;
#pragma omp parallel for schedule(dynamic)
    for (int tile = 0; tile < tiles_x * tiles_y; ++tile) {
        const int tx = tile % tiles_x, ty = tile / tiles_x;
        for (int y = ty * 16; y < std::min(H, ty * 16 + 16); ++y)
            for (int x = tx * 16; x < std::min(W, tx * 16 + 16); ++x) {
                V3<T> o, d, c;
                T A;
                RayOut ro;
                const size_t rid = (size_t)y * W + x;
                make_ray<T>(cam, x, y, o, d);
                trace<T>(sc, o, d, D, c, A, ro, seq ? seq + rid * seq_cap : nullptr, seq_cap, false, T(0), tc);
                const size_t pid = (size_t)(H - y - 1) * W + x; // framebuffer row flip (cudaRenderer.cu:2367)
                if (out_rgbt) {
                    out_rgbt[4 * pid + 0] = c.x;
                    out_rgbt[4 * pid + 1] = c.y;
                    out_rgbt[4 * pid + 2] = c.z;
                    out_rgbt[4 * pid + 3] = T(1) - A;
                }
                if (counters) std::memcpy(counters + 8 * rid, ro.cnt, sizeof(ro.cnt));
                if (hashes) hashes[rid] = ro.hash;
            }
    }
}

template <class T>
void render_rays(const Scene& sc, const float* origins, const float* dirs, uint64_t n, int D, T* out_rgbt,
                 int32_t* counters, uint64_t* hashes, int threads, const Touch* tc = nullptr) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long i = 0; i < (long long)n; ++i) {
        const V3<T> o = {T(origins[3 * i]), T(origins[3 * i + 1]), T(origins[3 * i + 2])};
        const V3<T> d = {T(dirs[3 * i]), T(dirs[3 * i + 1]), T(dirs[3 * i + 2])};
        V3<T> c;
        T A;
        RayOut ro;
        trace<T>(sc, o, d, D, c, A, ro, nullptr, 0, false, T(0), tc);
        if (out_rgbt) {
            out_rgbt[4 * i + 0] = c.x;
            out_rgbt[4 * i + 1] = c.y;
            out_rgbt[4 * i + 2] = c.z;
            out_rgbt[4 * i + 3] = T(1) - A;
        }
        if (counters) std::memcpy(counters + 8 * i, ro.cnt, sizeof(ro.cnt));
        if (hashes) hashes[i] = ro.hash;
    }
}

// ---- hybrid render (renderKernelGaussian_Mesh_Hybrid, cudaRenderer.cu:2217-2311) ----

struct SecSphere { // DemoSphere (cudaRenderer.cu:2153-2160); same layout as kdgrt_secondary_sphere
    float c[3];
    float r;
    int32_t mat;
    float refl, ior;
    float col[3];
};

struct SecScene {
    int prim; // 1 spheres, 0 triangles (SECONDARY_PRIM, Kd-treeConverter.h:28-30)
    int count;
    const SecSphere* sph;
    const float* tri; // 9 floats per triangle
    int32_t mesh_mat;
    float mesh_refl, mesh_ior, mesh_col[3];
    int max_bounces; // MAX_BOUNCES
};

template <class T>
struct MeshHit { // cudaRenderer.cu:2050-2060
    bool hit = false;
    T t = T(1e8f);
    V3<T> normal{}, hitPoint{};
    int mat = 0;
    T refl = T(0), ior = T(0);
    V3<T> color{};
    bool front = true;
};

const float kEps = 0.00001f; // EPSILON (DefineConst.h:24)

// rayTriIntersect (cudaRenderer.cu:1017-1065).
template <class T>
bool ray_tri(V3<T> o, V3<T> d, V3<T> v0, V3<T> v1, V3<T> v2, T& t_out) {
    const V3<T> e1 = sub(v1, v0), e2 = sub(v2, v0);
    const V3<T> pvec = cross(d, e2);
    const T det = dot(e1, pvec);
    if (det > T(-kEps) && det < T(kEps)) return false;
    const T inv_det = T(1) / det;
    const V3<T> tvec = sub(o, v0);
    const T u = dot(tvec, pvec) * inv_det;
    if (u < T(0) || u > T(1)) return false;
    const V3<T> qvec = cross(tvec, e1);
    const T v = dot(d, qvec) * inv_det;
    if (v < T(0) || u + v > T(1)) return false;
    const T t = dot(e2, qvec) * inv_det;
    if (t > T(kEps)) {
        t_out = t;
        return true;
    }
    return false;
}

// intersectSphere (cudaRenderer.cu:2170-2215) / intersectTriangleMesh (:2111-2151).
template <class T>
MeshHit<T> sec_intersect(const SecScene& s, V3<T> o, V3<T> d) {
    MeshHit<T> h;
    if (s.prim == 1) {
        for (int i = 0; i < s.count; ++i) {
            const SecSphere& q = s.sph[i];
            const V3<T> c = {T(q.c[0]), T(q.c[1]), T(q.c[2])};
            const V3<T> oc = sub(o, c);
            const T a = dot(d, d);
            const T b = T(2) * dot(oc, d);
            const T cc = dot(oc, oc) - T(q.r) * T(q.r);
            const T disc = b * b - T(4) * a * cc;
            if (disc >= T(0)) {
                const T sq = tsqrt(disc);
                T root = (-b - sq) / (T(2) * a);
                if (root < T(kEps)) root = (-b + sq) / (T(2) * a);
                if (root > T(kEps) && root < h.t) {
                    h.hit = true;
                    h.t = root;
                    h.hitPoint = add(o, muls(d, h.t));
                    h.normal = normalize(sub(h.hitPoint, c));
                    if (dot(d, h.normal) > T(0)) {
                        h.normal = muls(h.normal, T(-1));
                        h.front = false;
                    } else {
                        h.front = true;
                    }
                    h.mat = q.mat;
                    h.refl = T(q.refl);
                    h.ior = T(q.ior);
                    h.color = {T(q.col[0]), T(q.col[1]), T(q.col[2])};
                }
            }
        }
    } else {
        for (int i = 0; i < s.count; ++i) {
            const float* v = s.tri + 9 * (size_t)i;
            const V3<T> v0 = {T(v[0]), T(v[1]), T(v[2])}, v1 = {T(v[3]), T(v[4]), T(v[5])},
                        v2 = {T(v[6]), T(v[7]), T(v[8])};
            T t;
            if (ray_tri(o, d, v0, v1, v2, t) && t < h.t) {
                h.hit = true;
                h.t = t;
                h.hitPoint = add(o, muls(d, t));
                h.normal = normalize(cross(sub(v1, v0), sub(v2, v0)));
                if (dot(d, h.normal) > T(0)) {
                    h.normal = muls(h.normal, T(-1));
                    h.front = false;
                } else {
                    h.front = true;
                }
                h.mat = s.mesh_mat;
                h.refl = T(s.mesh_refl);
                h.ior = T(s.mesh_ior);
                h.color = {T(s.mesh_col[0]), T(s.mesh_col[1]), T(s.mesh_col[2])};
            }
        }
    }
    return h;
}

// The per-pixel bounce loop (cudaRenderer.cu:2228-2304); returns the
// unclamped colour and the number of segments traced.  path (optional):
// per segment origin, direction and max_t (7 values).
template <class T>
V3<T> hybrid_pixel(const Scene& sc, const SecScene& s, V3<T> o, V3<T> d, int D, int& segments,
                   T* path = nullptr) {
    V3<T> fin = {T(0), T(0), T(0)}, thr = {T(1), T(1), T(1)}, acc = {T(0), T(0), T(0)};
    T A = T(0);
    segments = 0;
    for (int bounce = 0; bounce < s.max_bounces; ++bounce) {
        const MeshHit<T> mh = sec_intersect<T>(s, o, d);
        const T max_t = mh.hit ? mh.t : T(1e8f);
        if (path) {
            const T rec[7] = {o.x, o.y, o.z, d.x, d.y, d.z, max_t};
            std::memcpy(path + 7 * segments, rec, sizeof(rec));
        }
        RayOut ro;
        trace<T>(sc, o, d, D, acc, A, ro, nullptr, 0, true, max_t);
        ++segments;
        fin = add(fin, mul(thr, acc));
        thr = muls(thr, T(1) - A);
        if (thr.x < T(0.01f) && thr.y < T(0.01f) && thr.z < T(0.01f)) break;
        if (!mh.hit) break;
        const V3<T> N = mh.normal, I = d;
        if (mh.mat == 1) { // reflection
            const V3<T> rd = sub(I, muls(muls(N, T(2)), dot(I, N)));
            o = add(mh.hitPoint, muls(N, T(1e-4f)));
            d = normalize(rd);
            V3<T> mc = mh.color;
            if (bounce > 0) mc = {T(0), T(0), T(0)};
            fin = add(fin, mul(muls(thr, T(1) - mh.refl), mc));
            thr = muls(thr, mh.refl);
        } else if (mh.mat == 2) { // refraction
            const T eta = mh.front ? T(1) / mh.ior : mh.ior;
            const T cosI = -dot(N, I);
            const T sinT2 = eta * eta * (T(1) - cosI * cosI);
            if (sinT2 > T(1)) { // total internal reflection
                const V3<T> rd = sub(I, muls(muls(N, T(2)), dot(I, N)));
                o = add(mh.hitPoint, muls(N, T(1e-4f)));
                d = normalize(rd);
            } else {
                const T cosT = tsqrt(T(1) - sinT2);
                const V3<T> rd = add(muls(I, eta), muls(N, eta * cosI - cosT));
                o = sub(mh.hitPoint, muls(N, T(1e-4f)));
                d = normalize(rd);
            }
            thr = mul(thr, mh.color);
        } else {
            fin = add(fin, mul(thr, mh.color));
            break;
        }
    }
    return fin;
}

template <class T>
void render_hybrid(const Scene& sc, const SecScene& s, const CamInfo& cam, int W, int H, int D, T* out_rgb,
                   int32_t* segments, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            V3<T> o, d;
            make_ray<T>(cam, x, y, o, d);
            int nseg = 0;
            const V3<T> c = hybrid_pixel<T>(sc, s, o, d, D, nseg);
            const size_t pid = (size_t)(H - y - 1) * W + x;
            out_rgb[3 * pid + 0] = tmin(c.x, T(1));
            out_rgb[3 * pid + 1] = tmin(c.y, T(1));
            out_rgb[3 * pid + 2] = tmin(c.z, T(1));
            if (segments) segments[(size_t)y * W + x] = nseg;
        }
}

// Hybrid paths from explicit primary rays (f32 origins and directions, used
// as given): out_rgb n*3, min(c, 1).
template <class T>
void render_hybrid_rays(const Scene& sc, const SecScene& s, const float* origins, const float* dirs, uint64_t n,
                        int D, T* out_rgb, int32_t* segments, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long i = 0; i < (long long)n; ++i) {
        const V3<T> o = {T(origins[3 * i]), T(origins[3 * i + 1]), T(origins[3 * i + 2])};
        const V3<T> d = {T(dirs[3 * i]), T(dirs[3 * i + 1]), T(dirs[3 * i + 2])};
        int nseg = 0;
        const V3<T> c = hybrid_pixel<T>(sc, s, o, d, D, nseg);
        out_rgb[3 * i + 0] = tmin(c.x, T(1));
        out_rgb[3 * i + 1] = tmin(c.y, T(1));
        out_rgb[3 * i + 2] = tmin(c.z, T(1));
        if (segments) segments[i] = nseg;
    }
}

Scene make_scene(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb) {
    Scene sc;
    sc.nodes = (const Node*)nodes;
    sc.offsets = offsets;
    sc.g = (const Gauss256*)g;
    for (int k = 0; k < 3; ++k) {
        sc.bmin[k] = aabb[k];
        sc.bmax[k] = aabb[3 + k];
    }
    return sc;
}

} // namespace

extern "C" {

// g: 256-B Gaussian records with inverted scales; aabb in .kdt order.
// out_rgbt: W*H*4 (row-flipped), counters: W*H*8 int32 by (y*W+x),
// hashes: W*H uint64, seq: W*H*seq_cap uint32 (caller fills with 0xFFFFFFFF).
void oracle_render_f32(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                       const void* cam, int W, int H, int D, float* out_rgbt, int32_t* counters, uint64_t* hashes,
                       uint32_t* seq, int seq_cap, int threads) {
    render_view<float>(make_scene(nodes, offsets, g, aabb), *(const CamInfo*)cam, W, H, D, out_rgbt, counters,
                       hashes, seq, seq_cap, threads);
}
void oracle_render_f64(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                       const void* cam, int W, int H, int D, double* out_rgbt, int32_t* counters, uint64_t* hashes,
                       int threads) {
    render_view<double>(make_scene(nodes, offsets, g, aabb), *(const CamInfo*)cam, W, H, D, out_rgbt, counters,
                        hashes, nullptr, 0, threads);
}
void oracle_render_rays_f32(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                            const float* origins, const float* dirs, uint64_t n, int D, float* out_rgbt,
                            int32_t* counters, uint64_t* hashes, int threads) {
    render_rays<float>(make_scene(nodes, offsets, g, aabb), origins, dirs, n, D, out_rgbt, counters, hashes,
                       threads);
}
void oracle_render_rays_f64(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                            const float* origins, const float* dirs, uint64_t n, int D, double* out_rgbt,
                            int32_t* counters, uint64_t* hashes, int threads) {
    render_rays<double>(make_scene(nodes, offsets, g, aabb), origins, dirs, n, D, out_rgbt, counters, hashes,
                        threads);
}
float oracle_expf(float x) { return kd_expf(x); }

// Hybrid render.  prim 1: `count` spheres (40-B records: centre, radius,
// material, reflectivity, ior, colour); prim 0: `count` triangles (9 floats)
// with the mesh material `mesh` = {material, reflectivity, ior, r, g, b}.
// out_rgb: W*H*3 (row-flipped, min(c, 1)); segments (optional): traced
// segments per ray by y*W+x.
static SecScene make_sec(int prim, int count, const void* spheres, const float* tri, const float* mesh,
                         int max_bounces) {
    SecScene s;
    s.prim = prim;
    s.count = count;
    s.sph = (const SecSphere*)spheres;
    s.tri = tri;
    s.mesh_mat = mesh ? (int32_t)mesh[0] : 0;
    s.mesh_refl = mesh ? mesh[1] : 0.f;
    s.mesh_ior = mesh ? mesh[2] : 0.f;
    for (int k = 0; k < 3; ++k) s.mesh_col[k] = mesh ? mesh[3 + k] : 0.f;
    s.max_bounces = max_bounces;
    return s;
}
void oracle_render_hybrid_f32(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                              const void* cam, int W, int H, int D, int prim, int count, const void* spheres,
                              const float* tri, const float* mesh, int max_bounces, float* out_rgb,
                              int32_t* segments, int threads) {
    render_hybrid<float>(make_scene(nodes, offsets, g, aabb), make_sec(prim, count, spheres, tri, mesh, max_bounces),
                         *(const CamInfo*)cam, W, H, D, out_rgb, segments, threads);
}
void oracle_render_hybrid_f64(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                              const void* cam, int W, int H, int D, int prim, int count, const void* spheres,
                              const float* tri, const float* mesh, int max_bounces, double* out_rgb,
                              int32_t* segments, int threads) {
    render_hybrid<double>(make_scene(nodes, offsets, g, aabb),
                          make_sec(prim, count, spheres, tri, mesh, max_bounces), *(const CamInfo*)cam, W, H, D,
                          out_rgb, segments, threads);
}
void oracle_render_hybrid_rays_f32(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                                   const float* origins, const float* dirs, uint64_t n, int D, int prim, int count,
                                   const void* spheres, const float* tri, const float* mesh, int max_bounces,
                                   float* out_rgb, int32_t* segments, int threads) {
    render_hybrid_rays<float>(make_scene(nodes, offsets, g, aabb),
                              make_sec(prim, count, spheres, tri, mesh, max_bounces), origins, dirs, n, D, out_rgb,
                              segments, threads);
}
void oracle_render_hybrid_rays_f64(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                                   const float* origins, const float* dirs, uint64_t n, int D, int prim, int count,
                                   const void* spheres, const float* tri, const float* mesh, int max_bounces,
                                   double* out_rgb, int32_t* segments, int threads) {
    render_hybrid_rays<double>(make_scene(nodes, offsets, g, aabb),
                               make_sec(prim, count, spheres, tri, mesh, max_bounces), origins, dirs, n, D, out_rgb,
                               segments, threads);
}
// Segments of one hybrid path through the secondary primitives alone (no
// Gaussians): up to max_bounces records of origin, direction and max_t
// (7 floats each); returns the number of segments and writes the colour.
int oracle_hybrid_path_f32(int prim, int count, const void* spheres, const float* tri, const float* mesh,
                           int max_bounces, const float* o, const float* d, float* path, float* rgb) {
    const SecScene s = make_sec(prim, count, spheres, tri, mesh, max_bounces);
    // a scene box the rays never enter: the traces add nothing
    static const Node empty_leaf = {3u, 0u};
    static const uint32_t no_offsets[1] = {0};
    Scene sc;
    sc.nodes = &empty_leaf;
    sc.offsets = no_offsets;
    sc.g = nullptr;
    for (int k = 0; k < 3; ++k) {
        sc.bmin[k] = 1e30f;
        sc.bmax[k] = 1e30f;
    }
    int nseg = 0;
    const V3<float> c = hybrid_pixel<float>(sc, s, {o[0], o[1], o[2]}, {d[0], d[1], d[2]}, 12, nseg, path);
    rgb[0] = c.x;
    rgb[1] = c.y;
    rgb[2] = c.z;
    return nseg;
}
// One secondary intersection (tests of the closed forms): writes
// {hit, t, normal xyz, hit point xyz, front} as floats.
void oracle_sec_intersect_f32(int prim, int count, const void* spheres, const float* tri, const float* o,
                              const float* d, float* out9) {
    const SecScene s = make_sec(prim, count, spheres, tri, nullptr, 1);
    const MeshHit<float> h = sec_intersect<float>(s, {o[0], o[1], o[2]}, {d[0], d[1], d[2]});
    const float r[9] = {h.hit ? 1.f : 0.f, h.t, h.normal.x, h.normal.y, h.normal.z, h.hitPoint.x, h.hitPoint.y,
                        h.hitPoint.z, h.front ? 1.f : 0.f};
    std::memcpy(out9, r, sizeof(r));
}

// Ray directions of a view (for tests of D1/D2).
void oracle_ray_dirs(const void* cam, int W, int H, float* dirs) {
    const CamInfo& c = *(const CamInfo*)cam;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            V3<float> o, d;
            make_ray<float>(c, x, y, o, d);
            float* p = dirs + 3 * ((size_t)y * W + x);
            p[0] = d.x;
            p[1] = d.y;
            p[2] = d.z;
        }
}
// Touch maps of a view or a ray list (FP32): node records visited, leaf
// slots tested, leaf slots accepted, Gaussians evaluated, Gaussians blended;
// byte maps the caller zeroes.  Several rays store the same value 1 into a
// byte, so the concurrent stores agree.
void oracle_touch_view_f32(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                           const void* cam, int W, int H, int D, uint8_t* node, uint8_t* slot, uint8_t* hit_slot,
                           uint8_t* eval, uint8_t* blend, int threads) {
    const Touch tc{node, slot, hit_slot, eval, blend};
    render_view<float>(make_scene(nodes, offsets, g, aabb), *(const CamInfo*)cam, W, H, D, nullptr, nullptr, nullptr,
                       nullptr, 0, threads, &tc);
}
void oracle_touch_rays_f32(const void* nodes, const uint32_t* offsets, const void* g, const float* aabb,
                           const float* origins, const float* dirs, uint64_t n, int D, uint8_t* node, uint8_t* slot,
                           uint8_t* hit_slot, uint8_t* eval, uint8_t* blend, int threads) {
    const Touch tc{node, slot, hit_slot, eval, blend};
    render_rays<float>(make_scene(nodes, offsets, g, aabb), origins, dirs, n, D, nullptr, nullptr, nullptr, threads,
                       &tc);
}

// Unit entry points of the intersection and SH restatements, for the
// closed-form pins.  g: one 256-B record with inverted scales.
int oracle_ray_gaussian_f32(const float* o, const float* d, const void* g, float t_near, float t_far, float* out) {
    float t, gd;
    const bool acc = ray_gaussian<float>({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, *(const Gauss256*)g, t_near,
                                         t_far, t, gd);
    out[0] = t;
    out[1] = gd;
    return acc ? 1 : 0;
}
int oracle_ray_gaussian_f64(const double* o, const double* d, const void* g, double t_near, double t_far,
                            double* out) {
    double t, gd;
    const bool acc = ray_gaussian<double>({o[0], o[1], o[2]}, {d[0], d[1], d[2]}, *(const Gauss256*)g, t_near,
                                          t_far, t, gd);
    out[0] = t;
    out[1] = gd;
    return acc ? 1 : 0;
}
void oracle_sh_basis_f32(const float* d, float* s) { sh_basis<float>({d[0], d[1], d[2]}, s); }
void oracle_eval_sh_f32(const float* d, const void* g, float* rgb) {
    const V3<float> c = eval_sh<float>({d[0], d[1], d[2]}, *(const Gauss256*)g);
    rgb[0] = c.x;
    rgb[1] = c.y;
    rgb[2] = c.z;
}
void oracle_eval_sh_f64(const double* d, const void* g, double* rgb) {
    const V3<double> c = eval_sh<double>({d[0], d[1], d[2]}, *(const Gauss256*)g);
    rgb[0] = c.x;
    rgb[1] = c.y;
    rgb[2] = c.z;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

} // extern "C"
