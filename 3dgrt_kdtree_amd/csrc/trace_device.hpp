// Per-ray device path for gfx950: short-stack kd-tree traversal, per-leaf
// ray/Gaussian tests, depth-ordered front-to-back compositing with early
// termination.  One ray per lane of a 64-wide wavefront.
//
// Semantics follow cudaRenderer.cu (paths relative to
// Kd-treeConverter_v0.9/Kd-treeConverter/):
//   ray generation            :2333-2339
//   scene slab test           :379-399, 1536-1537
//   traversal + ring stack    :1542-1573, 1690-1708, 189-210
//   rayPrimIntersect (volume) :1069-1135
//   sortHits (stable by t)    :415-426
//   composite loop            :1608-1688, particleResponse<4> :681-685,
//                              eval_sh_final(3) :487-558
// with the documented substitutions __fdividef(a, b) -> a * (1/b) (IEEE
// reciprocal of the direction component, once per ray), rsqrtf -> 1/sqrtf,
// expf -> kd_expf, no FMA contraction (-ffp-contract=off).
//
// Machine plan (not the reference's):
//  * the short stack is the reference's ring (overwrite the oldest when
//    full, restart from the root when empty) held in LDS, laid out
//    entry-major across the wave's 64 lanes; an entry holds the pushed node
//    itself (both children are fetched at each step, the pair is adjacent)
//    and its t_far, so a pop costs one LDS access and no global load;
//  * the exact test reads its 80-B geometry record by leaf slot (records are
//    stored per slot with the Gaussian ID in the last word), so it does not
//    wait for an ID load first;
//  * the per-leaf hit list is a fixed-capacity register buffer of K records
//    kept sorted by (t, leaf slot), which is the stable insertion-sort
//    order; a leaf with more than K hits is finished by exact re-scans that
//    select the next K records above the last composited key, so any leaf
//    size is handled with no scratch memory;
//  * the composite step reuses the test's grayDist (bit-identical to the
//    reference's recomputation: same operands, normalize == safe_normalize
//    for non-zero length, and zero length is rejected by the t test) and
//    the opacity fetched with the geometry; a buffer entry is (t, grayDist,
//    opacity, ID), a skipped sample reads no SH data and a blended one
//    reads 192 B;
//  * the 15 SH basis scalars depend only on the ray direction; they are
//    formed (float products, the reference's operator order) when a sample
//    is blended (about five times per ray) instead of occupying registers.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kd_expf.h"


namespace kdgrt_dev {

struct DevScene {
    const uint2* __restrict__ nodes;
    const uint4* __restrict__ quads;  // per child pair c: nodes c, c+1, then the child pairs of both (48 B)
    const uint32_t* __restrict__ offs; // device number of the Gaussian in each leaf slot
    const uint32_t* __restrict__ ext;  // the caller's ID of each device number (debug outputs only)
    const float4* __restrict__ shade; // 12 x float4 per Gaussian, channel-major (f_dc[c], f_rest[15c..15c+14])
    const float4* __restrict__ lsph;  // bound spheres in leaf-slot order: centre, packed (radius, kappa) as 2 x f16
    const float4* __restrict__ vcone; // per-view cone records in leaf-slot order (views only, else null)
    const float4* __restrict__ sgeom; // geometry records in leaf-slot order, Gaussian ID in the last word
    float bmin[3];
    float bmax[3];
};

struct CamInfo {
    float eye[3], u[3], v[3], start[3];
    float step_x, step_y;
};

struct RayDbg {
    int32_t cnt[8];
    uint64_t hash;
    uint32_t* seq;
    int seq_cap;
    int nseq;
    // phase profile (PROF variant): shader-clock cycles between marks, by phase
    // 0 descend, 1 pre-test windows, 2 exact tests + hit buffer, 3 composite,
    // 4 stack pop / restart / loop control
    uint64_t cyc[5];
    uint64_t tlast;
    uint32_t pretests, survivors; // PROF variant: candidates pre-tested / passed on to the exact test
    // Touch maps (DBG variant, each may be null): a byte set to 1 for every
    // record the ray reads -- node records visited (by node index), child-pair
    // quad records loaded (by first-child index), leaf slots pre-tested and
    // exact-tested (by slot), Gaussians entering the composite loop and
    // Gaussians blended (by ID).  Lanes store the same value, so the races
    // are benign.
    uint8_t* t_node;
    uint8_t* t_quad;
    uint8_t* t_pre;
    uint8_t* t_exact;
    uint8_t* t_eval;
    uint8_t* t_blend;
};

#define KD_PROF_MARK(ph)                                                                                        \
    do {                                                                                                        \
        if (PROF) {                                                                                             \
            const uint64_t now_ = clock64();                                                                    \
            dbg->cyc[ph] += now_ - dbg->tlast;                                                                  \
            dbg->tlast = now_;                                                                                  \
        }                                                                                                       \
    } while (0)

// A validated tree visits each leaf at most once per ray; this cap is a
// guard against malformed input only and is never reached otherwise.
constexpr int kLeafGuard = 1 << 22;

// Zero records after the last leaf-slot sphere: the pre-test reads groups of
// four slots and may run up to three past the end of the last leaf.
constexpr uint64_t kLsphPad = 4;

__device__ __forceinline__ void make_ray(const CamInfo& c, int x, int y, float o[3], float d[3]) {
    const float sx = (float)x + 0.5f, sy = (float)y + 0.5f;
    float w[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = c.start[k] + (c.u[k] * sx) * c.step_x;
        float b = (c.v[k] * sy) * c.step_y;
        w[k] = (a - b) - c.eye[k];
        o[k] = c.eye[k];
    }
    const float inv = 1.0f / sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = w[k] * inv;
}

// SH basis scalars of eval_sh_final (cudaRenderer.cu:517-553).  Every
// SH_C* constant carries an f suffix (cudaRenderer.h:33-48), so each factor
// is a float product in the reference's left-to-right order: SH_C1 * y
// rounds once, (SH_C3_0 * y) * (3xx - yy) twice.
__device__ __forceinline__ void sh_basis(const float d[3], float s[16]) {
    const float C1 = 0.4886025119029199f, C2_0 = 1.0925484305920792f, C2_1 = -1.0925484305920792f,
                C2_2 = 0.31539156525252005f, C2_3 = -1.0925484305920792f, C2_4 = 0.5462742152960396f,
                C3_0 = -0.5900435899266435f, C3_1 = 2.890611442640554f, C3_2 = -0.4570457994644658f,
                C3_3 = 0.3731763325901154f, C3_4 = -0.4570457994644658f, C3_5 = 1.445305721320277f,
                C3_6 = -0.5900435899266435f;
    const float x = d[0], y = d[1], z = d[2];
    s[0] = 0.28209479177387814f;
    s[1] = C1 * y;
    s[2] = C1 * z;
    s[3] = C1 * x;
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    s[4] = C2_0 * xy;
    s[5] = C2_1 * yz;
    s[6] = C2_2 * (2.0f * zz - xx - yy);
    s[7] = C2_3 * xz;
    s[8] = C2_4 * (xx - yy);
    s[9] = (C3_0 * y) * (3.0f * xx - yy);
    s[10] = (C3_1 * xy) * z;
    s[11] = (C3_2 * y) * (4.0f * zz - xx - yy);
    s[12] = (C3_3 * z) * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    s[13] = (C3_4 * x) * (4.0f * zz - xx - yy);
    s[14] = (C3_5 * z) * (xx - yy);
    s[15] = (C3_6 * x) * (xx - 3.0f * yy);
}

// Radiance of one Gaussian: the 16 coefficient float3s times the basis
// scalars, accumulated in the reference's left-to-right order with the signs
// of degree 1 (minus, plus, minus), +0.5, clamped below at 0.  The device SH
// record is channel-major: channel c occupies float4s 4c..4c+3 as
// (f_dc[c], f_rest[15c+0 .. 15c+14]), so one channel is loaded at a time.
template <class SH>
__device__ __forceinline__ void sh_radiance(const SH& sh, const float s[16], float rad[3]) {
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float c[16];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float4 q = sh[4 * ch + i];
            c[4 * i + 0] = q.x;
            c[4 * i + 1] = q.y;
            c[4 * i + 2] = q.z;
            c[4 * i + 3] = q.w;
        }
        float r = c[0] * s[0];
        r = r - c[1] * s[1];
        r = r + c[2] * s[2];
        r = r - c[3] * s[3];
#pragma unroll
        for (int i = 4; i < 16; ++i) r = r + c[i] * s[i];
        r = r + 0.5f;
        rad[ch] = fmaxf(r, 0.0f);
    }
}

// Conservative pre-test.  Any Gaussian that the exact test below can accept
// lies within radius r of the ray in world space (r bounds sqrt(8) normalised
// units mapped back through diag(1/scale)*M, with a relative margin), and its
// t lies within r of the projection of the centre; kappa*|a|_1 covers the
// float rounding of both the exact test and this one.  r = inf disables the
// pre-test for a Gaussian (singular or non-orthogonal rotation).  The exact
// test alone decides acceptance, so the pre-test changes no result.  The
// pre-tests use fused multiply-adds: their margins bound the rounding of
// separate products and sums, and a fused operation rounds once where those
// round twice, so its error stays within the same bounds.
__device__ __forceinline__ bool sphere_pass(const float o[3], const float d[3], const float4 sp, float t_near,
                                            float t_far) {
    const float ax = sp.x - o[0], ay = sp.y - o[1], az = sp.z - o[2];
    const uint32_t w = __float_as_uint(sp.w);
    const float r = __half2float(__ushort_as_half((unsigned short)(w & 0xFFFFu)));
    const float kap = __half2float(__ushort_as_half((unsigned short)(w >> 16)));
    const float R = __builtin_fmaf(kap, fabsf(ax) + fabsf(ay) + fabsf(az), r);
    const float tp = __builtin_fmaf(az, d[2], __builtin_fmaf(ay, d[1], ax * d[0]));
    const float cx = __builtin_fmaf(ay, d[2], -(az * d[1])), cy = __builtin_fmaf(az, d[0], -(ax * d[2])),
                cz = __builtin_fmaf(ax, d[1], -(ay * d[0]));
    const float dist2 = __builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx));
    return (dist2 <= R * R) & (tp + R >= t_near) & (tp - R <= t_far);
}

// Pre-test for rays that share the origin o of a view.  The record holds
// a = centre - o (the f32 difference the sphere test forms) and a threshold
// q: every direction the exact test can accept has a.d >= q, so testing
// a.d >= q keeps all of them.  q (k_gaussian_cone, formed in double and
// rounded down) is the tighter of the bound sphere's cone and the cone
// around the Gaussian's ellipsoid projected across the view direction; its
// margins cover the rounding of this dot product.
__device__ __forceinline__ bool cone_pass(const float d[3], const float4 c) {
    const float tp = __builtin_fmaf(c.z, d[2], __builtin_fmaf(c.y, d[1], c.x * d[0]));
    return tp >= c.w;
}

struct TestOut {
    float t;
    float gd;
    float op;
    bool hit;
};

// rayPrimIntersect (cudaRenderer.cu:1069-1135) on the 80-byte geometry
// record: g0 = pos|opacity, g1 = inverse scale|pscale.x, g2 = pscale.y,
// pscale.z, M00, M01, g3 = M02, M10, M11, M12, g4 = M20, M21, M22, 0, where
// M = quaternionWXYZToMatrixTranspose(rot) (cudaRenderer.cu:808-828) and
// pscale = 1 / (inverse scale), both formed on the host with the same f32
// operations.  grayDist < 8 is tested before t: a hit needs both, so the
// order changes no outcome and most non-hits skip the t square root.
__device__ __forceinline__ TestOut ray_gaussian(const float o[3], const float d[3], const float4 g0, const float4 g1,
                                                const float4 g2, const float4 g3, const float4 g4, float t_near,
                                                float t_far) {
    TestOut r;
    r.hit = false;
    r.t = 0.0f;
    r.gd = 0.0f;
    r.op = g0.w;
    const float giscl[3] = {g1.x, g1.y, g1.z};
    const float pscale[3] = {g1.w, g2.x, g2.y};
    const float m[3][3] = {{g2.z, g2.w, g3.x}, {g3.y, g3.z, g3.w}, {g4.x, g4.y, g4.z}};
    const float gposc[3] = {o[0] - g0.x, o[1] - g0.y, o[2] - g0.z};
    float gro[3], grdu[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float pr = m[k][0] * gposc[0] + m[k][1] * gposc[1] + m[k][2] * gposc[2];
        const float dr = m[k][0] * d[0] + m[k][1] * d[1] + m[k][2] * d[2];
        gro[k] = giscl[k] * pr;
        grdu[k] = dr * giscl[k];
    }
    const float l = grdu[0] * grdu[0] + grdu[1] * grdu[1] + grdu[2] * grdu[2];
    float grd[3];
    if (l > 0.0f) {
        const float inv = 1.0f / sqrtf(l);
#pragma unroll
        for (int k = 0; k < 3; ++k) grd[k] = grdu[k] * inv;
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) grd[k] = grdu[k];
    }
    const float c0 = grd[1] * gro[2] - grd[2] * gro[1];
    const float c1 = grd[2] * gro[0] - grd[0] * gro[2];
    const float c2 = grd[0] * gro[1] - grd[1] * gro[0];
    const float gd = c0 * c0 + c1 * c1 + c2 * c2;
    if (!(gd < 8.f)) return r;
    const float grp = -(grd[0] * gro[0] + grd[1] * gro[1] + grd[2] * gro[2]);
    float grds[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) grds[k] = (pscale[k] * grd[k]) * grp;
    const float t = (grp < 0.f ? -1.f : 1.f) * sqrtf(grds[0] * grds[0] + grds[1] * grds[1] + grds[2] * grds[2]);
    if (t < t_near || t > t_far) return r;
    r.t = t;
    r.gd = gd;
    r.hit = true;
    return r;
}

// Inserts a hit into the K-entry register buffer kept sorted by (t, slot):
// slots arrive in increasing order, so "after every entry with t <= r.t" is
// the stable insertion-sort position.  An entry holds t, grayDist, opacity
// and the Gaussian ID.  A hit that does not fit (or evicts the last entry)
// marks the buffer as overflowed.  Entries no lane of the wave holds are
// skipped with wave-uniform branches: most leaf visits hold zero or one hit,
// so most insertions touch only the first entries.
template <int K>
__device__ __forceinline__ void insert_hit(float (&kt)[K], float (&kg)[K], float (&ko)[K], uint32_t (&kid)[K],
                                           int& n, bool& overflow, const TestOut& r, uint32_t gid) {
    int pos = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        if (!__any(i < n)) break; // no lane of the wave holds entry i
        pos += (i < n && kt[i] <= r.t) ? 1 : 0;
    }
    if (pos >= K) {
        overflow = true;
        return;
    }
    if (n == K) overflow = true;
#pragma unroll
    for (int i = K - 1; i > 0; --i) {
        if (!__any(i <= n)) continue; // entry i stays unused in every lane
        if (i > pos) {
            kt[i] = kt[i - 1];
            kg[i] = kg[i - 1];
            ko[i] = ko[i - 1];
            kid[i] = kid[i - 1];
        } else if (i == pos) {
            kt[i] = r.t;
            kg[i] = r.gd;
            ko[i] = r.op;
            kid[i] = gid;
        }
    }
    if (pos == 0) {
        kt[0] = r.t;
        kg[0] = r.gd;
        ko[0] = r.op;
        kid[0] = gid;
    }
    n = n + 1 < K ? n + 1 : K;
}

// Value select by split axis.  The operands are values, not lvalues: a
// conditional over adjacent struct fields becomes an indexed load and moves
// the whole ray state to scratch memory.
__device__ __forceinline__ float sel_axis(uint32_t axis, float x, float y, float z) {
    return axis == 0 ? x : (axis == 1 ? y : z);
}

// State of one ray between leaf steps (registers of its lane).
struct RayState {
    float ox, oy, oz, dx, dy, dz;
    float ix, iy, iz; // 1 / direction
    uint32_t signs;   // bit k = sign bit of direction component k
    float t_near, t_far, t_scene_far;
    uint2 node;
    uint32_t top, quant; // ring stack index and fill (cudaRenderer.cu:189-210)
    uint32_t nidx;       // index of st.node (DBG variant only: touch maps)
    float rgb[3];
    float A;
    int guard;
    float max_t;    // hybrid segments: the secondary hit that ends the segment
    uint32_t cone;  // hybrid segments: pre-test with the view's cone records
};

// Starts a ray: scene-box slab test (cudaRenderer.cu:379-399, 1536-1537),
// root node, empty ring.  Returns false when the ray misses the box; its
// result is then black with T = 1.
__device__ __forceinline__ bool ray_begin(RayState& st, const DevScene& sc, const float o[3], const float d[3],
                                          int depth) {
    st.rgb[0] = st.rgb[1] = st.rgb[2] = 0.0f;
    st.A = 0.0f;
    st.ox = o[0];
    st.oy = o[1];
    st.oz = o[2];
    st.dx = d[0];
    st.dy = d[1];
    st.dz = d[2];
    const float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    st.ix = inv[0];
    st.iy = inv[1];
    st.iz = inv[2];
    st.signs = (__float_as_uint(d[0]) >> 31) | ((__float_as_uint(d[1]) >> 31) << 1) |
               ((__float_as_uint(d[2]) >> 31) << 2);
    float t_near = 1e-2f, t_far = 3.402823466e+38f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float l1 = (sc.bmin[k] - o[k]) * inv[k];
        const float l2 = (sc.bmax[k] - o[k]) * inv[k];
        t_near = fmaxf(fminf(l1, l2), t_near);
        t_far = fminf(fmaxf(l1, l2), t_far);
    }
    st.t_near = t_near;
    st.t_far = t_far;
    st.t_scene_far = t_far;
    st.top = (uint32_t)depth - 1u;
    st.quant = 0;
    st.guard = 0;
    st.nidx = 0;
    if (!((t_far >= t_near) & (t_far >= 0.f))) return false;
    st.node = sc.nodes[0];
    return true;
}

// Descent from st.node to the next leaf (cudaRenderer.cu:1555-1573) with
// the ring short stack (:189-210).  stk is this lane's ring in LDS, entry e
// at stk[64 * e] (entry-major over the wave's 64 lanes).
// Stack entry e of this lane: node word x at stk[64 e], node word y at
// stk[64 (depth + e)], t_far at stk[64 (2 depth + e)] (32-bit words, entry-
// major over the wave's lanes).  An entry holds the pushed node itself, so a
// pop needs no global load.
// One inner-node step (cudaRenderer.cu:1555-1573) with the node's children
// n0, n1 at hand; returns which child is next (0 near-side n0, 1 n1).
template <bool DBG>
__device__ __forceinline__ uint32_t node_step(RayState& st, const uint2 n0, const uint2 n1, int depth,
                                              uint32_t* __restrict__ stk, RayDbg* dbg) {
    if (DBG && dbg->t_node) dbg->t_node[st.nidx] = 1;
    const uint32_t axis = st.node.x & 3u;
    const float oa = sel_axis(axis, st.ox, st.oy, st.oz);
    const float ia = sel_axis(axis, st.ix, st.iy, st.iz);
    const float t_split = (__uint_as_float(st.node.y) - oa) * ia;
    const uint32_t sign = (st.signs >> (axis > 2u ? 2u : axis)) & 1u;
    const uint32_t pick = sign ^ (t_split <= st.t_near ? 1u : 0u);
    if (st.t_near < t_split && t_split < st.t_far) {
        if (++st.top == (uint32_t)depth) st.top = 0; // push overwrites the oldest when full
        st.quant = st.quant + 1 < (uint32_t)depth ? st.quant + 1 : (uint32_t)depth;
        const uint2 far = sign ? n0 : n1; // child + (sign ^ 1)
        stk[64 * st.top] = far.x;
        stk[64 * (depth + st.top)] = far.y;
        stk[64 * (2 * depth + st.top)] = __float_as_uint(st.t_far);
        // DBG: a fourth plane holds the pushed node's index (child + (sign ^ 1))
        if (DBG) stk[64 * (3 * depth + st.top)] = (st.node.x >> 3) + (sign ^ 1u);
        st.t_far = t_split;
    }
    if (DBG) st.nidx = (st.node.x >> 3) + pick;
    st.node = pick ? n1 : n0;
    return pick;
}

// Descent to the next leaf, two levels per memory round trip: the 48-B quad
// record of a child pair holds the pair and both of its child pairs, so the
// second step needs no load of its own.  The steps, pushes and counters are
// those of the one-level loop.
template <bool DBG>
__device__ __forceinline__ void descend(RayState& st, const DevScene& sc, int depth, uint32_t* __restrict__ stk,
                                        RayDbg* dbg) {
    while ((st.node.x & 7u) != 3u) {
        const uint4* qp = sc.quads + 3 * (size_t)(st.node.x >> 3);
        if (DBG && dbg->t_quad) dbg->t_quad[st.node.x >> 3] = 1;
        const uint4 q0 = qp[0], q1 = qp[1], q2 = qp[2];
        const uint32_t pick = node_step<DBG>(st, make_uint2(q0.x, q0.y), make_uint2(q0.z, q0.w), depth, stk, dbg);
        if (DBG) dbg->cnt[0]++;
        if ((st.node.x & 7u) == 3u) break;
        const uint4 g = pick ? q2 : q1;
        node_step<DBG>(st, make_uint2(g.x, g.y), make_uint2(g.z, g.w), depth, stk, dbg);
        if (DBG) dbg->cnt[0]++;
    }
    if (DBG && dbg->t_node) dbg->t_node[st.nidx] = 1; // the leaf
}

// After a leaf (cudaRenderer.cu:1690-1708): termination test, then pop or
// restart at the root.  Returns true while the ray continues.
template <bool DBG>
__device__ __forceinline__ bool next_leaf(RayState& st, const DevScene& sc, int depth, const uint32_t* __restrict__ stk,
                                          RayDbg* dbg) {
    if ((st.A > 0.99f) | (st.t_far >= st.t_scene_far)) return false;
    if (st.quant == 0) {
        st.node = sc.nodes[0]; // restart from the root
        st.t_near = st.t_far;
        st.t_far = st.t_scene_far;
        if (DBG) {
            dbg->cnt[7]++;
            st.nidx = 0;
        }
    } else {
        st.node = make_uint2(stk[64 * st.top], stk[64 * (depth + st.top)]);
        if (DBG) st.nidx = stk[64 * (3 * depth + st.top)];
        const float tf = __uint_as_float(stk[64 * (2 * depth + st.top)]);
        if (st.top == 0) st.top = (uint32_t)depth;
        --st.top;
        --st.quant;
        st.t_near = st.t_far;
        st.t_far = tf;
    }
    return (++st.guard <= kLeafGuard) && (st.A < 0.99f);
}

// The step after a leaf without the opacity test (cudaRenderer.cu:1690-1708):
// false when the leaf's segment reaches the scene's far end (no leaf follows)
// or the leaf guard runs out; otherwise pop or restart at the root.  The
// opacity test, which needs the leaf composited, is made after it.
__device__ __forceinline__ bool advance(RayState& st, const DevScene& sc, int depth, const uint32_t* __restrict__ stk) {
    if (st.t_far >= st.t_scene_far) return false;
    if (st.quant == 0) {
        st.node = sc.nodes[0];
        st.t_near = st.t_far;
        st.t_far = st.t_scene_far;
    } else {
        st.node = make_uint2(stk[64 * st.top], stk[64 * (depth + st.top)]);
        const float tf = __uint_as_float(stk[64 * (2 * depth + st.top)]);
        if (st.top == 0) st.top = (uint32_t)depth;
        --st.top;
        --st.quant;
        st.t_near = st.t_far;
        st.t_far = tf;
    }
    return ++st.guard <= kLeafGuard;
}

// One iteration of the descent (one quad record, up to two steps).
__device__ __forceinline__ void descend_once(RayState& st, const DevScene& sc, int depth, uint32_t* __restrict__ stk) {
    const uint4* qp = sc.quads + 3 * (size_t)(st.node.x >> 3);
    const uint4 q0 = qp[0], q1 = qp[1], q2 = qp[2];
    const uint32_t pick = node_step<false>(st, make_uint2(q0.x, q0.y), make_uint2(q0.z, q0.w), depth, stk, nullptr);
    if ((st.node.x & 7u) == 3u) return;
    const uint4 g = pick ? q2 : q1;
    node_step<false>(st, make_uint2(g.x, g.y), make_uint2(g.z, g.w), depth, stk, nullptr);
}

// The leaf in st.node (cudaRenderer.cu:1576-1688): the leaf's candidates in
// stored order, exact test, stable order by t, composite with the skip rule
// and early termination.  The K-entry register buffer holds the leaf's hits
// by (t, slot); a leaf with more than K hits is finished by further passes
// that keep only keys above the last composited one.
// EARLY_SH: request the SH record before the response is formed (below).
template <int K, bool DBG, bool PROF, bool CONE, bool HYB, bool EARLY_SH = CONE && !HYB>
__device__ __forceinline__ void leaf(RayState& st, const DevScene& sc, RayDbg* dbg) {
    const uint32_t base = st.node.y, count = st.node.x >> 3;
    if (DBG) {
        dbg->cnt[1]++;
        dbg->cnt[2] += (int32_t)count;
        if (dbg->t_pre)
            for (uint32_t j = 0; j < count; ++j) dbg->t_pre[base + j] = 1;
    }
    const float o[3] = {st.ox, st.oy, st.oz}, d[3] = {st.dx, st.dy, st.dz};
    bool have_lb = false, done = false;
    float lb_t = 0.f;
    uint32_t lb_s = 0;
    int first_hits = 0;
    while (true) {
        float kt[K], kg[K], ko[K];
        uint32_t kid[K];
#pragma unroll
        for (int i = 0; i < K; ++i) {
            kt[i] = 0.f;
            kg[i] = 0.f;
            ko[i] = 0.f;
            kid[i] = 0;
        }
        int n = 0;
        bool overflow = false;
        // Windows of 32 slots: the cheap pre-test runs for every slot and
        // builds a survivor mask; the exact test then runs only for
        // survivors, in slot order, so a wavefront executes the exact test as
        // often as its busiest lane needs it rather than once for every slot
        // some lane still needs.
        for (uint32_t w0 = 0; w0 < count; w0 += 32) {
            const uint32_t wn = (count - w0) < 32u ? (count - w0) : 32u;
            uint32_t mask = 0;
            // four slots per step, one contiguous read; the group may run past
            // the leaf (the record arrays are padded) and bits at or above wn
            // are cleared afterwards
            if (!CONE && !HYB) {
                // sphere pre-test only (ray lists, batches): the next group's
                // records are requested before this group is tested, which
                // hides their latency behind the test's arithmetic
                const float4* rp = sc.lsph + base + w0;
                float4 r0 = rp[0], r1 = rp[1], r2 = rp[2], r3 = rp[3];
                for (uint32_t j0 = 0; j0 < wn; j0 += 4) {
                    float4 n0 = r0, n1 = r1, n2 = r2, n3 = r3;
                    if (j0 + 4 < wn) {
                        const float4* np = rp + j0 + 4;
                        n0 = np[0];
                        n1 = np[1];
                        n2 = np[2];
                        n3 = np[3];
                    }
                    const uint32_t bits = (sphere_pass(o, d, r0, st.t_near, st.t_far) ? 1u : 0u) |
                                          (sphere_pass(o, d, r1, st.t_near, st.t_far) ? 2u : 0u) |
                                          (sphere_pass(o, d, r2, st.t_near, st.t_far) ? 4u : 0u) |
                                          (sphere_pass(o, d, r3, st.t_near, st.t_far) ? 8u : 0u);
                    mask |= bits << j0;
                    r0 = n0;
                    r1 = n1;
                    r2 = n2;
                    r3 = n3;
                }
            } else {
                for (uint32_t j0 = 0; j0 < wn; j0 += 4) {
                    uint32_t bits;
                    if (CONE || (HYB && st.cone)) {
                        const float4* cp = sc.vcone + base + w0 + j0;
                        const float4 c0 = cp[0], c1 = cp[1], c2 = cp[2], c3 = cp[3];
                        bits = (cone_pass(d, c0) ? 1u : 0u) | (cone_pass(d, c1) ? 2u : 0u) |
                               (cone_pass(d, c2) ? 4u : 0u) | (cone_pass(d, c3) ? 8u : 0u);
                    } else {
                        const float4* sp = sc.lsph + base + w0 + j0;
                        const float4 s0 = sp[0], s1 = sp[1], s2 = sp[2], s3 = sp[3];
                        bits = (sphere_pass(o, d, s0, st.t_near, st.t_far) ? 1u : 0u) |
                               (sphere_pass(o, d, s1, st.t_near, st.t_far) ? 2u : 0u) |
                               (sphere_pass(o, d, s2, st.t_near, st.t_far) ? 4u : 0u) |
                               (sphere_pass(o, d, s3, st.t_near, st.t_far) ? 8u : 0u);
                    }
                    mask |= bits << j0;
                }
            }
            if (wn < 32) mask &= (1u << wn) - 1u;
            if (PROF) dbg->pretests += wn;
            KD_PROF_MARK(1);
            if (PROF) dbg->survivors += (uint32_t)__popc(mask);
            // exact tests, two survivors per step so that their loads overlap
            while (mask) {
                const uint32_t j1 = (uint32_t)__builtin_ctz(mask);
                mask &= mask - 1;
                const bool two = mask != 0;
                const uint32_t j2 = two ? (uint32_t)__builtin_ctz(mask) : j1;
                if (two) mask &= mask - 1;
                const uint32_t s1 = w0 + j1, s2 = w0 + j2;
                if (DBG && dbg->t_exact) {
                    dbg->t_exact[base + s1] = 1;
                    dbg->t_exact[base + s2] = 1;
                }
                const float4* r1p = sc.sgeom + 5 * (size_t)(base + s1);
                const float4* r2p = sc.sgeom + 5 * (size_t)(base + s2);
                const float4 a1 = r1p[0], b1 = r1p[1], c1 = r1p[2], e1 = r1p[3], f1 = r1p[4];
                // a lane with one survivor left reads no second record: the
                // texture path, not the ALU, is the busier unit here
                float4 a2 = a1, b2 = b1, c2 = c1, e2 = e1, f2 = f1;
                if (two) {
                    a2 = r2p[0];
                    b2 = r2p[1];
                    c2 = r2p[2];
                    e2 = r2p[3];
                    f2 = r2p[4];
                }
                const uint32_t g1 = __float_as_uint(f1.w), g2 = __float_as_uint(f2.w);
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (u == 1 && !two) break;
                    const uint32_t sl = u ? s2 : s1, gid = u ? g2 : g1;
                    const TestOut r = u ? ray_gaussian(o, d, a2, b2, c2, e2, f2, st.t_near, st.t_far)
                                        : ray_gaussian(o, d, a1, b1, c1, e1, f1, st.t_near, st.t_far);
                    if (!r.hit) continue;
                    if (have_lb && !(r.t > lb_t || (r.t == lb_t && sl > lb_s))) continue;
                    if (DBG && !have_lb) first_hits++;
                    insert_hit<K>(kt, kg, ko, kid, n, overflow, r, gid);
                }
            }
            KD_PROF_MARK(2);
        }
        if (DBG && !have_lb) {
            dbg->cnt[3] += first_hits;
            if (first_hits > dbg->cnt[5]) dbg->cnt[5] = first_hits;
        }
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if (i >= n || done) break;
            // hybrid segments stop at the secondary hit (cudaRenderer.cu:1612-1616)
            if (HYB && kt[i] > st.max_t) {
                done = true;
                break;
            }
            if (DBG) {
                dbg->cnt[6]++;
                if (dbg->t_eval) dbg->t_eval[sc.ext[kid[i]]] = 1;
            }
            // With EARLY_SH (every kernel but the ray-list debug kernel) the
            // 192-B SH record is requested before the response is formed, so
            // its latency hides behind kd_expf and the basis; only samples
            // that the skip rule drops (alpha < 1/255; grayDist < 8 keeps the
            // response above 0.0113) waste the load.  Without it the record
            // is read after the skip rule.
            constexpr bool kEarlySH = EARLY_SH;
            const uint32_t gid = kid[i];
            const float4* shp = sc.shade + 12 * (size_t)gid;
            float4 shq[kEarlySH ? 12 : 1];
            if (kEarlySH) {
#pragma unroll
                for (int j = 0; j < (kEarlySH ? 12 : 1); ++j) shq[j] = shp[j];
            }
            // particleResponse<4> and the skip rule (cudaRenderer.cu:1655-1671)
            const float gd = kg[i];
            const float gres = kd_expf(-0.0555555555556f * gd * gd);
            const float alpha = fminf(0.99f, ko[i] * gres);
            if (gres < 0.0113f || alpha < 1.0f / 255.0f) continue;
            float c[3];
            // The basis depends only on the ray direction; it is formed here,
            // per blended sample, rather than held in 16 registers or 4 KB of
            // LDS per wave (either costs occupancy, which this path needs).
            float bx = st.dx, by = st.dy, bz = st.dz;
            asm volatile("" : "+v"(bx), "+v"(by), "+v"(bz));
            const float dv[3] = {bx, by, bz};
            float shs[16];
            sh_basis(dv, shs);
            if (kEarlySH) sh_radiance(shq, shs, c);
            else sh_radiance(shp, shs, c); // channel by channel from memory
            const float w = 1.0f - st.A;
            st.rgb[0] = st.rgb[0] + (c[0] * alpha) * w;
            st.rgb[1] = st.rgb[1] + (c[1] * alpha) * w;
            st.rgb[2] = st.rgb[2] + (c[2] * alpha) * w;
            st.A = st.A + alpha * (1.0f - st.A);
            if (DBG) {
                dbg->cnt[4]++;
                const uint32_t xid = sc.ext[gid]; // the caller's ID
                if (dbg->t_blend) dbg->t_blend[xid] = 1;
                dbg->hash = (dbg->hash ^ (uint64_t)xid) * 0x100000001b3ull;
                if (dbg->seq && dbg->nseq < dbg->seq_cap) dbg->seq[dbg->nseq] = xid;
                dbg->nseq++;
            }
            if (st.A > 0.99f) done = true;
        }
        KD_PROF_MARK(3);
        if (done || !overflow) break;
        // Next pass: keys above the last composited (t, slot).  Leaf ID lists
        // hold each ID once (validated), so the slot is the ID's position.
        lb_t = kt[K - 1];
        lb_s = 0;
        while (lb_s < count && sc.offs[base + lb_s] != kid[K - 1]) ++lb_s;
        have_lb = true;
    }
}

// Per-wavefront LDS of the trace: the lanes' short-stack rings, depth
// entries of 12 B (node words x, y and t_far, each entry-major over the 64
// lanes): 768 * depth bytes per wavefront, plus a 16-B stash per lane
// (1024 B) where a lane parks the traversal state of its next leaf while it
// composites the current one; the DBG variant, which does not look ahead,
// has a fourth plane instead (the pushed nodes' indices), 1024 * depth bytes.
__host__ __device__ constexpr size_t wave_lds_bytes(int depth, bool dbg = false) {
    return dbg ? (size_t)depth * 1024u : (size_t)depth * 768u + 1024u;
}

// Whole ray: outer loop of cudaRenderer.cu:1552-1709.  CONE: the ray starts
// at the view origin that sc.vcone was prepared for.  wave_lds: this wave's
// LDS (wave_lds_bytes(depth)); lane: the lane index within the wave.
// HYB: one segment of a hybrid path (SECONDARY_RAY, cudaRenderer.cu:1538-1541,
// 1612-1616): rgb and A come in as the path's running colour and opacity and
// are accumulated further (the reference passes them by reference and does
// not reset them between segments), t_scene_far is clamped to max_t, and
// cone_rt selects the cone pre-test at run time (the first segment of a view
// ray) while later segments use the sphere pre-test.
template <int K, bool DBG, bool PROF = false, bool CONE = false, bool HYB = false, bool AHEAD = true,
          bool EARLY_SH = CONE && !HYB>
__device__ __forceinline__ void trace_ray(const DevScene& sc, const float o[3], const float d[3], int depth,
                                          uint2* __restrict__ wave_lds, int lane, float rgb[3], float& A,
                                          RayDbg* dbg, float max_t = 0.0f, bool cone_rt = false) {
    uint32_t* const stk = reinterpret_cast<uint32_t*>(wave_lds) + lane;
    RayState st;
    bool go = ray_begin(st, sc, o, d, depth);
    if (HYB) {
        st.rgb[0] = rgb[0];
        st.rgb[1] = rgb[1];
        st.rgb[2] = rgb[2];
        st.A = A;
        st.max_t = max_t;
        st.cone = cone_rt ? 1u : 0u;
        if (go) {
            const float tf = fminf(st.t_scene_far, max_t);
            st.t_scene_far = tf;
            st.t_far = tf;
            go = !(st.t_near > tf) && (st.A < 0.99f);
        }
    }
    // Look-ahead (AHEAD; never in the counting and profiling variants): the
    // lanes of a wave descend in lock step and wait for the lane with the
    // longest descent before any of them composites its leaf.  A lane that
    // has found its leaf takes it (its node and t range), makes the step
    // after it at once -- the pop or root restart depends only on the stack
    // and the leaf's t_far, not on the composite -- and goes on descending
    // towards its next leaf while other lanes still descend, so the next
    // descent phase starts part way or not at all.  The opacity test follows
    // the composite as before; a ray that terminates has only spent
    // iterations in which its lane would have idled.  Each lane visits the
    // same leaves in the same order, so results are unchanged.  While it
    // composites, a lane parks the state of its next descent in its LDS stash.
    constexpr bool kAhead = AHEAD && !DBG && !PROF;
    if (go && !kAhead) {
        while (true) {
            descend<DBG>(st, sc, depth, stk, dbg);
            KD_PROF_MARK(0);
            leaf<K, DBG, PROF, CONE, HYB, EARLY_SH>(st, sc, dbg);
            const bool more = next_leaf<DBG>(st, sc, depth, stk, dbg);
            KD_PROF_MARK(4);
            if (!more) break;
        }
    } else if (kAhead) {
        // the stash holds whichever state is not being worked on: the taken
        // leaf while the lane descends, the next descent while it composites
        uint32_t* const stash = stk + 64 * (3 * depth);
        bool alive = go, have = false, last = false;
        while (__any(alive)) {
            while (true) {
                if (alive && !have && (st.node.x & 7u) == 3u) {
                    stash[0] = st.node.x;
                    stash[64] = st.node.y;
                    stash[128] = __float_as_uint(st.t_near);
                    stash[192] = __float_as_uint(st.t_far);
                    have = true;
                    last = !advance(st, sc, depth, stk);
                }
                if (!__any(alive && !have)) break;
                if (alive && !(have && last) && (st.node.x & 7u) != 3u) descend_once(st, sc, depth, stk);
            }
            if (alive) {
                const uint32_t lx = stash[0], ly = stash[64], ltn = stash[128], ltf = stash[192];
                stash[0] = st.node.x;
                stash[64] = st.node.y;
                stash[128] = __float_as_uint(st.t_near);
                stash[192] = __float_as_uint(st.t_far);
                st.node = make_uint2(lx, ly);
                st.t_near = __uint_as_float(ltn);
                st.t_far = __uint_as_float(ltf);
                // the parked values are read back from LDS, not carried in
                // registers through the composite
                asm volatile("" ::: "memory");
                leaf<K, DBG, PROF, CONE, HYB, EARLY_SH>(st, sc, dbg);
                asm volatile("" ::: "memory");
                st.node = make_uint2(stash[0], stash[64]);
                st.t_near = __uint_as_float(stash[128]);
                st.t_far = __uint_as_float(stash[192]);
                have = false;
                if (!(st.A < 0.99f) || last) alive = false;
            }
        }
    }
    rgb[0] = st.rgb[0];
    rgb[1] = st.rgb[1];
    rgb[2] = st.rgb[2];
    A = st.A;
}

} // namespace kdgrt_dev
