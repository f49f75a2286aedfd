// SAH kd-tree builder over icosahedral Gaussian proxies, leaves holding the
// sorted unique Gaussian IDs (the reference's ELLIPSOID_BY_TRI build with
// COUNT_BY_GID=1, SAH_MODE=1, MINCOST planar side).
//
// Reference: Kd-treeConstructor.cpp
//   build_kd_tree_recursive      :1399-1635  (leaf cost, forced split, axis loop, leaf lists)
//   try_to_split                 :579-868    (GID-aware sweep, SAH_MODE 0/1 cost, planar side)
//   set_bound_edge / compare     :122-151
//   push_triangles_to_child_vector :545-577
//   clip_triangle + intersect_edge_plane :364-486, 153-179
//   processInternalNode          :1361-1391
//   setInnerNode / setLeafNode   :108-120
//
// Same tree, different machine plan:
//  * The reference sweeps 2T sorted triangle edges per axis with a hash map
//    per position.  Its counts depend only on, per Gaussian ID g, the first
//    START (a_g = min of triangle AABB minima) and the last END (b_g = max of
//    maxima): at a candidate p,
//        open      = #{a_g <  p}
//        close     = #{b_g <  p}
//        local_cls = #{b_g == p, a_g != p}
//        planars   = #{a_g == b_g == p}
//    so the sweep here runs over sorted per-ID endpoints, and the candidate
//    set is still every distinct triangle-edge value, evaluated in ascending
//    order with the reference's strict-< first-minimum rule.  Edge order
//    within equal values never mattered (the reference groups equal t), and
//    child triangle order never matters (leaf lists are sets), so the result
//    is independent of the reference's qsort tie order.
//  * Subtrees are built as OpenMP tasks, then emitted in the reference's
//    allocation order (children pair taken when the parent is processed,
//    left subtree completely before right), so node and offset indices are
//    those of the sequential recursion.
//
// Compiled with -ffp-contract=off; float and double expressions follow the
// reference's evaluation order and types.

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <omp.h>

#include "internal.hpp"

namespace kdgrt {

void proxy_mesh(const kdgrt_gaussian* g, uint64_t n, std::vector<float>& tri_xyz, std::vector<uint32_t>& tri_gid,
                float aabb[6], uint64_t* n_kept);

namespace {

const float kEps = 0.00001f; // KD_TREE_EPSILON, Kd-treeConstructor.h:10

struct Tri {
    uint32_t gid; // Gaussian ID (PrimList.offset = material_ID)
    uint32_t tri; // index of the original triangle vertices
    float lo[3];  // current (clipped) AABB
    float hi[3];
};

struct BNode {
    int32_t axis = -1;
    float split = 0.0f;
    std::unique_ptr<BNode> kid[2];
    std::vector<uint32_t> gids; // leaf: sorted unique Gaussian IDs
};

struct Split {
    double cost;
    float pos;
    int axis;
    int planar_side; // 0 = planar triangles go left (BoundEdge::START), 1 = right
};

// Monotone key of a float for radix sorting (-0 sorts just before +0).
inline uint32_t fkey(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float kfloat(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Sorts floats ascending (total order by key).  LSD radix with 11-bit
// digits for large arrays.
void sort_floats(std::vector<float>& v) {
    const size_t n = v.size();
    if (n < 2048) {
        std::sort(v.begin(), v.end(), [](float a, float b) { return fkey(a) < fkey(b); });
        return;
    }
    std::vector<uint32_t> a(n), b(n);
    for (size_t i = 0; i < n; ++i) a[i] = fkey(v[i]);
    for (int pass = 0; pass < 3; ++pass) {
        const int shift = pass * 11;
        size_t cnt[2048] = {0};
        for (size_t i = 0; i < n; ++i) cnt[(a[i] >> shift) & 2047]++;
        size_t s = 0;
        for (int d = 0; d < 2048; ++d) {
            size_t c = cnt[d];
            cnt[d] = s;
            s += c;
        }
        for (size_t i = 0; i < n; ++i) b[cnt[(a[i] >> shift) & 2047]++] = a[i];
        a.swap(b);
    }
    for (size_t i = 0; i < n; ++i) v[i] = kfloat(a[i]);
}

struct Params {
    float travl, isect, bonus;
    uint32_t max_level, min_prims, force_split;
    int sah_mode;
};

class Builder {
  public:
    Builder(const float* verts, const Params& p) : V(verts), P(p) {}

    std::unique_ptr<BNode> build(std::vector<Tri>&& tris, const float bmin[3], const float bmax[3], uint32_t level);

    std::atomic<uint32_t> max_level_seen{0};

  private:
    const float* V;
    Params P;

    static uint32_t count_gids(const std::vector<Tri>& t) {
        uint32_t g = 0;
        for (size_t i = 0; i < t.size(); ++i)
            if (i == 0 || t[i].gid != t[i - 1].gid) ++g;
        return g;
    }
    Split try_split(int axis, const std::vector<Tri>& tris, uint32_t G, const float bmin[3], const float bmax[3],
                    const Split& init) const;
    void clip(Tri& t, int axis, float split, int side) const;
};

Split Builder::try_split(int axis, const std::vector<Tri>& tris, uint32_t G, const float bmin[3], const float bmax[3],
                         const Split& init) const {
    static const int modulo[] = {0, 1, 2, 0, 1};
    const int axis1 = modulo[axis + 1], axis2 = modulo[axis + 2];
    float ext[3];
    ext[0] = bmax[0] - bmin[0];
    ext[1] = bmax[1] - bmin[1];
    ext[2] = bmax[2] - bmin[2];
    // Kd-treeConstructor.cpp:584-593; the half area is a float expression.
    const double cell_area = (ext[0] * ext[1]) + (ext[1] * ext[2]) + (ext[0] * ext[2]);
    const double cell_area_rcp = 1 / cell_area;
    const double area_mul = double(ext[axis1]) + double(ext[axis2]);
    const double area_add = double(ext[axis1]) * double(ext[axis2]);
    const float cell_min = bmin[axis];
    const float cell_max = bmax[axis];
    const float lo_lim = cell_min + kEps;
    const float hi_lim = cell_max - kEps;

    const size_t T = tris.size();
    std::vector<float> edges(2 * T);
    for (size_t i = 0; i < T; ++i) {
        edges[2 * i] = tris[i].lo[axis];
        edges[2 * i + 1] = tris[i].hi[axis];
    }
    // Per-ID endpoints: triangles are grouped by ID (the node list is kept
    // sorted by ID).
    std::vector<float> A, B, Bnp, Pl;
    A.reserve(G);
    B.reserve(G);
    for (size_t i = 0; i < T;) {
        size_t j = i;
        float a = tris[i].lo[axis], b = tris[i].hi[axis];
        for (++j; j < T && tris[j].gid == tris[i].gid; ++j) {
            a = std::min(a, tris[j].lo[axis]);
            b = std::max(b, tris[j].hi[axis]);
        }
        A.push_back(a);
        B.push_back(b);
        if (a == b) Pl.push_back(a);
        else Bnp.push_back(b);
        i = j;
    }
    sort_floats(edges);
    sort_floats(A);
    sort_floats(B);
    sort_floats(Bnp);
    sort_floats(Pl);

    Split best = init;
    size_t ia = 0, ib = 0, ibn = 0, ipl = 0;
    const int Gi = (int)G;
    for (size_t e = 0; e < edges.size();) {
        const float p = edges[e];
        size_t e2 = e + 1;
        while (e2 < edges.size() && edges[e2] == p) ++e2; // the reference's group of equal t
        e = e2;
        while (ia < A.size() && A[ia] < p) ++ia;
        while (ib < B.size() && B[ib] < p) ++ib;
        while (ibn < Bnp.size() && Bnp[ibn] < p) ++ibn;
        while (ipl < Pl.size() && Pl[ipl] < p) ++ipl;
        int local_close = 0, num_planars = 0;
        for (size_t k = ibn; k < Bnp.size() && Bnp[k] == p; ++k) ++local_close;
        for (size_t k = ipl; k < Pl.size() && Pl[k] == p; ++k) ++num_planars;

        if (p <= lo_lim) continue;
        if (p >= hi_lim) break;

        const int open = (int)ia, close = (int)ib;
        const double extent_l = double(p) - cell_min, extent_r = cell_max - double(p);
        const double prob_l = (extent_l * area_mul + area_add) * cell_area_rcp,
                     prob_r = (extent_r * area_mul + area_add) * cell_area_rcp;
        const int n_leftOnly = close + local_close;
        const int n_cross = open - n_leftOnly;
        const int n_rightOnly = Gi - (n_leftOnly + n_cross + num_planars);

        const int tl[2] = {n_leftOnly + n_cross + num_planars, n_leftOnly + n_cross};
        const int tr[2] = {n_rightOnly + n_cross, n_rightOnly + n_cross + num_planars};
        const float bonus[2] = {(tl[0] == 0 || tr[0] == 0) ? P.bonus : 1.0f,
                                (tl[1] == 0 || tr[1] == 0) ? P.bonus : 1.0f};
        double SAH[2];
        for (int s = 0; s < 2; ++s) {
            if (P.sah_mode == 1) {
                const double N_total = double(tl[s] + tr[s]);
                const double diff = std::abs(double(tl[s]) - double(tr[s]));
                const double balanceRatio = (N_total > 0.0) ? (diff / N_total) : 0.0;
                const double BALANCE_WEIGHT = 0.1;
                const double balancePenalty = 1.0 + (balanceRatio * BALANCE_WEIGHT);
                SAH[s] = P.travl + P.isect * (double(tl[s]) * prob_l + double(tr[s]) * prob_r) * bonus[s] *
                                       balancePenalty;
            } else {
                SAH[s] = P.travl + P.isect * (double(tl[s]) * prob_l + double(tr[s]) * prob_r) * bonus[s];
            }
        }
        double cost;
        int side;
        if (SAH[0] <= SAH[1]) {
            cost = SAH[0];
            side = 0;
        } else {
            cost = SAH[1];
            side = 1;
        }
        if (cost < best.cost) {
            best.cost = cost;
            best.pos = p;
            best.axis = axis;
            best.planar_side = side;
        }
    }
    return best;
}

// intersect_edge_plane (Kd-treeConstructor.cpp:153-179).
static inline bool edge_plane(const float* p0, const float* p1, const float* pp, const float* pn, float* hit) {
    float u[3], w[3];
    u[0] = p1[0] - p0[0];
    u[1] = p1[1] - p0[1];
    u[2] = p1[2] - p0[2];
    w[0] = p0[0] - pp[0];
    w[1] = p0[1] - pp[1];
    w[2] = p0[2] - pp[2];
    float D = pn[0] * u[0] + pn[1] * u[1] + pn[2] * u[2];
    float N = -(pn[0] * w[0] + pn[1] * w[1] + pn[2] * w[2]);
    if (std::fabs(D) < kEps) return false;
    float sI = N / D;
    if (sI < 0.f || sI > 1.f) return false;
    hit[0] = p0[0] + u[0] * sI;
    hit[1] = p0[1] + u[1] * sI;
    hit[2] = p0[2] + u[2] * sI;
    return true;
}

// clip_triangle for one triangle already known to straddle the plane
// (Kd-treeConstructor.cpp:364-486).  side 0 keeps the part below the split.
void Builder::clip(Tri& t, int axis, float split, int side) const {
    const float norm[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const float* pn = norm[axis];
    const float pp[3] = {pn[0] * split, pn[1] * split, pn[2] * split};
    const float* p = V + 9 * (size_t)t.tri;
    float L[7][3], R[7][3];
    int nl = 0, nr = 0;
    for (int v = 0; v < 3; ++v) {
        float* dst = (p[3 * v + axis] < split) ? L[nl++] : R[nr++];
        dst[0] = p[3 * v];
        dst[1] = p[3 * v + 1];
        dst[2] = p[3 * v + 2];
    }
    const int nl0 = nl, nr0 = nr;
    for (int j = 0; j < nl0; ++j)
        for (int k = 0; k < nr0; ++k) {
            float h[3];
            if (edge_plane(L[j], R[k], pp, pn, h)) {
                std::memcpy(L[nl++], h, sizeof(h));
                std::memcpy(R[nr++], h, sizeof(h));
            }
        }
    float(*S)[3] = side == 0 ? L : R;
    const int ns = side == 0 ? nl : nr;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (ns > 0) {
        for (int k = 0; k < 3; ++k) lo[k] = hi[k] = S[0][k];
        for (int j = 1; j < ns; ++j)
            for (int k = 0; k < 3; ++k) {
                lo[k] = (lo[k] < S[j][k]) ? lo[k] : S[j][k]; // MyMIN
                hi[k] = (hi[k] > S[j][k]) ? hi[k] : S[j][k]; // MyMAX
            }
        for (int k = 0; k < 3; ++k) {
            lo[k] = (lo[k] > t.lo[k]) ? lo[k] : t.lo[k];
            hi[k] = (hi[k] < t.hi[k]) ? hi[k] : t.hi[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        t.lo[k] = lo[k];
        t.hi[k] = hi[k];
    }
}

std::unique_ptr<BNode> Builder::build(std::vector<Tri>&& tris_in, const float bmin[3], const float bmax[3],
                                      uint32_t level) {
    std::vector<Tri> tris = std::move(tris_in);
    {
        uint32_t cur = max_level_seen.load(std::memory_order_relaxed);
        while (level > cur && !max_level_seen.compare_exchange_weak(cur, level)) {
        }
    }
    const uint32_t G = count_gids(tris);
    Split best{double(G) * P.isect, FLT_MAX, -1, -1};
    if (P.force_split && tris.size() > P.force_split) best.cost = DBL_MAX;

    if (level < P.max_level && G > P.min_prims) {
        Split ax[3];
        const bool par = tris.size() > 20000;
#pragma omp taskgroup
        {
            for (int a = 0; a < 3; ++a) {
#pragma omp task shared(ax, tris) firstprivate(a) if (par)
                ax[a] = try_split(a, tris, G, bmin, bmax, best);
            }
        }
        for (int a = 0; a < 3; ++a)
            if (ax[a].cost < best.cost) best = ax[a];
    }

    auto node = std::make_unique<BNode>();
    if (best.axis < 0) {
        node->gids.reserve(G);
        for (size_t i = 0; i < tris.size(); ++i)
            if (i == 0 || tris[i].gid != tris[i - 1].gid) node->gids.push_back(tris[i].gid);
        return node;
    }
    node->axis = best.axis;
    node->split = best.pos;
    const int axis = best.axis;
    const float sp = best.pos;

    // push_triangles_to_child_vector + clip_triangle, order-preserving so the
    // children stay grouped by ID.
    std::vector<Tri> left, right;
    {
        size_t nl = 0, nr = 0;
        for (const Tri& t : tris) {
            const float lo = t.lo[axis], hi = t.hi[axis];
            if (lo != hi) {
                nl += lo < sp;
                nr += hi > sp;
            } else if (lo < sp) nl++;
            else if (lo > sp) nr++;
            else (best.planar_side == 0 ? nl : nr)++;
        }
        left.reserve(nl);
        right.reserve(nr);
        for (const Tri& t : tris) {
            const float lo = t.lo[axis], hi = t.hi[axis];
            if (lo != hi) {
                if (lo < sp) left.push_back(t);
                if (hi > sp) right.push_back(t);
            } else if (lo < sp) left.push_back(t);
            else if (lo > sp) right.push_back(t);
            else if (best.planar_side == 0) left.push_back(t);
            else right.push_back(t);
        }
    }
    std::vector<Tri>().swap(tris);
    for (Tri& t : left)
        if (t.lo[axis] < sp && t.hi[axis] > sp) clip(t, axis, sp, 0);
    for (Tri& t : right)
        if (t.lo[axis] < sp && t.hi[axis] > sp) clip(t, axis, sp, 1);

    float lmax[3] = {bmax[0], bmax[1], bmax[2]};
    float rmin[3] = {bmin[0], bmin[1], bmin[2]};
    lmax[axis] = sp;
    rmin[axis] = sp;
    const bool spawn = left.size() + right.size() > 4096;
    BNode* self = node.get();
#pragma omp task shared(self, left, bmin, lmax) if (spawn)
    self->kid[0] = build(std::move(left), bmin, lmax, level + 1);
    self->kid[1] = build(std::move(right), rmin, bmax, level + 1);
#pragma omp taskwait
    return node;
}

// Emits nodes in the reference's allocation order: the root is node 0; an
// inner node takes the next free pair for its children when it is processed,
// then its left subtree is emitted completely before its right subtree
// (Kd-treeConstructor.cpp:1605-1622); leaves append their ID lists in the
// same order (:1560-1586).
void emit(const BNode* root, kdgrt_tree& out) {
    struct Item {
        const BNode* n;
        uint32_t slot;
    };
    out.nodes.assign(1, kdgrt_node{0, 0});
    out.offsets.clear();
    std::vector<Item> stack;
    stack.push_back({root, 0});
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const BNode* n = it.n;
        if (n->axis < 0) {
            const uint32_t cnt = (uint32_t)n->gids.size();
            out.nodes[it.slot].x = (cnt << 3) | 3u;
            out.nodes[it.slot].y = (uint32_t)out.offsets.size();
            out.offsets.insert(out.offsets.end(), n->gids.begin(), n->gids.end());
        } else {
            const uint32_t c = (uint32_t)out.nodes.size();
            out.nodes.push_back({0, 0});
            out.nodes.push_back({0, 0});
            uint32_t bits;
            std::memcpy(&bits, &n->split, 4);
            out.nodes[it.slot].x = (c << 3) | (uint32_t)n->axis;
            out.nodes[it.slot].y = bits;
            stack.push_back({n->kid[1].get(), c + 1});
            stack.push_back({n->kid[0].get(), c});
        }
    }
}

} // namespace
} // namespace kdgrt

using namespace kdgrt;

extern "C" {

kdgrt_status kdgrt_build_params_default(kdgrt_build_params* p) {
    if (!p) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "build_params_default: null");
    // Kd-treeConverter.h:43-44 (TRAVL_COST, EMTPY_BONUS), :64 (MIN_TRI);
    // Kd-treeConverterMain.cpp:46-61 (ISCET_COST 5, MAX_LEVEL 256,
    // FORCE_SPLIT_THRESHOLD 256, SAH_MODE 1).
    p->traversal_cost = 1.0f;
    p->isect_cost = 5.0f;
    p->empty_bonus = 0.9f;
    p->max_level = 256;
    p->min_prims = 16;
    p->force_split = 256;
    p->sah_mode = 1;
    p->threads = 0;
    return KDGRT_OK;
}

kdgrt_status kdgrt_tree_build(const kdgrt_gaussian* g, uint64_t n, const kdgrt_build_params* params,
                              kdgrt_tree** tree) {
    if (!tree) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "tree_build: null output");
    if (!g || n == 0) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "tree_build: no Gaussians");
    if (n >= (1ull << 31)) return set_error(KDGRT_ERR_CAPACITY, "tree_build: too many Gaussians");
    kdgrt_build_params bp;
    if (params) bp = *params;
    else kdgrt_build_params_default(&bp);
    if (bp.sah_mode != 0 && bp.sah_mode != 1) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "tree_build: sah_mode");
    const int prev_threads = omp_get_max_threads();
    if (bp.threads > 0) omp_set_num_threads(bp.threads);

    const double t0 = wall_seconds();
    std::vector<float> verts;
    std::vector<uint32_t> gids;
    float aabb[6];
    uint64_t kept = 0;
    proxy_mesh(g, n, verts, gids, aabb, &kept);
    const uint64_t T = gids.size();
    const double t1 = wall_seconds();
    if (T == 0) {
        if (bp.threads > 0) omp_set_num_threads(prev_threads);
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "tree_build: every Gaussian was culled by opacity");
    }
    if (T >= (1ull << 32)) {
        if (bp.threads > 0) omp_set_num_threads(prev_threads);
        return set_error(KDGRT_ERR_CAPACITY, "tree_build: too many proxy triangles");
    }

    // Initial triangle AABBs (initialize_kd_tree, Kd-treeConstructor.cpp:1040-1056).
    std::vector<Tri> tris(T);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)T; ++i) {
        const float* p = &verts[9 * (size_t)i];
        Tri& t = tris[i];
        t.gid = gids[i];
        t.tri = (uint32_t)i;
        for (int k = 0; k < 3; ++k) {
            float a = p[k], b = p[3 + k], c = p[6 + k];
            float mn = (a < b) ? a : b;
            mn = (mn < c) ? mn : c;
            float mx = (a > b) ? a : b;
            mx = (mx > c) ? mx : c;
            t.lo[k] = mn;
            t.hi[k] = mx;
        }
    }
    std::vector<uint32_t>().swap(gids);

    Params P{bp.traversal_cost, bp.isect_cost, bp.empty_bonus, bp.max_level, bp.min_prims, bp.force_split,
             bp.sah_mode};
    Builder builder(verts.data(), P);
    const float bmin[3] = {aabb[0], aabb[1], aabb[2]};
    const float bmax[3] = {aabb[3], aabb[4], aabb[5]};
    std::unique_ptr<BNode> root;
#pragma omp parallel
#pragma omp single
    root = builder.build(std::move(tris), bmin, bmax, 0);

    auto* t = new kdgrt_tree();
    emit(root.get(), *t);
    std::memcpy(t->aabb, aabb, sizeof(aabb));
    root.reset();
    const double t2 = wall_seconds();
    std::string err;
    compute_tree_stats(*t, t->stats, &err);
    t->stats.n_triangles = T;
    t->stats.n_kept_gaussians = kept;
    t->stats.mesh_seconds = t1 - t0;
    t->stats.build_seconds = t2 - t1;
    if (bp.threads > 0) omp_set_num_threads(prev_threads);
    if (t->nodes.size() >= (1u << 29) || t->offsets.size() >= (1ull << 32)) {
        delete t;
        return set_error(KDGRT_ERR_CAPACITY, "tree_build: tree exceeds 32-bit node/offset encoding");
    }
    *tree = t;
    return KDGRT_OK;
}

} // extern "C"
