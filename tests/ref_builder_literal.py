"""Literal, slow restatement of the reference's ELLIPSOID_BY_TRI kd-tree build
(test infrastructure: the fast builder is compared against it on small scenes).

Follows Kd-treeConstructor.cpp statement by statement (paths relative to
Kd-treeConverter_v0.9/Kd-treeConverter/):
  build_kd_tree_recursive         :1399-1635
  try_to_split (COUNT_BY_GID, SAH_MODE 0/1, MINCOST planar side) :579-868
  set_bound_edge + compare_bound_edge (qsort on (t, offset))   :122-151
  push_triangles_to_child_vector  :545-577
  clip_triangle, intersect_edge_plane :364-486, 153-179
  setInnerNode / setLeafNode      :108-120
Float expressions use numpy float32 scalars (IEEE binary32, one rounding per
operation), double expressions Python floats.  The edge sort is Python's
stable sort on (t, offset), one of the orders qsort may produce; the
reference's result does not depend on which.
"""
import numpy as np

f32 = np.float32
EPS = f32(0.00001)
FLT_MAX = float(np.finfo(np.float32).max)
DBL_MAX = 1.7976931348623157e308


class Prim:
    __slots__ = ("offset", "lo", "hi", "pts")

    def __init__(self, offset, lo, hi, pts):
        self.offset, self.lo, self.hi, self.pts = offset, lo, hi, pts


def _split_cost(axis, box_lo, box_hi, prims, init, P):
    axis1, axis2 = (axis + 1) % 3, (axis + 2) % 3
    ext = [f32(box_hi[k] - box_lo[k]) for k in range(3)]
    cell_area = float(f32(f32(f32(ext[0] * ext[1]) + f32(ext[1] * ext[2])) + f32(ext[0] * ext[2])))
    cell_area_rcp = 1 / cell_area
    area_mul = float(ext[axis1]) + float(ext[axis2])
    area_add = float(ext[axis1]) * float(ext[axis2])
    cell_min, cell_max = box_lo[axis], box_hi[axis]
    # set_bound_edge: START then END per primitive, sorted by (t, offset)
    edges = []
    for i, p in enumerate(prims):
        planar = p.lo[axis] == p.hi[axis]
        edges.append((p.lo[axis], p.offset, 0, planar, i))
        edges.append((p.hi[axis], p.offset, 1, planar, i))
    edges.sort(key=lambda e: (float(e[0]), e[1]))
    n = len(edges)
    total_ends = {}
    for e in edges:
        if e[2] == 1:
            total_ends[e[1]] = total_ends.get(e[1], 0) + 1
    seen_starts, seen_ends = set(), {}
    best = dict(init)
    open_, close = 0, 0
    local_open = local_close = num_planars = 0
    i = 0
    while i < n:
        open_ += local_open + num_planars
        close += local_close + num_planars
        local_open = local_close = num_planars = 0
        cur = edges[i][0]
        status = {}
        j = i
        while j < n:
            e = edges[j]
            if e[0] != cur:
                break
            gid = e[1]
            if e[2] == 0:
                if gid not in seen_starts:
                    seen_starts.add(gid)
                    status[gid] = status.get(gid, 0) | 1
            else:
                seen_ends[gid] = seen_ends.get(gid, 0) + 1
                if seen_ends[gid] == total_ends[gid]:
                    status[gid] = status.get(gid, 0) | 2
            i = j
            j += 1
        for st in status.values():
            if st == 3:
                num_planars += 1
            elif st == 1:
                local_open += 1
            elif st == 2:
                local_close += 1
        i += 1
        if cur <= f32(cell_min + EPS):
            continue
        if cur >= f32(cell_max - EPS):
            break
        extent_l = float(cur) - float(cell_min)
        extent_r = float(cell_max) - float(cur)
        prob_l = (extent_l * area_mul + area_add) * cell_area_rcp
        prob_r = (extent_r * area_mul + area_add) * cell_area_rcp
        n_left_only = close + local_close
        n_cross = open_ - n_left_only
        n_right_only = len(total_ends) - (n_left_only + n_cross + num_planars)
        tl = [n_left_only + n_cross + num_planars, n_left_only + n_cross]
        tr = [n_right_only + n_cross, n_right_only + n_cross + num_planars]
        bonus = [float(P["bonus"]) if (tl[s] == 0 or tr[s] == 0) else 1.0 for s in range(2)]
        sah = []
        for s in range(2):
            base = float(tl[s]) * prob_l + float(tr[s]) * prob_r
            if P["sah_mode"] == 1:
                ntot = float(tl[s] + tr[s])
                diff = abs(float(tl[s]) - float(tr[s]))
                ratio = diff / ntot if ntot > 0.0 else 0.0
                pen = 1.0 + ratio * 0.1
                sah.append(float(P["travl"]) + float(P["isect"]) * base * bonus[s] * pen)
            else:
                sah.append(float(P["travl"]) + float(P["isect"]) * base * bonus[s])
        if sah[0] <= sah[1]:
            cost, side = sah[0], 0
        else:
            cost, side = sah[1], 1
        if cost < best["cost"]:
            best = dict(cost=cost, pos=cur, axis=axis, side=side)
    return best


def _edge_plane(p0, p1, pp, pn):
    u = [f32(p1[k] - p0[k]) for k in range(3)]
    w = [f32(p0[k] - pp[k]) for k in range(3)]
    D = f32(f32(f32(pn[0] * u[0]) + f32(pn[1] * u[1])) + f32(pn[2] * u[2]))
    N = -f32(f32(f32(pn[0] * w[0]) + f32(pn[1] * w[1])) + f32(pn[2] * w[2]))
    if abs(D) < EPS:
        return None
    sI = f32(N / D)
    if sI < 0 or sI > 1:
        return None
    return [f32(p0[k] + f32(u[k] * sI)) for k in range(3)]


def _clip(p, axis, split, side):
    pn = [f32(1.0 if k == axis else 0.0) for k in range(3)]
    pp = [f32(pn[k] * split) for k in range(3)]
    L, R = [], []
    for v in p.pts:
        (L if v[axis] < split else R).append(v)
    nl, nr = len(L), len(R)
    for j in range(nl):
        for k in range(nr):
            h = _edge_plane(L[j], R[k], pp, pn)
            if h is not None:
                L.append(h)
                R.append(h)
    S = L if side == 0 else R
    lo, hi = [f32(0)] * 3, [f32(0)] * 3
    if S:
        lo = list(S[0])
        hi = list(S[0])
        for v in S[1:]:
            for k in range(3):
                lo[k] = lo[k] if lo[k] < v[k] else v[k]
                hi[k] = hi[k] if hi[k] > v[k] else v[k]
        for k in range(3):
            lo[k] = lo[k] if lo[k] > p.lo[k] else p.lo[k]
            hi[k] = hi[k] if hi[k] < p.hi[k] else p.hi[k]
    return Prim(p.offset, lo, hi, p.pts)


def build(tri_xyz, tri_gid, aabb, P=None):
    """tri_xyz: (T,3,3) float32, tri_gid: (T,), aabb: .kdt order.  Returns (nodes list of (x,y), offsets)."""
    P = P or dict(travl=f32(1.0), isect=f32(5.0), bonus=f32(0.9), max_level=256, min_prims=16, force=256,
                  sah_mode=1)
    prims = []
    for t in range(len(tri_gid)):
        pts = [[f32(tri_xyz[t, v, k]) for k in range(3)] for v in range(3)]
        lo, hi = [], []
        for k in range(3):
            a, b, c = pts[0][k], pts[1][k], pts[2][k]
            mn = a if a < b else b
            mn = mn if mn < c else c
            mx = a if a > b else b
            mx = mx if mx > c else c
            lo.append(mn)
            hi.append(mx)
        prims.append(Prim(int(tri_gid[t]), lo, hi, pts))
    nodes = [None]
    offsets = []
    counter = [1]

    def rec(prims, blo, bhi, level, slot):
        gids = sorted({p.offset for p in prims})
        best = dict(cost=float(len(gids)) * float(P["isect"]), pos=None, axis=-1, side=-1)
        if P["force"] and len(prims) > P["force"]:
            best["cost"] = DBL_MAX
        if level < P["max_level"] and len(gids) > P["min_prims"]:
            per_axis = [_split_cost(a, blo, bhi, prims, best, P) for a in range(3)]
            for a in range(3):
                if per_axis[a]["cost"] < best["cost"]:
                    best = per_axis[a]
        if best["axis"] < 0:
            nodes[slot] = ((len(gids) << 3) | 3, len(offsets))
            offsets.extend(gids)
            return
        axis, sp = best["axis"], best["pos"]
        c = counter[0]
        counter[0] += 2
        nodes.extend([None, None])
        nodes[slot] = ((c << 3) | axis, int(np.float32(sp).view(np.uint32)))
        edges = []
        for i, p in enumerate(prims):
            planar = p.lo[axis] == p.hi[axis]
            edges.append((p.lo[axis], p.offset, 0, planar, p))
            edges.append((p.hi[axis], p.offset, 1, planar, p))
        edges.sort(key=lambda e: (float(e[0]), e[1]))
        left, right = [], []
        for t, _, typ, planar, p in edges:
            if not planar:
                if t < sp and typ == 0:
                    left.append(p)
                elif t > sp and typ == 1:
                    right.append(p)
            elif typ == 0:
                if t < sp:
                    left.append(p)
                elif t > sp:
                    right.append(p)
                elif best["side"] == 0:
                    left.append(p)
                else:
                    right.append(p)
        left = [(_clip(p, axis, sp, 0) if (p.lo[axis] < sp and p.hi[axis] > sp) else p) for p in left]
        right = [(_clip(p, axis, sp, 1) if (p.lo[axis] < sp and p.hi[axis] > sp) else p) for p in right]
        lhi = list(bhi)
        lhi[axis] = sp
        rlo = list(blo)
        rlo[axis] = sp
        rec(left, blo, lhi, level + 1, c)
        rec(right, rlo, bhi, level + 1, c + 1)

    blo = [f32(aabb[0]), f32(aabb[1]), f32(aabb[2])]
    bhi = [f32(aabb[3]), f32(aabb[4]), f32(aabb[5])]
    rec(prims, blo, bhi, 0, 0)
    return nodes, offsets
