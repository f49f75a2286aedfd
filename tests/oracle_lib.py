"""ctypes loader for the CPU oracle (oracle/liboracle.so), which device results are compared against.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_L = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.oracle_expf.restype = C.c_float
        L.oracle_expf.argtypes = [C.c_float]
        L.oracle_max_threads.restype = C.c_int
        _L = L
    return _L


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def render(nodes, offsets, g_inv, aabb, ci, w, h, depth=12, f64=False, seq_cap=0, threads=0):
    """Renders one view; g_inv holds inverted scales.  Returns dict with
    rgbt (h, w, 4, row-flipped like the device output), counters (h, w, 8),
    hash (h, w) and optionally seq (h, w, seq_cap)."""
    nodes = np.ascontiguousarray(nodes)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    if len(offsets) == 0:
        offsets = np.zeros(1, np.uint32)
    aabb = np.ascontiguousarray(aabb, np.float32)
    cam = np.frombuffer(bytes(ci), dtype=np.uint8).copy()
    out = np.zeros((h, w, 4), np.float64 if f64 else np.float32)
    cnt = np.zeros((h, w, 8), np.int32)
    hsh = np.zeros((h, w), np.uint64)
    if f64:
        lib().oracle_render_f64(_p(nodes), _p(offsets), _p(g_inv), _p(aabb), _p(cam), C.c_int(w), C.c_int(h),
                                C.c_int(depth), _p(out), _p(cnt), _p(hsh), C.c_int(threads))
        return {"rgbt": out, "counters": cnt, "hash": hsh}
    seq = None
    if seq_cap:
        seq = np.full((h, w, seq_cap), 0xFFFFFFFF, np.uint32)
    lib().oracle_render_f32(_p(nodes), _p(offsets), _p(g_inv), _p(aabb), _p(cam), C.c_int(w), C.c_int(h),
                            C.c_int(depth), _p(out), _p(cnt), _p(hsh), _p(seq), C.c_int(seq_cap), C.c_int(threads))
    r = {"rgbt": out, "counters": cnt, "hash": hsh}
    if seq is not None:
        r["seq"] = seq
    return r


def render_rays(nodes, offsets, g_inv, aabb, origins, dirs, depth=12, f64=False, threads=0):
    nodes = np.ascontiguousarray(nodes)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    if len(offsets) == 0:
        offsets = np.zeros(1, np.uint32)
    aabb = np.ascontiguousarray(aabb, np.float32)
    origins = np.ascontiguousarray(origins, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    n = len(origins)
    out = np.zeros((n, 4), np.float64 if f64 else np.float32)
    cnt = np.zeros((n, 8), np.int32)
    hsh = np.zeros(n, np.uint64)
    fn = lib().oracle_render_rays_f64 if f64 else lib().oracle_render_rays_f32
    fn(_p(nodes), _p(offsets), _p(g_inv), _p(aabb), _p(origins), _p(dirs), C.c_uint64(n), C.c_int(depth), _p(out),
       _p(cnt), _p(hsh), C.c_int(threads))
    return {"rgbt": out, "counters": cnt, "hash": hsh}


def ray_dirs(ci, w, h):
    cam = np.frombuffer(bytes(ci), dtype=np.uint8).copy()
    d = np.zeros((h, w, 3), np.float32)
    lib().oracle_ray_dirs(_p(cam), C.c_int(w), C.c_int(h), _p(d))
    return d


def expf(x):
    return lib().oracle_expf(float(x))


def render_hybrid(nodes, offsets, g_inv, aabb, ci, w, h, spheres=None, triangles=None, mesh=(1, 0.95, 1.1, 0.95, 0.95,
                  0.95), max_bounces=5, depth=12, f64=False, threads=0):
    """Hybrid render (renderKernelGaussian_Mesh_Hybrid).  spheres: array of
    the 40-B sphere records; triangles: (n, 9) float32 with the mesh
    material (type, reflectivity, ior, r, g, b).  Returns rgb (h, w, 3),
    row-flipped and clamped to 1 like the device output, and the number of
    segments traced per ray (h, w)."""
    nodes = np.ascontiguousarray(nodes)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    if len(offsets) == 0:
        offsets = np.zeros(1, np.uint32)
    aabb = np.ascontiguousarray(aabb, np.float32)
    cam = np.frombuffer(bytes(ci), dtype=np.uint8).copy()
    out = np.zeros((h, w, 3), np.float64 if f64 else np.float32)
    seg = np.zeros((h, w), np.int32)
    if spheres is not None:
        sph = np.ascontiguousarray(spheres)
        assert sph.dtype.itemsize == 40
        prim, count, tri, mesh_arr = 1, len(sph), None, None
    else:
        sph = None
        tri = np.ascontiguousarray(triangles, np.float32).reshape(-1, 9)
        prim, count = 0, len(tri)
        mesh_arr = np.asarray(mesh, np.float32)
    fn = lib().oracle_render_hybrid_f64 if f64 else lib().oracle_render_hybrid_f32
    fn(_p(nodes), _p(offsets), _p(g_inv), _p(aabb), _p(cam), C.c_int(w), C.c_int(h), C.c_int(depth), C.c_int(prim),
       C.c_int(count), _p(sph), _p(tri), _p(mesh_arr), C.c_int(max_bounces), _p(out), _p(seg), C.c_int(threads))
    return {"rgb": out, "segments": seg}


def sec_intersect(origin, direction, spheres=None, triangles=None):
    """One secondary intersection: (hit, t, normal, hit point, front)."""
    o = np.ascontiguousarray(origin, np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    out = np.zeros(9, np.float32)
    if spheres is not None:
        sph = np.ascontiguousarray(spheres)
        lib().oracle_sec_intersect_f32(C.c_int(1), C.c_int(len(sph)), _p(sph), None, _p(o), _p(d), _p(out))
    else:
        tri = np.ascontiguousarray(triangles, np.float32).reshape(-1, 9)
        lib().oracle_sec_intersect_f32(C.c_int(0), C.c_int(len(tri)), None, _p(tri), _p(o), _p(d), _p(out))
    return bool(out[0]), float(out[1]), out[2:5].copy(), out[5:8].copy(), bool(out[8])


def hybrid_path(origin, direction, spheres=None, triangles=None, mesh=(1, 0.95, 1.1, 0.95, 0.95, 0.95),
                max_bounces=5):
    """Segments of one hybrid path through the secondary primitives alone:
    (segments (n, 7) = origin, direction, max_t; colour (3,))."""
    o = np.ascontiguousarray(origin, np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    path = np.zeros((max_bounces, 7), np.float32)
    rgb = np.zeros(3, np.float32)
    if spheres is not None:
        sph = np.ascontiguousarray(spheres)
        n = lib().oracle_hybrid_path_f32(C.c_int(1), C.c_int(len(sph)), _p(sph), None, None, C.c_int(max_bounces),
                                         _p(o), _p(d), _p(path), _p(rgb))
    else:
        tri = np.ascontiguousarray(triangles, np.float32).reshape(-1, 9)
        m = np.asarray(mesh, np.float32)
        n = lib().oracle_hybrid_path_f32(C.c_int(0), C.c_int(len(tri)), None, _p(tri), _p(m), C.c_int(max_bounces),
                                         _p(o), _p(d), _p(path), _p(rgb))
    return path[:n], rgb


def render_hybrid_rays(nodes, offsets, g_inv, aabb, origins, dirs, spheres, max_bounces=5, depth=12, f64=False,
                       threads=0):
    """Hybrid paths from explicit primary rays (secondary spheres only):
    rgb (n, 3) clamped to 1, segments (n,)."""
    nodes = np.ascontiguousarray(nodes)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    aabb = np.ascontiguousarray(aabb, np.float32)
    origins = np.ascontiguousarray(origins, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    sph = np.ascontiguousarray(spheres)
    n = len(origins)
    out = np.zeros((n, 3), np.float64 if f64 else np.float32)
    seg = np.zeros(n, np.int32)
    fn = lib().oracle_render_hybrid_rays_f64 if f64 else lib().oracle_render_hybrid_rays_f32
    fn(_p(nodes), _p(offsets), _p(g_inv), _p(aabb), _p(origins), _p(dirs), C.c_uint64(n), C.c_int(depth), C.c_int(1),
       C.c_int(len(sph)), _p(sph), None, None, C.c_int(max_bounces), _p(out), _p(seg), C.c_int(threads))
    return {"rgb": out, "segments": seg}


def ray_gaussian(o, d, g_rec, t_near=1e-2, t_far=3.4e38, f64=False):
    """rayPrimIntersect on one inverted-scale record: (accepted, t, grayDist)."""
    g = np.ascontiguousarray(g_rec).reshape(-1)[:1]
    if f64:
        out = np.zeros(2, np.float64)
        acc = lib().oracle_ray_gaussian_f64(_p(np.asarray(o, np.float64)), _p(np.asarray(d, np.float64)), _p(g),
                                            C.c_double(t_near), C.c_double(t_far), _p(out))
    else:
        out = np.zeros(2, np.float32)
        acc = lib().oracle_ray_gaussian_f32(_p(np.asarray(o, np.float32)), _p(np.asarray(d, np.float32)), _p(g),
                                            C.c_float(t_near), C.c_float(t_far), _p(out))
    return bool(acc), out[0], out[1]


def sh_basis(d):
    """The 16 SH basis scalars of eval_sh_final for direction d (float32)."""
    s = np.zeros(16, np.float32)
    lib().oracle_sh_basis_f32(_p(np.asarray(d, np.float32)), _p(s))
    return s


def eval_sh(d, g_rec, f64=False):
    """eval_sh_final(3, d, g) of one record: rgb (3,)."""
    g = np.ascontiguousarray(g_rec).reshape(-1)[:1]
    if f64:
        out = np.zeros(3, np.float64)
        lib().oracle_eval_sh_f64(_p(np.asarray(d, np.float64)), _p(g), _p(out))
    else:
        out = np.zeros(3, np.float32)
        lib().oracle_eval_sh_f32(_p(np.asarray(d, np.float32)), _p(g), _p(out))
    return out


def touch(nodes, offsets, g_inv, aabb, n_gaussians, ci=None, w=0, h=0, origins=None, dirs=None, depth=12):
    """Touch maps of a view (ci, w, h) or a ray list (origins, dirs): bool
    arrays node (n_nodes), slot / hit_slot (n_offsets), eval / blend
    (n_gaussians)."""
    nodes = np.ascontiguousarray(nodes)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    aabb = np.ascontiguousarray(aabb, np.float32)
    m = {"node": np.zeros(len(nodes), np.uint8), "slot": np.zeros(max(1, len(offsets)), np.uint8),
         "hit_slot": np.zeros(max(1, len(offsets)), np.uint8), "eval": np.zeros(n_gaussians, np.uint8),
         "blend": np.zeros(n_gaussians, np.uint8)}
    maps = [_p(m[k]) for k in ("node", "slot", "hit_slot", "eval", "blend")]
    if ci is not None:
        cam = np.frombuffer(bytes(ci), dtype=np.uint8).copy()
        lib().oracle_touch_view_f32(_p(nodes), _p(offsets), _p(g_inv), _p(aabb), _p(cam), C.c_int(w), C.c_int(h),
                                    C.c_int(depth), *maps, C.c_int(0))
    else:
        origins = np.ascontiguousarray(origins, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        lib().oracle_touch_rays_f32(_p(nodes), _p(offsets), _p(g_inv), _p(aabb), _p(origins), _p(dirs),
                                    C.c_uint64(len(origins)), C.c_int(depth), *maps, C.c_int(0))
    return {k: v[:len(offsets)].astype(bool) if k in ("slot", "hit_slot") else v.astype(bool) for k, v in m.items()}
