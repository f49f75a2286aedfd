"""ctypes loader for the CPU oracle (oracle/liboracle.so) — the checker.

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
