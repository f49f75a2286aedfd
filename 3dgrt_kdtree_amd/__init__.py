"""Python host side of the MI355X kd-tree 3D-Gaussian ray-render path.

Thin ctypes layer over lib/libkdgrt.so (C ABI in include/kdgrt.h): scene
generation, proxy meshing and SAH tree build on the host; device scenes,
plain/debug/explicit-ray renders, the persistent batched launch, the
autotuner and the serving loop on the GPU.  The GPU entry points raise
KdgrtError when the HIP library or a device is missing; there is no CPU
fallback in this package.

Reference entry points mirrored (Kd-treeConverter_v0.9/Kd-treeConverter/):
  Renderer.init()             <- initCuda()                     cudaRenderer.cu:2438
  Renderer.setup(...)         <- renderGaussianWithCudaSetup()  cudaRenderer.cu:2557
  Renderer.frame(...)         <- renderGaussianWithCudaFrame()  cudaRenderer.cu:3099
  Renderer.warm_up(...)       <- warmUp()                       cudaRenderer.cu:2532
  Renderer.cleanup()          <- cleanupCudaResources()         cudaRenderer.cu:3486
  build_tree(...)             <- build_kd_tree_for_composite_object  Kd-treeConverter.cpp:22
  Tree.write_kdt / read_kdt   <- dump_/read_kd_tree_...          Kd-treeConverter.cpp:297, 444
  Device.render_hybrid        <- renderKernelGaussian_Mesh_Hybrid cudaRenderer.cu:2217-2311
  obj_read_secondary          <- loadSecondaryMesh()             Kd-treeConverterMain.cpp:178-226
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("KDGRT_LIB_PATH") or os.path.join(_HERE, "lib", "libkdgrt.so")

GAUSSIAN_DTYPE = np.dtype([
    ("pos", "<f4", 3), ("opacity", "<f4"), ("scale", "<f4", 3), ("k_scale", "<f4"), ("rot", "<f4", 4),
    ("f_dc", "<f4", 3), ("f_rest", "<f4", 45), ("pad", "<f4", 4)])
assert GAUSSIAN_DTYPE.itemsize == 256
NODE_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4")])

OUT_RGBT = 0
OUT_RGB = 1
N_COUNTERS = 8
COUNTER_NAMES = ["node_visits", "leaf_visits", "isect_tests", "hits_found", "blend_ops", "max_sort", "n_eval",
                 "restarts"]

# Secondary primitives of the hybrid render (SECONDARY_PRIM, Kd-treeConverter.h:28-30)
# and their materials (MeshHit.materialType, cudaRenderer.cu:2050-2060).
SECONDARY_TRI = 0
SECONDARY_SPHERE = 1
MATERIAL_DIFFUSE = 0
MATERIAL_MIRROR = 1
MATERIAL_GLASS = 2
SECONDARY_SPHERE_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("material", "<i4"),
                                   ("reflectivity", "<f4"), ("ior", "<f4"), ("color", "<f4", 3)])
assert SECONDARY_SPHERE_DTYPE.itemsize == 40

STATUS = {0: "OK", 1: "INVALID_ARGUMENT", 2: "OUT_OF_MEMORY", 3: "DEVICE", 4: "IO", 5: "FORMAT", 6: "CAPACITY",
          7: "NOT_READY", 8: "NO_DEVICE", 9: "CANCELLED", 10: "TIMEOUT"}


class KdgrtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3), ("n", C.c_float * 3),
                ("fovy_deg", C.c_double), ("aspect", C.c_double), ("near_c", C.c_double)]


class CameraInfo(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3), ("start", C.c_float * 3),
                ("step_x", C.c_float), ("step_y", C.c_float)]


class Secondary(C.Structure):
    _fields_ = [("prim", C.c_int32), ("count", C.c_int32), ("spheres", C.c_void_p), ("d_vertices", C.c_void_p),
                ("mesh_material", C.c_int32), ("mesh_reflectivity", C.c_float), ("mesh_ior", C.c_float),
                ("mesh_color", C.c_float * 3), ("max_bounces", C.c_int32)]


class BuildParams(C.Structure):
    _fields_ = [("traversal_cost", C.c_float), ("isect_cost", C.c_float), ("empty_bonus", C.c_float),
                ("max_level", C.c_uint32), ("min_prims", C.c_uint32), ("force_split", C.c_uint32),
                ("sah_mode", C.c_int32), ("threads", C.c_int32)]


class TreeStats(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("n_leaves", C.c_uint32), ("n_empty_leaves", C.c_uint32),
                ("max_depth", C.c_uint32), ("max_leaf_size", C.c_uint32), ("n_offsets", C.c_uint64),
                ("n_triangles", C.c_uint64), ("n_kept_gaussians", C.c_uint64), ("build_seconds", C.c_double),
                ("mesh_seconds", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("tile_w", C.c_int32), ("tile_h", C.c_int32), ("stack_depth", C.c_int32),
                ("hit_capacity", C.c_int32), ("waves_per_cu", C.c_int32), ("block_threads", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class View(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("cam", CameraInfo), ("width", C.c_int32), ("height", C.c_int32),
                ("out_rgbt", C.c_void_p)]


class RaySet(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("origins", C.c_void_p), ("dirs", C.c_void_p), ("n_rays", C.c_uint64),
                ("out_rgbt", C.c_void_p)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch_views", C.c_int32),
                ("max_inflight_batches", C.c_int32), ("plain_launches", C.c_int32), ("tuning", Tuning),
                ("batch_window_ms", C.c_double), ("simulate_mrays_per_s", C.c_double),
                ("persistent_kernel", C.c_int32), ("persistent_max_requests", C.c_int32),
                ("epoch_idle_ms", C.c_double), ("persistent_spin_limit", C.c_uint32), ("idle_plain", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("scene_id", C.c_int32), ("cam", CameraInfo), ("width", C.c_int32), ("height", C.c_int32),
                ("arrival_s", C.c_double), ("deadline_s", C.c_double)]


class ResultInfo(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("status", C.c_int32), ("arrival_s", C.c_double),
                ("admitted_s", C.c_double), ("launched_s", C.c_double), ("done_s", C.c_double),
                ("batch_id", C.c_uint32), ("batch_views", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("cancelled", C.c_uint64), ("batches", C.c_uint64), ("rays", C.c_uint64), ("busy_s", C.c_double),
                ("wall_s", C.c_double), ("bytes_resident", C.c_uint64), ("bytes_peak", C.c_uint64)]


_lib = None


def lib():
    """Loads lib/libkdgrt.so (built by build.py); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KdgrtError(3, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.kdgrt_last_error.restype = C.c_char_p
        L.kdgrt_version.restype = C.c_char_p
        L.kdgrt_tree_nodes.restype = C.c_void_p
        L.kdgrt_tree_offsets.restype = C.c_void_p
        L.kdgrt_scene_bytes_estimate.restype = C.c_uint64
        L.kdgrt_scene_bytes_estimate.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64]
        if hasattr(L, "kdgrt_server_now"):
            L.kdgrt_server_now.restype = C.c_double
        _lib = L
    return _lib


def check(st):
    if st != 0:
        raise KdgrtError(st, lib().kdgrt_last_error().decode(errors="replace"))


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


# ------------------------------------------------------------------ host side
def generate_gaussians(kind, n, seed):
    """Synthetic activated Gaussians (reference 256-B layout, scales not inverted)."""
    g = np.zeros(n, dtype=GAUSSIAN_DTYPE)
    check(lib().kdgrt_generate_gaussians(C.c_int(kind), C.c_uint64(n), C.c_uint64(seed), _p(g)))
    return g


def invert_scales(g):
    g = g.copy()
    check(lib().kdgrt_gaussians_invert_scales(_p(g), C.c_uint64(len(g))))
    return g


def default_build_params():
    p = BuildParams()
    check(lib().kdgrt_build_params_default(C.byref(p)))
    return p


def proxy_mesh(g):
    n = C.c_uint64()
    check(lib().kdgrt_proxy_mesh(_p(g), C.c_uint64(len(g)), None, None, C.c_uint64(0), C.byref(n), None))
    xyz = np.zeros((n.value, 3, 3), np.float32)
    gid = np.zeros(n.value, np.uint32)
    aabb = np.zeros(6, np.float32)
    check(lib().kdgrt_proxy_mesh(_p(g), C.c_uint64(len(g)), _p(xyz), _p(gid), C.c_uint64(n.value), C.byref(n),
                                 _p(aabb)))
    return xyz, gid, aabb


class Tree:
    """Host kd-tree (nodes, leaf Gaussian-ID lists, scene AABB in .kdt order)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @staticmethod
    def build(g, params=None, threads=0):
        p = params or default_build_params()
        p.threads = threads
        h = C.c_void_p()
        check(lib().kdgrt_tree_build(_p(g), C.c_uint64(len(g)), C.byref(p), C.byref(h)))
        return Tree(h)

    @staticmethod
    def from_arrays(nodes, offsets, aabb):
        nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        aabb = np.ascontiguousarray(aabb, dtype=np.float32)
        h = C.c_void_p()
        check(lib().kdgrt_tree_from_arrays(_p(nodes), C.c_uint32(len(nodes)), _p(offsets) if len(offsets) else None,
                                           C.c_uint64(len(offsets)), _p(aabb), C.byref(h)))
        return Tree(h)

    @staticmethod
    def read_kdt(path):
        h = C.c_void_p()
        check(lib().kdgrt_kdt_read(path.encode(), C.byref(h)))
        return Tree(h)

    def write_kdt(self, path):
        check(lib().kdgrt_kdt_write(self.h, path.encode()))

    def stats(self):
        s = TreeStats()
        check(lib().kdgrt_tree_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in TreeStats._fields_}

    @property
    def nodes(self):
        n = self.stats()["n_nodes"]
        ptr = lib().kdgrt_tree_nodes(self.h)
        return np.ctypeslib.as_array((C.c_uint32 * (2 * n)).from_address(ptr)).view(NODE_DTYPE).copy()

    @property
    def offsets(self):
        n = self.stats()["n_offsets"]
        if n == 0:
            return np.zeros(0, np.uint32)
        ptr = lib().kdgrt_tree_offsets(self.h)
        return np.ctypeslib.as_array((C.c_uint32 * n).from_address(ptr)).copy()

    @property
    def aabb(self):
        a = np.zeros(6, np.float32)
        check(lib().kdgrt_tree_aabb(self.h, _p(a)))
        return a

    def validate(self, n_gaussians=0):
        check(lib().kdgrt_tree_validate(self.h, C.c_uint64(n_gaussians)))

    def free(self):
        if self.h:
            lib().kdgrt_tree_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def ply_read(path):
    """Activated Gaussians of a 3DGS PLY file (kdgrt_ply_read: sigmoid opacity,
    exp scales, normalised quaternions, k_scale), reference 256-B layout."""
    n = C.c_uint64()
    check(lib().kdgrt_ply_count(str(path).encode(), C.byref(n)))
    g = np.zeros(n.value, GAUSSIAN_DTYPE)
    check(lib().kdgrt_ply_read(str(path).encode(), _p(g), C.c_uint64(n.value), C.byref(n)))
    return g[:n.value]


def cameras_from_nerf_json(path, aspect):
    """Camera list of a NeRF transforms.json (kdgrt_cameras_from_nerf_json)."""
    n = C.c_int()
    check(lib().kdgrt_cameras_from_nerf_json(path.encode(), C.c_double(aspect), None, C.c_int(0), C.byref(n)))
    arr = (Camera * max(1, n.value))()
    check(lib().kdgrt_cameras_from_nerf_json(path.encode(), C.c_double(aspect), arr, C.c_int(n.value), C.byref(n)))
    return list(arr)[:n.value]


def demo_spheres():
    """The reference's two demo spheres (cudaRenderer.cu:2162-2167)."""
    out = np.zeros(2, SECONDARY_SPHERE_DTYPE)
    check(lib().kdgrt_secondary_demo_spheres(_p(out)))
    return out


def obj_read_secondary(path):
    """Triangles of an OBJ file as the hybrid render's secondary mesh,
    shape (n, 9) float32, vertices stored as (x, -z, y)."""
    n = C.c_uint64()
    check(lib().kdgrt_obj_read_secondary(str(path).encode(), None, C.c_uint64(0), C.byref(n)))
    out = np.zeros((n.value, 9), np.float32)
    if n.value:
        check(lib().kdgrt_obj_read_secondary(str(path).encode(), _p(out), C.c_uint64(n.value), C.byref(n)))
    return out


def write_png(path, pixels):
    """PNG file of 8-bit pixels, shape (h, w, 3 or 4), rows top to bottom."""
    px = np.ascontiguousarray(pixels, np.uint8)
    h, w, c = px.shape
    check(lib().kdgrt_png_write(str(path).encode(), C.c_int(w), C.c_int(h), C.c_int(c), _p(px)))


def write_stats(counters, width, height, directory):
    """Writes the reference's per-ray statistics files from render_debug
    counters (shape (height, width, 8) or (height*width, 8), ray order)."""
    c = np.ascontiguousarray(counters, dtype=np.int32).reshape(-1, N_COUNTERS)
    assert len(c) == width * height
    check(lib().kdgrt_stats_write(_p(c), C.c_int(width), C.c_int(height), str(directory).encode()))


def camera_make(cam, width, height):
    ci = CameraInfo()
    check(lib().kdgrt_camera_make(C.byref(cam), C.c_int(width), C.c_int(height), C.byref(ci)))
    return ci


def look_at(eye, target, fovy_deg=60.0, aspect=1.0, near=0.005, up=(0.0, 1.0, 0.0)):
    """Pinhole camera at `eye` looking at `target` (forward = -n)."""
    eye = np.asarray(eye, np.float64)
    n = eye - np.asarray(target, np.float64)
    n /= np.linalg.norm(n)
    up = np.asarray(up, np.float64)
    if abs(np.dot(up, n)) > 0.999:
        up = np.array([0.0, 0.0, 1.0])
    u = np.cross(up, n)
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    c = Camera()
    for k in range(3):
        c.pos[k], c.u[k], c.v[k], c.n[k] = eye[k], u[k], v[k], n[k]
    c.fovy_deg, c.aspect, c.near_c = fovy_deg, aspect, near
    return c


def cam_info_bytes(ci):
    return bytes(ci)


# ------------------------------------------------------- viewer camera controls
KEY_W, KEY_S, KEY_A, KEY_D, KEY_Q, KEY_E = 1, 2, 4, 8, 16, 32
# the reference viewer's defaults (DefineConst.h:16-17, updateOrbitCamera)
CAM_MOVE_SPEED, CAM_ROT_SPEED, ORBIT_SPEED = 0.5, 0.1, 0.8
ORBIT_CENTRE, ORBIT_AXIS = (0.393, 1.093, 0.698), (0.0138, 0.9038, 0.4278)


def camera_copy(cam):
    c = Camera()
    C.memmove(C.byref(c), C.byref(cam), C.sizeof(Camera))
    return c


def camera_move(cam, keys, speed=CAM_MOVE_SPEED, dt=1.0 / 60.0):
    """Held-key move (KEY_* bits) in place; returns whether a key was held."""
    moved = C.c_int()
    check(lib().kdgrt_camera_move(C.byref(cam), C.c_uint32(keys), C.c_float(speed), C.c_float(dt),
                                  C.byref(moved)))
    return bool(moved.value)


def camera_orbit(cam, angle, centre=ORBIT_CENTRE, axis=ORBIT_AXIS):
    """Orbit about the line (centre, axis) by `angle` radians, in place."""
    ce, ax = (C.c_float * 3)(*centre), (C.c_float * 3)(*axis)
    check(lib().kdgrt_camera_orbit(C.byref(cam), ce, ax, C.c_float(angle)))


def camera_drag(cam, delx, dely, rot_speed=CAM_ROT_SPEED, zoom=False):
    """Mouse drag: yaw/pitch (or fovy zoom with zoom=True), in place."""
    check(lib().kdgrt_camera_drag(C.byref(cam), C.c_int(delx), C.c_int(dely), C.c_float(rot_speed),
                                  C.c_int(1 if zoom else 0)))


# ------------------------------------------------------------------ device side
class Device:
    """A kdgrt context on one HIP device plus helpers for device buffers."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().kdgrt_init(C.c_int(device), C.byref(self.h)))
        self.device = device

    def info(self):
        name = C.create_string_buffer(256)
        ncu, mem, lds = C.c_int(), C.c_size_t(), C.c_size_t()
        check(lib().kdgrt_device_info(self.h, name, C.c_size_t(256), C.byref(ncu), C.byref(mem), C.byref(lds)))
        return {"name": name.value.decode(), "n_cu": ncu.value, "total_mem": mem.value, "lds_per_block": lds.value}

    def alloc(self, nbytes):
        p = C.c_void_p()
        check(lib().kdgrt_device_alloc(self.h, C.c_size_t(nbytes), C.byref(p)))
        return DeviceBuffer(self, p, nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        b = self.alloc(arr.nbytes)
        check(lib().kdgrt_memcpy_h2d(self.h, b.ptr, _p(arr), C.c_size_t(arr.nbytes)))
        return b

    def stream(self):
        s = C.c_void_p()
        check(lib().kdgrt_stream_create(self.h, C.byref(s)))
        return s

    def sync(self, stream=None):
        if stream is None:
            check(lib().kdgrt_device_sync(self.h))
        else:
            check(lib().kdgrt_stream_sync(self.h, stream))

    def timer_begin(self, stream=None):
        check(lib().kdgrt_timer_begin(self.h, stream))

    def timer_end(self, stream=None):
        ms = C.c_float()
        check(lib().kdgrt_timer_end(self.h, stream, C.byref(ms)))
        return ms.value

    def scene(self, tree, g, scales_inverted=False):
        return Scene(self, tree, g, scales_inverted)

    def render(self, scene, ci, w, h, out, depth=12, fmt=OUT_RGBT, stream=None):
        check(lib().kdgrt_render(self.h, scene.h, C.byref(ci), C.c_int(w), C.c_int(h), C.c_int(depth), C.c_int(fmt),
                                 out.ptr, stream))

    def render_image(self, scene, ci, w, h, depth=12, fmt=OUT_RGBT):
        ch = 4 if fmt == OUT_RGBT else 3
        out = self.alloc(w * h * ch * 4)
        self.render(scene, ci, w, h, out, depth, fmt)
        self.sync()
        img = out.download(np.float32, (h, w, ch))
        out.free()
        return img

    def render_hybrid(self, scene, ci, w, h, spheres=None, triangles=None, mesh_material=MATERIAL_MIRROR,
                      mesh_reflectivity=0.95, mesh_ior=1.1, mesh_color=(0.95, 0.95, 0.95), max_bounces=5, depth=12,
                      out=None, stream=None):
        """Hybrid view (kdgrt_render_hybrid): secondary spheres (an array of
        SECONDARY_SPHERE_DTYPE) or a triangle mesh ((n, 9) float32, one
        material).  Returns the (h, w, 3) image, or launches into the device
        buffer `out` without waiting when one is given."""
        sec = Secondary()
        sec.max_bounces = max_bounces
        keep = []
        if spheres is not None:
            sph = np.ascontiguousarray(spheres, SECONDARY_SPHERE_DTYPE)
            keep.append(sph)
            sec.prim, sec.count, sec.spheres = SECONDARY_SPHERE, len(sph), sph.ctypes.data
        else:
            tri = np.ascontiguousarray(triangles if triangles is not None else np.zeros((0, 9)), np.float32)
            tri = tri.reshape(-1, 9)
            dv = self.upload(tri) if len(tri) else None
            keep.append(dv)
            sec.prim, sec.count, sec.d_vertices = SECONDARY_TRI, len(tri), dv.ptr if dv else None
            sec.mesh_material, sec.mesh_reflectivity, sec.mesh_ior = mesh_material, mesh_reflectivity, mesh_ior
            sec.mesh_color[:] = mesh_color
        buf = out if out is not None else self.alloc(w * h * 12)
        try:
            check(lib().kdgrt_render_hybrid(self.h, scene.h, C.byref(ci), C.c_int(w), C.c_int(h), C.c_int(depth),
                                            C.byref(sec), buf.ptr, stream))
            if out is not None:
                return None
            self.sync()
            return buf.download(np.float32, (h, w, 3))
        finally:
            if out is None:
                buf.free()
            for k in keep:
                if isinstance(k, DeviceBuffer):
                    self.sync(stream)
                    k.free()

    def framebuffer_to_u8(self, fb, w, h, in_channels, out_channels=3, flip_rows=True, stream=None):
        """8-bit pixels (h, w, out_channels) of a device framebuffer, converted
        on the device as the reference's PNG dump does."""
        out = self.alloc(w * h * out_channels)
        try:
            check(lib().kdgrt_framebuffer_to_u8(self.h, fb.ptr, C.c_int(w), C.c_int(h), C.c_int(in_channels), out.ptr,
                                                C.c_int(out_channels), C.c_int(1 if flip_rows else 0), stream))
            self.sync(stream)
            return out.download(np.uint8, (h, w, out_channels))
        finally:
            out.free()

    def render_debug(self, scene, ci, w, h, depth=12, seq_cap=0):
        out = self.alloc(w * h * 16)
        cnt = self.alloc(w * h * N_COUNTERS * 4)
        hsh = self.alloc(w * h * 8)
        seq = None
        if seq_cap:
            seq = self.alloc(w * h * seq_cap * 4)
            check(lib().kdgrt_memset(self.h, seq.ptr, C.c_int(0xFF), C.c_size_t(w * h * seq_cap * 4)))
        check(lib().kdgrt_render_debug(self.h, scene.h, C.byref(ci), C.c_int(w), C.c_int(h), C.c_int(depth), out.ptr,
                                       cnt.ptr, hsh.ptr, seq.ptr if seq else None, C.c_int(seq_cap), None))
        self.sync()
        res = {"rgbt": out.download(np.float32, (h, w, 4)),
               "counters": cnt.download(np.int32, (h, w, N_COUNTERS)),
               "hash": hsh.download(np.uint64, (h, w))}
        if seq:
            res["seq"] = seq.download(np.uint32, (h, w, seq_cap))
            seq.free()
        for b in (out, cnt, hsh):
            b.free()
        return res

    def render_rays(self, scene, origins, dirs, depth=12, debug=False):
        origins = np.ascontiguousarray(origins, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(origins)
        o, d = self.upload(origins), self.upload(dirs)
        out = self.alloc(n * 16)
        res = {}
        if debug:
            cnt, hsh = self.alloc(n * 32), self.alloc(n * 8)
            check(lib().kdgrt_render_rays_debug(self.h, scene.h, o.ptr, d.ptr, C.c_uint64(n), C.c_int(depth), out.ptr,
                                                cnt.ptr, hsh.ptr, None))
            self.sync()
            res["counters"] = cnt.download(np.int32, (n, N_COUNTERS))
            res["hash"] = hsh.download(np.uint64, (n,))
            cnt.free()
            hsh.free()
        else:
            check(lib().kdgrt_render_rays(self.h, scene.h, o.ptr, d.ptr, C.c_uint64(n), C.c_int(depth), out.ptr, None))
            self.sync()
        res["rgbt"] = out.download(np.float32, (n, 4))
        for b in (o, d, out):
            b.free()
        return res

    def _touch_maps(self, scene):
        sizes = {"nodes": scene.n_nodes, "quads": scene.n_nodes, "slots_pretested": scene.n_offsets,
                 "slots_tested": scene.n_offsets, "gaussians_evaluated": scene.n_gaussians,
                 "gaussians_blended": scene.n_gaussians}
        bufs = {k: self.alloc(max(1, n)) for k, n in sizes.items()}
        for k, b in bufs.items():
            check(lib().kdgrt_memset(self.h, b.ptr, C.c_int(0), C.c_size_t(max(1, sizes[k]))))
        m = TouchMaps(*[bufs[k].ptr for k, _ in TouchMaps._fields_])
        return m, bufs, sizes

    def _touch_result(self, bufs, sizes):
        res = {k: b.download(np.uint8, (sizes[k],)).astype(bool) for k, b in bufs.items()}
        for b in bufs.values():
            b.free()
        return res

    def valu_probe(self, iters=1 << 14, stream=None, reps=3):
        """The VALU-issue ceiling (k_valu_probe): fused multiply-adds issued per
        second over `reps` timed launches after one warm-up."""
        waves, per = C.c_uint64(), C.c_uint64()
        check(lib().kdgrt_valu_probe(self.h, C.c_int(iters), stream, C.byref(waves), C.byref(per)))
        self.sync(stream)
        self.timer_begin(stream)
        for _ in range(reps):
            check(lib().kdgrt_valu_probe(self.h, C.c_int(iters), stream, C.byref(waves), C.byref(per)))
        ms = self.timer_end(stream) / reps
        self.sync(stream)
        return {"kernel": "k_valu_probe", "waves": waves.value, "fma_per_wave": per.value, "ms": ms,
                "valu_instr_per_s": waves.value * per.value / (ms * 1e-3)}

    def view_cone_table(self, scene, eye):
        """The cone records a plain view from `eye` pre-tests
        (kdgrt_view_cone_table): float32 (n_offsets + 4, 4), rows (a, q) in
        leaf-slot order, then four zero padding rows."""
        e = (C.c_float * 3)(*[float(v) for v in eye])
        n = C.c_uint64()
        check(lib().kdgrt_view_cone_table(self.h, scene.h, e, None, C.c_uint64(0), C.byref(n), None))
        out = self.alloc(n.value * 16)
        try:
            check(lib().kdgrt_view_cone_table(self.h, scene.h, e, out.ptr, C.c_uint64(n.value), C.byref(n), None))
            self.sync()
            return out.download(np.float32, (n.value, 4))
        finally:
            out.free()

    def render_touch(self, scene, ci, w, h, depth=12):
        """Touch maps of one plain view (kdgrt_render_touch): which node,
        quad, leaf-slot and Gaussian records the launch reads."""
        out = self.alloc(w * h * 16)
        m, bufs, sizes = self._touch_maps(scene)
        check(lib().kdgrt_render_touch(self.h, scene.h, C.byref(ci), C.c_int(w), C.c_int(h), C.c_int(depth), out.ptr,
                                       C.byref(m), None))
        self.sync()
        res = self._touch_result(bufs, sizes)
        res["rgbt"] = out.download(np.float32, (h, w, 4))
        out.free()
        return res

    def render_rays_touch(self, scene, origins, dirs, depth=12):
        """Touch maps of a ray list (kdgrt_render_rays_touch)."""
        origins = np.ascontiguousarray(origins, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(origins)
        o, d, out = self.upload(origins), self.upload(dirs), self.alloc(n * 16)
        m, bufs, sizes = self._touch_maps(scene)
        check(lib().kdgrt_render_rays_touch(self.h, scene.h, o.ptr, d.ptr, C.c_uint64(n), C.c_int(depth), out.ptr,
                                            C.byref(m), None))
        self.sync()
        res = self._touch_result(bufs, sizes)
        res["rgbt"] = out.download(np.float32, (n, 4))
        for b in (o, d, out):
            b.free()
        return res

    def close(self):
        if self.h:
            lib().kdgrt_shutdown(self.h)
            self.h = None


class TouchMaps(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("quads", C.c_void_p), ("slots_pretested", C.c_void_p),
                ("slots_tested", C.c_void_p), ("gaussians_evaluated", C.c_void_p),
                ("gaussians_blended", C.c_void_p)]


def compulsory_bytes(touch, offsets, n_rays, explicit_rays=False):
    """Compulsory bytes of a launch: every record its rays read, counted once.

    reference layout (cudaRenderer.cu:292-307, 1079-1085, 1595): 8-B node
    records visited, 4-B leaf-list entries, the 48-B geometry of each Gaussian
    tested, the rest of the 256-B record of each Gaussian evaluated, 16 B of
    output per ray (+24 B of ray input for explicit rays).
    this layout: the root node, 48-B child-pair quads loaded, a 16-B pre-test
    record per slot pre-tested (cone records for views, bound spheres for ray
    lists), the 80-B geometry record of each slot exact-tested, the 192-B SH
    record of each Gaussian blended, the same output and input."""
    offsets = np.asarray(offsets)
    io = n_rays * (16 + (24 if explicit_rays else 0))
    tested = np.unique(offsets[np.flatnonzero(touch["slots_pretested"])])
    evaluated = int(touch["gaussians_evaluated"].sum())
    ref = {"nodes": 8 * int(touch["nodes"].sum()), "leaf_lists": 4 * int(touch["slots_pretested"].sum()),
           "gaussian_geometry": 48 * len(tested), "gaussian_rest": (256 - 48) * evaluated, "io": io}
    ours = {"root": 8, "quads": 48 * int(touch["quads"].sum()), "pretest": 16 * int(touch["slots_pretested"].sum()),
            "geometry": 80 * int(touch["slots_tested"].sum()), "sh": 192 * int(touch["gaussians_blended"].sum()),
            "io": io}
    return {"reference_layout": sum(ref.values()), "this_layout": sum(ours.values()),
            "reference_layout_parts": ref, "this_layout_parts": ours}


class DeviceBuffer:
    def __init__(self, dev, ptr, nbytes):
        self.dev, self.ptr, self.nbytes = dev, ptr, nbytes

    def download(self, dtype, shape):
        a = np.empty(shape, dtype)
        assert a.nbytes <= self.nbytes
        check(lib().kdgrt_memcpy_d2h(self.dev.h, _p(a), self.ptr, C.c_size_t(a.nbytes)))
        return a

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().kdgrt_memcpy_h2d(self.dev.h, self.ptr, _p(arr), C.c_size_t(arr.nbytes)))

    def free(self):
        if self.ptr:
            lib().kdgrt_device_free(self.dev.h, self.ptr)
            self.ptr = None


class Scene:
    """Device-resident scene (repacked tree + Gaussians)."""

    def __init__(self, dev, tree, g, scales_inverted=False):
        self.dev = dev
        self.h = C.c_void_p()
        g = np.ascontiguousarray(g, dtype=GAUSSIAN_DTYPE)
        check(lib().kdgrt_scene_create_from_tree(dev.h, tree.h, _p(g), C.c_uint64(len(g)),
                                                 C.c_int(1 if scales_inverted else 0), C.byref(self.h)))
        st = tree.stats()
        self.n_nodes, self.n_offsets, self.n_gaussians = int(st["n_nodes"]), int(st["n_offsets"]), len(g)

    def device_bytes(self):
        b = C.c_uint64()
        check(lib().kdgrt_scene_device_bytes(self.h, C.byref(b)))
        return b.value

    def set_tile_schedule(self, enable):
        """Tile order of Device.render / render_hybrid (on by default): tiles
        start longest-first, by the previous launch's durations when the
        camera stays put, else by the projected density of Gaussian centres.
        Off, launches run in raster order.  Results are the same either way."""
        check(lib().kdgrt_scene_set_tile_schedule(self.h, C.c_int(1 if enable else 0)))

    def free(self):
        if self.h:
            lib().kdgrt_scene_destroy(self.h)
            self.h = None


class Renderer:
    """The reference renderer's host API (cudaRenderer.h:60-109) for Python
    callers, on the C ABI: init() <- initCuda, setup() <-
    renderGaussianWithCudaSetup (Gaussians carry inverted scales, the
    caller contract of Kd-treeConverterMain.cpp:3636-3642; with `secondary`
    it is the SECONDARY_RAY set-up), frame() <- renderGaussianWithCudaFrame
    (w*h*3 float RGB into a caller-owned device buffer, rows flipped, blocks,
    returns kernel-only FPS), warm_up() <- warmUp, cleanup() <-
    cleanupCudaResources."""

    def __init__(self):
        self.dev = None
        self.scene = None
        self.secondary = None

    def init(self, device=0):
        if self.dev is None:
            self.dev = Device(device)
        return True

    def setup(self, tree, gaussians, secondary=None):
        """secondary: None (plain frames), or a dict for hybrid frames:
        {"spheres": array of SECONDARY_SPHERE_DTYPE} or {"triangles": (n, 9)
        float32, "mesh_material", "mesh_reflectivity", "mesh_ior",
        "mesh_color"}; "max_bounces" defaults to the reference's 5."""
        if self.dev is None:
            raise KdgrtError(1, "setup: init() was not called")
        if len(gaussians) == 0:
            raise KdgrtError(1, "setup: Kd-tree or Gaussian data is empty")
        if self.scene is not None:
            self.scene.free()
        self.scene = Scene(self.dev, tree, gaussians, scales_inverted=True)
        self.secondary = dict(secondary) if secondary is not None else None

    def frame(self, camera, width, height, fb, stream=None):
        if self.scene is None:
            raise KdgrtError(1, "frame: no scene set up")
        ci = camera_make(camera, width, height)
        self.dev.timer_begin(stream)
        if self.secondary is None:
            self.dev.render(self.scene, ci, width, height, fb, fmt=OUT_RGB, stream=stream)
        else:
            self.dev.render_hybrid(self.scene, ci, width, height, out=fb, stream=stream, **self.secondary)
        ms = self.dev.timer_end(stream)
        self.dev.sync(stream)
        return 1000.0 / ms if ms > 0 else 0.0

    def warm_up(self, stream=None):
        """One plain frame from a corner of the scene box into a scratch
        buffer (warmUp renders the plain kernel once to load the code object
        and fault in the scene)."""
        if self.scene is None:
            raise KdgrtError(1, "warm_up: no scene set up")
        check(lib().kdgrt_warm_up(self.dev.h, self.scene.h, stream))

    def cleanup(self):
        if self.scene is not None:
            self.scene.free()
            self.scene = None
        if self.dev is not None:
            self.dev.close()
            self.dev = None


def default_tuning(dev=None):
    t = Tuning()
    check(lib().kdgrt_tuning_default(dev.h if dev else None, C.byref(t)))
    return t


class Batch:
    """A persistent batched launch over many views / ray sets."""

    def __init__(self, dev, views, ray_sets=(), tuning=None):
        self.dev = dev
        nv, ns = len(views), len(ray_sets)
        va = (View * max(1, nv))(*views)
        ra = (RaySet * max(1, ns))(*ray_sets)
        self.h = C.c_void_p()
        check(lib().kdgrt_batch_create(dev.h, va if nv else None, C.c_int(nv), ra if ns else None, C.c_int(ns),
                                       C.byref(tuning) if tuning is not None else None, C.byref(self.h)))

    def run(self, stream=None, use_graph=False):
        check(lib().kdgrt_batch_run(self.h, stream, C.c_int(1 if use_graph else 0)))

    def info(self):
        r, t, g = C.c_uint64(), C.c_uint32(), C.c_uint32()
        check(lib().kdgrt_batch_info(self.h, C.byref(r), C.byref(t), C.byref(g)))
        return {"n_rays": r.value, "n_tiles": t.value, "grid_blocks": g.value}

    def free(self):
        if self.h:
            lib().kdgrt_batch_destroy(self.h)
            self.h = None


class Server:
    """Serving loop (kdgrt_server_*): resident scenes, HBM-budget admission,
    continuous batching.  dev=None with simulate_mrays_per_s > 0 runs the same
    scheduler against a modelled device."""

    def __init__(self, dev, budget_bytes, max_batch_views=16, max_inflight=2, plain=False, tuning=None,
                 window_ms=1.0, simulate_mrays_per_s=0.0, persistent=False, epoch_max_requests=0, epoch_idle_ms=5.0,
                 spin_limit=0, idle_plain=False):
        cfg = ServerConfig()
        cfg.hbm_budget_bytes = int(budget_bytes)
        cfg.max_batch_views = int(max_batch_views)
        cfg.max_inflight_batches = int(max_inflight)
        cfg.plain_launches = 1 if plain else 0
        cfg.tuning = tuning if tuning is not None else default_tuning(dev)
        cfg.batch_window_ms = float(window_ms)
        cfg.simulate_mrays_per_s = float(simulate_mrays_per_s)
        cfg.persistent_kernel = 1 if persistent else 0
        cfg.persistent_max_requests = int(epoch_max_requests)
        cfg.epoch_idle_ms = float(epoch_idle_ms)
        cfg.persistent_spin_limit = int(spin_limit)
        cfg.idle_plain = 1 if idle_plain else 0
        self.h = C.c_void_p()
        check(lib().kdgrt_server_create(dev.h if dev is not None else None, C.byref(cfg), C.byref(self.h)))

    def add_scene(self, tree, g, scales_inverted=False):
        g = np.ascontiguousarray(g, dtype=GAUSSIAN_DTYPE)
        sid = C.c_int32()
        check(lib().kdgrt_server_add_scene(self.h, tree.h, _p(g), C.c_uint64(len(g)),
                                           C.c_int(1 if scales_inverted else 0), C.byref(sid)))
        return sid.value

    def add_resident_scene(self, scene):
        """Serves a device Scene the caller keeps (it must outlive the server)."""
        sid = C.c_int32()
        check(lib().kdgrt_server_add_resident_scene(self.h, scene.h, C.byref(sid)))
        return sid.value

    def start(self):
        check(lib().kdgrt_server_start(self.h))

    def submit(self, scene_id, cam_info, width, height, arrival_s=0.0, deadline_s=0.0):
        r = Request()
        r.scene_id, r.cam, r.width, r.height = scene_id, cam_info, width, height
        r.arrival_s, r.deadline_s = arrival_s, deadline_s
        rid = C.c_uint64()
        check(lib().kdgrt_server_submit(self.h, C.byref(r), C.byref(rid)))
        return rid.value

    def wait(self, request_id, timeout_ms=-1.0, out=None):
        """Waits for a request; `out` (float32 array of width*height*4) receives
        the RGBT image if given.  Returns the ResultInfo."""
        info = ResultInfo()
        ptr = None
        if out is not None:
            assert out.dtype == np.float32 and out.flags["C_CONTIGUOUS"]
            ptr = out.ctypes.data_as(C.POINTER(C.c_float))
        check(lib().kdgrt_server_wait(self.h, C.c_uint64(request_id), C.c_double(timeout_ms), ptr,
                                      C.byref(info)))
        return info

    def now(self):
        return lib().kdgrt_server_now(self.h)

    def stats(self):
        st = ServerStats()
        check(lib().kdgrt_server_stats_get(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def stop(self):
        check(lib().kdgrt_server_stop(self.h))

    def close(self):
        if self.h:
            lib().kdgrt_server_stop(self.h)
            lib().kdgrt_server_destroy(self.h)
            self.h = None


FRAME_RGB_F32, FRAME_RGBA8, FRAME_RGB8 = 0, 1, 2


class ViewerFrame(C.Structure):
    _fields_ = [("frame_id", C.c_uint64), ("pixels", C.c_void_p), ("bytes", C.c_uint64), ("width", C.c_int),
                ("height", C.c_int), ("format", C.c_int), ("render_ms", C.c_float), ("fps", C.c_float),
                ("camera", Camera)]


class ViewerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("delivered", C.c_uint64), ("mean_render_ms", C.c_double),
                ("window_avg_fps", C.c_double), ("window_frames", C.c_uint64), ("wall_fps", C.c_double)]


class Viewer:
    """Frame loop of the interactive viewer (kdgrt_viewer_*): frames rendered
    into a ring of slots and copied to pinned host memory while the next
    frame renders; acquire() returns them in submission order."""

    def __init__(self, dev, scene, width, height, fmt=FRAME_RGBA8, slots=3, stack_depth=12, window=(500, 1000)):
        self.dev, self.width, self.height, self.format = dev, width, height, fmt
        self.h = C.c_void_p()
        check(lib().kdgrt_viewer_create(dev.h, scene.h, C.c_int(width), C.c_int(height), C.c_int(fmt),
                                        C.c_int(slots), C.c_int(stack_depth), C.c_int(window[0]), C.c_int(window[1]),
                                        C.byref(self.h)))

    def submit(self, cam, timeout_ms=-1):
        fid = C.c_uint64()
        check(lib().kdgrt_viewer_submit(self.h, C.byref(cam), C.c_int(timeout_ms), C.byref(fid)))
        return fid.value

    def acquire(self, timeout_ms=-1, copy=True):
        """(frame info, pixels): pixels is a copy (copy=True) or a view of the
        slot's pinned memory, valid until release(frame_id)."""
        f = ViewerFrame()
        check(lib().kdgrt_viewer_acquire(self.h, C.c_int(timeout_ms), C.byref(f)))
        if f.format == FRAME_RGB_F32:
            dt, shape = np.float32, (f.height, f.width, 3)
        else:
            dt, shape = np.uint8, (f.height, f.width, 4 if f.format == FRAME_RGBA8 else 3)
        buf = (C.c_char * f.bytes).from_address(f.pixels)
        px = np.frombuffer(buf, dtype=dt).reshape(shape)
        return f, (px.copy() if copy else px)

    def release(self, frame_id):
        check(lib().kdgrt_viewer_release(self.h, C.c_uint64(frame_id)))

    def stats(self):
        st = ViewerStats()
        check(lib().kdgrt_viewer_stats_get(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ViewerStats._fields_}

    def close(self):
        if self.h:
            lib().kdgrt_viewer_destroy(self.h)
            self.h = None


def autotune(dev, views, ray_sets=(), max_trials=24, reps=3):
    nv, ns = len(views), len(ray_sets)
    va = (View * max(1, nv))(*views)
    ra = (RaySet * max(1, ns))(*ray_sets)
    best, ms = Tuning(), C.c_float()
    check(lib().kdgrt_autotune(dev.h, va if nv else None, C.c_int(nv), ra if ns else None, C.c_int(ns),
                               C.c_int(max_trials), C.c_int(reps), C.byref(best), C.byref(ms)))
    return best, ms.value
