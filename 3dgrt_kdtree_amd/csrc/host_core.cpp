// Host core of the kd-tree Gaussian ray-render path: errors, synthetic scene
// generators, PLY ingest, proxy meshing, .kdt I/O, camera set-up and tree
// validation.  The SAH builder lives in sah_builder.cpp.
//
// Floating-point contract: this file is compiled with -ffp-contract=off so
// that every float expression is evaluated operation by operation, in the
// order the reference writes it (MSVC /fp:precise does not contract either).

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace kdgrt {

static thread_local std::string g_last_error;

kdgrt_status set_error(kdgrt_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}
void clear_error() { g_last_error.clear(); }

double wall_seconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// Random numbers: xoshiro256** (Blackman & Vigna 2018) seeded by SplitMix64,
// Box-Muller normals.  Only double-precision libm calls (log, sqrt, cos, exp)
// enter, so a scene is identical wherever the same glibc runs.
// ---------------------------------------------------------------------------
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        const double u1 = 1.0 - uniform(); // (0, 1]
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

static const float kShC0 = 0.28209479177387814f; // SH_C0, cudaRenderer.h:33

// Activation as loadGaussiansFromPly does it (Kd-treeConverterMain.cpp:1729-1747).
static inline float sigmoid_f(float x) { return 1.0f / (1.0f + expf(-x)); }

// fMyVecNormalize4D (Kd-treeConverterMain.cpp:1568-1577).
static inline void normalize4(float v[4]) {
    float len_sq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
    if (len_sq > 0.00001f) {
        float len_inv = 1.0f / sqrtf(len_sq);
        v[0] *= len_inv;
        v[1] *= len_inv;
        v[2] *= len_inv;
        v[3] *= len_inv;
    }
}

// calcKernelScale (Gaussian.h:6-16); the EBT render path never reads it, the
// value is kept for layout fidelity.
static inline float kernel_scale(float density) {
    const float min_resp = std::min(0.0113f / density, 0.97f);
    const float b = 4.0f;
    const float a = -4.5f / std::pow(3.0f, b);
    return std::pow(std::log(min_resp) / a, 1.0f / b);
}

static void finish_gaussian(kdgrt_gaussian& g, const float pos[3], const float log_scale[3], const float quat[4],
                            float opacity_logit, const float rgb[3]) {
    std::memset(&g, 0, sizeof(g));
    for (int k = 0; k < 3; ++k) g.pos[k] = pos[k];
    g.opacity = sigmoid_f(opacity_logit);
    for (int k = 0; k < 3; ++k) g.scale[k] = expf(log_scale[k]);
    for (int k = 0; k < 4; ++k) g.rot[k] = quat[k];
    normalize4(g.rot);
    g.k_scale = kernel_scale(g.opacity);
    for (int k = 0; k < 3; ++k) g.f_dc[k] = (rgb[k] - 0.5f) / kShC0;
}

} // namespace kdgrt

using namespace kdgrt;

extern "C" {

const char* kdgrt_last_error(void) { return g_last_error.c_str(); }

kdgrt_status kdgrt_generate_gaussians(int kind, uint64_t n, uint64_t seed, kdgrt_gaussian* out) {
    if (!out && n) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "generate: null output");
    if (kind != 0 && kind != 1) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "generate: kind must be 0 or 1");
    if (kind == 0) {
        // Uniform cube scene: mu ~ U[-1,1]^3, s = exp(N(ln 0.01, 0.5)),
        // uniform unit quaternion, o = sigmoid(N(0,1)), rgb ~ U[0,1].
        const double ln_s = std::log(0.01);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            Rng r(seed_for(seed, (uint64_t)i));
            float pos[3], ls[3], q[4], rgb[3];
            for (int k = 0; k < 3; ++k) pos[k] = (float)r.uniform(-1.0, 1.0);
            for (int k = 0; k < 3; ++k) ls[k] = (float)(ln_s + 0.5 * r.normal());
            for (int k = 0; k < 4; ++k) q[k] = (float)r.normal();
            float ol = (float)r.normal();
            for (int k = 0; k < 3; ++k) rgb[k] = (float)r.uniform();
            finish_gaussian(out[i], pos, ls, q, ol, rgb);
        }
        return KDGRT_OK;
    }
    // Clustered scene: 64 anisotropic clusters (centres U[-10,10]^3, per-axis
    // std log-U[0.2,2], random rotation) for 90% of points, 10% uniform
    // background in the same cube; s = exp(N(ln 0.02, 0.7)),
    // o = sigmoid(N(0,1.5)), rgb ~ U[0,1].
    const int kClusters = 64;
    double centre[kClusters][3], stdv[kClusters][3], rot[kClusters][9];
    {
        Rng rc(seed_for(seed, 0xC1C1C1C1C1C1ull));
        for (int c = 0; c < kClusters; ++c) {
            for (int k = 0; k < 3; ++k) centre[c][k] = rc.uniform(-10.0, 10.0);
            for (int k = 0; k < 3; ++k) stdv[c][k] = std::exp(rc.uniform(std::log(0.2), std::log(2.0)));
            double q[4];
            double nn = 0;
            for (int k = 0; k < 4; ++k) {
                q[k] = rc.normal();
                nn += q[k] * q[k];
            }
            nn = 1.0 / std::sqrt(nn);
            for (double& v : q) v *= nn;
            const double w = q[0], x = q[1], y = q[2], z = q[3];
            double* m = rot[c];
            m[0] = 1 - 2 * (y * y + z * z); m[1] = 2 * (x * y - w * z);     m[2] = 2 * (x * z + w * y);
            m[3] = 2 * (x * y + w * z);     m[4] = 1 - 2 * (x * x + z * z); m[5] = 2 * (y * z - w * x);
            m[6] = 2 * (x * z - w * y);     m[7] = 2 * (y * z + w * x);     m[8] = 1 - 2 * (x * x + y * y);
        }
    }
    const double ln_s = std::log(0.02);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        Rng r(seed_for(seed, (uint64_t)i));
        float pos[3], ls[3], q[4], rgb[3];
        if (r.uniform() < 0.1) {
            for (int k = 0; k < 3; ++k) pos[k] = (float)r.uniform(-10.0, 10.0);
        } else {
            int c = (int)(r.uniform() * kClusters);
            if (c >= kClusters) c = kClusters - 1;
            double z[3];
            for (int k = 0; k < 3; ++k) z[k] = stdv[c][k] * r.normal();
            const double* m = rot[c];
            for (int k = 0; k < 3; ++k)
                pos[k] = (float)(centre[c][k] + m[3 * k] * z[0] + m[3 * k + 1] * z[1] + m[3 * k + 2] * z[2]);
        }
        for (int k = 0; k < 3; ++k) ls[k] = (float)(ln_s + 0.7 * r.normal());
        for (int k = 0; k < 4; ++k) q[k] = (float)r.normal();
        float ol = (float)(1.5 * r.normal());
        for (int k = 0; k < 3; ++k) rgb[k] = (float)r.uniform();
        finish_gaussian(out[i], pos, ls, q, ol, rgb);
    }
    return KDGRT_OK;
}

kdgrt_status kdgrt_gaussians_invert_scales(kdgrt_gaussian* g, uint64_t n) {
    if (!g && n) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "invert_scales: null");
    // Kd-treeConverterMain.cpp:3636-3642: scale <- 1/scale, k_scale <- 1/k_scale (f32).
    for (uint64_t i = 0; i < n; ++i) {
        g[i].scale[0] = 1 / g[i].scale[0];
        g[i].scale[1] = 1 / g[i].scale[1];
        g[i].scale[2] = 1 / g[i].scale[2];
        g[i].k_scale = 1 / g[i].k_scale;
    }
    return KDGRT_OK;
}

// ---------------------------------------------------------------------------
// PLY ingest (Kd-treeConverterMain.cpp:1652-1769): properties located by
// name, opacity through a sigmoid, scales through exp, quaternion normalised.
// ---------------------------------------------------------------------------
struct PlyHeader {
    uint64_t n = 0;
    int stride = 0;
    std::map<std::string, int> off;
    std::map<std::string, std::string> type;
    std::streamoff data_start = 0;
};

static kdgrt_status parse_ply_header(std::ifstream& f, PlyHeader& h) {
    std::string line;
    if (!std::getline(f, line) || line.rfind("ply", 0) != 0)
        return set_error(KDGRT_ERR_FORMAT, "ply: missing magic");
    bool in_vertex = false, binary_le = false, seen_end = false;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line == "end_header") {
            seen_end = true;
            break;
        }
        std::stringstream ss(line);
        std::string tok;
        ss >> tok;
        if (tok == "format") {
            std::string fmt;
            ss >> fmt;
            binary_le = (fmt == "binary_little_endian");
        } else if (tok == "element") {
            std::string name;
            uint64_t cnt = 0;
            ss >> name >> cnt;
            in_vertex = (name == "vertex");
            if (in_vertex) h.n = cnt;
        } else if (tok == "property" && in_vertex) {
            std::string ty, name;
            ss >> ty >> name;
            int sz = 0;
            if (ty == "float" || ty == "float32" || ty == "int" || ty == "uint") sz = 4;
            else if (ty == "double" || ty == "float64") sz = 8;
            else if (ty == "uchar" || ty == "char" || ty == "uint8") sz = 1;
            else if (ty == "short" || ty == "ushort") sz = 2;
            else return set_error(KDGRT_ERR_FORMAT, "ply: unsupported property type " + ty);
            h.off[name] = h.stride;
            h.type[name] = ty;
            h.stride += sz;
        }
    }
    if (!seen_end) return set_error(KDGRT_ERR_FORMAT, "ply: no end_header");
    if (!binary_le) return set_error(KDGRT_ERR_FORMAT, "ply: only binary_little_endian is supported");
    for (const char* req : {"x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2",
                            "rot_3", "f_dc_0", "f_dc_1", "f_dc_2"}) {
        if (!h.off.count(req)) return set_error(KDGRT_ERR_FORMAT, std::string("ply: missing property ") + req);
        const std::string& ty = h.type[req];
        if (ty != "float" && ty != "float32")
            return set_error(KDGRT_ERR_FORMAT, std::string("ply: property must be float: ") + req);
    }
    h.data_start = f.tellg();
    return KDGRT_OK;
}

kdgrt_status kdgrt_ply_count(const char* path, uint64_t* n) {
    if (!path || !n) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "ply_count: null");
    std::ifstream f(path, std::ios::binary);
    if (!f) return set_error(KDGRT_ERR_IO, std::string("ply: cannot open ") + path);
    PlyHeader h;
    kdgrt_status st = parse_ply_header(f, h);
    if (st != KDGRT_OK) return st;
    *n = h.n;
    return KDGRT_OK;
}

kdgrt_status kdgrt_ply_read(const char* path, kdgrt_gaussian* out, uint64_t capacity, uint64_t* n) {
    if (!path || !n) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "ply_read: null");
    std::ifstream f(path, std::ios::binary);
    if (!f) return set_error(KDGRT_ERR_IO, std::string("ply: cannot open ") + path);
    PlyHeader h;
    kdgrt_status st = parse_ply_header(f, h);
    if (st != KDGRT_OK) return st;
    *n = h.n;
    if (!out) return KDGRT_OK;
    if (capacity < h.n) return set_error(KDGRT_ERR_CAPACITY, "ply_read: capacity too small");
    std::vector<char> buf(h.stride);
    auto getf = [&](const std::string& name) {
        float v;
        std::memcpy(&v, &buf[h.off[name]], 4);
        return v;
    };
    const bool has_rest = h.off.count("f_rest_0") && h.off.count("f_rest_44");
    for (uint64_t i = 0; i < h.n; ++i) {
        f.read(buf.data(), h.stride);
        if (!f) return set_error(KDGRT_ERR_IO, "ply: truncated vertex data at " + std::to_string(i));
        kdgrt_gaussian& g = out[i];
        std::memset(&g, 0, sizeof(g));
        g.pos[0] = getf("x");
        g.pos[1] = getf("y");
        g.pos[2] = getf("z");
        for (int k = 0; k < 3; ++k) g.f_dc[k] = getf("f_dc_" + std::to_string(k));
        if (has_rest)
            for (int k = 0; k < 45; ++k) g.f_rest[k] = getf("f_rest_" + std::to_string(k));
        g.opacity = sigmoid_f(getf("opacity"));
        for (int k = 0; k < 3; ++k) g.scale[k] = expf(getf("scale_" + std::to_string(k)));
        for (int k = 0; k < 4; ++k) g.rot[k] = getf("rot_" + std::to_string(k));
        normalize4(g.rot);
        g.k_scale = kernel_scale(g.opacity);
    }
    return KDGRT_OK;
}

// ---------------------------------------------------------------------------
// Camera (cudaRenderer.cu:3108-3127).  M_PI there is the float literal of
// cudaRenderer.h:31; fovy, aspect and near are doubles (OpenGLStuffs.h:27).
// ---------------------------------------------------------------------------
kdgrt_status kdgrt_camera_make(const kdgrt_camera* cam, int width, int height, kdgrt_camera_info* out) {
    if (!cam || !out) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "camera_make: null");
    if (width <= 0 || height <= 0) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "camera_make: bad size");
    const float kPi = 3.14159265358979323846f;
    float fov_rad = (float)(cam->fovy_deg * (double)(kPi / 180.0f));
    float plane_height = (float)(2.0f * cam->near_c * (double)tanf(fov_rad * 0.5f));
    float plane_width = (float)(plane_height * cam->aspect);
    out->step_x = plane_width / (float)width;
    out->step_y = plane_height / (float)height;
    const float near_f = (float)cam->near_c;
    const float hw = plane_width * 0.5f, hh = plane_height * 0.5f;
    for (int k = 0; k < 3; ++k) {
        out->eye[k] = cam->pos[k];
        out->u[k] = cam->u[k];
        out->v[k] = cam->v[k];
        float s = cam->pos[k] - cam->n[k] * near_f;
        s = s - cam->u[k] * hw;
        s = s + cam->v[k] * hh;
        out->start[k] = s;
    }
    return KDGRT_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Viewer camera controls (Kd-treeConverterMain.cpp:422-480, 1193-1296,
// 3463-3500), f32 host arithmetic as there.
// ---------------------------------------------------------------------------
namespace {

// fMyVecNormalize (MyMathUtility.cpp:77-90): float squares summed in float,
// the length in double; a vector shorter than 1e-8 becomes zero.
void vec_normalize(float* v) {
    const double len = std::sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    if (len < 0.00000001) {
        v[0] = v[1] = v[2] = 0.0f;
        return;
    }
    for (int i = 0; i < 3; ++i) v[i] = (float)((double)v[i] / len);
}

// fMyVecCrossProduct (MyMathUtility.cpp:59-63): out = a x b.
void vec_cross(const float* a, const float* b, float* out) {
    out[0] = (a[1] * b[2]) - (a[2] * b[1]);
    out[1] = -((a[0] * b[2]) - (a[2] * b[0]));
    out[2] = (a[0] * b[1]) - (a[1] * b[0]);
}

// n normalised, v = n x u, u = v x n, each normalised: the orthonormalisation
// that follows every rotation in the reference viewer.
void reorthonormalise(kdgrt_camera& c) {
    vec_normalize(c.n);
    vec_cross(c.n, c.u, c.v);
    vec_normalize(c.v);
    vec_cross(c.v, c.n, c.u);
    vec_normalize(c.u);
}

// rotateVectorAroundArbitraryAxis (Kd-treeConverterMain.cpp:422-436).
void rotate_about(float* vec, const float* axis, float cos_a, float sin_a) {
    const float cr[3] = {axis[1] * vec[2] - axis[2] * vec[1], axis[2] * vec[0] - axis[0] * vec[2],
                         axis[0] * vec[1] - axis[1] * vec[0]};
    const float dot = axis[0] * vec[0] + axis[1] * vec[1] + axis[2] * vec[2];
    for (int i = 0; i < 3; ++i) vec[i] = vec[i] * cos_a + cr[i] * sin_a + axis[i] * dot * (1.0f - cos_a);
}

// The matrix of glRotatef(angle_deg, x, y, z) as the OpenGL specification
// defines it (axis normalised, m[r][c]); a zero axis gives the identity.
void gl_rotation(float angle_deg, float x, float y, float z, float m[3][3]) {
    const float mag = std::sqrt(x * x + y * y + z * z);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r][c] = r == c ? 1.0f : 0.0f;
    if (!(mag > 0.0f)) return;
    x /= mag;
    y /= mag;
    z /= mag;
    const float a = angle_deg * (3.14159265358979323846f / 180.0f);
    const float c = std::cos(a), s = std::sin(a), oc = 1.0f - c;
    m[0][0] = x * x * oc + c;
    m[0][1] = x * y * oc - z * s;
    m[0][2] = x * z * oc + y * s;
    m[1][0] = y * x * oc + z * s;
    m[1][1] = y * y * oc + c;
    m[1][2] = y * z * oc - x * s;
    m[2][0] = x * z * oc - y * s;
    m[2][1] = y * z * oc + x * s;
    m[2][2] = z * z * oc + c;
}

// vec := m vec, the reference's R[0] x + R[4] y + R[8] z ... on the
// column-major GL matrix.
void apply(const float m[3][3], float* vec) {
    const float t[3] = {vec[0], vec[1], vec[2]};
    for (int r = 0; r < 3; ++r) vec[r] = m[r][0] * t[0] + m[r][1] * t[1] + m[r][2] * t[2];
}

} // namespace

extern "C" {

kdgrt_status kdgrt_camera_move(kdgrt_camera* cam, uint32_t keys, float speed, float dt, int* moved) {
    if (!cam) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "camera_move: null camera");
    const float s = speed * dt;
    struct Step {
        uint32_t key;
        const float* axis;
        float sign;
    };
    const Step steps[6] = {{KDGRT_KEY_W, cam->n, -1.0f}, {KDGRT_KEY_S, cam->n, 1.0f},
                           {KDGRT_KEY_A, cam->u, -1.0f}, {KDGRT_KEY_D, cam->u, 1.0f},
                           {KDGRT_KEY_Q, cam->v, 1.0f},  {KDGRT_KEY_E, cam->v, -1.0f}};
    int any = 0;
    for (const Step& st : steps) {
        if (!(keys & st.key)) continue;
        any = 1;
        for (int i = 0; i < 3; ++i) {
            const float d = st.axis[i] * s;
            cam->pos[i] = st.sign < 0.0f ? cam->pos[i] - d : cam->pos[i] + d;
        }
    }
    if (moved) *moved = any;
    return KDGRT_OK;
}

kdgrt_status kdgrt_camera_orbit(kdgrt_camera* cam, const float centre[3], const float axis[3], float angle) {
    if (!cam || !centre || !axis) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "camera_orbit: null argument");
    const float c = std::cos(angle), s = std::sin(angle);
    float p[3] = {cam->pos[0] - centre[0], cam->pos[1] - centre[1], cam->pos[2] - centre[2]};
    rotate_about(p, axis, c, s);
    for (int i = 0; i < 3; ++i) cam->pos[i] = p[i] + centre[i];
    rotate_about(cam->u, axis, c, s);
    rotate_about(cam->v, axis, c, s);
    rotate_about(cam->n, axis, c, s);
    reorthonormalise(*cam);
    return KDGRT_OK;
}

kdgrt_status kdgrt_camera_drag(kdgrt_camera* cam, int delx, int dely, float rot_speed, int zoom) {
    if (!cam) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "camera_drag: null camera");
    if (zoom) {
        // ZOOM_SENSITIVITY 0.25, MIN_FOVY 0.5, MAX_FOVY 100 (Kd-treeConverterMain.h:16-19)
        cam->fovy_deg += delx * 0.25;
        if (cam->fovy_deg < 0.5) cam->fovy_deg = 0.5;
        if (cam->fovy_deg > 100.0) cam->fovy_deg = 100.0;
        return KDGRT_OK;
    }
    const float yaw = delx * rot_speed;
    float pitch = dely * rot_speed;
    float m[3][3];
    gl_rotation(yaw, 0.0f, 1.0f, 0.0f, m);
    apply(m, cam->u);
    apply(m, cam->v);
    apply(m, cam->n);
    const float ny = cam->n[1];
    if ((pitch > 0.0f && ny > 0.98f) || (pitch < 0.0f && ny < -0.98f)) pitch = 0.0f;
    if (pitch != 0.0f) {
        gl_rotation(pitch, cam->u[0], cam->u[1], cam->u[2], m);
        apply(m, cam->v);
        apply(m, cam->n);
    }
    reorthonormalise(*cam);
    return KDGRT_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Proxy meshing (Kd-treeConverterMain.cpp:1860-2099, icosahedron
// Kd-treeConverterMain.h:128-153): each Gaussian with opacity >= 0.0113 and
// >= 1/255 becomes 20 triangles whose vertices are the 12 icosahedron
// vertices (-1, phi, 0)... scaled per axis by scale * k_iso, rotated by the
// quaternion and translated; k_iso = sqrt(2 ln(o / 0.0113)) * 0.5257311.
// ---------------------------------------------------------------------------
namespace kdgrt {

static const float kPhi = 1.618033988749895f;
static const float kIcoV[12][3] = {{-1, kPhi, 0}, {1, kPhi, 0},  {0, 1, -kPhi}, {-kPhi, 0, -1},
                                   {-kPhi, 0, 1}, {0, 1, kPhi},  {kPhi, 0, 1},  {0, -1, kPhi},
                                   {-1, -kPhi, 0}, {0, -1, -kPhi}, {kPhi, 0, -1}, {1, -kPhi, 0}};
static const int kIcoF[20][3] = {{0, 1, 2},  {0, 2, 3},  {0, 3, 4},  {0, 4, 5},  {0, 5, 1},
                                 {6, 1, 5},  {6, 5, 7},  {6, 7, 11}, {6, 11, 10}, {6, 10, 1},
                                 {8, 4, 3},  {8, 3, 9},  {8, 9, 11}, {8, 11, 7}, {8, 7, 4},
                                 {9, 3, 2},  {9, 2, 10}, {9, 10, 11}, {5, 4, 7}, {1, 10, 2}};
static const float kUnitSphereFactor = 0.5257311121191335703f;
static const float kKernelMinResponse = 0.0113f;

static inline bool gaussian_kept(float sigma) {
    return !(sigma < kKernelMinResponse || sigma < 1 / 255.0f);
}

// fMyVecCrossProduct (MyMathUtility.cpp:59-63), evaluated as written.
static inline void cross_ref(const float* a, const float* b, float* o) {
    o[0] = ((a[1] * b[2]) - (a[2] * b[1]));
    o[1] = -((a[0] * b[2]) - (a[2] * b[0]));
    o[2] = ((a[0] * b[1]) - (a[1] * b[0]));
}

// Writes the 20 triangles (9 floats each) of one Gaussian.
static void mesh_one(const kdgrt_gaussian& g, float* tri) {
    const float sigma = g.opacity;
    float k_iso = 0.0f;
    if (sigma / kKernelMinResponse > 1.0f) k_iso = sqrtf(2.0f * logf(sigma / kKernelMinResponse)) * kUnitSphereFactor;
    const float fs[3] = {g.scale[0] * k_iso, g.scale[1] * k_iso, g.scale[2] * k_iso};
    const float qv[3] = {g.rot[1], g.rot[2], g.rot[3]};
    float vx[12][3];
    for (int j = 0; j < 12; ++j) {
        float vs[3] = {kIcoV[j][0] * fs[0], kIcoV[j][1] * fs[1], kIcoV[j][2] * fs[2]};
        // rotate_vector_by_quaternion (Kd-treeConverterMain.cpp:1623-1649)
        float t[3], tt[3];
        cross_ref(qv, vs, t);
        for (int i = 0; i < 3; ++i) t[i] *= 2.0f;
        cross_ref(qv, t, tt);
        for (int i = 0; i < 3; ++i) {
            float r = vs[i] + g.rot[0] * t[i] + tt[i];
            vx[j][i] = r + g.pos[i];
        }
    }
    for (int f = 0; f < 20; ++f)
        for (int l = 0; l < 3; ++l)
            for (int i = 0; i < 3; ++i) tri[9 * f + 3 * l + i] = vx[kIcoF[f][l]][i];
}

// Meshes all kept Gaussians.  tri_xyz: 9 floats per triangle; aabb in
// .kdt order (min xyz, max xyz), reduced with fminf/fmaxf from +-FLT_MAX.
void proxy_mesh(const kdgrt_gaussian* g, uint64_t n, std::vector<float>& tri_xyz, std::vector<uint32_t>& tri_gid,
                float aabb[6], uint64_t* n_kept) {
    std::vector<uint64_t> base(n + 1, 0);
    for (uint64_t i = 0; i < n; ++i) base[i + 1] = base[i] + (gaussian_kept(g[i].opacity) ? 20 : 0);
    const uint64_t T = base[n];
    tri_xyz.assign(T * 9, 0.0f);
    tri_gid.assign(T, 0);
    float bmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
#pragma omp parallel
    {
        float lmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
#pragma omp for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            if (base[i + 1] == base[i]) continue;
            float* t = &tri_xyz[base[i] * 9];
            mesh_one(g[i], t);
            for (int f = 0; f < 20; ++f) tri_gid[base[i] + f] = (uint32_t)i;
            for (int v = 0; v < 60; ++v)
                for (int k = 0; k < 3; ++k) {
                    lmin[k] = fminf(lmin[k], t[3 * v + k]);
                    lmax[k] = fmaxf(lmax[k], t[3 * v + k]);
                }
        }
#pragma omp critical
        for (int k = 0; k < 3; ++k) {
            bmin[k] = fminf(bmin[k], lmin[k]);
            bmax[k] = fmaxf(bmax[k], lmax[k]);
        }
    }
    for (int k = 0; k < 3; ++k) {
        aabb[k] = bmin[k];
        aabb[3 + k] = bmax[k];
    }
    if (n_kept) *n_kept = T / 20;
}

// ---------------------------------------------------------------------------
// Tree validation and statistics.
// ---------------------------------------------------------------------------
kdgrt_status validate_tree(const kdgrt_node* nodes, uint32_t n_nodes, uint64_t n_offsets, const uint32_t* offsets,
                           uint64_t n_gaussians) {
    if (n_nodes == 0 || !nodes) return set_error(KDGRT_ERR_FORMAT, "tree: no nodes");
    if (n_nodes >= (1u << 29)) return set_error(KDGRT_ERR_FORMAT, "tree: too many nodes");
    for (uint32_t i = 0; i < n_nodes; ++i) {
        const kdgrt_node nd = nodes[i];
        if ((nd.x & 7u) == 3u) {
            uint64_t cnt = nd.x >> 3, off = nd.y;
            if (off + cnt > n_offsets)
                return set_error(KDGRT_ERR_FORMAT, "tree: leaf " + std::to_string(i) + " offset range out of bounds");
        } else {
            if ((nd.x & 7u) > 2u)
                return set_error(KDGRT_ERR_FORMAT, "tree: node " + std::to_string(i) + " has an invalid tag");
            uint64_t c = nd.x >> 3;
            // Children are allocated after their parent in the reference
            // builder (Kd-treeConstructor.cpp:1605-1611); requiring c > i also
            // guarantees every descent terminates.
            if (c <= i || c + 1 >= n_nodes)
                return set_error(KDGRT_ERR_FORMAT, "tree: node " + std::to_string(i) + " child index out of range");
            float sp;
            std::memcpy(&sp, &nd.y, 4);
            if (std::isnan(sp)) return set_error(KDGRT_ERR_FORMAT, "tree: NaN split at node " + std::to_string(i));
        }
    }
    if (n_gaussians && offsets) {
        for (uint64_t k = 0; k < n_offsets; ++k)
            if (offsets[k] >= n_gaussians)
                return set_error(KDGRT_ERR_FORMAT, "tree: leaf GID " + std::to_string(offsets[k]) + " >= " +
                                                       std::to_string(n_gaussians));
    }
    if (offsets) {
        // Each leaf lists a Gaussian at most once (the reference's leaf lists
        // are std::set contents, Kd-treeConstructor.cpp:1543-1586).
        std::vector<uint32_t> ids;
        for (uint32_t i = 0; i < n_nodes; ++i) {
            const kdgrt_node nd = nodes[i];
            if ((nd.x & 7u) != 3u) continue;
            const uint32_t cnt = nd.x >> 3;
            if (cnt < 2) continue;
            ids.assign(offsets + nd.y, offsets + nd.y + cnt);
            std::sort(ids.begin(), ids.end());
            if (std::adjacent_find(ids.begin(), ids.end()) != ids.end())
                return set_error(KDGRT_ERR_FORMAT, "tree: leaf " + std::to_string(i) + " lists a Gaussian twice");
        }
    }
    return KDGRT_OK;
}

bool compute_tree_stats(const kdgrt_tree& t, kdgrt_tree_stats& st, std::string* err) {
    const uint64_t nn = t.nodes.size();
    st.n_nodes = (uint32_t)nn;
    st.n_offsets = t.offsets.size();
    st.n_leaves = st.n_empty_leaves = st.max_depth = st.max_leaf_size = 0;
    if (nn == 0) return true;
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    stack.push_back({0, 0});
    uint64_t visited = 0;
    while (!stack.empty()) {
        auto [i, d] = stack.back();
        stack.pop_back();
        if (++visited > nn) {
            if (err) *err = "tree: cycle";
            return false;
        }
        st.max_depth = std::max(st.max_depth, d);
        const kdgrt_node nd = t.nodes[i];
        if ((nd.x & 7u) == 3u) {
            st.n_leaves++;
            uint32_t cnt = nd.x >> 3;
            if (cnt == 0) st.n_empty_leaves++;
            st.max_leaf_size = std::max(st.max_leaf_size, cnt);
        } else {
            uint64_t c = nd.x >> 3;
            if (c + 1 >= nn) {
                if (err) *err = "tree: child out of range";
                return false;
            }
            stack.push_back({(uint32_t)c + 1, d + 1});
            stack.push_back({(uint32_t)c, d + 1});
        }
    }
    return true;
}

} // namespace kdgrt

extern "C" {

kdgrt_status kdgrt_proxy_mesh(const kdgrt_gaussian* g, uint64_t n, float* tri_xyz, uint32_t* tri_gid,
                              uint64_t capacity, uint64_t* n_triangles, float aabb[6]) {
    if (!g && n) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "proxy_mesh: null gaussians");
    if (!n_triangles) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "proxy_mesh: null count");
    uint64_t T = 0;
    for (uint64_t i = 0; i < n; ++i) T += gaussian_kept(g[i].opacity) ? 20 : 0;
    *n_triangles = T;
    if (!tri_xyz && !tri_gid && !aabb) return KDGRT_OK;
    if ((tri_xyz || tri_gid) && capacity < T) return set_error(KDGRT_ERR_CAPACITY, "proxy_mesh: capacity");
    std::vector<float> xyz;
    std::vector<uint32_t> gid;
    float bb[6];
    proxy_mesh(g, n, xyz, gid, bb, nullptr);
    if (tri_xyz) std::memcpy(tri_xyz, xyz.data(), xyz.size() * sizeof(float));
    if (tri_gid) std::memcpy(tri_gid, gid.data(), gid.size() * sizeof(uint32_t));
    if (aabb) std::memcpy(aabb, bb, sizeof(bb));
    return KDGRT_OK;
}

// ---------------------------------------------------------------------------
// Trees: arrays, stats, files.  .kdt (Kd-treeConverter.cpp:320-342, 458-486):
// f32 version 0.1, f32 aabb[6], i32 n_nodes, n_nodes x 8 B, i32 n_offsets,
// n_offsets x u32, all little-endian.
// ---------------------------------------------------------------------------
kdgrt_status kdgrt_tree_from_arrays(const kdgrt_node* nodes, uint32_t n_nodes, const uint32_t* offsets,
                                    uint64_t n_offsets, const float aabb[6], kdgrt_tree** tree) {
    if (!tree || !aabb) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "tree_from_arrays: null");
    kdgrt_status st = validate_tree(nodes, n_nodes, n_offsets, offsets, 0);
    if (st != KDGRT_OK) return st;
    auto* t = new kdgrt_tree();
    t->nodes.assign(nodes, nodes + n_nodes);
    if (n_offsets) t->offsets.assign(offsets, offsets + n_offsets);
    std::memcpy(t->aabb, aabb, sizeof(t->aabb));
    std::string err;
    if (!compute_tree_stats(*t, t->stats, &err)) {
        delete t;
        return set_error(KDGRT_ERR_FORMAT, err);
    }
    *tree = t;
    return KDGRT_OK;
}

kdgrt_status kdgrt_tree_stats_get(const kdgrt_tree* tree, kdgrt_tree_stats* st) {
    if (!tree || !st) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "tree_stats: null");
    *st = tree->stats;
    return KDGRT_OK;
}
const kdgrt_node* kdgrt_tree_nodes(const kdgrt_tree* tree) { return tree ? tree->nodes.data() : nullptr; }
const uint32_t* kdgrt_tree_offsets(const kdgrt_tree* tree) { return tree ? tree->offsets.data() : nullptr; }
kdgrt_status kdgrt_tree_aabb(const kdgrt_tree* tree, float aabb[6]) {
    if (!tree || !aabb) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "tree_aabb: null");
    std::memcpy(aabb, tree->aabb, sizeof(tree->aabb));
    return KDGRT_OK;
}
kdgrt_status kdgrt_tree_validate(const kdgrt_tree* tree, uint64_t n_gaussians) {
    if (!tree) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "tree_validate: null");
    return validate_tree(tree->nodes.data(), (uint32_t)tree->nodes.size(), tree->offsets.size(),
                         tree->offsets.data(), n_gaussians);
}
kdgrt_status kdgrt_tree_free(kdgrt_tree* tree) {
    delete tree;
    return KDGRT_OK;
}

kdgrt_status kdgrt_kdt_write(const kdgrt_tree* tree, const char* path) {
    if (!tree || !path) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "kdt_write: null");
    FILE* fp = fopen(path, "wb");
    if (!fp) return set_error(KDGRT_ERR_IO, std::string("kdt_write: cannot open ") + path);
    const float version = 0.1f;
    const int32_t nn = (int32_t)tree->nodes.size();
    const int32_t no = (int32_t)tree->offsets.size();
    bool ok = fwrite(&version, 4, 1, fp) == 1 && fwrite(tree->aabb, 4, 6, fp) == 6 && fwrite(&nn, 4, 1, fp) == 1 &&
              fwrite(tree->nodes.data(), 8, nn, fp) == (size_t)nn && fwrite(&no, 4, 1, fp) == 1 &&
              (no == 0 || fwrite(tree->offsets.data(), 4, no, fp) == (size_t)no);
    ok = (fclose(fp) == 0) && ok;
    if (!ok) return set_error(KDGRT_ERR_IO, std::string("kdt_write: write failed ") + path);
    return KDGRT_OK;
}

kdgrt_status kdgrt_kdt_read(const char* path, kdgrt_tree** tree) {
    if (!tree || !path) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "kdt_read: null");
    FILE* fp = fopen(path, "rb");
    if (!fp) return set_error(KDGRT_ERR_IO, std::string("kdt_read: cannot open ") + path);
    fseek(fp, 0, SEEK_END);
    long size = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    auto fail = [&](const std::string& m) {
        fclose(fp);
        return set_error(KDGRT_ERR_FORMAT, "kdt_read: " + m + " (" + path + ")");
    };
    float version, aabb[6];
    int32_t nn = 0, no = 0;
    if (size < 32 || fread(&version, 4, 1, fp) != 1 || fread(aabb, 4, 6, fp) != 6 || fread(&nn, 4, 1, fp) != 1)
        return fail("truncated header");
    if (nn <= 0 || (int64_t)32 + (int64_t)nn * 8 + 4 > size) return fail("bad node count");
    std::vector<kdgrt_node> nodes(nn);
    if (fread(nodes.data(), 8, nn, fp) != (size_t)nn) return fail("truncated nodes");
    if (fread(&no, 4, 1, fp) != 1 || no < 0) return fail("bad offset count");
    if ((int64_t)32 + (int64_t)nn * 8 + 4 + (int64_t)no * 4 != size) return fail("size mismatch");
    std::vector<uint32_t> offs(no);
    if (no && fread(offs.data(), 4, no, fp) != (size_t)no) return fail("truncated offsets");
    fclose(fp);
    return kdgrt_tree_from_arrays(nodes.data(), (uint32_t)nn, offs.data(), (uint64_t)no, aabb, tree);
}

const char* kdgrt_version(void) { return "kdgrt-mi355x 0.1 (gfx950)"; }

} // extern "C"
