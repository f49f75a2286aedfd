// Drives the host side of the library (C ABI, include/kdgrt.h) under
// AddressSanitizer and UndefinedBehaviorSanitizer: synthetic scenes, proxy
// meshing, the OpenMP SAH builder at several thread counts, .kdt round trips
// and corrupt .kdt files, PLY / OBJ / NeRF-JSON parsers on valid, truncated
// and mutated files, the PNG and statistics writers.  Malformed inputs must
// come back as error codes; any memory error or undefined behaviour aborts
// the process (the sanitizers' halt-on-error settings).
//
// usage: sanitize_driver <scratch directory>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "kdgrt.h"

static int failures = 0;
#define EXPECT(cond)                                                                                     \
    do {                                                                                                 \
        if (!(cond)) {                                                                                   \
            std::fprintf(stderr, "EXPECT failed at line %d: %s (%s)\n", __LINE__, #cond, kdgrt_last_error()); \
            ++failures;                                                                                  \
        }                                                                                                \
    } while (0)

static std::string dir;

static std::string path(const char* name) { return dir + "/" + name; }

static void write_file(const std::string& p, const std::string& bytes) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) std::abort();
    std::fwrite(bytes.data(), 1, bytes.size(), f);
    std::fclose(f);
}

static std::string read_file(const std::string& p) {
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) return {};
    std::string s;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    std::fclose(f);
    return s;
}

// Every prefix length in `cuts` and a set of single-byte mutations of
// `good`, written to `p` and fed to `parse`; the result may be OK or an
// error, but the call must be memory-safe.
template <class F>
static void mangle(const std::string& p, const std::string& good, F&& parse, uint64_t seed) {
    std::mt19937_64 rng(seed);
    for (size_t cut : {size_t(0), size_t(1), good.size() / 7, good.size() / 3, good.size() / 2, good.size() - 1}) {
        write_file(p, good.substr(0, std::min(cut, good.size())));
        parse();
    }
    for (int k = 0; k < 60; ++k) {
        std::string bad = good;
        const int edits = 1 + (int)(rng() % 4);
        for (int e = 0; e < edits && !bad.empty(); ++e) {
            const size_t at = rng() % bad.size();
            const char repl[] = {'\0', '-', '9', 'e', ' ', '\n', '{', '[', '"', (char)0xFF};
            bad[at] = repl[rng() % sizeof repl];
        }
        write_file(p, bad);
        parse();
    }
}

static void scenes_and_builder() {
    const uint64_t n = 3000;
    std::vector<kdgrt_gaussian> g(n);
    EXPECT(kdgrt_generate_gaussians(1, n, 77, g.data()) == KDGRT_OK);
    EXPECT(kdgrt_generate_gaussians(9, n, 77, g.data()) != KDGRT_OK);
    uint64_t ntri = 0;
    float aabb[6];
    EXPECT(kdgrt_proxy_mesh(g.data(), n, nullptr, nullptr, 0, &ntri, nullptr) == KDGRT_OK);
    std::vector<float> xyz(9 * ntri);
    std::vector<uint32_t> gid(ntri);
    EXPECT(kdgrt_proxy_mesh(g.data(), n, xyz.data(), gid.data(), ntri, &ntri, aabb) == KDGRT_OK);
    kdgrt_build_params p;
    kdgrt_build_params_default(&p);
    kdgrt_tree* ref = nullptr;
    for (int threads : {1, 2, 8}) {
        p.threads = threads;
        kdgrt_tree* t = nullptr;
        EXPECT(kdgrt_tree_build(g.data(), n, &p, &t) == KDGRT_OK);
        EXPECT(kdgrt_tree_validate(t, n) == KDGRT_OK);
        kdgrt_tree_stats st;
        kdgrt_tree_stats_get(t, &st);
        if (!ref) {
            ref = t;
            continue;
        }
        kdgrt_tree_stats rs;
        kdgrt_tree_stats_get(ref, &rs);
        EXPECT(st.n_nodes == rs.n_nodes && st.n_offsets == rs.n_offsets);
        EXPECT(std::memcmp(kdgrt_tree_nodes(t), kdgrt_tree_nodes(ref), 8ull * st.n_nodes) == 0);
        kdgrt_tree_free(t);
    }
    // a dense, heavily clipped scene and a flat one
    for (auto& x : g) {
        for (int k = 0; k < 3; ++k) x.pos[k] *= 0.02f;
    }
    p.threads = 4;
    kdgrt_tree* dense = nullptr;
    EXPECT(kdgrt_tree_build(g.data(), 800, &p, &dense) == KDGRT_OK);
    for (auto& x : g) x.scale[2] = 0.0f;
    kdgrt_tree* flat = nullptr;
    EXPECT(kdgrt_tree_build(g.data(), 400, &p, &flat) == KDGRT_OK);
    kdgrt_tree_free(flat);
    kdgrt_tree_free(dense);
    EXPECT(kdgrt_gaussians_invert_scales(g.data(), n) == KDGRT_OK);

    // .kdt round trip, then truncated and mutated files
    const std::string kp = path("t.kdt");
    EXPECT(kdgrt_kdt_write(ref, kp.c_str()) == KDGRT_OK);
    kdgrt_tree* back = nullptr;
    EXPECT(kdgrt_kdt_read(kp.c_str(), &back) == KDGRT_OK);
    kdgrt_tree_stats a, b;
    kdgrt_tree_stats_get(ref, &a);
    kdgrt_tree_stats_get(back, &b);
    EXPECT(a.n_nodes == b.n_nodes && a.n_offsets == b.n_offsets);
    kdgrt_tree_free(back);
    const std::string good = read_file(kp);
    mangle(kp, good, [&] {
        kdgrt_tree* t = nullptr;
        if (kdgrt_kdt_read(kp.c_str(), &t) == KDGRT_OK) {
            kdgrt_tree_validate(t, n);
            kdgrt_tree_free(t);
        }
    }, 5);
    // lying node / offset counts in the header
    std::string lie = good;
    const int32_t huge = 0x7FFFFFF0;
    std::memcpy(&lie[28], &huge, 4);
    write_file(kp, lie);
    kdgrt_tree* t = nullptr;
    EXPECT(kdgrt_kdt_read(kp.c_str(), &t) != KDGRT_OK);
    EXPECT(kdgrt_kdt_read(path("missing.kdt").c_str(), &t) != KDGRT_OK);
    // trees from arrays that fail validation
    const kdgrt_node bad_nodes[3] = {{(5u << 3) | 0u, 0}, {(1u << 3) | 3u, 0}, {(1u << 3) | 3u, 1}};
    const uint32_t offs[2] = {0, 1};
    const float box[6] = {-1, -1, -1, 1, 1, 1};
    EXPECT(kdgrt_tree_from_arrays(bad_nodes, 3, offs, 2, box, &t) != KDGRT_OK);
    kdgrt_tree_free(ref);
}

static void ply_files() {
    const int n = 300;
    std::string hdr = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(n) + "\n";
    const char* names[] = {"x", "y", "z", "nx", "ny", "nz", "opacity", "scale_0", "scale_1", "scale_2",
                           "rot_0", "rot_1", "rot_2", "rot_3", "f_dc_0", "f_dc_1", "f_dc_2"};
    int np = 0;
    for (const char* nm : names) {
        hdr += std::string("property float ") + nm + "\n";
        ++np;
    }
    for (int i = 0; i < 45; ++i, ++np) hdr += "property float f_rest_" + std::to_string(i) + "\n";
    hdr += "end_header\n";
    std::mt19937 rng(3);
    std::normal_distribution<float> N(0.f, 1.f);
    std::string body(4ull * np * n, '\0');
    for (int i = 0; i < np * n; ++i) {
        const float v = N(rng);
        std::memcpy(&body[4ull * i], &v, 4);
    }
    const std::string good = hdr + body;
    const std::string pp = path("s.ply");
    write_file(pp, good);
    uint64_t cnt = 0;
    EXPECT(kdgrt_ply_count(pp.c_str(), &cnt) == KDGRT_OK && cnt == (uint64_t)n);
    std::vector<kdgrt_gaussian> g(n);
    EXPECT(kdgrt_ply_read(pp.c_str(), g.data(), n, &cnt) == KDGRT_OK);
    EXPECT(kdgrt_ply_read(pp.c_str(), g.data(), n / 2, &cnt) != KDGRT_OK || cnt <= (uint64_t)n / 2);
    mangle(pp, good, [&] {
        uint64_t c = 0;
        if (kdgrt_ply_count(pp.c_str(), &c) == KDGRT_OK && c <= (uint64_t)n) kdgrt_ply_read(pp.c_str(), g.data(), n, &c);
    }, 11);
    // header that declares more vertices than the file holds, and odd types
    std::string more = good;
    const size_t at = more.find(std::to_string(n));
    more.replace(at, std::to_string(n).size(), "100000");
    write_file(pp, more);
    EXPECT(kdgrt_ply_count(pp.c_str(), &cnt) != KDGRT_OK || kdgrt_ply_read(pp.c_str(), g.data(), n, &cnt) != KDGRT_OK);
    std::string dbl = good;
    dbl.replace(dbl.find("property float x"), 16, "property double x");
    write_file(pp, dbl);
    kdgrt_ply_count(pp.c_str(), &cnt);
    write_file(pp, "ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nend_header\n1\n2\n3\n");
    kdgrt_ply_count(pp.c_str(), &cnt);
}

static void obj_files() {
    const std::string good =
        "# mesh\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0 0 1\nvn 0 0 1\nvt 0 0\n"
        "f 1 2 3\nf 1/1 3/1 4/1\nf 1//1 2//1 5//1\nf 1 2 3 4\nf -1 -2 -3\ns off\ng group\nusemtl m\n";
    const std::string op = path("m.obj");
    write_file(op, good);
    uint64_t nt = 0;
    EXPECT(kdgrt_obj_read_secondary(op.c_str(), nullptr, 0, &nt) == KDGRT_OK && nt == 6);
    std::vector<float> xyz(9 * 64);
    EXPECT(kdgrt_obj_read_secondary(op.c_str(), xyz.data(), 64, &nt) == KDGRT_OK);
    EXPECT(kdgrt_obj_read_secondary(op.c_str(), xyz.data(), 2, &nt) == KDGRT_OK);  // capacity below the count
    mangle(op, good, [&] {
        uint64_t c = 0;
        kdgrt_obj_read_secondary(op.c_str(), xyz.data(), 64, &c);
    }, 17);
    write_file(op, "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0 0 1\nf 1 2 3 4 5\n");
    EXPECT(kdgrt_obj_read_secondary(op.c_str(), xyz.data(), 64, &nt) == KDGRT_ERR_FORMAT);
    write_file(op, "v 0 0 0\nf 1 2 3\nf 0 1 1\nf 99 1 1\n");
    kdgrt_obj_read_secondary(op.c_str(), xyz.data(), 64, &nt);
}

static void json_files() {
    std::string good = "{\"camera_angle_x\": 0.69, \"frames\": [";
    for (int i = 0; i < 5; ++i) {
        if (i) good += ",";
        good += "{\"file_path\": \"./r_" + std::to_string(i) +
                "\", \"rotation\": 0.01, \"transform_matrix\": [[1, 0, 0, 0.5], [0, 0.8, -0.6, 2e0], "
                "[0, 0.6, 0.8, -3.25E-1], [0, 0, 0, 1]]}";
    }
    good += "], \"extra\": {\"nested\": [1, [2, [3, {\"k\": \"v\\u00e9\\n\"}]]], \"b\": true, \"n\": null}}";
    const std::string jp = path("transforms.json");
    write_file(jp, good);
    kdgrt_camera cams[8];
    int count = 0;
    EXPECT(kdgrt_cameras_from_nerf_json(jp.c_str(), 1.0, cams, 8, &count) == KDGRT_OK && count == 5);
    EXPECT(kdgrt_cameras_from_nerf_json(jp.c_str(), 1.0, cams, 2, &count) == KDGRT_OK && count == 5);
    mangle(jp, good, [&] {
        int c = 0;
        kdgrt_cameras_from_nerf_json(jp.c_str(), 1.0, cams, 8, &c);
    }, 23);
    std::string deep(200000, '[');
    write_file(jp, "{\"frames\": " + deep);
    EXPECT(kdgrt_cameras_from_nerf_json(jp.c_str(), 1.0, cams, 8, &count) != KDGRT_OK);
}

static void writers() {
    std::vector<uint8_t> px(37 * 23 * 4);
    for (size_t i = 0; i < px.size(); ++i) px[i] = (uint8_t)(i * 7);
    EXPECT(kdgrt_png_write(path("a.png").c_str(), 37, 23, 4, px.data()) == KDGRT_OK);
    EXPECT(kdgrt_png_write(path("b.png").c_str(), 37, 23, 3, px.data()) == KDGRT_OK);
    EXPECT(kdgrt_png_write(path("c.png").c_str(), 0, 23, 3, px.data()) != KDGRT_OK);
    std::vector<int32_t> cnt(8 * 29 * 17);
    for (size_t i = 0; i < cnt.size(); ++i) cnt[i] = (int32_t)((i * 2654435761u) % 97);
    EXPECT(kdgrt_stats_write(cnt.data(), 29, 17, dir.c_str()) == KDGRT_OK);
    EXPECT(kdgrt_stats_write(cnt.data(), 29, 17, path("no/such/dir").c_str()) != KDGRT_OK);
    kdgrt_camera cam{};
    cam.u[0] = cam.v[1] = cam.n[2] = 1;
    cam.pos[2] = 3;
    cam.fovy_deg = 60;
    cam.aspect = 1.5;
    cam.near_c = 0.005;
    kdgrt_camera_info ci;
    EXPECT(kdgrt_camera_make(&cam, 300, 200, &ci) == KDGRT_OK);
    EXPECT(kdgrt_camera_make(&cam, -1, 200, &ci) != KDGRT_OK);
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    dir = argv[1];
    scenes_and_builder();
    ply_files();
    obj_files();
    json_files();
    writers();
    std::printf("sanitize driver done, %d failed expectation(s)\n", failures);
    return failures ? 1 : 0;
}
