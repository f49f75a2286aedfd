// Host products around the render path:
//  * NeRF transforms.json camera lists (loadCameraJson,
//    Kd-treeConverter_v0.9/Kd-treeConverter/LoadCamera.hpp:25-62);
//  * the per-ray statistics files of HIT_AND_NODE_COUNT_DEBUG renders
//    (cudaRenderer.cu:3173-3400): renderStats.txt, render_stats.csv,
//    heatmap_*.csv, isectCnt_kd.bin, blendCnt_kd.bin;
//  * the secondary mesh of the hybrid render from a Wavefront OBJ file
//    (loadSecondaryMesh, Kd-treeConverterMain.cpp:178-226);
//  * PNG files of rendered frames (the reference's saveDeviceFramebufferToPNG,
//    cudaRenderer.cu:3066-3096, writes them with stb_image_write).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

using kdgrt::set_error;

namespace {

// Minimal JSON reader: enough of RFC 8259 to walk objects and arrays and read
// numbers; strings are skipped over (escapes honoured) and compared raw.
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    std::string str;
    std::vector<Json> items;                              // array elements / object values
    std::vector<std::string> keys;                        // object keys
    const Json* get(const char* k) const {
        for (size_t i = 0; i < keys.size(); ++i)
            if (keys[i] == k) return &items[i];
        return nullptr;
    }
};

class Parser {
  public:
    Parser(const char* p, const char* e) : p_(p), e_(e) {}
    bool parse(Json& out) {
        if (!value(out, 0)) return false;
        ws();
        return p_ == e_;
    }
    std::string error;

  private:
    const char *p_, *e_;
    void ws() {
        while (p_ < e_ && std::isspace((unsigned char)*p_)) ++p_;
    }
    bool fail(const char* m) {
        if (error.empty()) error = m;
        return false;
    }
    bool lit(const char* s) {
        const size_t n = std::strlen(s);
        if ((size_t)(e_ - p_) < n || std::strncmp(p_, s, n) != 0) return fail("bad literal");
        p_ += n;
        return true;
    }
    bool string(std::string& s) {
        if (p_ >= e_ || *p_ != '"') return fail("expected string");
        ++p_;
        while (p_ < e_ && *p_ != '"') {
            if (*p_ == '\\') {
                if (++p_ >= e_) return fail("bad escape");
                if (*p_ == 'u') {
                    if (e_ - p_ < 5) return fail("bad escape");
                    s += '?';
                    p_ += 5;
                    continue;
                }
                const char c = *p_;
                s += c == 'n' ? '\n' : c == 't' ? '\t' : c == 'r' ? '\r' : c == 'b' ? '\b' : c == 'f' ? '\f' : c;
                ++p_;
                continue;
            }
            s += *p_++;
        }
        if (p_ >= e_) return fail("unterminated string");
        ++p_;
        return true;
    }
    bool value(Json& v, int depth) {
        if (depth > 64) return fail("nesting too deep");
        ws();
        if (p_ >= e_) return fail("unexpected end");
        const char c = *p_;
        if (c == '{') {
            v.kind = Json::Obj;
            ++p_;
            ws();
            if (p_ < e_ && *p_ == '}') {
                ++p_;
                return true;
            }
            while (true) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (p_ >= e_ || *p_ != ':') return fail("expected ':'");
                ++p_;
                v.keys.push_back(k);
                v.items.emplace_back();
                if (!value(v.items.back(), depth + 1)) return false;
                ws();
                if (p_ < e_ && *p_ == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < e_ && *p_ == '}') {
                    ++p_;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Json::Arr;
            ++p_;
            ws();
            if (p_ < e_ && *p_ == ']') {
                ++p_;
                return true;
            }
            while (true) {
                v.items.emplace_back();
                if (!value(v.items.back(), depth + 1)) return false;
                ws();
                if (p_ < e_ && *p_ == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < e_ && *p_ == ']') {
                    ++p_;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Json::Str;
            return string(v.str);
        }
        if (c == 't') {
            v.kind = Json::Bool;
            v.num = 1;
            return lit("true");
        }
        if (c == 'f') {
            v.kind = Json::Bool;
            return lit("false");
        }
        if (c == 'n') return lit("null");
        // number
        std::string tok;
        while (p_ < e_ && (std::isdigit((unsigned char)*p_) || *p_ == '-' || *p_ == '+' || *p_ == '.' || *p_ == 'e' ||
                           *p_ == 'E'))
            tok += *p_++;
        if (tok.empty()) return fail("unexpected character");
        char* end = nullptr;
        errno = 0;
        v.num = std::strtod(tok.c_str(), &end);
        if (end != tok.c_str() + tok.size() || errno == ERANGE) return fail("bad number");
        v.kind = Json::Num;
        return true;
    }
};

bool read_file(const char* path, std::string& out) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return false;
    std::fseek(fp, 0, SEEK_END);
    const long n = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    out.resize(n > 0 ? (size_t)n : 0);
    const bool ok = n <= 0 || std::fread(&out[0], 1, (size_t)n, fp) == (size_t)n;
    std::fclose(fp);
    return ok;
}

// One OBJ face vertex reference "v", "v/t", "v//n" or "v/t/n": the position
// index, 1-based, negative counting back from the last vertex read.
bool obj_vertex_index(const char* tok, size_t n_vertices, size_t& out) {
    char* end = nullptr;
    errno = 0;
    const long v = std::strtol(tok, &end, 10);
    if (end == tok || errno || v == 0) return false;
    const long long idx = v > 0 ? (long long)v - 1 : (long long)n_vertices + v;
    if (idx < 0 || (size_t)idx >= n_vertices) return false;
    out = (size_t)idx;
    return true;
}

// CRC-32 (ISO-HDLC, reflected polynomial 0xEDB88320) of PNG chunks.
uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0) {
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        ready = true;
    }
    c = ~c;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return ~c;
}

void put_be32(std::vector<uint8_t>& b, uint32_t v) {
    for (int k = 3; k >= 0; --k) b.push_back((uint8_t)(v >> (8 * k)));
}

void png_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put_be32(out, (uint32_t)data.size());
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    put_be32(out, crc32(td.data(), td.size()));
}

} // namespace

extern "C" {

kdgrt_status kdgrt_png_write(const char* path, int width, int height, int channels, const uint8_t* pixels) {
    if (!path || !pixels || width <= 0 || height <= 0 || (channels != 3 && channels != 4))
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "png_write: bad argument");
    const size_t row = (size_t)width * channels;
    // scanlines with filter type 0, as one zlib stream of stored blocks
    std::vector<uint8_t> raw;
    raw.reserve((row + 1) * height);
    for (int y = 0; y < height; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), pixels + (size_t)y * row, pixels + (size_t)(y + 1) * row);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + len >= raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(len & 0xFF));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xFF));
        z.push_back((uint8_t)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        if (last) break;
    }
    uint32_t a = 1, b = 0; // Adler-32
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    ihdr.push_back(8);                       // bit depth
    ihdr.push_back(channels == 4 ? 6 : 2);   // colour type RGBA / RGB
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return set_error(KDGRT_ERR_IO, std::string("png_write: cannot open ") + path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), fp) == out.size();
    std::fclose(fp);
    return ok ? KDGRT_OK : set_error(KDGRT_ERR_IO, std::string("png_write: short write to ") + path);
}

kdgrt_status kdgrt_obj_read_secondary(const char* path, float* out_xyz, uint64_t capacity, uint64_t* n_triangles) {
    if (!path || !n_triangles) return set_error(KDGRT_ERR_INVALID_ARGUMENT, "obj_read_secondary: null argument");
    std::string text;
    if (!read_file(path, text)) return set_error(KDGRT_ERR_IO, std::string("obj_read_secondary: cannot read ") + path);
    std::vector<float> vx;         // positions as read, 3 per vertex
    std::vector<size_t> tri;       // 3 vertex indices per triangle, faces in file order
    size_t line_no = 0, pos = 0;
    while (pos < text.size()) {
        size_t eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        std::string line = text.substr(pos, eol - pos);
        pos = eol + 1;
        ++line_no;
        for (char& c : line)
            if (c == '\r' || c == '\t') c = ' ';
        size_t b = line.find_first_not_of(' ');
        if (b == std::string::npos) continue;
        const char* p = line.c_str() + b;
        if (p[0] == 'v' && p[1] == ' ') {
            // "v x y z [w | r g b]": the first three numbers; missing ones are 0
            float xyz[3] = {0.0f, 0.0f, 0.0f};
            const char* q = p + 2;
            for (int k = 0; k < 3; ++k) {
                char* end = nullptr;
                const double val = std::strtod(q, &end);
                if (end == q) break;
                xyz[k] = (float)val;
                q = end;
            }
            vx.insert(vx.end(), xyz, xyz + 3);
        } else if (p[0] == 'f' && p[1] == ' ') {
            std::vector<size_t> face;
            const char* q = p + 2;
            while (*q) {
                while (*q == ' ') ++q;
                if (!*q) break;
                const char* tok = q;
                while (*q && *q != ' ') ++q;
                size_t idx;
                if (!obj_vertex_index(tok, vx.size() / 3, idx))
                    return set_error(KDGRT_ERR_FORMAT, "obj_read_secondary: bad vertex index on line " +
                                                           std::to_string(line_no));
                face.push_back(idx);
            }
            if (face.size() < 3) continue; // degenerate face, skipped as tinyobjloader does
            if (face.size() == 3) {
                tri.insert(tri.end(), face.begin(), face.end());
            } else if (face.size() == 4) {
                // quad: split along the shorter diagonal, (0,1,2)(0,2,3) when
                // |v2 - v0|^2 < |v3 - v1|^2, else (0,1,3)(1,2,3)
                const float* v0 = &vx[3 * face[0]];
                const float* v1 = &vx[3 * face[1]];
                const float* v2 = &vx[3 * face[2]];
                const float* v3 = &vx[3 * face[3]];
                const float a0 = v2[0] - v0[0], a1 = v2[1] - v0[1], a2 = v2[2] - v0[2];
                const float b0 = v3[0] - v1[0], b1 = v3[1] - v1[1], b2 = v3[2] - v1[2];
                const float s02 = a0 * a0 + a1 * a1 + a2 * a2, s13 = b0 * b0 + b1 * b1 + b2 * b2;
                const size_t order02[6] = {0, 1, 2, 0, 2, 3}, order13[6] = {0, 1, 3, 1, 2, 3};
                const size_t* o = s02 < s13 ? order02 : order13;
                for (int k = 0; k < 6; ++k) tri.push_back(face[o[k]]);
            } else {
                return set_error(KDGRT_ERR_FORMAT, "obj_read_secondary: face of " + std::to_string(face.size()) +
                                                       " vertices on line " + std::to_string(line_no) +
                                                       " (triangles and quads only)");
            }
        }
    }
    const uint64_t n = tri.size() / 3;
    if (out_xyz) {
        const uint64_t m = std::min<uint64_t>(n, capacity);
        for (uint64_t i = 0; i < 3 * m; ++i) {
            // stored as (x, -z, y)
            const float* v = &vx[3 * tri[i]];
            out_xyz[3 * i + 0] = v[0];
            out_xyz[3 * i + 1] = -v[2];
            out_xyz[3 * i + 2] = v[1];
        }
    }
    *n_triangles = n;
    return KDGRT_OK;
}

kdgrt_status kdgrt_cameras_from_nerf_json(const char* path, double aspect, kdgrt_camera* out, int capacity,
                                          int* count) {
    if (!path || !count || (capacity > 0 && !out) || capacity < 0)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "cameras_from_nerf_json: null argument");
    std::string text;
    if (!read_file(path, text)) return set_error(KDGRT_ERR_IO, std::string("cameras_from_nerf_json: cannot read ") + path);
    Json root;
    Parser ps(text.data(), text.data() + text.size());
    if (!ps.parse(root)) return set_error(KDGRT_ERR_FORMAT, "cameras_from_nerf_json: " + ps.error);
    const Json* frames = root.kind == Json::Obj ? root.get("frames") : nullptr;
    if (!frames || frames->kind != Json::Arr)
        return set_error(KDGRT_ERR_FORMAT, "cameras_from_nerf_json: no \"frames\" array");
    int n = 0;
    for (const Json& f : frames->items) {
        const Json* m = f.kind == Json::Obj ? f.get("transform_matrix") : nullptr;
        if (!m || m->kind != Json::Arr || m->items.size() < 3)
            return set_error(KDGRT_ERR_FORMAT, "cameras_from_nerf_json: frame " + std::to_string(n) +
                                                   " has no 3x4 transform_matrix");
        float a[3][4];
        for (int r = 0; r < 3; ++r) {
            const Json& row = m->items[r];
            if (row.kind != Json::Arr || row.items.size() < 4)
                return set_error(KDGRT_ERR_FORMAT, "cameras_from_nerf_json: short matrix row");
            for (int c = 0; c < 4; ++c) {
                if (row.items[c].kind != Json::Num)
                    return set_error(KDGRT_ERR_FORMAT, "cameras_from_nerf_json: non-numeric matrix entry");
                a[r][c] = (float)row.items[c].num;
            }
        }
        if (n < capacity) {
            // columns 0, 1, 2 are the u, v, n axes and column 3 the position;
            // fovy 39.6 degrees and near 0.005 are the loader's constants
            kdgrt_camera& cam = out[n];
            for (int k = 0; k < 3; ++k) {
                cam.u[k] = a[k][0];
                cam.v[k] = a[k][1];
                cam.n[k] = a[k][2];
                cam.pos[k] = a[k][3];
            }
            cam.fovy_deg = 39.6;
            cam.aspect = aspect;
            cam.near_c = 0.005;
        }
        ++n;
    }
    *count = n;
    return KDGRT_OK;
}

kdgrt_status kdgrt_stats_write(const int32_t* counters, int width, int height, const char* dir) {
    if (!counters || !dir || width <= 0 || height <= 0)
        return set_error(KDGRT_ERR_INVALID_ARGUMENT, "stats_write: bad argument");
    const size_t np = (size_t)width * height;
    std::string base(dir);
    if (!base.empty() && base.back() != '/') base += '/';
    // counters: 8 per ray in ray order y * width + x (the reference's debug
    // buffer order); 0 node visits, 1 leaf visits, 2 tests, 3 hits, 4 blends,
    // 5 max sort size
    int mx[6] = {0, 0, 0, 0, 0, 0};
    double tot[6] = {0, 0, 0, 0, 0, 0};
    uint64_t blending = 0;
    std::vector<uint64_t> hist_node(1024, 0), hist_leaf(256, 0);
    auto add = [](std::vector<uint64_t>& h, int v) {
        if (v < 0) return;
        if ((size_t)v >= h.size()) h.resize((size_t)v + 1, 0);
        h[v]++;
    };
    for (size_t i = 0; i < np; ++i) {
        const int32_t* c = counters + 8 * i;
        for (int k = 0; k < 6; ++k) mx[k] = std::max(mx[k], c[k]);
        if (c[4] != 0) { // averages and histograms over rays that blended
            for (int k = 0; k < 6; ++k) tot[k] += c[k];
            ++blending;
            add(hist_node, c[0]);
            add(hist_leaf, c[1]);
        }
    }
    auto open = [&](const char* name, const char* mode) { return std::fopen((base + name).c_str(), mode); };
    FILE* fp = open("renderStats.txt", "w");
    if (!fp) return set_error(KDGRT_ERR_IO, "stats_write: cannot write " + base + "renderStats.txt");
    const double b = (double)blending;
    const char* names[6] = {"node_visits\t\t", "leaf_visits\t\t", "intersection_tests\t", "hits_found\t\t",
                            "blend_ops\t\t", "max_sort_size\t"};
    std::fprintf(fp, "\n===========================================\n");
    for (int k = 0; k < 6; ++k) std::fprintf(fp, "avg_%s: %f\n", names[k], tot[k] / b);
    std::fprintf(fp, "===========================================\n");
    const char* mnames[6] = {"node_visits\t\t", "leaf_visits\t\t", "intersection_tests\t", "hits_found\t\t",
                             "blend_ops\t\t", "max_sort_size\t"};
    for (int k = 0; k < 6; ++k) std::fprintf(fp, "max_%s: %d\n", mnames[k], mx[k]);
    std::fprintf(fp, "===========================================\n");
    for (int k = 0; k < 6; ++k) std::fprintf(fp, "tot_%s: %f\n", names[k], tot[k]);
    std::fprintf(fp, "===========================================\n");
    std::fclose(fp);

    fp = open("render_stats.csv", "w");
    if (!fp) return set_error(KDGRT_ERR_IO, "stats_write: cannot write render_stats.csv");
    std::fprintf(fp, "Value,Node_Visits,Leaf_Visits\n");
    const size_t range = std::max(hist_node.size(), hist_leaf.size());
    for (size_t v = 0; v < range; ++v) {
        const uint64_t a = v < hist_node.size() ? hist_node[v] : 0, c = v < hist_leaf.size() ? hist_leaf[v] : 0;
        if (a || c) std::fprintf(fp, "%zu,%llu,%llu\n", v, (unsigned long long)a, (unsigned long long)c);
    }
    std::fclose(fp);

    const char* maps[6] = {"heatmap_node_visits.csv", "heatmap_leaf_visits.csv", "heatmap_intersection.csv",
                           "heatmap_hits_found.csv",  "heatmap_blend_ops.csv",   "heatmap_max_sort_size.csv"};
    for (int k = 0; k < 6; ++k) {
        fp = open(maps[k], "w");
        if (!fp) return set_error(KDGRT_ERR_IO, std::string("stats_write: cannot write ") + maps[k]);
        for (int y = 0; y < height; ++y) {
            for (int x = 0; x < width; ++x)
                std::fprintf(fp, x + 1 < width ? "%d," : "%d", counters[8 * ((size_t)y * width + x) + k]);
            std::fputc('\n', fp);
        }
        std::fclose(fp);
    }

    // isectCnt_kd.bin and blendCnt_kd.bin: one uint32 per ray; the reference
    // writes hits_found into both (cudaRenderer.cu:3309-3339)
    std::vector<uint32_t> hits(np);
    for (size_t i = 0; i < np; ++i) hits[i] = (uint32_t)counters[8 * i + 3];
    for (const char* name : {"isectCnt_kd.bin", "blendCnt_kd.bin"}) {
        fp = open(name, "wb");
        if (!fp) return set_error(KDGRT_ERR_IO, std::string("stats_write: cannot write ") + name);
        const bool ok = std::fwrite(hits.data(), 4, np, fp) == np;
        std::fclose(fp);
        if (!ok) return set_error(KDGRT_ERR_IO, std::string("stats_write: short write ") + name);
    }
    return KDGRT_OK;
}

} // extern "C"
