// rt_obj.cpp -- Wavefront OBJ input: the triangle array of one model, as
// GpuModel::getTriangles makes it (src/ray-tracing/RtScene.h:13-30) from
// Mesh::Mesh (src/scene/mesh.cpp:96-139).  The reference reads the file with
// tinyobjloader (the external/obj-loader submodule, not vendored, no pinned
// version) and de-duplicates (position, normal) vertices; that changes the
// index buffer only, so triangle k is the positions of face k's corners in
// file order.  Records read here:
//   v x y z [...]   a position (further numbers, w or colours, ignored)
//   f a b c [...]   a face; each corner "p", "p/t", "p//n" or "p/t/n"; p is
//                   1-based, or negative for "counted back from the last
//                   position read"; faces of k > 3 corners are split into
//                   the fan (0, i, i+1), i = 1..k-2 (tinyobjloader's
//                   triangulation of convex faces)
// Every other record (vt, vn, o, g, s, usemtl, mtllib, comments) is ignored.
// Numbers are parsed as double (strtod) and rounded to float.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rt_common.hpp"

namespace {

struct ObjError {
    std::string msg;
};

bool is_space(char c) { return c == ' ' || c == '\t'; }

const char* skip_space(const char* p, const char* end) {
    while (p < end && is_space(*p)) ++p;
    return p;
}

float parse_float(const char*& p, const char* end, size_t line) {
    p = skip_space(p, end);
    char buf[64];
    size_t n = 0;
    while (p + n < end && !is_space(p[n]) && n < sizeof buf - 1) {
        buf[n] = p[n];
        ++n;
    }
    buf[n] = 0;
    if (p + n < end && !is_space(p[n])) throw ObjError{"line " + std::to_string(line) + ": number too long"};
    char* stop = nullptr;
    errno = 0;
    const double v = std::strtod(buf, &stop);
    if (n == 0 || stop != buf + n) throw ObjError{"line " + std::to_string(line) + ": bad number"};
    p += n;
    return (float)v;
}

void parse(const char* data, size_t size, uint32_t material, std::vector<rt_triangle>& tris) {
    std::vector<float> pos;
    const char* p = data;
    const char* end = data + size;
    size_t line = 0;
    std::vector<int64_t> corners;
    while (p < end) {
        ++line;
        const char* eol = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!eol) eol = end;
        const char* q = skip_space(p, eol);
        const char* stop = eol;
        if (stop > q && stop[-1] == '\r') --stop;
        if (stop - q >= 2 && q[0] == 'v' && is_space(q[1])) {
            q += 2;
            for (int k = 0; k < 3; ++k) pos.push_back(parse_float(q, stop, line));
        } else if (stop - q >= 2 && q[0] == 'f' && is_space(q[1])) {
            q += 2;
            corners.clear();
            const int64_t nv = (int64_t)(pos.size() / 3);
            for (;;) {
                q = skip_space(q, stop);
                if (q >= stop) break;
                char* after = nullptr;
                const long long idx = std::strtoll(q, &after, 10);
                if (after == q) throw ObjError{"line " + std::to_string(line) + ": bad face corner"};
                int64_t k = idx > 0 ? idx - 1 : nv + idx;          // idx < 0: relative
                if (idx == 0 || k < 0 || k >= nv)
                    throw ObjError{"line " + std::to_string(line) + ": position index out of range"};
                corners.push_back(k);
                q = after;
                while (q < stop && !is_space(*q)) ++q;              // skip "/t/n"
            }
            if (corners.size() < 3) throw ObjError{"line " + std::to_string(line) + ": face with fewer than 3 corners"};
            for (size_t i = 1; i + 1 < corners.size(); ++i) {
                rt_triangle t;
                std::memset(&t, 0, sizeof t);
                const int64_t c[3] = {corners[0], corners[i], corners[i + 1]};
                float* dst[3] = {t.v0, t.v1, t.v2};
                for (int j = 0; j < 3; ++j)
                    for (int a = 0; a < 3; ++a) dst[j][a] = pos[3 * (size_t)c[j] + a];
                t.materialIndex = material;
                tris.push_back(t);
            }
        }
        p = eol < end ? eol + 1 : end;
    }
}

}  // namespace

namespace rtamd {

int obj_triangles(const char* path, uint32_t material, std::vector<rt_triangle>& tris, std::string& err) {
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open ") + path;
        return RT_ERR_NOT_FOUND;
    }
    std::string data;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) data.append(buf, got);
    std::fclose(f);
    try {
        parse(data.data(), data.size(), material, tris);
    } catch (const ObjError& e) {
        err = std::string(path) + ": " + e.msg;
        return RT_ERR_INVALID;
    } catch (const std::bad_alloc&) {
        return RT_ERR_NOMEM;
    }
    return RT_OK;
}

}  // namespace rtamd

namespace {
thread_local std::string g_obj_error;
}

extern "C" int rt_obj_load(const char* path, uint32_t material_index, rt_triangle** out, uint32_t* count) {
    if (!path || !out || !count) return RT_ERR_INVALID;
    *out = nullptr;
    *count = 0;
    std::vector<rt_triangle> tris;
    g_obj_error.clear();
    int rc = rtamd::obj_triangles(path, material_index, tris, g_obj_error);
    if (rc) return rc;
    if (tris.size() > 0xffffffffull) return RT_ERR_INVALID;
    rt_triangle* a = (rt_triangle*)std::malloc(std::max<size_t>(1, tris.size()) * sizeof(rt_triangle));
    if (!a) return RT_ERR_NOMEM;
    if (!tris.empty()) std::memcpy(a, tris.data(), tris.size() * sizeof(rt_triangle));
    *out = a;
    *count = (uint32_t)tris.size();
    return RT_OK;
}

extern "C" const char* rt_obj_last_error(void) { return g_obj_error.c_str(); }

extern "C" void rt_free(void* p) { std::free(p); }
