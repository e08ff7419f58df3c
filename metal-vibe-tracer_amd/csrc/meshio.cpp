// meshio.cpp -- host readers and writers of the mesh formats that feed the renderer: the indexed mesh
// payload VTMESH01 (Sources/MeshGeometry.swift:61-92) and the OBJ reader (Sources/StudioModel.swift:
// 444-684, OBJMesh.parts).  Both produce what vt_scene_upload_indexed takes: vertices of 8 floats
// (position, normal, uv) and faces of 4 u32 (three vertex indices, subset).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "scene.h"
#include "vt_mesh.h"

namespace {

constexpr char kMagic[8] = {'V', 'T', 'M', 'E', 'S', 'H', '0', '1'};

uint32_t load_u32(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
void store_u32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

struct V3 {
    float x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float length2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
float length(V3 a) { return std::sqrt(length2(a)); }
V3 normalize(V3 a) {
    const float r = 1.0f / std::sqrt(length2(a));
    return {a.x * r, a.y * r, a.z * r};
}

// A part's welded geometry (MeshGeometry.Builder: corners with bit-identical attributes share a vertex).
struct Builder {
    struct Key {
        uint32_t w[8];
        bool operator==(const Key &o) const { return std::memcmp(w, o.w, sizeof w) == 0; }
    };
    struct Hash {
        size_t operator()(const Key &k) const {
            uint64_t h = 1469598103934665603ull;
            for (uint32_t v : k.w) h = (h ^ v) * 1099511628211ull;
            return (size_t)h;
        }
    };
    std::unordered_map<Key, uint32_t, Hash> index;
    std::vector<float> vertices;
    std::vector<uint32_t> faces;
    uint32_t vertex(const float v[8]) {
        Key k;
        std::memcpy(k.w, v, sizeof k.w);
        auto it = index.find(k);
        if (it != index.end()) return it->second;
        const uint32_t i = (uint32_t)(vertices.size() / 8);
        vertices.insert(vertices.end(), v, v + 8);
        index.emplace(k, i);
        return i;
    }
};

}  // namespace

struct vt_obj {
    struct Part {
        std::string object, group;
        std::vector<std::string> subsets;
        Builder geometry;
    };
    std::vector<Part> parts;
    std::vector<std::string> libraries;
    uint32_t skipped = 0;
};

namespace {

class ObjReader {
public:
    ObjReader(vt_obj &out, uint32_t node_budget, uint64_t triangle_limit)
        : out(out), budget(node_budget ? node_budget : UINT32_MAX), limit(triangle_limit ? triangle_limit : (1ull << 30)) {}

    void run(const uint8_t *bytes, size_t n) {
        // Physical lines end at LF, CR, CR LF, VT or FF; a trailing backslash outside a comment joins
        // the next physical line (with a space) to the statement, which keeps its first line's number.
        std::string pending;
        bool continuing = false;
        size_t physical = 0, start = 0;
        for (;;) {
            size_t end = start;
            while (end < n && bytes[end] != 0x0A && bytes[end] != 0x0D && bytes[end] != 0x0B && bytes[end] != 0x0C) ++end;
            ++physical;
            if (!continuing) line_number = physical;
            const uint8_t *raw = bytes + start;
            long last = (long)(end - start) - 1;
            while (last >= 0 && (raw[last] == 0x20 || raw[last] == 0x09)) --last;
            const bool joins = last >= 0 && raw[last] == 0x5C && std::memchr(raw, 0x23, (size_t)last) == nullptr;
            if (joins) {
                pending.append((const char *)raw, (size_t)last);
                pending.push_back(' ');
                continuing = true;
            } else if (!continuing) {
                statement(raw, end - start);
            } else {
                pending.append((const char *)raw, end - start);
                statement((const uint8_t *)pending.data(), pending.size());
                pending.clear();
                continuing = false;
            }
            if (end == n) break;
            start = end + (bytes[end] == 0x0D && end + 1 < n && bytes[end + 1] == 0x0A ? 2 : 1);
        }
        if (continuing) statement((const uint8_t *)pending.data(), pending.size());
        if (out.parts.empty()) {
            if (out.skipped)
                throw std::invalid_argument("OBJ contains no usable faces (" + std::to_string(out.skipped) +
                                            " degenerate faces skipped).");
            throw std::invalid_argument("OBJ contains no usable faces.");
        }
    }

private:
    struct Range {
        size_t a, b;
        size_t size() const { return b - a; }
    };
    vt_obj &out;
    uint32_t budget;
    uint64_t limit;
    size_t line_number = 0;
    std::vector<V3> positions, normals;
    std::vector<float> uvs;  // pairs
    std::string object = "Object", group = "Mesh", material = "Default";
    std::unordered_map<std::string, std::unordered_map<std::string, uint32_t>> part_index;
    std::vector<std::unordered_map<std::string, uint32_t>> subset_index;
    std::unordered_map<std::string, bool> objects;
    bool have_current = false;
    uint32_t cur_part = 0, cur_subset = 0;
    uint64_t total = 0;
    std::vector<Range> fields;

    [[noreturn]] void failure(const std::string &message) const {
        // The message's final period becomes " on line N.".
        throw std::invalid_argument(message.substr(0, message.size() - 1) + " on line " + std::to_string(line_number) + ".");
    }

    float number(const uint8_t *line, Range r) const {
        char s[64];
        bool ok = false;
        float v = 0.0f;
        if (r.size() < sizeof s) {
            std::memcpy(s, line + r.a, r.size());
            s[r.size()] = 0;
            char *end = nullptr;
            v = std::strtof(s, &end);
            const unsigned char c0 = (unsigned char)s[0];
            ok = end == s + r.size() && !(c0 == 0x20 || (c0 >= 0x09 && c0 <= 0x0D));
        }
        if (!ok || !std::isfinite(v) || !(std::fabs(v) < 1e8f)) failure("OBJ contains an invalid number.");
        return v;
    }

    size_t index(const uint8_t *line, Range r, size_t count) const {
        size_t i = r.a;
        bool negative = false;
        if (i < r.b && (line[i] == 0x2B || line[i] == 0x2D)) {
            negative = line[i] == 0x2D;
            ++i;
        }
        if (!(i < r.b) || r.b - i > 18) failure("Invalid OBJ index.");
        int64_t n = 0;
        for (; i < r.b; ++i) {
            const int d = (int)line[i] - 0x30;
            if (d < 0 || d > 9) failure("Invalid OBJ index.");
            n = n * 10 + d;
        }
        if (n == 0) failure("Invalid OBJ index.");
        const int64_t k = negative ? (int64_t)count - n : n - 1;
        if (k < 0 || k >= (int64_t)count) failure("OBJ index is out of bounds.");
        return (size_t)k;
    }

    std::string rest(const uint8_t *line) const {
        std::string s;
        for (size_t k = 1; k < fields.size(); ++k) {
            if (k > 1) s.push_back(' ');
            s.append((const char *)line + fields[k].a, fields[k].size());
        }
        return s;
    }

    bool keyword(const uint8_t *line, const char *word) const {
        const size_t len = std::strlen(word);
        return fields[0].size() == len && std::memcmp(line + fields[0].a, word, len) == 0;
    }

    void statement(const uint8_t *line, size_t n) {
        fields.clear();
        size_t i = 0;
        while (i < n && line[i] != 0x23) {
            if (line[i] == 0x20 || line[i] == 0x09) {
                ++i;
                continue;
            }
            const size_t a = i;
            while (i < n && line[i] != 0x20 && line[i] != 0x09 && line[i] != 0x23) ++i;
            fields.push_back({a, i});
        }
        if (fields.empty()) return;
        if (keyword(line, "o")) {
            object = rest(line);
            group = "Mesh";
            have_current = false;
        } else if (keyword(line, "g")) {
            group = rest(line);
            have_current = false;
        } else if (keyword(line, "usemtl")) {
            material = rest(line);
            have_current = false;
        } else if (keyword(line, "mtllib")) {
            if (fields.size() > 1) out.libraries.push_back(rest(line));
        } else if (keyword(line, "v") || keyword(line, "vn")) {
            if (fields.size() < 4) failure("Incomplete OBJ vertex.");
            const V3 v{number(line, fields[1]), number(line, fields[2]), number(line, fields[3])};
            if (fields[0].size() == 1) positions.push_back(v);
            else normals.push_back(length2(v) > 1e-12f ? normalize(v) : V3{0.0f, 1.0f, 0.0f});
        } else if (keyword(line, "vt")) {
            if (fields.size() < 2) failure("Incomplete OBJ UV.");
            const float u = number(line, fields[1]);
            const float v = fields.size() >= 3 ? number(line, fields[2]) : 0.0f;
            uvs.push_back(u);
            uvs.push_back(1.0f - v);
        } else if (keyword(line, "f")) {
            face(line);
        }
    }

    struct Corner {
        V3 p;
        float u, v;
        bool has_n;
        V3 n;
    };

    void face(const uint8_t *line) {
        if (fields.size() < 4 || fields.size() > 4097) failure("OBJ face must contain 3\xE2\x80\x93" "4096 vertices.");
        std::vector<Corner> corners;
        corners.reserve(fields.size() - 1);
        for (size_t k = 1; k < fields.size(); ++k) {
            const Range f = fields[k];
            size_t slash[3];
            int ns = 0;
            for (size_t q = f.a; q < f.b; ++q)
                if (line[q] == 0x2F) {
                    if (ns == 2) failure("Invalid OBJ face.");
                    slash[ns++] = q;
                }
            size_t ends[3];
            for (int q = 0; q < ns; ++q) ends[q] = slash[q];
            ends[ns] = f.b;
            Corner c{};
            c.p = positions[index(line, {f.a, ends[0]}, positions.size())];
            if (ns >= 1 && slash[0] + 1 < ends[1]) {
                const size_t t = index(line, {slash[0] + 1, ends[1]}, uvs.size() / 2);
                c.u = uvs[2 * t];
                c.v = uvs[2 * t + 1];
            }
            if (ns == 2 && slash[1] + 1 < ends[2]) {
                c.n = normals[index(line, {slash[1] + 1, ends[2]}, normals.size())];
                c.has_n = true;
            }
            corners.push_back(c);
        }
        // Convex polygons: fan triangulation.
        for (size_t i = 1; i + 1 < corners.size(); ++i) {
            const Corner &a = corners[0], &b = corners[i], &c = corners[i + 1];
            const V3 e1 = sub(b.p, a.p), e2 = sub(c.p, a.p), cr = cross(e1, e2);
            // Relative test: only collinear or coincident edges are skipped.
            const float area = length(cr);
            if (!(std::isfinite(area) && area > 1e-7f * length(e1) * length(e2))) {
                ++out.skipped;
                continue;
            }
            const V3 n = normalize(cr);
            if (!have_current) select_part();
            vt_obj::Part &part = out.parts[cur_part];
            uint32_t idx[3];
            const Corner *cs[3] = {&a, &b, &c};
            for (int q = 0; q < 3; ++q) {
                const V3 nn = cs[q]->has_n ? cs[q]->n : n;
                const float vert[8] = {cs[q]->p.x, cs[q]->p.y, cs[q]->p.z, nn.x, nn.y, nn.z, cs[q]->u, cs[q]->v};
                idx[q] = part.geometry.vertex(vert);
            }
            part.geometry.faces.insert(part.geometry.faces.end(), {idx[0], idx[1], idx[2], cur_subset});
            if (++total > limit)
                failure("OBJ exceeds the " + grouped(limit) + "-triangle import limit.");
        }
    }

    static std::string grouped(uint64_t v) {
        std::string d = std::to_string(v), s;
        for (size_t i = 0; i < d.size(); ++i) {
            if (i && (d.size() - i) % 3 == 0) s.push_back(',');
            s.push_back(d[i]);
        }
        return s;
    }

    void select_part() {
        auto &groups = part_index[object];
        auto it = groups.find(group);
        if (it != groups.end()) {
            cur_part = it->second;
        } else {
            const bool fresh = objects.find(object) == objects.end();
            if ((uint64_t)objects.size() + out.parts.size() + (fresh ? 2 : 1) > budget)
                failure("OBJ has too many objects and groups: the scene holds at most 65535 nodes (" +
                        std::to_string(budget) + " available). Merge groups before importing.");
            objects[object] = true;
            cur_part = (uint32_t)out.parts.size();
            out.parts.emplace_back();
            out.parts.back().object = object;
            out.parts.back().group = group;
            subset_index.emplace_back();
            groups[group] = cur_part;
        }
        auto &subs = subset_index[cur_part];
        auto s = subs.find(material);
        if (s != subs.end()) {
            cur_subset = s->second;
        } else {
            cur_subset = (uint32_t)out.parts[cur_part].subsets.size();
            out.parts[cur_part].subsets.push_back(material);
            subs[material] = cur_subset;
        }
        have_current = true;
    }
};

template <class F>
vt_status guarded(F &&f) {
    return vt::guarded_call(std::function<vt_status()>(std::forward<F>(f)));
}

}  // namespace

extern "C" {

uint64_t vt_mesh_payload_bytes(uint32_t vertex_count, uint32_t face_count) {
    return 16ull + 32ull * vertex_count + 16ull * face_count;
}

vt_status vt_mesh_payload_write(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t face_count,
                                uint8_t *out, uint64_t capacity) {
    return guarded([&]() -> vt_status {
        if (!out || (vertex_count && !vertices) || (face_count && !faces))
            return vt::fail(VT_ERROR_INVALID_ARGUMENT, "buffers must not be NULL");
        if (capacity < vt_mesh_payload_bytes(vertex_count, face_count))
            return vt::fail(VT_ERROR_INVALID_ARGUMENT, "the output is smaller than the payload");
        std::memcpy(out, kMagic, 8);
        store_u32(out + 8, vertex_count);
        store_u32(out + 12, face_count);
        uint8_t *p = out + 16;
        for (uint64_t i = 0; i < 8ull * vertex_count; ++i, p += 4) {
            uint32_t w;
            std::memcpy(&w, vertices + i, 4);
            store_u32(p, w);
        }
        for (uint64_t i = 0; i < 4ull * face_count; ++i, p += 4) store_u32(p, faces[i]);
        return VT_OK;
    });
}

vt_status vt_mesh_from_triangles(const float *triangles, uint32_t count, float **vertices, uint32_t *vertex_count,
                                 uint32_t **faces, uint32_t *face_count) {
    return guarded([&]() -> vt_status {
        if (!vertices || !vertex_count || !faces || !face_count || (count && !triangles))
            return vt::fail(VT_ERROR_INVALID_ARGUMENT, "arguments must not be NULL");
        Builder b;
        b.faces.reserve((size_t)count * 4);
        for (uint32_t i = 0; i < count; ++i) {
            // MeshTriangle: a, b, c, na, nb, nc (float4 each), uvab (uv of a, uv of b), uvc (uv of c, subset).
            const float *t = triangles + 32 * (size_t)i;
            const float *uvab = t + 24, *uvc = t + 28;
            const float corner[3][8] = {{t[0], t[1], t[2], t[12], t[13], t[14], uvab[0], uvab[1]},
                                        {t[4], t[5], t[6], t[16], t[17], t[18], uvab[2], uvab[3]},
                                        {t[8], t[9], t[10], t[20], t[21], t[22], uvc[0], uvc[1]}};
            const float z = uvc[2];
            const uint32_t subset = std::isfinite(z) && z >= 0.0f && z < 4e9f ? (uint32_t)z : 0u;
            for (int k = 0; k < 3; ++k) b.faces.push_back(b.vertex(corner[k]));
            b.faces.push_back(subset);
        }
        const size_t nv = b.vertices.size() / 8, nf = b.faces.size() / 4;
        float *v = (float *)std::malloc(nv ? 32 * nv : 1);
        uint32_t *f = (uint32_t *)std::malloc(nf ? 16 * nf : 1);
        if (!v || !f) {
            std::free(v);
            std::free(f);
            throw std::bad_alloc();
        }
        if (nv) std::memcpy(v, b.vertices.data(), 32 * nv);
        if (nf) std::memcpy(f, b.faces.data(), 16 * nf);
        *vertices = v;
        *faces = f;
        *vertex_count = (uint32_t)nv;
        *face_count = (uint32_t)nf;
        return VT_OK;
    });
}

vt_status vt_mesh_payload_read(const uint8_t *data, uint64_t length, uint32_t maximum_faces, float **vertices,
                               uint32_t *vertex_count, uint32_t **faces, uint32_t *face_count) {
    return guarded([&]() -> vt_status {
        if (!vertices || !vertex_count || !faces || !face_count || (length && !data))
            return vt::fail(VT_ERROR_INVALID_ARGUMENT, "arguments must not be NULL");
        if (length < 16 || std::memcmp(data, kMagic, 8) != 0)
            return vt::fail(VT_ERROR_INVALID_ARGUMENT, "Invalid mesh data: missing header.");
        const uint32_t nv = load_u32(data + 8), nf = load_u32(data + 12);
        if (nf > maximum_faces || (uint64_t)nv > 3ull * nf || length != vt_mesh_payload_bytes(nv, nf))
            return vt::fail(VT_ERROR_INVALID_ARGUMENT, "Invalid mesh data: counts do not match the payload.");
        float *v = (float *)std::malloc(nv ? 32ull * nv : 1);
        uint32_t *f = (uint32_t *)std::malloc(nf ? 16ull * nf : 1);
        if (!v || !f) {
            std::free(v);
            std::free(f);
            throw std::bad_alloc();
        }
        const uint8_t *p = data + 16;
        for (uint64_t i = 0; i < 8ull * nv; ++i, p += 4) {
            const uint32_t w = load_u32(p);
            std::memcpy(v + i, &w, 4);
        }
        for (uint64_t i = 0; i < 4ull * nf; ++i, p += 4) f[i] = load_u32(p);
        *vertices = v;
        *faces = f;
        *vertex_count = nv;
        *face_count = nf;
        return VT_OK;
    });
}

vt_status vt_obj_parse(const char *text, uint64_t length, uint32_t node_budget, uint64_t triangle_limit, vt_obj **out) {
    return guarded([&]() -> vt_status {
        if (!out || (length && !text)) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "arguments must not be NULL");
        *out = nullptr;
        auto obj = std::make_unique<vt_obj>();
        ObjReader reader(*obj, node_budget, triangle_limit);
        reader.run((const uint8_t *)text, (size_t)length);
        *out = obj.release();
        return VT_OK;
    });
}

void vt_obj_destroy(vt_obj *obj) { delete obj; }
uint32_t vt_obj_part_count(const vt_obj *obj) { return obj ? (uint32_t)obj->parts.size() : 0u; }
uint32_t vt_obj_skipped(const vt_obj *obj) { return obj ? obj->skipped : 0u; }
uint32_t vt_obj_library_count(const vt_obj *obj) { return obj ? (uint32_t)obj->libraries.size() : 0u; }
const char *vt_obj_library(const vt_obj *obj, uint32_t index) {
    return obj && index < obj->libraries.size() ? obj->libraries[index].c_str() : nullptr;
}

vt_status vt_obj_part(const vt_obj *obj, uint32_t part, const char **object, const char **group, uint32_t *subset_count,
                      const float **vertices, uint32_t *vertex_count, const uint32_t **faces, uint32_t *face_count) {
    if (!obj || part >= obj->parts.size()) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "no such OBJ part");
    const vt_obj::Part &p = obj->parts[part];
    if (object) *object = p.object.c_str();
    if (group) *group = p.group.c_str();
    if (subset_count) *subset_count = (uint32_t)p.subsets.size();
    if (vertices) *vertices = p.geometry.vertices.data();
    if (vertex_count) *vertex_count = (uint32_t)(p.geometry.vertices.size() / 8);
    if (faces) *faces = p.geometry.faces.data();
    if (face_count) *face_count = (uint32_t)(p.geometry.faces.size() / 4);
    return VT_OK;
}

const char *vt_obj_subset(const vt_obj *obj, uint32_t part, uint32_t subset) {
    if (!obj || part >= obj->parts.size() || subset >= obj->parts[part].subsets.size()) return nullptr;
    return obj->parts[part].subsets[subset].c_str();
}

}  // extern "C"
