// workloads.cpp -- seeded synthetic scenes and ray sets (include/vt_workloads.h).
#include "vt_workloads.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <vector>

namespace {

// Seeded stream: the PCG-RXS-M-XS hash chained on its own output, as main.swift:91-101.
struct Rng {
    uint32_t s;
    static uint32_t hash(uint32_t x) {
        uint32_t state = x * 747796405u + 2891336453u;
        uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
        return (word >> 22u) ^ word;
    }
    explicit Rng(uint32_t seed) : s(hash(seed ^ 0x9e3779b9u)) {}
    double uniform() {  // [0, 1) with 32 bits
        s = hash(s);
        return s * (1.0 / 4294967296.0);
    }
    double normal() {  // Box-Muller
        double u1 = 1.0 - uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

struct V3 { double x, y, z; };
V3 normalize(V3 v) { double l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); return {v.x / l, v.y / l, v.z / l}; }

void icosphere(int subdivisions, std::vector<V3> &verts, std::vector<uint32_t> &tris) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    verts = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
             {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto &v : verts) v = normalize(v);
    tris = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
            3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int s = 0; s < subdivisions; ++s) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            V3 m = normalize({(verts[a].x + verts[b].x) / 2, (verts[a].y + verts[b].y) / 2, (verts[a].z + verts[b].z) / 2});
            uint32_t i = (uint32_t)verts.size();
            verts.push_back(m);
            mid[key] = i;
            return i;
        };
        std::vector<uint32_t> next;
        next.reserve(tris.size() * 4);
        for (size_t k = 0; k < tris.size(); k += 3) {
            uint32_t a = tris[k], b = tris[k + 1], c = tris[k + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            next.insert(next.end(), q, q + 12);
        }
        tris.swap(next);
    }
}

// Value noise on an integer lattice, smoothstep-interpolated; hashed lattice values in [-1, 1].
double lattice(int64_t x, int64_t z, uint32_t seed) {
    uint32_t h = Rng::hash((uint32_t)x * 0x8da6b343u ^ Rng::hash((uint32_t)z * 0xd8163841u ^ seed));
    return h * (2.0 / 4294967296.0) - 1.0;
}
double value_noise(double x, double z, uint32_t seed) {
    double fx = std::floor(x), fz = std::floor(z);
    int64_t ix = (int64_t)fx, iz = (int64_t)fz;
    double u = x - fx, v = z - fz;
    u = u * u * (3 - 2 * u);
    v = v * v * (3 - 2 * v);
    double a = lattice(ix, iz, seed), b = lattice(ix + 1, iz, seed), c = lattice(ix, iz + 1, seed), d = lattice(ix + 1, iz + 1, seed);
    return (a * (1 - u) + b * u) * (1 - v) + (c * (1 - u) + d * u) * v;
}

void parallel_rows(uint32_t rows, const std::function<void(uint32_t, uint32_t)> &f) {
    int T = (int)std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t)
        pool.emplace_back([&, t]() { f((uint32_t)((uint64_t)rows * t / T), (uint32_t)((uint64_t)rows * (t + 1) / T)); });
    for (auto &th : pool) th.join();
}
}  // namespace

extern "C" {

int vt_gen_icosphere_field(uint32_t spheres, uint32_t subdivisions, float mu, float sigma, float half_extent,
                           uint32_t seed, float *vertices, uint32_t *vertex_count, uint32_t *faces,
                           uint32_t *face_count) {
    if (subdivisions > 6) return -1;
    std::vector<V3> unit;
    std::vector<uint32_t> tris;
    icosphere((int)subdivisions, unit, tris);
    const uint32_t nv = (uint32_t)unit.size(), nf = (uint32_t)(tris.size() / 3);
    if (vertex_count) *vertex_count = nv * spheres;
    if (face_count) *face_count = nf * spheres;
    if (!vertices || !faces) return 0;
    Rng rng(seed);
    for (uint32_t s = 0; s < spheres; ++s) {
        double r = std::exp((double)mu + (double)sigma * rng.normal());
        double c[3];
        for (int a = 0; a < 3; ++a) c[a] = (2.0 * rng.uniform() - 1.0) * half_extent;
        for (uint32_t v = 0; v < nv; ++v) {
            float *o = vertices + 8 * ((size_t)s * nv + v);
            const V3 &u = unit[v];
            o[0] = (float)(c[0] + r * u.x);
            o[1] = (float)(c[1] + r * u.y);
            o[2] = (float)(c[2] + r * u.z);
            o[3] = (float)u.x; o[4] = (float)u.y; o[5] = (float)u.z;
            o[6] = (float)(0.5 + std::atan2(u.z, u.x) / (2 * M_PI));
            o[7] = (float)(std::acos(std::max(-1.0, std::min(1.0, u.y))) / M_PI);
        }
        for (uint32_t f = 0; f < nf; ++f) {
            uint32_t *o = faces + 4 * ((size_t)s * nf + f);
            o[0] = s * nv + tris[3 * f];
            o[1] = s * nv + tris[3 * f + 1];
            o[2] = s * nv + tris[3 * f + 2];
            o[3] = s;
        }
    }
    return 0;
}

int vt_gen_heightfield(uint32_t grid, uint32_t octaves, float gain, float half_extent, float height, uint32_t seed,
                       float *vertices, uint32_t *vertex_count, uint32_t *faces, uint32_t *face_count) {
    if (grid < 2) return -1;
    const uint64_t nv = (uint64_t)grid * grid, nf = 2ull * (grid - 1) * (grid - 1);
    if (nv > 0xffffffffull || nf > 0xffffffffull) return -1;
    if (vertex_count) *vertex_count = (uint32_t)nv;
    if (face_count) *face_count = (uint32_t)nf;
    if (!vertices || !faces) return 0;
    const double step = 2.0 * half_extent / (grid - 1);
    const double base = 1.0 / (0.2 * half_extent);  // lowest octave: ~5 features across the field
    auto h = [&](double x, double z) {
        double sum = 0, amp = 1, freq = base;
        for (uint32_t o = 0; o < octaves; ++o) {
            sum += amp * value_noise(x * freq, z * freq, seed + 977u * o);
            amp *= gain;
            freq *= 2.0;
        }
        return sum * height;
    };
    parallel_rows(grid, [&](uint32_t r0, uint32_t r1) {
        for (uint32_t j = r0; j < r1; ++j) {
            for (uint32_t i = 0; i < grid; ++i) {
                double x = -half_extent + i * step, z = -half_extent + j * step;
                double y = h(x, z);
                double gx = (h(x + 0.5 * step, z) - h(x - 0.5 * step, z)) / step;
                double gz = (h(x, z + 0.5 * step) - h(x, z - 0.5 * step)) / step;
                V3 n = normalize({-gx, 1.0, -gz});
                float *o = vertices + 8 * ((size_t)j * grid + i);
                o[0] = (float)x; o[1] = (float)y; o[2] = (float)z;
                o[3] = (float)n.x; o[4] = (float)n.y; o[5] = (float)n.z;
                o[6] = (float)i / (grid - 1); o[7] = (float)j / (grid - 1);
            }
        }
    });
    parallel_rows(grid - 1, [&](uint32_t r0, uint32_t r1) {
        for (uint32_t j = r0; j < r1; ++j) {
            for (uint32_t i = 0; i < grid - 1; ++i) {
                uint32_t a = j * grid + i, b = a + 1, c = a + grid, d = c + 1;
                uint32_t *o = faces + 8 * ((size_t)j * (grid - 1) + i);
                o[0] = a; o[1] = c; o[2] = d; o[3] = 0;
                o[4] = a; o[5] = d; o[6] = b; o[7] = 0;
            }
        }
    });
    return 0;
}

int vt_gen_cornell(float light_size, float *vertices, uint32_t *vertex_count, uint32_t *faces, uint32_t *face_count) {
    struct P { float x, y, z; };
    std::vector<P> tri;
    std::vector<uint32_t> tags;
    auto quad = [&](P a, P b, P c, P d, uint32_t tag) {  // a b c d around the quad
        tri.insert(tri.end(), {a, b, c, a, c, d});
        tags.insert(tags.end(), {tag, tag});
    };
    // Walls of the [-1, 1]^3 box, open towards -z (main.swift:1428-1432).
    quad({-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1}, 0);   // floor
    quad({-1, 1, -1}, {-1, 1, 1}, {1, 1, 1}, {1, 1, -1}, 0);       // ceiling
    quad({1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1}, 0);       // back
    quad({-1, 1, 1}, {-1, 1, -1}, {-1, -1, -1}, {-1, -1, 1}, 1);   // left, red
    quad({1, 1, -1}, {1, 1, 1}, {1, -1, 1}, {1, -1, -1}, 2);       // right, green
    // Boxes rotated by yaw about y: world = centre + R(yaw) local (main.swift:1214-1219 inverted).
    auto box = [&](P c, P h, float yaw) {
        double cs = std::cos((double)yaw), sn = std::sin((double)yaw);
        auto w = [&](float sx, float sy, float sz) {
            double lx = sx * h.x, ly = sy * h.y, lz = sz * h.z;
            return P{(float)(c.x + cs * lx - sn * lz), (float)(c.y + ly), (float)(c.z + sn * lx + cs * lz)};
        };
        quad(w(-1, -1, -1), w(1, -1, -1), w(1, 1, -1), w(-1, 1, -1), 0);
        quad(w(1, -1, 1), w(-1, -1, 1), w(-1, 1, 1), w(1, 1, 1), 0);
        quad(w(-1, -1, 1), w(-1, -1, -1), w(-1, 1, -1), w(-1, 1, 1), 0);
        quad(w(1, -1, -1), w(1, -1, 1), w(1, 1, 1), w(1, 1, -1), 0);
        quad(w(-1, 1, -1), w(1, 1, -1), w(1, 1, 1), w(-1, 1, 1), 0);
        quad(w(-1, -1, 1), w(1, -1, 1), w(1, -1, -1), w(-1, -1, -1), 0);
    };
    box({-0.33f, -0.4f, 0.35f}, {0.3f, 0.6f, 0.3f}, 0.32f);
    box({0.33f, -0.7f, -0.25f}, {0.3f, 0.3f, 0.3f}, -0.30f);
    // Ceiling light: geometric normal cross(b - a, c - a) points down (-y), so it emits downwards.
    const float h = 0.5f * light_size, y = 0.999f;
    tri.insert(tri.end(), {P{-h, y, -h}, P{h, y, h}, P{-h, y, h}, P{-h, y, -h}, P{h, y, -h}, P{h, y, h}});
    tags.insert(tags.end(), {3u, 3u});
    const uint32_t nf = (uint32_t)tags.size();
    if (vertex_count) *vertex_count = 3 * nf;
    if (face_count) *face_count = nf;
    if (!vertices || !faces) return 0;
    for (uint32_t f = 0; f < nf; ++f) {
        P a = tri[3 * f], b = tri[3 * f + 1], c = tri[3 * f + 2];
        double e1[3] = {(double)b.x - a.x, (double)b.y - a.y, (double)b.z - a.z};
        double e2[3] = {(double)c.x - a.x, (double)c.y - a.y, (double)c.z - a.z};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) {
            P p = tri[3 * f + k];
            float *o = vertices + 8 * (3 * (size_t)f + k);
            o[0] = p.x; o[1] = p.y; o[2] = p.z;
            o[3] = (float)(n[0] / l); o[4] = (float)(n[1] / l); o[5] = (float)(n[2] / l);
            o[6] = 0.0f; o[7] = 0.0f;
        }
        faces[4 * f] = 3 * f;
        faces[4 * f + 1] = 3 * f + 1;
        faces[4 * f + 2] = 3 * f + 2;
        faces[4 * f + 3] = tags[f];
    }
    return 0;
}

void vt_gen_camera_rays(const float eye[3], const float target[3], float fov_degrees, uint32_t width, uint32_t height,
                        uint32_t seed, float *ox, float *oy, float *oz, float *dx, float *dy, float *dz, float *tmin,
                        float *tmax) {
    double f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double &v : f) v /= fl;
    double up[3] = {0, 1, 0};
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double &v : r) v /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    const double th = std::tan(0.5 * fov_degrees * M_PI / 180.0), aspect = (double)width / height;
    parallel_rows(height, [&](uint32_t r0, uint32_t r1) {
        for (uint32_t y = r0; y < r1; ++y) {
            for (uint32_t x = 0; x < width; ++x) {
                size_t i = (size_t)y * width + x;
                Rng rng(seed * 0x9e3779b1u ^ (uint32_t)i);
                double sx = (2.0 * (x + rng.uniform()) / width - 1.0) * th * aspect;
                double sy = (1.0 - 2.0 * (y + rng.uniform()) / height) * th;
                double d[3];
                for (int a = 0; a < 3; ++a) d[a] = f[a] + sx * r[a] + sy * u[a];
                double l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                ox[i] = eye[0]; oy[i] = eye[1]; oz[i] = eye[2];
                dx[i] = (float)(d[0] / l); dy[i] = (float)(d[1] / l); dz[i] = (float)(d[2] / l);
                tmin[i] = 0.0f;
                tmax[i] = 1e20f;
            }
        }
    });
}

uint32_t vt_gen_bounce_rays(const float *vertices, const uint32_t *faces, uint32_t n, const float *in_dx,
                            const float *in_dy, const float *in_dz, const int32_t *hit_face, const float *hit_b1,
                            const float *hit_b2, uint32_t seed, float *ox, float *oy, float *oz, float *dx, float *dy,
                            float *dz, float *tmin, float *tmax) {
    std::vector<uint32_t> hits;
    hits.reserve(n);
    for (uint32_t i = 0; i < n; ++i)
        if (hit_face[i] >= 0) hits.push_back(i);
    if (hits.empty()) return 0;
    parallel_rows(n, [&](uint32_t r0, uint32_t r1) {
        for (uint32_t i = r0; i < r1; ++i) {
            Rng rng(seed * 0x85ebca6bu ^ i);
            uint32_t src = hit_face[i] >= 0 ? i : hits[(size_t)(rng.uniform() * hits.size())];
            const uint32_t *fc = faces + 4 * (size_t)hit_face[src];
            const float *a = vertices + 8 * (size_t)fc[0], *b = vertices + 8 * (size_t)fc[1], *c = vertices + 8 * (size_t)fc[2];
            // main.swift:2010: the hit point rebuilt from barycentrics.
            float b1 = hit_b1[src], b2 = hit_b2[src], b0 = 1.0f - b1 - b2;
            float p[3], e1[3], e2[3];
            for (int q = 0; q < 3; ++q) {
                p[q] = b0 * a[q] + b1 * b[q] + b2 * c[q];
                e1[q] = b[q] - a[q];
                e2[q] = c[q] - a[q];
            }
            double ng[3] = {(double)e1[1] * e2[2] - (double)e1[2] * e2[1], (double)e1[2] * e2[0] - (double)e1[0] * e2[2],
                            (double)e1[0] * e2[1] - (double)e1[1] * e2[0]};
            double l = std::sqrt(ng[0] * ng[0] + ng[1] * ng[1] + ng[2] * ng[2]);
            if (l > 0) for (double &v : ng) v /= l;
            else { ng[0] = 0; ng[1] = 1; ng[2] = 0; }
            double din[3] = {in_dx[src], in_dy[src], in_dz[src]};
            if (ng[0] * din[0] + ng[1] * din[1] + ng[2] * din[2] > 0) for (double &v : ng) v = -v;
            // main.swift:2462-2476 sample_cosine_hemisphere about the facing geometric normal.
            double phi = 2 * M_PI * rng.uniform();
            double ct = std::sqrt(rng.uniform() + 0.5 / 16777216.0), st = std::sqrt(std::max(0.0, 1 - ct * ct));
            double up[3] = {0, 0, 1};
            if (std::fabs(ng[2]) >= 0.999) { up[0] = 1; up[2] = 0; }
            double uu[3] = {up[1] * ng[2] - up[2] * ng[1], up[2] * ng[0] - up[0] * ng[2], up[0] * ng[1] - up[1] * ng[0]};
            double ul = std::sqrt(uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]);
            for (double &v : uu) v /= ul;
            double vv[3] = {ng[1] * uu[2] - ng[2] * uu[1], ng[2] * uu[0] - ng[0] * uu[2], ng[0] * uu[1] - ng[1] * uu[0]};
            double d[3];
            for (int q = 0; q < 3; ++q) d[q] = uu[q] * std::cos(phi) * st + vv[q] * std::sin(phi) * st + ng[q] * ct;
            double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            // main.swift:1343-1369 for imported meshes: ray_epsilon, mesh_hit_error, ray_origin, ray_t_min.
            auto max_abs = [](const float *v) { return std::max(std::fabs(v[0]), std::max(std::fabs(v[1]), std::fabs(v[2]))); };
            float eps = std::max(1e-7f, 9.536743e-7f * max_abs(p));
            float err = std::max(1e-7f, 1.907349e-6f * std::max(max_abs(a), std::max(max_abs(b), max_abs(c))));
            float off = std::max(err, eps);
            float o[3];
            for (int q = 0; q < 3; ++q) o[q] = p[q] + (float)ng[q] * off;
            ox[i] = o[0]; oy[i] = o[1]; oz[i] = o[2];
            dx[i] = (float)(d[0] / dl); dy[i] = (float)(d[1] / dl); dz[i] = (float)(d[2] / dl);
            tmin[i] = std::max(1e-7f, 9.536743e-7f * max_abs(o)) * 0.25f;
            tmax[i] = 1e20f;
        }
    });
    return (uint32_t)hits.size();
}

void vt_gen_random_rays(const float lo[3], const float hi[3], uint32_t n, uint32_t seed, float *ox, float *oy,
                        float *oz, float *dx, float *dy, float *dz, float *tmin, float *tmax) {
    parallel_rows(n, [&](uint32_t r0, uint32_t r1) {
        for (uint32_t i = r0; i < r1; ++i) {
            Rng rng(seed * 0xc2b2ae35u ^ i);
            ox[i] = (float)(lo[0] + (hi[0] - lo[0]) * rng.uniform());
            oy[i] = (float)(lo[1] + (hi[1] - lo[1]) * rng.uniform());
            oz[i] = (float)(lo[2] + (hi[2] - lo[2]) * rng.uniform());
            double z = 2 * rng.uniform() - 1, phi = 2 * M_PI * rng.uniform(), s = std::sqrt(std::max(0.0, 1 - z * z));
            dx[i] = (float)(s * std::cos(phi)); dy[i] = (float)(s * std::sin(phi)); dz[i] = (float)z;
            tmin[i] = 0.0f;
            tmax[i] = 1e20f;
        }
    });
}

}  // extern "C"
