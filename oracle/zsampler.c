/* Oracle (test infrastructure only): the Z++ sampler restated in C from main.swift:122-305 and
 * 497-515 (z_mix :132, z_shuffle :148, z_owen :162, z_owen_seed :171, z_sobol1 :178, z_o2m3_* :185-192,
 * z_float :193, z_morton :199, Sampler's Z draws :221-275, sampler_event / sampler_candidate :288-293,
 * z_pixel_key :497).  The structure is pinned by the reference's own properties (tests/
 * Fix_zsampling.swift: nets of every constituent, the shuffle's aligned-block bijection, per-pixel
 * temporal nets) and the second Sobol' dimension by the Pascal-matrix closed form; the scramble
 * constants are the reference's and only choose which randomisation is drawn. */
#include <stddef.h>
#include <stdint.h>

static uint32_t rev(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 32; ++b) r |= ((v >> b) & 1u) << (31 - b);
    return r;
}
uint32_t vo_z_mix(uint32_t v) {
    v ^= v >> 16; v *= 0x7feb352du; v ^= v >> 15; v *= 0x846ca68bu; v ^= v >> 16;
    return v;
}
uint32_t vo_z_shuffle(uint32_t index, uint32_t dimension) {
    const uint64_t perms = 0x6c9c78d8b4e4ull;
    uint32_t seed = vo_z_mix(dimension * 0x9e3779b9u + 0x632be5abu);
    uint32_t out = index & ~((1u << 24) - 1u);
    for (int d = 11; d >= 0; --d) {
        uint32_t shift = 2u * (uint32_t)d, digit = (index >> shift) & 3u;
        uint32_t h = vo_z_mix((index >> (shift + 2u)) ^ seed);
        uint32_t sigma = ((h >> 16) * 6u) >> 16;
        uint32_t image = (uint32_t)(perms >> (8u * sigma + 2u * digit)) & 3u;
        out |= (image ^ (h & 3u)) << shift;
    }
    return out;
}
uint32_t vo_z_owen(uint32_t v, uint32_t seed) {
    v = rev(v);
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return rev(v);
}
uint32_t vo_z_owen_seed(uint32_t dimension, uint32_t axis) {
    uint32_t h = (dimension ^ (axis << 24)) * 0x9e3779b9u;
    return (h ^ (h >> 16)) * 0x85ebca6bu;
}
uint32_t vo_z_sobol1(uint32_t i) {
    i ^= (i & 0xAAAAAAAAu) >> 1; i ^= (i & 0xCCCCCCCCu) >> 2; i ^= (i & 0xF0F0F0F0u) >> 4;
    i ^= (i & 0xFF00FF00u) >> 8; i ^= (i & 0xFFFF0000u) >> 16;
    return rev(i);
}
uint32_t vo_z_o2m3(uint32_t V, int axis) {
    if (axis == 1) V = ((V >> 1) & 0x45145145u) ^ (V & 0xB5EB5EB5u) ^ ((V << 1) & 0x8A28A28Au);
    if (axis == 2) V = ((V >> 1) & 0x45145145u) ^ (V & 0x7AD7AD7Au) ^ ((V << 1) & 0x8A28A28Au);
    V ^= ((V << 2) & 0xF3CF3CF3u) ^ ((V << 3) & 0x41041041u);
    V ^= (V << 4) & 0x3F03F03Fu;
    V ^= (V << 8) & 0x0FFF000Fu;
    V ^= (V << 16) & 0x00FFFFFFu;
    return V;
}
uint32_t vo_z_morton(uint32_t x, uint32_t y) {
    uint32_t m = 0;
    for (int b = 0; b < 16; ++b) m |= ((x >> b) & 1u) << (2 * b) | ((y >> b) & 1u) << (2 * b + 1);
    return m;
}
uint32_t vo_z_pixel_key(uint32_t x, uint32_t y, uint32_t frame_index, uint32_t sample_index, uint32_t model) {
    uint32_t j = (frame_index > 1u ? frame_index : 1u) - 1u, s0 = sample_index - j, t;
    if (model == 1u || model == 2u) t = s0 ^ ((s0 & 0xAAAAAAAAu) << 1);
    else if (model == 3u) t = vo_z_mix(s0 ^ 0x2c1b3c6du);
    else t = s0;
    uint32_t k = vo_z_shuffle(vo_z_morton(x, y), 0xffffffffu);
    if (model == 2u) k ^= k << 2;
    if ((j >> 14) != 0u) t ^= vo_z_mix(j >> 14);
    return ((k ^ t) << 14) | (j & ((1u << 14) - 1u));
}

typedef struct {
    uint32_t key, dimension, sub, sub_bits;
} zs;
static uint32_t code(const zs *s) { return s->key >> 14; }
static uint32_t frame(const zs *s) { return ((s->key & ((1u << 14) - 1u)) << s->sub_bits) | s->sub; }
static uint32_t zindex(const zs *s, uint32_t d) { return vo_z_owen(code(s), vo_z_owen_seed(d, 3u)); }
static uint32_t next_dim(zs *s) { return s->dimension++ | ((code(s) >> 12) << 17); }
static uint32_t frame_index(uint32_t J, uint32_t d) {
    uint32_t seed = vo_z_owen_seed(d, 4u);
    return vo_z_owen(J, seed) ^ vo_z_owen(0u, seed);
}
static uint32_t frame_offset(const zs *s, uint32_t v, uint32_t d, uint32_t axis) {
    uint32_t seed = vo_z_owen_seed(d, axis) ^ vo_z_mix(code(s) ^ 0x68e31da4u);
    return ((vo_z_owen(v, seed) ^ vo_z_owen(0u, seed)) & 0x3FFFFFFFu) | (v & 0xC0000000u);
}
static float zf(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

/* Raw 0.32 coordinates of one draw of `kind` (1..3) at the sampler's next dimension. */
static void draw(zs *s, int kind, uint32_t *x) {
    uint32_t d = next_dim(s), i = zindex(s, d), J = frame(s);
    if (kind == 1) {
        x[0] = vo_z_owen(rev(i), vo_z_owen_seed(d, 0u));
        if (J) x[0] ^= frame_offset(s, rev(frame_index(J, d)), d, 0u);
    } else if (kind == 2) {
        x[0] = vo_z_owen(rev(i), vo_z_owen_seed(d, 0u));
        x[1] = vo_z_owen(vo_z_sobol1(i), vo_z_owen_seed(d, 1u));
        if (J) {
            uint32_t k = frame_index(J, d);
            x[0] ^= frame_offset(s, rev(k), d, 0u);
            x[1] ^= frame_offset(s, vo_z_sobol1(k), d, 1u);
        }
    } else {
        uint32_t V = rev(i);
        for (int a = 0; a < 3; ++a) x[a] = vo_z_owen(vo_z_o2m3(V, a), vo_z_owen_seed(d, (uint32_t)a));
        if (J) {
            uint32_t W = rev(frame_index(J, d));
            for (int a = 0; a < 3; ++a) x[a] ^= frame_offset(s, vo_z_o2m3(W, a), d, (uint32_t)a);
        }
    }
}

/* Every pixel's event and draws, as vt_z_draws: out is pixel-major, draw by draw; raw (when not NULL)
 * receives the 0.32 coordinates. */
void vo_z_draws(uint32_t width, uint32_t height, uint32_t frame_index_, uint32_t sample_index, uint32_t model,
                uint32_t path_vertex, uint32_t stream, uint32_t candidate, uint32_t candidate_bits, uint32_t draws,
                uint32_t kind, float *out, uint32_t *raw) {
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            zs s = {vo_z_pixel_key(x, y, frame_index_, sample_index, model), (path_vertex * 16u + stream) * 32u,
                    candidate_bits ? candidate : 0u, candidate_bits};
            size_t base = ((size_t)y * width + x) * draws * kind;
            for (uint32_t k = 0; k < draws; ++k) {
                uint32_t v[3];
                draw(&s, (int)kind, v);
                for (uint32_t a = 0; a < kind; ++a) {
                    out[base + k * kind + a] = zf(v[a]);
                    if (raw) raw[base + k * kind + a] = v[a];
                }
            }
        }
}

/* One draw of `kind` for a sampler with the given key at *dimension (advanced), with the sampler's
 * candidate bits (used by the oracle integrator's Z mode). */
void vo_z_draw(uint32_t key, uint32_t *dimension, uint32_t sub, uint32_t sub_bits, int kind, float *out) {
    zs s = {key, *dimension, sub, sub_bits};
    uint32_t v[3];
    draw(&s, kind, v);
    *dimension = s.dimension;
    for (int a = 0; a < kind; ++a) out[a] = zf(v[a]);
}
