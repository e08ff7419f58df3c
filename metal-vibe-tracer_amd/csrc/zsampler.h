// zsampler.h -- the Z++ sampler (main.swift:122-305, 497-515): Owen-scrambled low-discrepancy
// constituents indexed along a recursively shuffled Morton curve of the image, with temporal
// indexing across an accumulation's frames.  Integer arithmetic only, so device and host agree bit
// for bit.  A pixel's key (z_pixel_key) identifies all of its samples for one frame; each sampling
// event (a path vertex and a stream) owns 32 consecutive dimensions, and each 1D / 2D / 3D draw
// takes the event's next one.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VT_ZHD __host__ __device__ __forceinline__
#else
#define VT_ZHD inline
#endif

namespace vt {

VT_ZHD uint32_t z_reverse(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bitreverse32(v);
#else
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
    v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
    return (v >> 16) | (v << 16);
#endif
}

// z_mix (main.swift:132-136): the lowbias32 integer finaliser.
VT_ZHD uint32_t z_mix(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

// z_shuffle (main.swift:139-158): each of the low 12 base-4 digits of a Z index is permuted by one of
// the 24 permutations of four quadrants (one of the six that fix quadrant 0, as 2-bit images packed
// in a constant, then an XOR with a digit), chosen by a hash of the digits above it and the dimension.
VT_ZHD uint32_t z_shuffle(uint32_t index, uint32_t dimension) {
    constexpr uint64_t kPermutations = 0x6c9c78d8b4e4ull;
    constexpr uint32_t kDigits = 12u;
    const uint32_t seed = z_mix(dimension * 0x9e3779b9u + 0x632be5abu);
    uint32_t result = index & ~((1u << (2u * kDigits)) - 1u);
    for (uint32_t d = kDigits; d-- > 0u;) {
        const uint32_t shift = 2u * d, digit = (index >> shift) & 3u;
        const uint32_t h = z_mix((index >> (shift + 2u)) ^ seed);
        const uint32_t sigma = ((h >> 16) * 6u) >> 16;
        const uint32_t image = (uint32_t)(kPermutations >> (8u * sigma + 2u * digit)) & 3u;
        result |= (image ^ (h & 3u)) << shift;
    }
    return result;
}

// z_owen (main.swift:162-170): nested uniform scrambling of a 0.32 fraction (Laine-Karras hash).
VT_ZHD uint32_t z_owen(uint32_t v, uint32_t seed) {
    v = z_reverse(v);
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return z_reverse(v);
}
VT_ZHD uint32_t z_owen_seed(uint32_t dimension, uint32_t axis) {
    const uint32_t h = (dimension ^ (axis << 24)) * 0x9e3779b9u;
    return (h ^ (h >> 16)) * 0x85ebca6bu;
}

// z_sobol1 (main.swift:175-182): the second Sobol' dimension of index i as a 0.32 fraction.
VT_ZHD uint32_t z_sobol1(uint32_t i) {
    i ^= (i & 0xAAAAAAAAu) >> 1;
    i ^= (i & 0xCCCCCCCCu) >> 2;
    i ^= (i & 0xF0F0F0F0u) >> 4;
    i ^= (i & 0xFF00FF00u) >> 8;
    i ^= (i & 0xFFFF0000u) >> 16;
    return z_reverse(i);
}

// The O2m3 sequence (main.swift:184-192): coordinates of the point with bit-reversed index V.
VT_ZHD uint32_t z_o2m3_x(uint32_t V) {
    V ^= ((V << 2) & 0xF3CF3CF3u) ^ ((V << 3) & 0x41041041u);
    V ^= (V << 4) & 0x3F03F03Fu;
    V ^= (V << 8) & 0x0FFF000Fu;
    V ^= (V << 16) & 0x00FFFFFFu;
    return V;
}
VT_ZHD uint32_t z_o2m3_y(uint32_t V) { return z_o2m3_x(((V >> 1) & 0x45145145u) ^ (V & 0xB5EB5EB5u) ^ ((V << 1) & 0x8A28A28Au)); }
VT_ZHD uint32_t z_o2m3_z(uint32_t V) { return z_o2m3_x(((V >> 1) & 0x45145145u) ^ (V & 0x7AD7AD7Au) ^ ((V << 1) & 0x8A28A28Au)); }
VT_ZHD float z_float(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

// z_morton (main.swift:199-204): x in the even bits.
VT_ZHD uint32_t z_morton(uint32_t x, uint32_t y) {
    auto spread = [](uint32_t v) {
        v &= 0xFFFFu;
        v = (v | (v << 8)) & 0x00FF00FFu;
        v = (v | (v << 4)) & 0x0F0F0F0Fu;
        v = (v | (v << 2)) & 0x33333333u;
        v = (v | (v << 1)) & 0x55555555u;
        return v;
    };
    return spread(x) | (spread(y) << 1);
}

constexpr uint32_t kZFrameBits = 14u;

// z_pixel_key (main.swift:497-506): bits 14-31 the pixel code (shuffled Morton index XOR the
// accumulation's temporal index), bits 0-13 the frame of the accumulation.  temporal_model: 0 plain,
// 1 and 2 the digit-interleaved variants, 3 hashed (sampler_z_temporal).
VT_ZHD uint32_t z_pixel_key(uint32_t x, uint32_t y, uint32_t frame_index, uint32_t sample_index, uint32_t temporal_model) {
    const uint32_t j = (frame_index > 1u ? frame_index : 1u) - 1u, s0 = sample_index - j;
    uint32_t t = temporal_model == 1u || temporal_model == 2u ? s0 ^ ((s0 & 0xAAAAAAAAu) << 1)
                 : temporal_model == 3u                       ? z_mix(s0 ^ 0x2c1b3c6du)
                                                              : s0;
    uint32_t k = z_shuffle(z_morton(x, y), 0xffffffffu);
    if (temporal_model == 2u) k ^= k << 2;
    if ((j >> kZFrameBits) != 0u) t ^= z_mix(j >> kZFrameBits);
    return ((k ^ t) << kZFrameBits) | (j & ((1u << kZFrameBits) - 1u));
}

// The Z-mode state of Sampler (main.swift:213-219) and its draws (main.swift:221-275).
struct ZSampler {
    uint32_t key, dimension, sub, sub_bits;

    // sampler_event / sampler_candidate (main.swift:288-293).
    VT_ZHD void event(uint32_t path_vertex, uint32_t stream) {
        dimension = (path_vertex * 16u + stream) * 32u;
        sub = 0u;
        sub_bits = 0u;
    }
    VT_ZHD void candidate(uint32_t path_vertex, uint32_t stream, uint32_t index, uint32_t bits) {
        event(path_vertex, stream);
        sub = index;
        sub_bits = bits;
    }
    VT_ZHD uint32_t code() const { return key >> kZFrameBits; }
    VT_ZHD uint32_t frame() const { return ((key & ((1u << kZFrameBits) - 1u)) << sub_bits) | sub; }
    VT_ZHD uint32_t index(uint32_t d) const { return z_owen(code(), z_owen_seed(d, 3u)); }
    VT_ZHD uint32_t next_dimension() { return dimension++ | ((code() >> 12) << 17); }
    VT_ZHD static uint32_t frame_index(uint32_t J, uint32_t d) {
        const uint32_t seed = z_owen_seed(d, 4u);
        return z_owen(J, seed) ^ z_owen(0u, seed);
    }
    VT_ZHD uint32_t frame_offset(uint32_t v, uint32_t d, uint32_t axis) const {
        const uint32_t seed = z_owen_seed(d, axis) ^ z_mix(code() ^ 0x68e31da4u);
        return ((z_owen(v, seed) ^ z_owen(0u, seed)) & 0x3FFFFFFFu) | (v & 0xC0000000u);
    }
    VT_ZHD float rand1() {
        const uint32_t d = next_dimension(), J = frame();
        uint32_t x = z_owen(z_reverse(index(d)), z_owen_seed(d, 0u));
        if (J != 0u) x ^= frame_offset(z_reverse(frame_index(J, d)), d, 0u);
        return z_float(x);
    }
    VT_ZHD void rand2(float &a, float &b) {
        const uint32_t d = next_dimension(), i = index(d), J = frame();
        uint32_t x = z_owen(z_reverse(i), z_owen_seed(d, 0u)), y = z_owen(z_sobol1(i), z_owen_seed(d, 1u));
        if (J != 0u) {
            const uint32_t k = frame_index(J, d);
            x ^= frame_offset(z_reverse(k), d, 0u);
            y ^= frame_offset(z_sobol1(k), d, 1u);
        }
        a = z_float(x);
        b = z_float(y);
    }
    VT_ZHD void rand3(float &a, float &b, float &c) {
        const uint32_t d = next_dimension(), V = z_reverse(index(d)), J = frame();
        uint32_t x = z_owen(z_o2m3_x(V), z_owen_seed(d, 0u)), y = z_owen(z_o2m3_y(V), z_owen_seed(d, 1u)),
                 z = z_owen(z_o2m3_z(V), z_owen_seed(d, 2u));
        if (J != 0u) {
            const uint32_t W = z_reverse(frame_index(J, d));
            x ^= frame_offset(z_o2m3_x(W), d, 0u);
            y ^= frame_offset(z_o2m3_y(W), d, 1u);
            z ^= frame_offset(z_o2m3_z(W), d, 2u);
        }
        a = z_float(x);
        b = z_float(y);
        c = z_float(z);
    }
};

}  // namespace vt
