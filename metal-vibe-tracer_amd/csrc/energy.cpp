// energy.cpp -- energy-compensation tables of the layered BSDF, prepared once per context.
//
// The ideal-metal table follows scripts/generate_metal_energy.py (129 x 129 nodes, 16384 Hammersley
// samples per node, baked on the device; the average complement integrates the table, piecewise
// linear in sqrt(cosine), exactly: :55-63), the table the reference splices over OpenPBR's own metal
// lookups (scripts/prepare_shaders.py:28-40, Shaders/MetalEnergy.metal).  The five dielectric tables
// are OpenPBR's constant data (Vendor/OpenPBR/impl/data/openpbr_*dielectric*_data.h, read through
// openpbr_microfacet_multiple_scattering_data.h:140-165): unorm16 values on 32-node axes (relative
// IOR, sqrt(alpha), cosine), stored here as floats v / 65535 in the order the lookups index them.
// The fuzz lobe's LTC table (Vendor/OpenPBR/impl/data/openpbr_ltc_data.h: 32 x 32 triples a_inv,
// b_inv, R over sheen roughness and cosine) follows them in the same buffer (kLtcOffset).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "render_kernels.h"
#include "scene.h"

namespace vt {
namespace {

#include "energy_data.inc"

void put_unorm16(float *dst, const uint16_t *src, int n) {
    for (int i = 0; i < n; ++i) dst[i] = (float)src[i] * (1.0f / 65535.0f);
}

}  // namespace

void ensure_energy_tables(vt_context *ctx) {
    if (ctx->energy_ready) return;
    check(hipSetDevice(ctx->device), "hipSetDevice");
    DeviceBuffer metal_raw;
    metal_raw.alloc((size_t)129 * 129 * 4);
    check(launch_bake_metal(metal_raw.as<float>(), nullptr), "bake metal energy");
    std::vector<float> metal(129 * 129 + 129);
    check(hipMemcpy(metal.data(), metal_raw.ptr, (size_t)129 * 129 * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    for (int row = 0; row < 129; ++row) {
        double integral = 0.0;
        for (int c = 0; c < 128; ++c) {
            double lo = c / 128.0, hi = (c + 1) / 128.0;
            double slope = ((double)metal[row * 129 + c + 1] - (double)metal[row * 129 + c]) * 128.0;
            double offset = (double)metal[row * 129 + c] - slope * lo;
            integral += offset * (std::pow(hi, 4) - std::pow(lo, 4)) + 0.8 * slope * (std::pow(hi, 5) - std::pow(lo, 5));
        }
        metal[129 * 129 + row] = (float)integral;
    }
    // Five slots of 32768 floats: ideal complement, ideal average, reflection ratio, opaque
    // complement, opaque average (the 2D tables use the first 1024 floats of their slot); then the LTC
    // table's 3072 floats.
    std::vector<float> luts((size_t)kLtcOffset + 3072, 0.0f);
    put_unorm16(luts.data() + 0 * 32768, kOpenPbrIdealDielectricComplement, 32768);
    put_unorm16(luts.data() + 1 * 32768, kOpenPbrIdealDielectricAvgComplement, 1024);
    put_unorm16(luts.data() + 2 * 32768, kOpenPbrIdealDielectricReflectionRatio, 1024);
    put_unorm16(luts.data() + 3 * 32768, kOpenPbrOpaqueDielectricComplement, 32768);
    put_unorm16(luts.data() + 4 * 32768, kOpenPbrOpaqueDielectricAvgComplement, 1024);
    for (int i = 0; i < 3072; ++i) luts[kLtcOffset + i] = kOpenPbrLtcSheen[i];
    ctx->energy_metal.upload(metal.data(), metal.size() * 4);
    ctx->energy_luts.upload(luts.data(), luts.size() * 4);
    ctx->energy_ready = true;
}

}  // namespace vt
