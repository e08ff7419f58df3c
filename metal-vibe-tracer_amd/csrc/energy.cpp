// energy.cpp -- energy-compensation tables of the layered BSDF, baked once per context.
//
// The ideal-metal table follows scripts/generate_metal_energy.py (129 x 129 nodes, 16384 Hammersley
// samples per node; the average complement integrates the table, piecewise linear in sqrt(cosine),
// exactly: :55-63).  The 32-node dielectric tables sit on OpenPBR's axes
// (Vendor/OpenPBR/impl/openpbr_microfacet_multiple_scattering_data.h:27-58); the upstream data
// files carry no generator, so they are the single-scattering integrals of the GGX dielectric
// interface: ideal dielectric complement 1 - (E_R + E_T), opaque dielectric complement 1 - E_R,
// their hemispherical averages 1 - 2 int E(mu) mu dmu (piecewise linear in mu), and the reflection
// ratio avg E_R / (avg E_R + avg E_T), stored as unorm16 like the upstream tables.  The device bakes
// the node integrals; the host forms averages and quantises.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <vector>

#include "render_kernels.h"
#include "scene.h"

namespace vt {
namespace {

float unorm16(double v) {
    uint32_t u = !(v > 0.0) ? 0u : v >= 1.0 ? 65535u : (uint32_t)std::lrint(v * 65535.0);
    return (float)u * (1.0f / 65535.0f);
}

}  // namespace

void ensure_energy_tables(vt_context *ctx) {
    if (ctx->energy_ready) return;
    check(hipSetDevice(ctx->device), "hipSetDevice");
    DeviceBuffer metal_raw, er, et;
    metal_raw.alloc((size_t)129 * 129 * 4);
    er.alloc((size_t)32768 * 4);
    et.alloc((size_t)32768 * 4);
    check(launch_bake_metal(metal_raw.as<float>(), nullptr), "bake metal energy");
    check(launch_bake_dielectric(er.as<float>(), et.as<float>(), nullptr), "bake dielectric energy");
    std::vector<float> metal(129 * 129 + 129), r(32768), t(32768);
    check(hipMemcpy(metal.data(), metal_raw.ptr, (size_t)129 * 129 * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    check(hipMemcpy(r.data(), er.ptr, (size_t)32768 * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    check(hipMemcpy(t.data(), et.ptr, (size_t)32768 * 4, hipMemcpyDeviceToHost), "hipMemcpy");
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
    std::vector<float> luts((size_t)5 * 32768, 0.0f);
    float *diel = luts.data(), *diel_avg = diel + 32768, *ratio = diel_avg + 32768, *opaque = ratio + 32768,
          *opaque_avg = opaque + 32768;
    for (int i = 0; i < 32; ++i) {
        for (int j = 0; j < 32; ++j) {
            double aR = 0, aT = 0;
            const float *R = &r[i * 1024 + j * 32], *T = &t[i * 1024 + j * 32];
            for (int k = 0; k < 31; ++k) {
                double lo = k / 31.0, hi = (k + 1) / 31.0;
                double sR = ((double)R[k + 1] - R[k]) * 31.0, oR = R[k] - sR * lo;
                double sT = ((double)T[k + 1] - T[k]) * 31.0, oT = T[k] - sT * lo;
                aR += 2.0 * (oR * (hi * hi - lo * lo) / 2.0 + sR * (hi * hi * hi - lo * lo * lo) / 3.0);
                aT += 2.0 * (oT * (hi * hi - lo * lo) / 2.0 + sT * (hi * hi * hi - lo * lo * lo) / 3.0);
            }
            for (int k = 0; k < 32; ++k) {
                diel[i * 1024 + j * 32 + k] = unorm16(1.0 - ((double)R[k] + T[k]));
                opaque[i * 1024 + j * 32 + k] = unorm16(1.0 - (double)R[k]);
            }
            diel_avg[i * 32 + j] = unorm16(1.0 - (aR + aT));
            opaque_avg[i * 32 + j] = unorm16(1.0 - aR);
            ratio[i * 32 + j] = unorm16(aR + aT > 0 ? aR / (aR + aT) : 1.0);
        }
    }
    ctx->energy_metal.upload(metal.data(), metal.size() * 4);
    ctx->energy_luts.upload(luts.data(), luts.size() * 4);
    ctx->energy_ready = true;
}

}  // namespace vt
