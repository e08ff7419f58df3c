// uplift.h -- the reflectance-uplift grid built on the device (uplift.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace vt {

// Device copies of the reflectance -> linear sRGB weights under D65 (3 rows of 471 samples, 360-830 nm)
// and the warped-phase cosines 2 cos(phi), 2 cos(2 phi), in double and rounded to float.
struct UpliftTables {
    const double *w, *c1, *c2;
    const float *wf, *c1f, *c2f;
};

// The 86^3 coarse nodes' exact moments into grid (float4 xyz, w = 0), in double.
hipError_t launch_uplift_coarse(const UpliftTables &t, float4 *grid, hipStream_t stream);
// Sets flags[cell] (85^3 bytes, cleared by the caller) for every cell with an 8-bit code the coarse
// conversion misses by more than 0.35 steps; nodes = the 86 linear node values.
hipError_t launch_uplift_flags(const UpliftTables &t, const float4 *grid, const float *nodes, uint8_t *flags,
                               hipStream_t stream);
// The exact moments of the 64 codes of each of `count` cells into refined[64 slot + entry].
hipError_t launch_uplift_refine(const UpliftTables &t, const float4 *grid, const float *nodes, const uint32_t *cells,
                                uint32_t count, float4 *refined, hipStream_t stream);

}  // namespace vt
