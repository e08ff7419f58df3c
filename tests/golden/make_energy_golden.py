#!/usr/bin/env python3
"""Writes tests/golden/openpbr_energy_tables.npz: the energy-compensation tables of the reference's
layered BSDF, as data, for the oracle and the parity tests.

  * the five dielectric tables vendored with Adobe OpenPBR (c91aad1, Apache-2.0),
    Vendor/OpenPBR/impl/data/openpbr_{ideal,opaque}_dielectric_*_data.h: unorm16 arrays, 32 x 32 x 32
    (relative IOR, sqrt(alpha), cosine) or 32 x 32, indexed x * 1024 + y * 32 + z
    (openpbr_energy_array_access.h:76-88);
  * the ideal-metal table the reference splices over OpenPBR's metal lookups
    (Shaders/MetalEnergy.metal: vibeMetalEnergy[129 * 129], vibeMetalAverageEnergy[129]);
  * the fuzz lobe's LTC table vendored with OpenPBR (Vendor/OpenPBR/impl/data/openpbr_ltc_data.h, from
    Zeltner, Burley and Chiang's ltc-sheen, Apache-2.0): 32 x 32 (sheen roughness, cosine) entries of
    (a_inv, b_inv, R), float32.
Reads the reference checkout in this container only; the npz travels, the reference does not:
    python3 tests/golden/make_energy_golden.py /root/reference
"""
import os
import re
import sys

import numpy as np

DIELECTRIC = (  # (npz key, vendored data file, shape)
    ("ideal_complement", "openpbr_ideal_dielectric_energy_complement_data.h", (32, 32, 32)),
    ("ideal_avg_complement", "openpbr_ideal_dielectric_avg_energy_complement_data.h", (32, 32)),
    ("reflection_ratio", "openpbr_ideal_dielectric_reflection_ratio_data.h", (32, 32)),
    ("opaque_complement", "openpbr_opaque_dielectric_energy_complement_data.h", (32, 32, 32)),
    ("opaque_avg_complement", "openpbr_opaque_dielectric_avg_energy_complement_data.h", (32, 32)),
)


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def metal_arrays(path):
    text = strip_comments(open(path).read())
    out = {}
    for name, n in (("vibeMetalEnergy", 16641), ("vibeMetalAverageEnergy", 129)):
        body = text[text.index(name + "["):]
        body = body[body.index("{") + 1:body.index("}")]
        vals = np.array([float(v.rstrip("f")) for v in body.replace("\n", "").split(",") if v.strip()], np.float32)
        assert vals.size == n, (name, vals.size)
        out[name] = vals
    return out


def ltc_array(path):
    """The 32 x 32 vec3 entries of openpbr_ltc_data.h, row (roughness) major."""
    text = strip_comments(open(path).read())
    vals = [float(v) for v in re.findall(r"vec3\(([^)]*)\)", text) for v in v.split(",")]
    assert len(vals) == 3072, len(vals)
    return np.array(vals, np.float32).reshape(32, 32, 3)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = os.path.join(ref, "Vendor", "OpenPBR", "impl", "data")
    arrays = {}
    for key, fname, shape in DIELECTRIC:
        vals = [int(v) for v in re.findall(r"\d+", strip_comments(open(os.path.join(data, fname)).read()))]
        assert len(vals) == int(np.prod(shape)), (fname, len(vals))
        arrays[key] = np.array(vals, np.uint16).reshape(shape)
    m = metal_arrays(os.path.join(ref, "Shaders", "MetalEnergy.metal"))
    arrays["metal_complement"] = m["vibeMetalEnergy"].reshape(129, 129)
    arrays["metal_avg_complement"] = m["vibeMetalAverageEnergy"]
    arrays["ltc_sheen"] = ltc_array(os.path.join(data, "openpbr_ltc_data.h"))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "openpbr_energy_tables.npz")
    np.savez_compressed(out, **arrays)
    print("wrote", out, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
