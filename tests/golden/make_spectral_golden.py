#!/usr/bin/env python3
"""Writes tests/golden/spectral_uplift.npz: the reference's reflectance uplift for a set of colours,
computed with the reference's own generator (scripts/generate_spectral_tables.py, stdlib only).

The renderer converts a linear sRGB colour c to a reflectance spectrum in four steps
(main.swift:842-903): the cell of c on the coarse grid of exact solves at 8-bit codes 0, 3, ..., 255
(FourierSRGB86.bin, made by the generator's solve_code :381-386 through solve_lagrange :308-340 and
moments_of_lagrange :227-239, stored as float32); trilinear interpolation of the cell's moments in
linear light (MSL spectral_moments_coarse :855-869, the generator's interpolate :412-425), or, in a
cell whose codes that interpolation reproduces worse than 0.35 8-bit steps (spectral_refine_flags_kernel
:8223-8233, the generator's code_error :357-366), the exact moments of the cell's codes at one-code
spacing (:870-884); then the bounded MESE's Lagrange multipliers (lagrange3 :188-220).  This script
runs those generator functions, for the steps the generator itself does, and restates the rest from
main.swift in double precision.

Run in the container that holds the reference checkout (the reference never travels):
    python3 tests/golden/make_spectral_golden.py /root/reference
Colours: the Cornell walls (main.swift:1422-1424), the six saturated corners of the emission basis,
the surface palettes of both terrain workloads, grid nodes, arbitrary 8-bit codes and saturated
colours in refined cells.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    sys.path.insert(0, os.path.join(ref, "scripts"))
    import generate_spectral_tables as g  # noqa: E402  (the reference's generator; this container only)

    data = g.Data()
    node_cache = {}

    def exact(code):
        code = tuple(int(v) for v in code)
        if code not in node_cache:
            c, _, err = g.solve_code(data, code)
            assert err < 1e-8, (code, err)
            node_cache[code] = np.array(c, np.float32)     # the grid file stores float32
        return node_cache[code]

    class LazyCoarse:
        """The generator's flat coarse array ((r * 86 + g) * 86 + b) * 3 + channel, solved on demand."""
        def __getitem__(self, idx):
            node, ch = divmod(idx, 3)
            i, rest = divmod(node, g.COARSE_N * g.COARSE_N)
            j, k = divmod(rest, g.COARSE_N)
            return float(exact((3 * i, 3 * j, 3 * k))[ch])

    coarse, axis = LazyCoarse(), g.interpolation_axis()
    flags = {}

    def refined(cell):
        if cell not in flags:
            ranges = [range(3 * i, 3 * i + (4 if i == 84 else 3)) for i in cell]
            worst = 0.0
            for r in ranges[0]:
                for gg in ranges[1]:
                    for b in ranges[2]:
                        if r == gg == b:
                            continue
                        worst = max(worst, g.code_error(data, g.interpolate(coarse, axis, (r, gg, b)), (r, gg, b)))
            flags[cell] = worst > g.FIX_THRESHOLD
        return flags[cell]

    nodes = np.array([g.srgb_eotf(3 * i / 255.0) for i in range(86)], np.float32).astype(np.float64)

    def uplift(rgb):
        c = np.clip(np.asarray(rgb, np.float32), 0, 1).astype(np.float64)
        if c.max() - c.min() <= 1e-4:
            return None
        code85 = np.array([g.srgb_oetf(v) * 85.0 for v in c])
        i = tuple(int(min(max(int(v), 0), 84)) for v in code85)
        ref_cell = refined(i)
        m = np.zeros(3)
        if not ref_cell:
            lo, hi = nodes[list(i)], nodes[[v + 1 for v in i]]
            t = np.clip((c - lo) / (hi - lo), 0, 1)
            base = [3 * v for v in i]
            step = 3
        else:
            s = [min(max(int(code85[k] * 3 - 3 * i[k]), 0), 2) for k in range(3)]
            base = [3 * i[k] + s[k] for k in range(3)]
            lo = np.array([g.srgb_eotf(b / 255.0) for b in base])
            hi = np.array([g.srgb_eotf((b + 1) / 255.0) for b in base])
            t = np.clip((c - lo) / (hi - lo), 0, 1)
            step = 1
        for a in (0, 1):
            for b in (0, 1):
                for d in (0, 1):
                    w = (t[0] if a else 1 - t[0]) * (t[1] if b else 1 - t[1]) * (t[2] if d else 1 - t[2])
                    if w == 0.0:
                        continue
                    code = (min(base[0] + a * step, 255), min(base[1] + b * step, 255), min(base[2] + d * step, 255))
                    m += w * exact(code).astype(np.float64)
        lam = g.lagrange3(*g.clamp_moments(m))
        return i, ref_cell, m, lam

    sys.path[:0] = [ROOT, os.path.join(ROOT, "metal-vibe-tracer_amd")]
    from vt import workloads  # the workload recipes, for their surface palettes

    colours = [(0.65, 0.05, 0.05), (0.12, 0.45, 0.15)]
    colours += [(1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)]
    for kind in ("restir", "openpbr"):
        colours += [tuple(m["albedo"]) for m in workloads.surface_palette(kind)]
    rng = np.random.default_rng(2024)
    colours += [tuple(g.srgb_eotf(3 * v / 255.0) for v in rng.integers(0, 86, 3)) for _ in range(16)]
    colours += [tuple(g.srgb_eotf(v / 255.0) for v in rng.integers(0, 256, 3)) for _ in range(24)]
    colours += [(0.98, 0.01, 0.0), (0.0, 0.9, 0.03), (0.02, 0.0, 0.85), (0.9, 0.95, 0.0), (0.7, 0.0, 0.6),
                (0.0, 0.45, 0.5), (0.99, 0.5, 0.002)]
    t0 = time.time()
    rows = []
    for col in colours:
        out = uplift(col)
        if out is None:
            continue
        i, ref_cell, m, lam = out
        if lam is None:
            continue
        rows.append((np.float32(col), i, ref_cell, m, lam))
    path = os.path.join(HERE, "spectral_uplift.npz")
    np.savez(path, rgb=np.array([r[0] for r in rows], np.float32), cell=np.array([r[1] for r in rows], np.int32),
             refined=np.array([r[2] for r in rows], bool), moments=np.array([r[3] for r in rows]),
             lagrange=np.array([r[4] for r in rows]))
    print(f"wrote {path}: {len(rows)} colours, {sum(r[2] for r in rows)} in refined cells, "
          f"{len(node_cache)} exact solves, {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
