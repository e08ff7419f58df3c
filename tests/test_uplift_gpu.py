"""GPU tests of the device-built reflectance-uplift grid (vt_uplift_grid_build: the reference's
refineSpectralGrid, main.swift:8187-8278) against the oracle's restatement of the reference's
generator (oracle/spectral.py, itself pinned to scripts/generate_spectral_tables.py's outputs by
tests/golden/spectral_uplift.npz) and against the reference's own count of refined cells ("5,278 of
614,125", main.swift:845).  Bars: coarse nodes within float storage of the oracle's exact moments;
the refined-cell count within 3% of the reference's (both kernels run in float, the reference's under
Metal's default fast math, so which cells near the 0.35-step threshold flag depends on the last bits
of pow, atan and the order of the 471-term sums; measured 5175, -2%); flags equal to the oracle's
double-precision flags on at
least 99% of sampled cells; refined codes' moments within 5e-5 of the oracle's exact moments (a float
solve to 2e-7 in linear sRGB, as the reference's)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid(gpu_ctx):
    g = gpu_ctx.uplift_grid()
    print(f"uplift grid: {g[4]} refined cells", flush=True)
    return g


def test_coarse_nodes_are_the_generators_moments(grid):
    coarse = grid[0]
    d, _ = oracle.spectral_data()
    rng = np.random.default_rng(3)
    nodes = [(0, 0, 0), (85, 85, 85), (85, 0, 0), (0, 85, 0), (0, 0, 85), (40, 3, 77)]
    nodes += [tuple(int(v) for v in rng.integers(0, 86, 3)) for _ in range(120)]
    worst = 0.0
    for n in nodes:
        want = d.code_moments([3 * v for v in n]).astype(np.float64)
        worst = max(worst, float(np.abs(coarse[n] - want).max()))
    print(f"coarse nodes: max |diff| {worst:.2e} over {len(nodes)} nodes", flush=True)
    assert worst < 2e-6


def test_refined_cell_count_is_the_references(grid):
    count = grid[4]
    assert abs(count - 5278) <= 0.03 * 5278, count
    flags = grid[1]
    assert int(flags.sum()) == count


def test_flags_agree_with_the_oracle(grid):
    flags, cells = grid[1], grid[2]
    d, _ = oracle.spectral_data()
    rng = np.random.default_rng(5)
    flagged = [tuple(int(v) for v in np.unravel_index(c, (85, 85, 85))) for c in rng.choice(cells, 60, replace=False)]
    others = [tuple(int(v) for v in rng.integers(0, 85, 3)) for _ in range(60)]
    sample = flagged + others
    agree = np.mean([bool(flags[c]) == d.refined(c) for c in sample])
    print(f"flags: {agree:.3f} of {len(sample)} sampled cells agree with the oracle", flush=True)
    assert agree >= 0.97


def test_refined_codes_are_exact_moments(grid):
    cells, refined = grid[2], grid[3]
    d, _ = oracle.spectral_data()
    rng = np.random.default_rng(7)
    worst = 0.0
    for slot in rng.choice(len(cells), 40, replace=False):
        i = np.unravel_index(int(cells[slot]), (85, 85, 85))
        for _ in range(3):
            e = rng.integers(0, 4, 3)
            code = [min(3 * int(i[k]) + int(e[k]), 255) for k in range(3)]
            want = d.code_moments(code).astype(np.float64)
            worst = max(worst, float(np.abs(refined[slot][tuple(e)] - want).max()))
    print(f"refined codes: max |diff| {worst:.2e}", flush=True)
    assert worst < 5e-5
