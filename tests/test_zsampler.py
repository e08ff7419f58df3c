"""CPU checks of the Z++ sampler oracle (oracle/zsampler.c, restating main.swift:122-305, 497-515),
following the reference's own properties (tests/Fix_zsampling.swift:190-292): O2m3 blocks of 4^m points
are (0, 2m, 3)-nets with (0, 2m, 2) pairwise projections and 2D Sobol' blocks (0, m, 2)-nets, scrambled
or not; the quadrant shuffle is a bijection mapping aligned blocks onto aligned blocks; each frame is a Z
mask (aligned pixel blocks hold stratified samples); a pixel's samples over an accumulation's frames
are nets.  The second Sobol' dimension is checked against the Pascal-matrix closed form (Lucas:
C(j, r) is odd iff r is a bit-subset of j).  The hash constants are the reference's; they choose the
randomisation and no property pins them (parity of the exact constants is unpinned)."""
import itertools

import numpy as np

import oracle


def _rev(v):
    return int(f"{v:032b}"[::-1], 2)


def _top(v, bits):
    return np.asarray(v, np.uint64) >> np.uint64(32 - bits) if bits else np.zeros(len(v), np.uint64)


def _is_net(coords, m):
    """Every elementary box of volume 2^-m (split over the axes) holds exactly one of the 2^m points."""
    d = len(coords)
    for split in itertools.product(range(m + 1), repeat=d):
        if sum(split) != m:
            continue
        cells = np.zeros(len(coords[0]), np.uint64)
        for c, b in zip(coords, split):
            cells = (cells << np.uint64(b)) | _top(c, b)
        if len(np.unique(cells)) != 2 ** m:
            return False
    return True


def test_sobol1_is_the_pascal_matrix():
    lib = oracle.load()
    rng = np.random.default_rng(1)
    for i in [0, 1, 2, 3, 0xFFFFFFFF] + rng.integers(0, 2 ** 32, 200).tolist():
        want = 0
        for r in range(32):
            bit = 0
            for j in range(32):
                if (i >> j) & 1 and (r & ~j) == 0:
                    bit ^= 1
            want |= bit << (31 - r)
        assert lib.vo_z_sobol1(i) == want, i


def test_constituents_are_nets_scrambled_or_not():
    lib = oracle.load()
    for m in (1, 2, 3, 4):
        for block in (0, 5):
            idx = [block * 4 ** m + k for k in range(4 ** m)]
            V = [_rev(i) for i in idx]
            pts = [[lib.vo_z_o2m3(v, a) for v in V] for a in range(3)]
            seeds = [lib.vo_z_owen_seed(7, a) for a in range(3)]
            scr = [[lib.vo_z_owen(x, seeds[a]) for x in pts[a]] for a in range(3)]
            for p in (pts, scr):
                assert _is_net(p, 2 * m)
                for a, b in ((0, 1), (0, 2), (1, 2)):
                    assert _is_net([p[a], p[b]], 2 * m)
            sob = [[_rev(i) for i in idx[:2 ** m]], [lib.vo_z_sobol1(i) for i in idx[:2 ** m]]]
            assert _is_net(sob, m)
            assert _is_net([[lib.vo_z_owen(x, seeds[0]) for x in sob[0]], [lib.vo_z_owen(x, seeds[1]) for x in sob[1]]], m)


def test_shuffle_maps_aligned_blocks_onto_aligned_blocks():
    lib = oracle.load()
    for dim in (0, 3, 0xFFFFFFFF):
        out = np.array([lib.vo_z_shuffle(i, dim) for i in range(4 ** 6)], np.int64)
        assert len(np.unique(out)) == len(out)
        for m in (1, 2, 3):
            blocks = out.reshape(-1, 4 ** m)
            lo = blocks.min(1)
            assert np.all(lo % 4 ** m == 0) and np.all(blocks.max(1) - lo == 4 ** m - 1)
    # Digits above the 12 shuffled ones are kept.
    assert lib.vo_z_shuffle(0xAB000000 | 5, 9) >> 24 == 0xAB


def test_morton_and_pixel_key_layout():
    lib = oracle.load()
    for x, y in ((0, 0), (1, 0), (0, 1), (3, 5), (65535, 1)):
        want = 0
        for b in range(16):
            want |= ((x >> b) & 1) << (2 * b) | ((y >> b) & 1) << (2 * b + 1)
        assert lib.vo_z_morton(x, y) == want
    # Bits 0-13 hold the frame of the accumulation (frame_index - 1).
    assert lib.vo_z_pixel_key(3, 4, 1, 1, 0) & 0x3FFF == 0
    assert lib.vo_z_pixel_key(3, 4, 9, 20, 0) & 0x3FFF == 8


def test_each_frame_is_a_z_mask_and_pixels_accumulate_nets():
    # Frame 0: aligned 4 x 4 pixel blocks hold one 1D sample per sixteenth and a (0, 4, 2)-net of 2D.
    for model in (0, 1, 3):
        f1, r1 = oracle.z_draws(16, 16, 1, 1, model, path_vertex=1, stream=0, draws=1, kind=1)
        f2, r2 = oracle.z_draws(16, 16, 1, 1, model, path_vertex=2, stream=1, draws=1, kind=2)
        for by in range(0, 16, 4):
            for bx in range(0, 16, 4):
                b1 = r1[by:by + 4, bx:bx + 4, 0, 0].ravel()
                assert len(np.unique(b1 >> np.uint32(28))) == 16, (model, bx, by)
                b2 = r2[by:by + 4, bx:bx + 4, 0].reshape(-1, 2)
                assert _is_net([b2[:, 0], b2[:, 1]], 4), (model, bx, by)
    # Over frames 1..16 of one accumulation a pixel's 1D, 2D and 3D draws are nets.
    for kind, m in ((1, 4), (2, 4), (3, 4)):
        words = np.stack([oracle.z_draws(8, 8, j, j, 0, path_vertex=0, stream=4, draws=2, kind=kind)[1]
                          for j in range(1, 2 ** m + 1)])  # (frames, h, w, draws, kind)
        for (y, x) in ((0, 0), (3, 5), (7, 2)):
            for d in range(2):
                pts = words[:, y, x, d, :]
                assert _is_net([pts[:, a] for a in range(kind)], m), (kind, y, x, d)
    f, _ = oracle.z_draws(64, 64, 3, 7, 0, path_vertex=0, stream=2, draws=4, kind=3)
    assert np.all((f >= 0) & (f < 1)) and abs(f.mean() - 0.5) < 0.01
