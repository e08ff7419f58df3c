"""CPU tests of the flat binary hierarchy (build option collapse = 2; the reference's OBJMesh.hierarchy):
the oracle's restatement (oracle/hierarchy.py) pinned by the properties the construction fixes, and the
library's C++ builder (vt_hierarchy_build_host) equal to it bit for bit, order and node records."""
import numpy as np
import pytest

from oracle import hierarchy as ohier
import scenes
import vt


def _meshes():
    out = []
    v, f = scenes.scattered_triangles(700, seed=3)
    out.append(("scattered", v, f))
    m = vt.icosphere_field(spheres=12, subdivisions=2, seed=5, half_extent=6)
    out.append(("spheres", m.vertices, m.faces))
    # A flat grid: every centre shares y, and rows share z -- ties the index order must break.
    g = 24
    xs, zs = np.meshgrid(np.arange(g + 1, dtype=np.float32), np.arange(g + 1, dtype=np.float32))
    vv = np.zeros(((g + 1) ** 2, 8), np.float32)
    vv[:, 0], vv[:, 2] = xs.ravel(), zs.ravel()
    ff = []
    for r in range(g):
        for c in range(g):
            a = r * (g + 1) + c
            ff += [[a, a + 1, a + g + 1, 0], [a + 1, a + g + 2, a + g + 1, 0]]
    out.append(("grid", vv, np.asarray(ff, np.uint32)))
    for n in (1, 4, 5, 9):
        vs, fs = scenes.scattered_triangles(n, seed=n)
        out.append((f"tiny{n}", vs, fs))
    return out


def _check_properties(v, f, order, lo, hi, links, depth):
    n = len(f)
    assert sorted(order.tolist()) == list(range(n))
    corners = np.asarray(v, np.float32)[:, :3][np.asarray(f)[:, :3].astype(np.int64)]
    centres = ((corners[:, 0] + corners[:, 1]) + corners[:, 2]) / np.float32(3)
    covered = []
    max_level = [0]

    def visit(idx, start, end, level):
        max_level[0] = max(max_level[0], level)
        pts = corners[order[start:end]].reshape(-1, 3)
        assert np.array_equal(lo[idx], pts.min(0)) and np.array_equal(hi[idx], pts.max(0)), idx
        l = links[idx]
        if end - start <= 4:
            assert l[0] == 0 and l[1] == 0 and l[2] == start and l[3] == end - start, (idx, l)
            covered.append((start, end))
            return idx + 1
        assert l[3] == 0 and l[0] == idx + 1, (idx, l)   # pre-order: the left child follows its parent
        e = hi[idx] - lo[idx]
        axis = (0 if e[0] > e[2] else 2) if e[0] > e[1] else (1 if e[1] > e[2] else 2)
        assert l[2] == axis
        mid = (start + end) // 2
        keys = [(float(centres[r, axis]), int(r)) for r in order[start:end]]
        assert max(keys[:mid - start]) < min(keys[mid - start:])   # the smaller half goes left
        nxt = visit(l[0], start, mid, level + 1)
        assert l[1] == nxt                                     # the right subtree follows the left one
        return visit(l[1], mid, end, level + 1)

    total = visit(0, 0, n, 1)
    assert total == len(lo)
    assert covered == sorted(covered) and covered[0][0] == 0 and covered[-1][1] == n
    assert all(covered[k][1] == covered[k + 1][0] for k in range(len(covered) - 1))
    assert depth == max_level[0]


@pytest.mark.parametrize("case", range(7))
def test_oracle_hierarchy_properties_and_library_equality(case):
    name, v, f = _meshes()[case]
    order, lo, hi, links, depth = ohier.build(v, f)
    _check_properties(v, f, order, lo, hi, links, depth)
    got_order, nodes, got_depth = vt.hierarchy_build_host(v, f)
    assert np.array_equal(got_order, order), name
    assert np.array_equal(nodes["lo"][:, :3], lo) and np.array_equal(nodes["hi"][:, :3], hi), name
    assert np.all(nodes["lo"][:, 3] == 0) and np.all(nodes["hi"][:, 3] == 0)
    assert np.array_equal(nodes["links"], links), name
    assert got_depth == depth


def test_planted_split_axis_mistake_breaks_the_properties():
    name, v, f = _meshes()[1]
    order, lo, hi, links, depth = ohier.build(v, f)
    bad = links.copy()
    inner = np.nonzero(bad[:, 3] == 0)[0]
    bad[inner[0], 2] = (bad[inner[0], 2] + 1) % 3
    with pytest.raises(AssertionError):
        _check_properties(v, f, order, lo, hi, bad, depth)
