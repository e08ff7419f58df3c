"""Oracle (test infrastructure only): the reference's flat binary hierarchy, OBJMesh.hierarchy
(Sources/StudioModel.swift:701-760), restated in Python for the checker of build option collapse = 2.

Steps, as the reference writes them: triangle centres (a + b + c) / 3 in float32; a node's bounds the
min / max of its triangles' corners; a node of more than four triangles splits at position
(start + end) // 2 on the axis of its largest extent (x if it beats y and z, else z; y if it beats z,
else z) by a quickselect of the triangle references -- a Hoare partition around the middle
reference's centre, references ordered by centre and then by index, at most 64 rounds before the
remaining range is sorted outright; the left subtree is built before the right, so nodes are in
pre-order.  Node records: lo, hi (w = 0), links = (left, right, axis, 0) inside, (0, 0, first, count)
at a leaf.

Pinned by tests/test_hierarchy.py through properties the construction fixes independently of this
restatement (a permutation of the faces; pre-order numbering; leaves of one to four triangles tiling
the order; bounds equal to the min / max of the corners below; the split axis the largest extent; the
left half holding exactly the smaller half of the range in the (centre, index) order), and against the
library's C++ builder bit for bit."""
from __future__ import annotations

import sys

import numpy as np


def build(vertices, faces):
    """vertices (V, >=3), faces (F, >=3) -> (order (F,) int64, lo (N, 3) float32, hi (N, 3) float32,
    links (N, 4) int64, depth)."""
    v = np.asarray(vertices, np.float32)[:, :3]
    f = np.asarray(faces)[:, :3].astype(np.int64)
    n = len(f)
    if n == 0:
        return np.zeros(0, np.int64), np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), \
            np.zeros((0, 4), np.int64), 0
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    centres = ((a + b) + c) / np.float32(3.0)
    corners = np.stack([a, b, c], 1)  # (F, 3, 3)
    order = list(range(n))
    lo, hi, links = [], [], []
    depth = [0]

    def key(i, axis):
        return (float(centres[i, axis]), i)

    def partition(start, end, middle, axis):
        low, high, rounds = start, end - 1, 0
        while low < high:
            if rounds >= 64:
                order[low:high + 1] = sorted(order[low:high + 1], key=lambda r: key(r, axis))
                return
            rounds += 1
            pivot = key(order[(low + high) // 2], axis)
            i, j = low, high
            while i <= j:
                while i <= high and key(order[i], axis) < pivot:
                    i += 1
                while j >= low and pivot < key(order[j], axis):
                    j -= 1
                if i <= j:
                    order[i], order[j] = order[j], order[i]
                    i += 1
                    j -= 1
            if middle <= j:
                high = j
            elif middle >= i:
                low = i
            else:
                return

    def node(start, end, level):
        depth[0] = max(depth[0], level)
        pts = corners[order[start:end]].reshape(-1, 3)
        nlo, nhi = pts.min(0), pts.max(0)
        index = len(lo)
        lo.append(nlo)
        hi.append(nhi)
        links.append([0, 0, 0, 0])
        if end - start <= 4:
            links[index] = [0, 0, start, end - start]
            return index
        e = nhi - nlo
        axis = (0 if e[0] > e[2] else 2) if e[0] > e[1] else (1 if e[1] > e[2] else 2)
        middle = (start + end) // 2
        partition(start, end, middle, axis)
        left = node(start, middle, level + 1)
        right = node(middle, end, level + 1)
        links[index] = [left, right, axis, 0]
        return index

    limit = sys.getrecursionlimit()
    sys.setrecursionlimit(max(limit, 10000))
    try:
        node(0, n, 1)
    finally:
        sys.setrecursionlimit(limit)
    return (np.asarray(order, np.int64), np.asarray(lo, np.float32), np.asarray(hi, np.float32),
            np.asarray(links, np.int64), depth[0])
