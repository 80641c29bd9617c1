"""Independent checkers for LBVH outputs (pure numpy / Python): brute-force
Karras topology, exact AABB unions, the collapse-mark characterization and a
decoder of Mufflon's 64-byte collapsed node layout (lbvh.hpp:13-44)."""
import numpy as np

TREE_HEAD = 0x10000000


def interleave21(q):
    """Naive bit interleave of three 21-bit ints: x at bit 3i+2, y 3i+1, z 3i."""
    qx, qy, qz = (np.asarray(a, np.uint64) for a in q)
    out = np.zeros_like(qx)
    for i in range(21):
        b = np.uint64(i)
        out |= ((qx >> b) & np.uint64(1)) << np.uint64(3 * i + 2)
        out |= ((qy >> b) & np.uint64(1)) << np.uint64(3 * i + 1)
        out |= ((qz >> b) & np.uint64(1)) << np.uint64(3 * i)
    return out


def spread21(x):
    """Naive bit spread of a 21-bit int to every third bit."""
    x = np.asarray(x, np.uint64)
    out = np.zeros_like(x)
    for i in range(21):
        out |= ((x >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i)
    return out


def _delta(keys, i, j):
    if keys[i] == keys[j]:
        return 64 + (32 - int(i ^ j).bit_length())
    return 64 - int(int(keys[i]) ^ int(keys[j])).bit_length()


def karras_bruteforce(keys):
    """Top-down construction by linear search for the split, with Karras's
    node numbering (an internal node is named after its first or last key)."""
    keys = [int(k) for k in keys]
    n = len(keys)
    parents = np.zeros(2 * n - 1, np.int64)
    parents[0] = TREE_HEAD
    stack = [(0, 0, n - 1)]  # (node, first, last)
    while stack:
        node, a, b = stack.pop()
        d = _delta(keys, a, b)
        s = a
        for i in range(a, b):
            if _delta(keys, a, i) > d:
                s = i
        left = s + n - 1 if s == a else s
        right = s + 1 + n - 1 if s + 1 == b else s + 1
        parents[left] = ~node
        parents[right] = node
        if s != a:
            stack.append((s, a, s))
        if s + 1 != b:
            stack.append((s + 1, s + 1, b))
    return parents.astype(np.int32)


def children_from_parents(parents, n):
    nint = n - 1
    left = np.full(nint, -1, np.int64)
    right = np.full(nint, -1, np.int64)
    for c in range(1, 2 * n - 1):
        p = int(parents[c])
        if p < 0:
            left[~p] = c
        else:
            right[p] = c
    return left, right


def parent_of(parents, x):
    p = int(parents[x])
    return ~p if p < 0 else p


def tri_boxes(verts, tris):
    v = np.asarray(verts, np.float32).reshape(-1, 3)[np.asarray(tris, np.int64).reshape(-1, 3)]
    return v.min(axis=1), v.max(axis=1)


def decode_nodes(nodes):
    """64-byte internal nodes -> (lo (m,2,3), hi (m,2,3), index (m,2), count (m,2))."""
    raw = np.frombuffer(np.ascontiguousarray(nodes).tobytes(), np.uint8).reshape(-1, 2, 32)
    f = raw[:, :, :24].copy().view(np.float32).reshape(-1, 2, 6)
    i = raw[:, :, 24:].copy().view(np.int32).reshape(-1, 2, 2)
    return f[:, :, 0:3], f[:, :, 3:6], i[:, :, 0], i[:, :, 1]


def walk_collapsed(nodes, num_internal, n):
    """Enumerate leaves of a collapsed BVH: list of (start, count, lo, hi)."""
    lo, hi, idx, cnt = decode_nodes(nodes)
    leaves = []
    seen_internal = np.zeros(num_internal, bool)
    stack = [0]
    seen_internal[0] = True
    while stack:
        nd = stack.pop()
        for s in range(2):
            c = int(idx[nd, s])
            if c >= num_internal:
                leaves.append((c - num_internal, int(cnt[nd, s]), lo[nd, s], hi[nd, s]))
            else:
                assert not seen_internal[c], "internal node reached twice"
                seen_internal[c] = True
                stack.append(c)
    assert seen_internal.all(), "unreachable internal nodes"
    return leaves


def expected_marks(parents, boxes, n, capped_fn, leaf_cost_fn):
    """Per-node collapse decision: keep=0 iff node is not the root, is below
    the count cap, and a node on the path to the first capped ancestor (or
    the root) is cheaper as a leaf."""
    nint = n - 1
    B = boxes.reshape(-1, 8)
    cnt = B[:, 3].view(np.int32)
    cost = B[:, 7]
    keep = np.ones(nint, np.int32)
    for x in range(1, nint):
        y = x
        while y != 0:
            if capped_fn(int(cnt[y])):
                break
            if np.float32(leaf_cost_fn(int(cnt[y]))) < cost[y]:
                keep[x] = 0
                break
            y = parent_of(parents, y)
    return keep


def lod_capped(e):
    return ((e >> 20) & 0x3FF) >= 1023 or ((e >> 10) & 0x3FF) >= 1023 or (e & 0x3FF) >= 1023


def lod_leaf_cost(e):
    x, y, z = np.float32((e >> 20) & 0x3FF), np.float32((e >> 10) & 0x3FF), np.float32(e & 0x3FF)
    return np.float32(2.0) + ((x * np.float32(1.2) + y * np.float32(4.0)) + z * np.float32(1.0))
