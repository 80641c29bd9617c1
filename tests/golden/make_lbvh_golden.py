"""Writes tests/golden/lbvh_grid.json: a known-answer LBVH case computed in
pure Python, without the oracle or the HIP library.

24 triangles with integer vertices v0=(a,b,c), v1=(a+3,b,c), v2=(a,b+3,c) in
a [0,16]^3 box. Their centroids (a+1, b+1, c) and the normalized coordinates
x/16 are exact in float32, so the 21-bit quantization is q = x * 2^17 exactly
and the Morton code is the plain bit interleave of q (x highest). Sorting is a
stable argsort; the hierarchy comes from the brute-force top-down Karras split.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import lbvh_check as LC  # noqa: E402


def main():
    rng = np.random.default_rng(2024)
    abc = rng.integers(0, 13, size=(24, 3))
    abc[5] = abc[2]  # equal centroids -> equal keys, ordered by id
    abc[17] = abc[2]
    verts, tris = [], []
    for i, (a, b, c) in enumerate(abc):
        verts += [[a, b, c], [a + 3, b, c], [a, b + 3, c]]
        tris.append([3 * i, 3 * i + 1, 3 * i + 2])
    cen = abc + np.array([1, 1, 0])
    q = (cen.astype(np.uint64) << np.uint64(17))
    morton = LC.interleave21((q[:, 0], q[:, 1], q[:, 2]))
    order = np.argsort(morton, kind="stable")
    parents = LC.karras_bruteforce(morton[order])
    out = dict(vertices=[[float(x) for x in v] for v in verts], triangles=tris,
               bbox=[0.0, 0.0, 0.0, 16.0, 16.0, 16.0], morton=[str(int(m)) for m in morton],
               sorted_ids=[int(i) for i in order], parents=[int(p) for p in parents])
    with open(os.path.join(HERE, "lbvh_grid.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
