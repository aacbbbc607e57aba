"""Writes tests/golden/known_answers.npz: closed-form ray/triangle cases.

Every triangle lies in a plane z = const with dyadic vertices and every ray is
axis-parallel (direction +-z) through a dyadic point, so the exact hit
distance and barycentrics are dyadic rationals computed here with Python
Fractions and representable exactly in float32. The set covers interior hits,
misses, hits exactly on a shared edge and on a shared vertex, back-face hits,
ties between overlapping triangles (lowest ID must win), and tmin/tmax
clipping. Run:  python tests/golden/make_golden.py
"""
from fractions import Fraction as Fr
import os

import numpy as np


def bary(p, a, b, c):
    """Exact barycentrics (u for b, v for c) of 2D point p in triangle abc, or None."""
    det = (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1])
    if det == 0:
        return None
    u = ((p[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (p[1] - a[1])) / det
    v = ((b[0] - a[0]) * (p[1] - a[1]) - (p[0] - a[0]) * (b[1] - a[1])) / det
    return u, v


def main():
    tris = []   # (v0, v1, v2) with z constant
    # 0,1: unit square at z=2 split along its diagonal (shared edge (0,0)-(1,1))
    tris.append(((0, 0, 2), (1, 0, 2), (1, 1, 2)))
    tris.append(((0, 0, 2), (1, 1, 2), (0, 1, 2)))
    # 2: a triangle at z=5 behind the square
    tris.append(((-2, -2, 5), (4, -2, 5), (-2, 4, 5)))
    # 3,4: two identical overlapping triangles at z=3.5 (tie: lower ID wins)
    tris.append(((2, 2, Fr(7, 2)), (3, 2, Fr(7, 2)), (2, 3, Fr(7, 2))))
    tris.append(((2, 2, Fr(7, 2)), (3, 2, Fr(7, 2)), (2, 3, Fr(7, 2))))
    # 5: a fan around vertex (6,6) shared with 6 and 7 at z=1.25
    tris.append(((6, 6, Fr(5, 4)), (7, 6, Fr(5, 4)), (6, 7, Fr(5, 4))))
    tris.append(((6, 6, Fr(5, 4)), (6, 7, Fr(5, 4)), (5, 6, Fr(5, 4))))
    tris.append(((6, 6, Fr(5, 4)), (5, 6, Fr(5, 4)), (6, 5, Fr(5, 4))))
    T = [[tuple(Fr(x) for x in v) for v in t] for t in tris]

    rays = []   # (o, d, tmin, tmax)
    pts = [(Fr(1, 4), Fr(1, 8)), (Fr(3, 4), Fr(1, 2)), (Fr(1, 2), Fr(1, 2)), (Fr(1, 8), Fr(7, 8)),
           (Fr(-1, 2), Fr(-1, 2)), (Fr(3, 1), Fr(3, 1)), (Fr(9, 4), Fr(9, 4)), (Fr(6), Fr(6)),
           (Fr(25, 4), Fr(6)), (Fr(6), Fr(23, 4)), (Fr(11, 2), Fr(11, 2)), (Fr(1), Fr(1)), (Fr(0), Fr(1, 2)),
           (Fr(7), Fr(7)), (Fr(-3), Fr(-3))]
    for (x, y) in pts:
        rays.append(((x, y, Fr(-1)), (0, 0, 1), Fr(0), Fr(10 ** 30)))          # toward +z from z=-1
        rays.append(((x, y, Fr(9)), (0, 0, -1), Fr(0), Fr(10 ** 30)))          # toward -z from z=9
    # clipped rays
    rays.append(((Fr(1, 4), Fr(1, 8), Fr(-1)), (0, 0, 1), Fr(4), Fr(10 ** 30)))   # skips z=2, hits z=5
    rays.append(((Fr(1, 4), Fr(1, 8), Fr(-1)), (0, 0, 1), Fr(0), Fr(3)))         # t=3 excluded: miss
    rays.append(((Fr(1, 4), Fr(1, 8), Fr(-1)), (0, 0, 1), Fr(3), Fr(6)))         # t=3 excluded: t=6 excluded

    exp_t, exp_id, exp_u, exp_v = [], [], [], []
    for o, d, tmin, tmax in rays:
        best = None
        for i, (a, b, c) in enumerate(T):
            bc = bary((o[0], o[1]), a, b, c)
            if bc is None:
                continue
            u, v = bc
            if u < 0 or v < 0 or u + v > 1:
                continue
            t = (a[2] - o[2]) / d[2]
            if not (tmin < t < tmax):
                continue
            if best is None or (t, i) < (best[0], best[1]):
                best = (t, i, u, v)
        if best is None:
            exp_t.append(1e30); exp_id.append(-1); exp_u.append(0.0); exp_v.append(0.0)
        else:
            exp_t.append(float(best[0])); exp_id.append(best[1]); exp_u.append(float(best[2])); exp_v.append(float(best[3]))

    tri_arr = np.array([[float(x) for v in t for x in v] for t in T], np.float32)
    ray_arr = np.array([[float(o[0]), float(o[1]), float(o[2]), d[0], d[1], d[2], float(tmin),
                         float(min(tmax, Fr(10 ** 30)))] for o, d, tmin, tmax in rays], np.float32).T.copy()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.npz")
    np.savez(out, tris=tri_arr, rays=ray_arr, t=np.array(exp_t, np.float32), id=np.array(exp_id, np.int32),
             u=np.array(exp_u, np.float32), v=np.array(exp_v, np.float32))
    print(f"wrote {out}: {len(T)} triangles, {len(rays)} rays, {sum(i >= 0 for i in exp_id)} hits")


if __name__ == "__main__":
    main()
