"""Numpy restatement of the secondary-ray generator (TEST INFRASTRUCTURE).

Checker for cudaBVH_genBounceRays: same float32 operation sequence (numpy
float32 arithmetic is IEEE single without contraction), PCG32 (O'Neill 2014,
pcg32_srandom_r / pcg32_random_r) in uint64, Malley's method by bounded
rejection, and the orthonormal basis of Duff et al. 2017. The reference's
diffuse-ray generator (cuda_rt_v37.cu:514-540) is the model: per-ray seeded
RNG, hemisphere around the hit normal.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32
MULT = np.uint64(6364136223846793005)


def _pcg_next(state, inc):
    old = state
    with np.errstate(over="ignore"):
        state = old * MULT + inc
    xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
    rot = (old >> np.uint64(59)).astype(np.uint32)
    out = (xs >> rot) | (xs << ((np.uint32(0) - rot) & np.uint32(31)))
    return state, out


def _pcg_seed(initstate, initseq):
    inc = (initseq << np.uint64(1)) | np.uint64(1)
    state = np.zeros_like(initseq)
    state, _ = _pcg_next(state, inc)
    with np.errstate(over="ignore"):
        state = state + initstate
    state, _ = _pcg_next(state, inc)
    return state, inc


def pinhole_rays(cam14, W, H, idx):
    """Primary rays of pixels idx (float32 sequence of the GPU generator)."""
    c = cam14.astype(f32)
    py = (idx // W).astype(np.int64)
    px = (idx - py * W).astype(np.int64)
    sx = (px.astype(f32) + f32(0.5)) / f32(W) * f32(2.0) - f32(1.0)
    sy = (py.astype(f32) + f32(0.5)) / f32(H) * f32(2.0) - f32(1.0)
    a, b = sx * c[12], sy * c[13]
    d = [(c[3 + k] + a * c[6 + k]) + b * c[9 + k] for k in range(3)]
    inv = f32(1.0) / np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
    o = [np.full(idx.shape, c[k], f32) for k in range(3)]
    return o, [d[k] * inv for k in range(3)]


def bounce_rays(cam14, W, H, hit_t, hit_id, tris, seed, light):
    """tris: original-order (n, 9) float32. Returns (diffuse (8,m), shadow (8,m), src (m,))."""
    idx = np.nonzero(hit_id >= 0)[0]
    o, d = pinhole_rays(cam14, W, H, idx)
    t = hit_t[idx].astype(f32)
    T = tris[hit_id[idx]].astype(f32)
    h = [o[k] + t * d[k] for k in range(3)]
    e1 = [T[:, 3 + k] - T[:, k] for k in range(3)]
    e2 = [T[:, 6 + k] - T[:, k] for k in range(3)]
    n = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
    nl = f32(1.0) / np.sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2])
    n = [n[k] * nl for k in range(3)]
    flip = ((n[0] * d[0] + n[1] * d[1]) + n[2] * d[2]) > f32(0.0)
    n = [np.where(flip, -n[k], n[k]) for k in range(3)]

    state, inc = _pcg_seed(np.uint64(seed), idx.astype(np.uint64))
    m = idx.shape[0]
    x = np.zeros(m, f32)
    y = np.zeros(m, f32)
    s = np.zeros(m, f32)
    ok = np.zeros(m, bool)
    for _ in range(64):
        act = ~ok
        if not act.any():
            break
        st, r1 = _pcg_next(state, inc)
        st, r2 = _pcg_next(st, inc)
        state = np.where(act, st, state)
        xx = f32(2.0) * ((r1 >> np.uint32(8)).astype(f32) * f32(2.0 ** -24)) - f32(1.0)
        yy = f32(2.0) * ((r2 >> np.uint32(8)).astype(f32) * f32(2.0 ** -24)) - f32(1.0)
        ss = xx * xx + yy * yy
        x = np.where(act, xx, x)
        y = np.where(act, yy, y)
        s = np.where(act, ss, s)
        ok = ok | (act & (ss < f32(1.0)))
    x = np.where(ok, x, f32(0))
    y = np.where(ok, y, f32(0))
    s = np.where(ok, s, f32(0))
    z = np.sqrt(f32(1.0) - s)
    sg = np.copysign(f32(1.0), n[2]).astype(f32)
    ca = f32(-1.0) / (sg + n[2])
    cb = n[0] * n[1] * ca
    t1 = [f32(1.0) + sg * n[0] * n[0] * ca, sg * cb, -sg * n[0]]
    t2 = [cb, sg + n[1] * n[1] * ca, -n[1]]
    dd = [(x * t1[k] + y * t2[k]) + z * n[k] for k in range(3)]
    dl = f32(1.0) / np.sqrt((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2])
    dd = [dd[k] * dl for k in range(3)]
    diffuse = np.stack(h + dd + [np.full(m, 1e-3, f32), np.full(m, 1e30, f32)]).astype(f32)
    L = [f32(light[k]) - h[k] for k in range(3)]
    dist = np.sqrt((L[0] * L[0] + L[1] * L[1]) + L[2] * L[2])
    il = f32(1.0) / dist
    shadow = np.stack(h + [L[k] * il for k in range(3)] + [np.full(m, 1e-3, f32), dist]).astype(f32)
    return diffuse, shadow, idx.astype(np.int32)
