"""The compressed 8-wide BVH: structure (every triangle once, every child box
contains its subtree when dequantized in float exactly as the kernel does)
and traversal parity (hits bit-identical to the oracle on the same rays,
any hit equal to closest-hit-exists), including flat rays."""
import os

import numpy as np
import pytest

from test_gpu_trace import GOLD, _assert_closest_equal, _gpu_any, _gpu_closest, _oracle_inputs

pytestmark = pytest.mark.gpu


def _decode(nodes):
    """Child slots of every node: (meta, lo[3], hi[3]) in float32, p + q * 2^e."""
    n = nodes.shape[0]
    p = nodes[:, 0:3].view(np.float32)
    e = np.stack([(nodes[:, 3] >> (8 * a)) & 0xFF for a in range(3)], 1).astype(np.uint8).view(np.int8)
    s = np.ldexp(np.float32(1.0), e.astype(np.int32)).astype(np.float32)
    meta = nodes[:, 6:8].copy().view(np.uint8).reshape(n, 8)
    q = nodes[:, 8:20].copy().view(np.uint8).reshape(n, 3, 2, 8)     # axis, lo/hi, slot
    lo = p[:, :, None] + q[:, :, 0, :].astype(np.float32) * s[:, :, None]
    hi = p[:, :, None] + q[:, :, 1, :].astype(np.float32) * s[:, :, None]
    return meta, lo.astype(np.float32), hi.astype(np.float32)


def _check_structure(tris, nodes, order, leaf_tris_of_bvh4):
    n = nodes.shape[0]
    meta, lo, hi = _decode(nodes)
    assert np.array_equal(np.sort(order), np.arange(len(order)))          # every triangle exactly once
    lt = leaf_tris_of_bvh4[order].reshape(-1, 3, 3)                          # CWBVH order, 3 vertices
    tlo, thi = lt.min(1), lt.max(1)
    childBase = nodes[:, 4].astype(np.int64)
    triBase = nodes[:, 5].astype(np.int64)
    boxes = {}                                                               # node -> bounds of its triangles
    for v in range(n - 1, -1, -1):
        m = meta[v]
        ulo, uhi = np.full(3, np.inf, np.float32), np.full(3, -np.inf, np.float32)
        for i in range(8):
            if m[i] == 0xFF:
                continue
            if m[i] & 0x80:
                c = childBase[v] + (m[i] & 0x1F)
                clo, chi = boxes[int(c)]
            else:
                first = triBase[v] + (m[i] & 0x1F)
                cnt = ((m[i] >> 5) & 3) + 1
                clo, chi = tlo[first:first + cnt].min(0), thi[first:first + cnt].max(0)
            assert np.all(lo[v, :, i] <= clo) and np.all(hi[v, :, i] >= chi), (v, i)
            ulo, uhi = np.minimum(ulo, clo), np.maximum(uhi, chi)
        boxes[v] = (ulo, uhi)


@pytest.mark.parametrize("kind", ["random", "terrain"])
def test_cwbvh_structure_and_parity(pkg, oracle, scenes, dev, kind):
    if kind == "random":
        tris = scenes.random_triangles(65536, seed=0)
        cam = scenes.parity_camera(256, 256)
    else:
        tris = scenes.terrain_scene(res=128, spheres=16, level=2)
        cam = scenes.terrain_camera(320, 180)
    print(f"cwbvh {kind}", flush=True)
    with pkg.BVH.build(tris) as b:
        nn = b.build_cwbvh()
        assert nn == b.num_cwbvh_nodes > 0
        nodes, order, depth = b.cwbvh()
        lt = np.ascontiguousarray(b.tri_data().T)
        _check_structure(tris, nodes, order, lt)
        assert depth <= b.depth
        bnodes, blt, bpid = _oracle_inputs(b)
        rays = oracle.gen_pinhole(scenes.cam_array(cam), *((256, 256) if kind == "random" else (320, 180)))
        rng = np.random.default_rng(3)
        m = 20000
        rnd = np.zeros((8, m), np.float32)
        rnd[0:3] = rng.uniform(-1.2, 1.2, (3, m)) if kind == "random" else rng.uniform(-8, 8, (3, m))
        d = rng.normal(size=(3, m))
        rnd[3:6] = d / np.linalg.norm(d, axis=0)
        rnd[6] = 1e-3
        rnd[7] = rng.uniform(0.05, 30.0, m)
        for rr in (rays, rnd):
            ref = oracle.trace_closest(bnodes, blt, bpid, rr)
            got = _gpu_closest(b, rr, dev, flags=pkg.FLAG_CWBVH)
            _assert_closest_equal(got, ref)
            occ = _gpu_any(b, rr, dev, flags=pkg.FLAG_CWBVH)
            assert np.array_equal(occ.astype(bool), ref["id"] >= 0)
        assert (ref["id"] >= 0).mean() > 0.05


def test_cwbvh_flat_rays_and_small(pkg, oracle, dev):
    g = np.load(GOLD)
    tris = np.concatenate([g["tris"], g["tris"] + np.float32(0.5)])
    rng = np.random.default_rng(9)
    n = 4096
    rays = np.zeros((8, n), np.float32)
    grid = np.float32(0.25) * rng.integers(-8, 40, (2, n)).astype(np.float32)
    rays[0], rays[1], rays[2] = grid[0], grid[1], -1.0
    rays[5] = 1.0
    mixed = rng.random(n) < 0.3
    d = rng.normal(size=(3, n)).astype(np.float32)
    d[2] = np.abs(d[2]) + 0.5
    rays[3:6, mixed] = (d / np.linalg.norm(d, axis=0))[:, mixed]
    rays[7] = 1e30
    with pkg.BVH.build(tris) as b:
        b.build_cwbvh()
        nodes, lt, pid = _oracle_inputs(b)
        ref = oracle.trace_closest(nodes, lt, pid, rays)
        _assert_closest_equal(_gpu_closest(b, rays, dev, flags=pkg.FLAG_CWBVH), ref)
    one = np.array([[0, 0, 1, 1, 0, 1, 0, 1, 1]], np.float32)          # a single-leaf tree
    with pkg.BVH.build(one) as b:
        rays = np.array([[0.25, 2.0], [0.25, 2.0], [0, 0], [0, 0], [0, 0], [1, 1], [0, 0], [1e30, 1e30]], np.float32)
        got = _gpu_closest(b, rays, dev, flags=pkg.FLAG_CWBVH)
        assert got["prim_id"].tolist() == [0, -1] and got["t"][0] == 1.0
        assert b.num_cwbvh_nodes == 1
    with pkg.BVH.build(tris) as b:
        with pytest.raises(pkg.CudaBVHError):                          # counted traces are BVH4 only
            import torch
            from conftest import to_dev, hit_buffers
            R = to_dev(rays, dev)
            H = hit_buffers(2, dev)
            c = torch.empty(2, dtype=torch.int32, device=dev)
            b.trace_closest_counted(R, H["t"], H["prim_id"], H["u"], H["v"], c, c, pkg.FLAG_CWBVH)
