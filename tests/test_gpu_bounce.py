"""GPU parity of the terrain workload chain: fused primary trace ->
secondary-ray generation (stable compaction) -> queued closest-hit and
any-hit traces, each against its CPU restatement."""
import numpy as np
import pytest

from conftest import RAY_KEYS

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("res,spheres,W,H", [(128, 16, 320, 180), (256, 32, 257, 131)])
def test_bounce_chain(pkg, oracle, scenes, dev, res, spheres, W, H):
    import torch
    import bounce_ref
    tris = scenes.terrain_scene(res=res, spheres=spheres, level=2)
    cam = scenes.terrain_camera(W, H)
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    n = W * H
    with pkg.BVH.build(tris) as bvh:
        nodes, lt, pid = bvh.node_data(), np.ascontiguousarray(bvh.tri_data().T), bvh.leaf_prim_ids()
        t = torch.empty(n, device=dev)
        ids = torch.empty(n, dtype=torch.int32, device=dev)
        bvh.trace_pinhole_closest(W, H, c, t, ids)
        D = {k: torch.full((n,), -7.0, device=dev) for k in RAY_KEYS}
        S = {k: torch.full((n,), -7.0, device=dev) for k in RAY_KEYS}
        src = torch.full((n,), -1, dtype=torch.int32, device=dev)
        cnt = torch.zeros(1, dtype=torch.int32, device=dev)
        bvh.gen_bounce_rays(W, H, c, t, ids, 1234, scenes.TERRAIN_LIGHT, D, S, src, cnt)
        torch.cuda.synchronize()
        m = int(cnt.item())
        ht, hid = t.cpu().numpy(), ids.cpu().numpy()
        # primary hits equal the oracle's
        rays = oracle.gen_pinhole(scenes.cam_array(cam), W, H)
        ref = oracle.trace_closest(nodes, lt, pid, rays)
        assert np.array_equal(hid, ref["id"]) and np.array_equal(_bits(ht), _bits(ref["t"]))
        # secondary rays equal the numpy restatement, in pixel order
        dref, sref, srcref = bounce_ref.bounce_rays(scenes.cam_array(cam), W, H, ht, hid, tris, 1234,
                                                   scenes.TERRAIN_LIGHT)
        assert m == (hid >= 0).sum() == dref.shape[1]
        assert np.array_equal(src.cpu().numpy()[:m], srcref)
        dg = np.stack([D[k].cpu().numpy()[:m] for k in RAY_KEYS])
        sg = np.stack([S[k].cpu().numpy()[:m] for k in RAY_KEYS])
        assert np.array_equal(_bits(dg), _bits(dref))
        assert np.array_equal(_bits(sg), _bits(sref))
        assert (D["ox"].cpu().numpy()[m:] == -7.0).all()              # nothing written past the count
        # diffuse directions are unit and lie in the hemisphere of the facing
        # normal (float64 normal of the hit triangle, turned against the ray)
        assert ((dref[3:6] * dref[3:6]).sum(0) > 0.9999).all()
        T = tris[hid[srcref]].astype(np.float64).reshape(-1, 3, 3)
        nrm = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
        d0 = rays[3:6, srcref].T.astype(np.float64)
        nrm[(nrm * d0).sum(1) > 0] *= -1
        assert ((dg[3:6].T.astype(np.float64) * nrm).sum(1) >= 0).all()
        # queued traces read the count from device memory
        dt = torch.empty(n, device=dev)
        did = torch.empty(n, dtype=torch.int32, device=dev)
        du = torch.empty(n, device=dev)
        dv = torch.empty(n, device=dev)
        bvh.trace_closest_queued(n, cnt, D, dt, did, du, dv)
        occ = torch.full((n,), 9, dtype=torch.uint8, device=dev)
        bvh.trace_any_queued(n, cnt, S, occ)
        torch.cuda.synchronize()
        rd = oracle.trace_closest(nodes, lt, pid, dref)
        assert np.array_equal(did.cpu().numpy()[:m], rd["id"])
        assert np.array_equal(_bits(dt.cpu().numpy()[:m]), _bits(rd["t"]))
        assert np.array_equal(_bits(du.cpu().numpy()[:m]), _bits(rd["u"]))
        rs = oracle.trace_any(nodes, lt, pid, sref)
        assert np.array_equal(occ.cpu().numpy()[:m], rs["occ"])
        assert (occ.cpu().numpy()[m:] == 9).all()
        assert 0.0 < rs["occ"].mean() < 1.0 and 0.0 < (rd["id"] >= 0).mean() < 1.0


def _tile_strip_order(W, H):
    """Pixels in the order of CUDABVH_FLAG_TILE_ORDER: complete 8x8 tiles in
    vertical strips of 8 tiles (row-major inside a strip and inside a tile),
    then the right margin and the bottom rows, row-major."""
    tx8, ty8 = W // 8, H // 8
    order = []
    for s0 in range(0, tx8, 8):
        for ty in range(ty8):
            for tx in range(s0, min(s0 + 8, tx8)):
                for k in range(64):
                    order.append((ty * 8 + k // 8) * W + tx * 8 + k % 8)
    for y in range(ty8 * 8):
        order.extend(y * W + x for x in range(tx8 * 8, W))
    for y in range(ty8 * 8, H):
        order.extend(y * W + x for x in range(W))
    assert sorted(order) == list(range(W * H))
    return np.array(order)


@pytest.mark.parametrize("W,H", [(320, 180), (257, 131)])
def test_bounce_tile_order(pkg, scenes, dev, W, H):
    """FLAG_TILE_ORDER emits the same rays as pixel order, permuted into the
    tile walk: srcIndex lists the hit pixels in that order, and every ray
    equals the pixel-order ray of its source pixel."""
    import torch
    tris = scenes.terrain_scene(res=128, spheres=16, level=2)
    cam = scenes.terrain_camera(W, H)
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    n = W * H
    with pkg.BVH.build(tris) as bvh:
        t = torch.empty(n, device=dev)
        ids = torch.empty(n, dtype=torch.int32, device=dev)
        bvh.trace_pinhole_closest(W, H, c, t, ids)
        out = []
        for flags in (0, pkg.FLAG_TILE_ORDER):
            D = {k: torch.full((n,), -7.0, device=dev) for k in RAY_KEYS}
            S = {k: torch.full((n,), -7.0, device=dev) for k in RAY_KEYS}
            src = torch.full((n,), -1, dtype=torch.int32, device=dev)
            cnt = torch.zeros(1, dtype=torch.int32, device=dev)
            bvh.gen_bounce_rays(W, H, c, t, ids, 99, scenes.TERRAIN_LIGHT, D, S, src, cnt, flags=flags)
            torch.cuda.synchronize()
            m = int(cnt.item())
            out.append((m, src.cpu().numpy()[:m], np.stack([D[k].cpu().numpy()[:m] for k in RAY_KEYS]),
                        np.stack([S[k].cpu().numpy()[:m] for k in RAY_KEYS])))
        hid = ids.cpu().numpy()
    (m0, s0, d0, h0), (m1, s1, d1, h1) = out
    assert m0 == m1 == (hid >= 0).sum()
    order = _tile_strip_order(W, H)
    assert np.array_equal(s1, order[hid[order] >= 0])
    pos = np.full(n, -1)
    pos[s0] = np.arange(m0)
    assert np.array_equal(_bits(d1), _bits(d0[:, pos[s1]]))
    assert np.array_equal(_bits(h1), _bits(h0[:, pos[s1]]))
