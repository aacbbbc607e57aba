"""GPU ray sort: keys equal the oracle's (cuda_rt_v29.cu:84-106 restated),
the permutation equals a stable argsort, and sorted traces give exactly the
unsorted answers (permutation invariance)."""
import numpy as np
import pytest

from conftest import RAY_KEYS, hit_buffers, to_dev

pytestmark = pytest.mark.gpu


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = rng.uniform(-1.2, 1.2, (3, n))
    d = rng.normal(size=(3, n))
    rays[3:6] = d / np.linalg.norm(d, axis=0)
    rays[6] = 1e-3
    rays[7] = rng.uniform(0.05, 3.0, n)
    rays[:, ::97] = rays[:, :1]                      # duplicates: stability matters
    return rays


@pytest.mark.parametrize("n", [1, 2047, 2048, 100_003])
def test_sort_keys_and_permutation(pkg, oracle, scenes, dev, n):
    import torch
    tris = scenes.random_triangles(3000, seed=5)
    rays = _rays(n, 40 + n % 7)
    with pkg.BVH.build(tris) as bvh:
        keys = torch.empty(n, dtype=torch.int32, device=dev)
        perm = torch.empty(n, dtype=torch.int32, device=dev)
        bvh.sort_rays(to_dev(rays, dev), keys, perm)
        torch.cuda.synchronize()
        k = keys.cpu().numpy().view(np.uint32)
        p = perm.cpu().numpy()
        pts = tris.reshape(-1, 3)
        lo, hi = pts.min(0), pts.max(0)
        inv = (np.float32(1.0) / (hi - lo)).astype(np.float32)
        ref = oracle.sort_keys(rays, lo, inv)
        order = np.argsort(ref, kind="stable")
        assert np.array_equal(p, order)
        assert np.array_equal(k, ref[order])


def test_sorted_trace_equals_unsorted(pkg, oracle, scenes, dev):
    tris = scenes.random_triangles(20000, seed=6)
    rays = _rays(150_000, 77)
    with pkg.BVH.build(tris) as bvh:
        R = to_dev(rays, dev)
        import torch
        a, b = hit_buffers(rays.shape[1], dev), hit_buffers(rays.shape[1], dev)
        bvh.trace_closest(R, a["t"], a["prim_id"], a["u"], a["v"])
        bvh.trace_closest(R, b["t"], b["prim_id"], b["u"], b["v"], flags=pkg.FLAG_SORT_RAYS)
        oa = torch.empty(rays.shape[1], dtype=torch.uint8, device=dev)
        ob = torch.empty(rays.shape[1], dtype=torch.uint8, device=dev)
        bvh.trace_any(R, oa)
        bvh.trace_any(R, ob, flags=pkg.FLAG_SORT_RAYS)
        torch.cuda.synchronize()
        for k in a:
            assert np.array_equal(a[k].cpu().numpy().view(np.uint32), b[k].cpu().numpy().view(np.uint32))
        assert np.array_equal(oa.cpu().numpy(), ob.cpu().numpy())
        nodes, lt, pid = bvh.node_data(), np.ascontiguousarray(bvh.tri_data().T), bvh.leaf_prim_ids()
        ref = oracle.trace_closest(nodes, lt, pid, rays)
        assert np.array_equal(b["prim_id"].cpu().numpy(), ref["id"])
