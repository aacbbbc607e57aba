"""GPU parity of batched scenes: the forest build (every tree equals the
oracle's tree of its scene, up to node and triangle offsets) and one
persistent launch over the rays of all scenes (hit IDs, t, u, v bit-exact
against the oracle per scene), direct and hipGraph-replayed."""
import numpy as np
import pytest

from conftest import RAY_KEYS, to_dev

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _specs():
    rng = np.random.default_rng(0)
    tris = [int(x) for x in np.exp(rng.uniform(np.log(300), np.log(20000), 12))]
    rays = [int(x) for x in np.exp(rng.uniform(np.log(2000), np.log(30000), 12))]
    rays[4] = 0                                      # a scene without rays
    return tris, rays


def test_batch_build_and_trace(pkg, oracle, scenes, dev):
    import torch
    tris_n, rays_n = _specs()
    tris = [scenes.random_triangles(n, seed=100 + k) for k, n in enumerate(tris_n)]
    b = pkg.Batch(tris)
    roots = b.roots()
    forest = b.forest()
    nodes = forest.node_data()
    toff = b.tri_offsets
    ends = list(roots[1:]) + [nodes.shape[0]]
    orc = []
    for k, t in enumerate(tris):
        ob = oracle.OracleBVH(t)
        orc.append(ob)
        sub = nodes[roots[k]:ends[k]].copy()
        ch = sub[:, 48:].view(np.int32)
        inner = ch >= 0
        leaf = ch < -1
        enc = -(ch + 2)
        ch[inner] -= roots[k]
        first = (enc >> 3) - toff[k]
        ch[leaf] = -((first[leaf] << 3) | (enc[leaf] & 7)) - 2
        assert np.array_equal(sub, ob.nodes()), k
    # rays of all scenes, concatenated
    rays = [scenes.batch_rays(n, seed=200 + k) for k, n in enumerate(rays_n)]
    allr = np.concatenate(rays, axis=1)
    roff = np.concatenate([[0], np.cumsum(rays_n)]).astype(np.int32)
    total = int(roff[-1])
    R = to_dev(allr, dev)
    d_off = torch.from_numpy(roff).to(dev)
    for flags in (0, pkg.FLAG_GRAPH, pkg.FLAG_GRAPH):
        t = torch.full((total,), -1.0, device=dev)
        pid = torch.full((total,), -7, dtype=torch.int32, device=dev)
        u = torch.empty(total, device=dev)
        v = torch.empty(total, device=dev)
        b.trace_closest(total, d_off, R, t, pid, u, v, flags=flags)
        occ = torch.full((total,), 9, dtype=torch.uint8, device=dev)
        b.trace_any(total, d_off, R, occ, flags=flags)
        x = torch.ones(100, device=dev).sum()          # eager work between replays
        torch.cuda.synchronize()
        assert float(x) == 100.0
        tg, ig, ug, vg, og = (a.cpu().numpy() for a in (t, pid, u, v, occ))
        hits = 0
        for k in range(len(tris)):
            if rays_n[k] == 0:
                continue
            sl = slice(roff[k], roff[k + 1])
            ob = orc[k]
            ref = oracle.trace_closest(ob.nodes(), ob.leaf_tris(), ob.order(), rays[k])
            assert np.array_equal(ig[sl], ref["id"]), (flags, k)
            for name, g in (("t", tg), ("u", ug), ("v", vg)):
                assert np.array_equal(_bits(g[sl]), _bits(ref[name])), (flags, k, name)
            assert np.array_equal(og[sl].astype(bool), ref["id"] >= 0), (flags, k)
            hits += int((ref["id"] >= 0).sum())
        assert hits > 0.02 * total
    b.destroy()


def test_batch_errors(pkg, scenes):
    with pytest.raises(pkg.CudaBVHError):
        pkg.Batch([scenes.random_triangles(100, seed=1), scenes.random_triangles(3, seed=2)])
