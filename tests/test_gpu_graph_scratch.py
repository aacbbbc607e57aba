"""Captured launches and the buffers they address: a hipGraph replayed after
the traversal scratch it was captured with has been reallocated must be
re-captured (the scratch generation is part of the graph key), each batch
mode keeps its own captured launch, and a ray sort never throws away the
device CWBVH. Every result is compared with the oracle bit for bit, and the
traversal status word reads 0 (no stack overflow) after each sequence."""
import numpy as np
import pytest

from conftest import to_dev
from test_gpu_pt import _bits, _render, _scene
from test_gpu_trace import _assert_closest_equal, _gpu_any, _gpu_closest, _oracle_inputs

pytestmark = pytest.mark.gpu


def test_path_trace_graph_after_autotune(pkg, oracle, scenes, dev):
    """pathTrace(GRAPH), sceneAutotune (its LDS-depth-4 trials grow the scene's
    spill region), pathTrace(GRAPH): the second render is re-captured and
    equals the eager render and the oracle."""
    import torch
    print("pt graph / autotune / graph", flush=True)
    sc = scenes.instanced_scene(mesh_res=(12, 20, 28, 36), n_inst=64, seed=9, span=6.0)
    cam = scenes.look_at((0.0, 4.0, -18.0), (0.0, 0.0, 0.0), 96, 64, 60.0)
    gs = _scene(pkg, sc)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], gs.world_to_object())
    op = oracle.path_params(96, 64, spp=2, seed=31)
    ref, raov, rcnt = oracle.path_trace_2l(osc, sc["albedo"], scenes.cam_array(cam), op)
    p = pkg.path_params(96, 64, spp=2, seed=31, tail_paths=0)
    rgb1, _, cnt1 = _render(pkg, gs, cam, p, dev, flags=pkg.FLAG_GRAPH)
    assert np.array_equal(_bits(rgb1), _bits(ref)) and np.array_equal(cnt1, rcnt)
    # rays for the autotuner: the camera's primaries
    rays = oracle.gen_pinhole(scenes.cam_array(cam), 96, 64)
    R = to_dev(rays, dev)
    for cls in (2, 1, 0):
        gs.autotune(cls, rays.shape[1], R, reps=1)
    torch.cuda.synchronize()
    rgb2, aov2, cnt2 = _render(pkg, gs, cam, p, dev, flags=pkg.FLAG_GRAPH)
    assert np.array_equal(_bits(rgb2), _bits(ref)) and np.array_equal(cnt2, rcnt)
    assert np.array_equal(_bits(aov2), _bits(raov))
    rgb3, _, _ = _render(pkg, gs, cam, p, dev, flags=pkg.FLAG_GRAPH)     # replay of the re-captured graph
    assert np.array_equal(_bits(rgb3), _bits(ref))
    rgbe, _, _ = _render(pkg, gs, cam, p, dev, flags=0)
    assert np.array_equal(_bits(rgbe), _bits(ref))
    assert gs.trace_status() == 0
    gs.destroy()


def _batch_check(oracle, orc, rays, roff, rays_n, t, pid, occ=None):
    tg, ig = t.cpu().numpy(), pid.cpu().numpy()
    for k, ob in enumerate(orc):
        if rays_n[k] == 0:
            continue
        sl = slice(roff[k], roff[k + 1])
        ref = oracle.trace_closest(ob.nodes(), ob.leaf_tris(), ob.order(), rays[k])
        assert np.array_equal(ig[sl], ref["id"]), k
        assert np.array_equal(_bits(tg[sl]), _bits(ref["t"])), k
        if occ is not None:
            assert np.array_equal(occ.cpu().numpy()[sl].astype(bool), ref["id"] >= 0), k


def test_batch_graph_per_mode_and_scratch_growth(pkg, oracle, scenes, dev):
    """closest(GRAPH) twice (the second a replay), any(eager) with a deeper
    spill region, closest(GRAPH) again (re-captured: the spill it addressed
    was freed), any(GRAPH) twice; all equal to the oracle."""
    import torch
    print("batch graph per mode", flush=True)
    tris_n, rays_n = [800, 5000, 12000, 3000], [4000, 9000, 20000, 6000]
    tris = [scenes.random_triangles(n, seed=300 + k) for k, n in enumerate(tris_n)]
    b = pkg.Batch(tris)
    orc = [oracle.OracleBVH(t) for t in tris]
    rays = [scenes.batch_rays(n, seed=400 + k) for k, n in enumerate(rays_n)]
    roff = np.concatenate([[0], np.cumsum(rays_n)]).astype(np.int32)
    total = int(roff[-1])
    R = to_dev(np.concatenate(rays, axis=1), dev)
    d_off = torch.from_numpy(roff).to(dev)

    def closest(flags):
        t = torch.full((total,), -1.0, device=dev)
        pid = torch.full((total,), -7, dtype=torch.int32, device=dev)
        b.trace_closest(total, d_off, R, t, pid, flags=flags)
        torch.cuda.synchronize()
        _batch_check(oracle, orc, rays, roff, rays_n, t, pid)

    def anyhit(flags):
        occ = torch.full((total,), 9, dtype=torch.uint8, device=dev)
        b.trace_any(total, d_off, R, occ, flags=flags)
        torch.cuda.synchronize()
        o = occ.cpu().numpy()
        for k, ob in enumerate(orc):
            sl = slice(roff[k], roff[k + 1])
            ref = oracle.trace_closest(ob.nodes(), ob.leaf_tris(), ob.order(), rays[k])
            assert np.array_equal(o[sl].astype(bool), ref["id"] >= 0), k

    closest(pkg.FLAG_GRAPH)
    closest(pkg.FLAG_GRAPH)
    forest = b.forest()
    forest.set_trace_config(pkg.CLASS_ANY, pkg.SCHEDULE_BATCH, 4, 0)     # more spill than the closest class
    anyhit(0)
    x = torch.arange(64, device=dev, dtype=torch.float32).sum()            # eager work between replays
    torch.cuda.synchronize()
    assert float(x) == 2016.0
    closest(pkg.FLAG_GRAPH)
    anyhit(pkg.FLAG_GRAPH)
    anyhit(pkg.FLAG_GRAPH)
    closest(pkg.FLAG_GRAPH | pkg.FLAG_SYNC)
    assert forest.trace_status() == 0
    b.destroy()


def test_cwbvh_survives_ray_sort(pkg, oracle, scenes, dev):
    """CWBVH trace, a sorted trace and a traceDiffuse that grow the sort
    workspace, CWBVH trace again: same hits, same CWBVH."""
    import torch
    print("cwbvh / sort growth / cwbvh", flush=True)
    tris = scenes.random_triangles(20000, seed=4)
    cam = scenes.parity_camera(64, 64)
    rays = oracle.gen_pinhole(scenes.cam_array(cam), 64, 64)
    with pkg.BVH.build(tris) as b:
        nn = b.build_cwbvh()
        nodes, lt, pid = _oracle_inputs(b)
        ref = oracle.trace_closest(nodes, lt, pid, rays)
        _assert_closest_equal(_gpu_closest(b, rays, dev, flags=pkg.FLAG_CWBVH), ref)
        big = scenes.batch_rays(200000, seed=5)
        refb = oracle.trace_closest(nodes, lt, pid, big)
        _assert_closest_equal(_gpu_closest(b, big, dev, flags=pkg.FLAG_SORT_RAYS), refb)
        o = np.ascontiguousarray(big[0:3].T)
        d = np.ascontiguousarray(big[3:6].T)
        more = np.concatenate([o, o]), np.concatenate([d, d])
        out = np.zeros(more[0].shape[0], np.float32)
        assert b.trace_diffuse(more[0], more[1], out) > 0
        assert b.num_cwbvh_nodes == nn
        nodes2, _, _ = b.cwbvh()
        _assert_closest_equal(_gpu_closest(b, rays, dev, flags=pkg.FLAG_CWBVH), ref)
        assert _gpu_any(b, rays, dev, flags=pkg.FLAG_CWBVH).astype(bool).tolist() == (ref["id"] >= 0).tolist()
        assert b.trace_status() == 0
        torch.cuda.synchronize()
