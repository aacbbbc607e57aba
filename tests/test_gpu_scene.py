"""GPU parity of two-level scenes (TLAS over instances of BLAS meshes):
TLAS build, world-to-object maps, closest-hit (t, instance, prim, u, v and
visit counters) and any-hit, direct and queued, against the oracle."""
import numpy as np
import pytest

from conftest import RAY_KEYS, to_dev
from test_oracle_scene import scene_rays, w2o_of

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _make(pkg, sc):
    blas = [pkg.BVH.build(m) for m in sc["meshes"]]
    return pkg.Scene(blas, sc["o2w"], sc["blas_of"], sc["albedo"])


@pytest.mark.parametrize("seed,res,n_inst", [(5, (6, 11, 17, 23), 24), (7, (20, 31, 45, 64), 64)])
def test_scene_closest_any_parity(pkg, oracle, scenes, dev, seed, res, n_inst):
    import torch
    sc = scenes.instanced_scene(mesh_res=res, n_inst=n_inst, seed=seed, span=6.0)
    print(f"scene seed={seed} res={res} inst={n_inst}", flush=True)
    gs = _make(pkg, sc)
    w2o = gs.world_to_object()
    ref_inv = w2o_of(sc["o2w"])
    assert np.allclose(w2o, ref_inv, rtol=1e-6, atol=1e-6)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], w2o)
    # the GPU-built TLAS equals the oracle's, byte for byte
    assert np.array_equal(gs.tlas().node_data(), osc.tlas_nodes)
    n = 40000
    rays = scene_rays(n, seed + 1, span=6.0)
    ref = osc.trace_closest(rays)
    assert (ref["id"] >= 0).mean() > 0.2
    R = to_dev(rays, dev)
    t = torch.empty(n, device=dev)
    inst = torch.empty(n, dtype=torch.int32, device=dev)
    pid = torch.empty(n, dtype=torch.int32, device=dev)
    u = torch.empty(n, device=dev)
    v = torch.empty(n, device=dev)
    cn = torch.empty(n, dtype=torch.int32, device=dev)
    ct = torch.empty(n, dtype=torch.int32, device=dev)
    gs.trace_closest_counted(R, t, inst, pid, u, v, cn, ct)
    torch.cuda.synchronize()
    assert np.array_equal(inst.cpu().numpy(), ref["inst"])
    assert np.array_equal(pid.cpu().numpy(), ref["id"])
    for k, g in (("t", t), ("u", u), ("v", v)):
        assert np.array_equal(_bits(g.cpu().numpy()), _bits(ref[k])), k
    assert np.array_equal(cn.cpu().numpy(), ref["nodes"])
    assert np.array_equal(ct.cpu().numpy(), ref["tris"])
    # world normals of the hits (misses give zero vectors), bit for bit
    nrm = [torch.empty(n, device=dev) for _ in range(3)]
    gs.hit_normals(inst, pid, *nrm)
    torch.cuda.synchronize()
    nref = osc.hit_normals(ref["inst"], ref["id"])
    for k in range(3):
        assert np.array_equal(_bits(nrm[k].cpu().numpy()), _bits(nref[k])), k
    # uncounted kernel: same hits
    t2 = torch.empty(n, device=dev)
    pid2 = torch.empty(n, dtype=torch.int32, device=dev)
    gs.trace_closest(R, t2, None, pid2)
    torch.cuda.synchronize()
    assert np.array_equal(pid2.cpu().numpy(), ref["id"]) and np.array_equal(_bits(t2.cpu().numpy()), _bits(ref["t"]))
    # any hit with finite lengths
    rng = np.random.default_rng(seed)
    rays[7] = rng.uniform(0.5, 40.0, n).astype(np.float32)
    occ_ref = osc.trace_any(rays)
    R = to_dev(rays, dev)
    occ = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    gs.trace_any(R, occ)
    torch.cuda.synchronize()
    assert np.array_equal(occ.cpu().numpy(), occ_ref)
    # queued: only the first count rays are traced, nothing written past them
    m = 12345
    cnt = torch.tensor([m], dtype=torch.int32, device=dev)
    occ.fill_(7)
    gs.trace_any_queued(n, cnt, R, occ)
    tq = torch.full((n,), -5.0, device=dev)
    iq = torch.full((n,), -5, dtype=torch.int32, device=dev)
    gs.trace_closest_queued(n, cnt, to_dev(scene_rays(n, seed + 1, span=6.0), dev), tq, iq)
    torch.cuda.synchronize()
    o = occ.cpu().numpy()
    assert np.array_equal(o[:m], occ_ref[:m]) and (o[m:] == 7).all()
    q = tq.cpu().numpy()
    assert np.array_equal(_bits(q[:m]), _bits(ref["t"][:m])) and (q[m:] == -5.0).all()
    assert np.array_equal(iq.cpu().numpy()[:m], ref["inst"][:m])
    gs.destroy()


def test_scene_edge_cases(pkg, oracle, scenes, dev):
    import torch
    sc = scenes.instanced_scene(mesh_res=(5, 7), n_inst=3, seed=11, span=2.0)
    gs = _make(pkg, sc)
    # zero rays is a no-op
    z = torch.empty(0, device=dev)
    gs.trace_closest({k: z for k in RAY_KEYS}, z)
    # rays pointing away from everything miss; axis-parallel rays are handled
    n = 4096
    rays = scene_rays(n, 3, span=2.0)
    rays[3:6, : n // 4] = -rays[0:3, : n // 4] / np.linalg.norm(rays[0:3, : n // 4], axis=0)
    rays[3:6, : n // 4] *= -1.0      # outward, from three times farther out
    rays[0:3, : n // 4] *= 3.0
    ax = slice(n // 4, n // 2)
    rays[3:6, ax] = 0.0
    rays[3 + (np.arange(n // 4) % 3), np.arange(n // 4, n // 2)] = np.where(np.arange(n // 4) % 2, 1.0, -1.0)
    rays[0:3, ax] = np.round(rays[0:3, ax])          # origins on integer planes
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], gs.world_to_object())
    ref = osc.trace_closest(rays)
    R = to_dev(rays, dev)
    t = torch.empty(n, device=dev)
    inst = torch.empty(n, dtype=torch.int32, device=dev)
    pid = torch.empty(n, dtype=torch.int32, device=dev)
    gs.trace_closest(R, t, inst, pid)
    torch.cuda.synchronize()
    assert (ref["id"][: n // 4] == -1).all()
    assert np.array_equal(pid.cpu().numpy(), ref["id"]) and np.array_equal(inst.cpu().numpy(), ref["inst"])
    assert np.array_equal(_bits(t.cpu().numpy()), _bits(ref["t"]))
    gs.destroy()
    # bad arguments are reported, not crashed on
    with pytest.raises(pkg.CudaBVHError):
        pkg.Scene([pkg.BVH.build(sc["meshes"][0])], sc["o2w"][:1], [3])


def test_scene_every_launch_configuration(pkg, oracle, scenes, dev):
    """Every schedule (while-while, batch, batch nearest-first for any hit,
    refill) and a range of LDS depths gives the oracle's hits; the autotuner
    keeps a configuration from its own table."""
    import torch
    sc = scenes.instanced_scene(mesh_res=(6, 11, 17, 23), n_inst=24, seed=5, span=6.0)
    gs = _make(pkg, sc)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], gs.world_to_object())
    n = 20000
    rays = scene_rays(n, 9, span=6.0)
    rng = np.random.default_rng(1)
    srays = rays.copy()
    srays[7] = rng.uniform(0.5, 40.0, n).astype(np.float32)
    ref = osc.trace_closest(rays)
    occ_ref = osc.trace_any(srays)
    R, S = to_dev(rays, dev), to_dev(srays, dev)
    for sched in (0, 1, 2, 3):
        for depth in (4, 16):
            print(f"scene config schedule={sched} lds={depth}", flush=True)
            if sched != 2:
                gs.set_trace_config(0, sched, depth, 0)
                t = torch.empty(n, device=dev)
                inst = torch.empty(n, dtype=torch.int32, device=dev)
                pid = torch.empty(n, dtype=torch.int32, device=dev)
                u = torch.empty(n, device=dev)
                v = torch.empty(n, device=dev)
                gs.trace_closest(R, t, inst, pid, u, v)
                torch.cuda.synchronize()
                assert np.array_equal(pid.cpu().numpy(), ref["id"]) and np.array_equal(inst.cpu().numpy(), ref["inst"])
                for k, g in (("t", t), ("u", u), ("v", v)):
                    assert np.array_equal(_bits(g.cpu().numpy()), _bits(ref[k])), (sched, depth, k)
            gs.set_trace_config(1, sched, depth, 4)
            occ = torch.full((n,), 9, dtype=torch.uint8, device=dev)
            gs.trace_any(S, occ)
            torch.cuda.synchronize()
            assert np.array_equal(occ.cpu().numpy(), occ_ref), (sched, depth)
    with pytest.raises(pkg.CudaBVHError):
        gs.set_trace_config(0, 2, 8, 0)          # nearest-first ordering is an any-hit schedule
    best, table = gs.autotune(1, n, S, reps=1)
    assert len(table) == 48 and best in [(s_, d_, b_) for s_, d_, b_, _ in table]
    # class 2 (the path tracer's secondary bounces) tunes closest-hit traces
    best2, table2 = gs.autotune(2, n, R, reps=1)
    assert len(table2) == 36 and {s_ for s_, _, _, _ in table2} == {0, 1, 3}
    assert gs.trace_config(2) == best2 and best2 in [(s_, d_, b_) for s_, d_, b_, _ in table2]
    gs.destroy()
