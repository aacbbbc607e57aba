"""GPU parity at the workloads' own sizes, against the CPU oracle.

* Terrain frame: the 3,407,872-triangle terrain (1024 x 1024-cell heightfield
  + 256 level-4 icospheres) at 1920 x 1080 — the GPU build byte-identical to
  the oracle's, then every ray of the frame's three classes bit-exact: fused
  pinhole primaries (closest hit, with node / triangle visit counts), the
  cosine-weighted diffuse bounce and the shadow ray of every hit (secondary
  rays equal the numpy restatement), diffuse closest hit and shadow any hit
  under the tuned launch schedules, the ray-sorted trace, and the
  reference's host-pointer calls; the cost-ordered primary launches equal
  the plain ones.
* Instanced path tracer: 64 instances of 4 meshes (10K / 50K / 200K / 500K
  triangles) at 1920 x 1080, 1 spp, depth 8, Russian roulette from ray 3 —
  image, first-hit buffers and per-bounce ray counts bit-exact against the
  oracle's path tracer, eager and replayed from a hipGraph.
* 256 scenes in one launch: every tree of the forest byte-equal to the
  oracle's tree of its scene, and every one of the ~87 M rays' hits bit-exact.

Every test also reads the traversal status word (no stack overflow)."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import RAY_KEYS
from test_gpu_pt import _render, _scene

pytestmark = pytest.mark.gpu

W, H = 1920, 1080


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _eq(a, b):
    return np.array_equal(_bits(a), _bits(b))


@pytest.fixture(scope="module")
def terrain(pkg, oracle, scenes, dev):
    print("full-size terrain: generate, GPU build, oracle build", flush=True)
    tris = scenes.terrain_scene()
    assert tris.shape[0] == 3407872
    bvh = pkg.BVH.build(tris)
    ob = oracle.OracleBVH(tris)
    yield tris, bvh, ob
    bvh.destroy()


def test_terrain_build_matches_oracle(terrain):
    tris, bvh, ob = terrain
    assert bvh.num_bvh4_nodes == ob.num_bvh4 and bvh.depth == ob.depth
    assert np.array_equal(bvh.node_data(), ob.nodes())
    assert np.array_equal(bvh.leaf_prim_ids(), ob.order())
    assert np.array_equal(np.ascontiguousarray(bvh.tri_data().T), ob.leaf_tris())


def test_terrain_frame_parity(pkg, oracle, scenes, dev, terrain):
    """Primary, diffuse and shadow rays of the 1080p frame, every ray."""
    import torch
    import bounce_ref
    tris, bvh, ob = terrain
    nodes, lt, pid = ob.nodes(), ob.leaf_tris(), ob.order()
    cam = scenes.terrain_camera(W, H)
    cam14 = scenes.cam_array(cam)
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    n = W * H
    print("terrain frame: primaries", flush=True)
    rays = oracle.gen_pinhole(cam14, W, H)
    ref = oracle.trace_closest(nodes, lt, pid, rays)
    pt = torch.empty(n, device=dev)
    pi = torch.empty(n, dtype=torch.int32, device=dev)
    pu = torch.empty(n, device=dev)
    pv = torch.empty(n, device=dev)
    bvh.trace_pinhole_closest(W, H, c, pt, pi, pu, pv)                  # tuned coherent schedule
    torch.cuda.synchronize()
    ht, hid = pt.cpu().numpy(), pi.cpu().numpy()
    assert np.array_equal(hid, ref["id"]), f"{(hid != ref['id']).sum()} primary IDs differ"
    assert _eq(ht, ref["t"]) and _eq(pu.cpu().numpy(), ref["u"]) and _eq(pv.cpu().numpy(), ref["v"])
    # counted launch: node visits and triangle tests equal the oracle's
    R = {k: torch.from_numpy(np.ascontiguousarray(rays[i])).to(dev) for i, k in enumerate(RAY_KEYS)}
    cn = torch.empty(n, dtype=torch.int32, device=dev)
    ct = torch.empty(n, dtype=torch.int32, device=dev)
    t2, u2, v2 = (torch.empty(n, device=dev) for _ in range(3))
    i2 = torch.empty(n, dtype=torch.int32, device=dev)
    bvh.trace_closest_counted(R, t2, i2, u2, v2, cn, ct)
    torch.cuda.synchronize()
    assert np.array_equal(i2.cpu().numpy(), ref["id"]) and _eq(t2.cpu().numpy(), ref["t"])
    assert np.array_equal(cn.cpu().numpy(), ref["nodes"]) and np.array_equal(ct.cpu().numpy(), ref["tris"])
    del R, cn, ct, t2, u2, v2, i2

    print("terrain frame: bounce rays", flush=True)
    D = {k: torch.full((n,), -7.0, device=dev) for k in RAY_KEYS}
    S = {k: torch.full((n,), -7.0, device=dev) for k in RAY_KEYS}
    src = torch.full((n,), -1, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    bvh.gen_bounce_rays(W, H, c, pt, pi, 1234, scenes.TERRAIN_LIGHT, D, S, src, cnt)
    torch.cuda.synchronize()
    m = int(cnt.item())
    dref, sref, srcref = bounce_ref.bounce_rays(cam14, W, H, ht, hid, tris, 1234, scenes.TERRAIN_LIGHT)
    assert m == dref.shape[1] == int((hid >= 0).sum())
    assert np.array_equal(src.cpu().numpy()[:m], srcref)
    assert _eq(np.stack([D[k].cpu().numpy()[:m] for k in RAY_KEYS]), dref)
    assert _eq(np.stack([S[k].cpu().numpy()[:m] for k in RAY_KEYS]), sref)

    print("terrain frame: diffuse and shadow traces", flush=True)
    rd = oracle.trace_closest(nodes, lt, pid, dref, counts=False)
    rs = oracle.trace_any(nodes, lt, pid, sref, counts=False)
    for flags in (0, pkg.FLAG_SORT_RAYS):
        dt = torch.empty(n, device=dev)
        di = torch.empty(n, dtype=torch.int32, device=dev)
        du = torch.empty(n, device=dev)
        dv = torch.empty(n, device=dev)
        if flags:
            Dm = {k: v[:m] for k, v in D.items()}
            bvh.trace_closest(Dm, dt[:m], di[:m], du[:m], dv[:m], flags=flags)
        else:
            bvh.trace_closest_queued(n, cnt, D, dt, di, du, dv)            # tuned incoherent schedule
        occ = torch.full((n,), 9, dtype=torch.uint8, device=dev)
        bvh.trace_any_queued(n, cnt, S, occ)                                # tuned shadow schedule
        torch.cuda.synchronize()
        gi = di.cpu().numpy()[:m]
        assert np.array_equal(gi, rd["id"]), f"{(gi != rd['id']).sum()} diffuse IDs differ (flags {flags})"
        assert _eq(dt.cpu().numpy()[:m], rd["t"]) and _eq(du.cpu().numpy()[:m], rd["u"])
        assert _eq(dv.cpu().numpy()[:m], rd["v"])
        assert np.array_equal(occ.cpu().numpy()[:m], rs["occ"])
    assert 0.2 < (rd["id"] >= 0).mean() < 0.6 and 0.05 < rs["occ"].mean() < 0.3

    print("terrain frame: host-pointer calls", flush=True)
    o_h = np.ascontiguousarray(dref[0:3].T)
    d_h = np.ascontiguousarray(dref[3:6].T)
    out = np.zeros(m, np.float32)
    assert bvh.trace_diffuse(o_h, d_h, out) > 0
    r0 = dref.copy()
    r0[6] = 0.0                                          # traceDiffuse traces (0, 1e30)
    rr = oracle.trace_closest(nodes, lt, pid, r0, counts=False)
    assert _eq(out, rr["t"])
    side = 1440
    hit = np.zeros(side * side, np.float32)
    assert bvh.trace_primary(side, (0.0, 2.0, -12.0), hit) > 0
    # the reference's primary camera: looking +z, 90 degree FOV, square
    cam90 = np.array([0.0, 2.0, -12.0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 1, 1], np.float32)
    rp = oracle.trace_closest(nodes, lt, pid, oracle.gen_pinhole(cam90, side, side), counts=False)
    assert _eq(hit, rp["t"])
    assert bvh.trace_status() == 0



def test_terrain_cost_ordered_primaries(pkg, oracle, scenes, dev, terrain):
    """CUDABVH_FLAG_COST_ORDER only reorders the 8x8 tiles: three launches in
    a row (the first in the plain order, the next two in the order of the
    previous launch's step counts), then a ragged image size (its costs
    start again from zero) and the 1080p frame again, each into outputs
    filled with a sentinel first, equal the oracle's hits (or the plain-order
    trace at the ragged size) bit for bit, so every tile is traced exactly
    once."""
    import torch
    tris, bvh, ob = terrain
    nodes, lt, pid = ob.nodes(), ob.leaf_tris(), ob.order()

    def run(w, h, flags):
        cam = scenes.terrain_camera(w, h)
        c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
        out = [torch.full((w * h,), -7.0, device=dev), torch.full((w * h,), -7, dtype=torch.int32, device=dev),
               torch.full((w * h,), -7.0, device=dev), torch.full((w * h,), -7.0, device=dev)]
        bvh.trace_pinhole_closest(w, h, c, *out, flags=flags)
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in out], scenes.cam_array(cam)

    def same(a, b):
        return np.array_equal(a[1], b[1]) and all(_eq(a[k], b[k]) for k in (0, 2, 3))

    print("cost-ordered primaries: 1080p x 3", flush=True)
    first, cam14 = run(W, H, 0)
    ref = oracle.trace_closest(nodes, lt, pid, oracle.gen_pinhole(cam14, W, H), counts=False)
    assert same(first, [ref["t"], ref["id"], ref["u"], ref["v"]])
    for k in range(3):
        got, _ = run(W, H, pkg.FLAG_COST_ORDER)
        assert same(got, first), f"cost-ordered launch {k} differs"
    print("cost-ordered primaries: ragged 1001 x 563, then 1080p", flush=True)
    plain, _ = run(1001, 563, 0)
    for k in range(2):
        got, _ = run(1001, 563, pkg.FLAG_COST_ORDER)
        assert same(got, plain)
    got, _ = run(W, H, pkg.FLAG_COST_ORDER)
    assert same(got, first)
    assert bvh.trace_status() == 0

def test_instanced_path_tracer_full_frame(pkg, oracle, scenes, dev):
    """The instanced path tracer's 1080p frame, bit-exact, eager and graph."""
    print("full-size path tracer: scene", flush=True)
    sc = scenes.instanced_scene()
    assert [m.shape[0] for m in sc["meshes"]] == [10082, 49928, 199712, 500000]
    gs = _scene(pkg, sc)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], gs.world_to_object())
    cam = scenes.instanced_camera(W, H)
    op = oracle.path_params(W, H, spp=1, max_depth=8, rr_start=3, seed=1)
    ref, raov, rcnt = oracle.path_trace_2l(osc, sc["albedo"], scenes.cam_array(cam), op)
    assert rcnt[0] == W * H and 0 < rcnt[7] < rcnt[3] < rcnt[1]          # uneven per-bounce populations
    for flags in (0, pkg.FLAG_GRAPH, pkg.FLAG_GRAPH):
        print(f"full-size path tracer: render flags {flags}", flush=True)
        p = pkg.path_params(W, H, spp=1, max_depth=8, rr_start=3, seed=1)
        rgb, aov, cnt = _render(pkg, gs, cam, p, dev, flags=flags)
        assert np.array_equal(cnt, rcnt), flags
        assert _eq(aov, raov), flags
        assert _eq(rgb, ref), f"{(_bits(rgb) != _bits(ref)).sum()} image values differ (flags {flags})"
    assert gs.trace_status() == 0
    gs.destroy()


def test_batched_256_scenes(pkg, oracle, scenes, dev):
    """The 256-scene forest: every tree byte-equal to its oracle tree, every
    ray's closest hit (ID, t, u, v) bit-exact and any hit consistent."""
    import torch
    specs = scenes.batch_specs(256)
    print(f"256 scenes: {sum(t for t, _ in specs)} triangles, {sum(r for _, r in specs)} rays", flush=True)
    tris = [scenes.random_triangles(nt, seed=5000 + k) for k, (nt, _) in enumerate(specs)]
    b = pkg.Batch(tris)
    roots = b.roots()
    forest = b.forest()
    nodes = forest.node_data()
    toff = b.tri_offsets
    ends = list(roots[1:]) + [nodes.shape[0]]
    threads = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(threads) as ex:                 # the oracle's builds release the GIL
        orc = list(ex.map(lambda t: oracle.OracleBVH(t), tris))
    print("256 scenes: trees", flush=True)
    for k, ob in enumerate(orc):
        sub = nodes[roots[k]:ends[k]].copy()
        ch = sub[:, 48:].view(np.int32)
        inner, leaf = ch >= 0, ch < -1
        enc = -(ch + 2)
        ch[inner] -= roots[k]
        first = (enc >> 3) - toff[k]
        ch[leaf] = -((first[leaf] << 3) | (enc[leaf] & 7)) - 2
        assert np.array_equal(sub, ob.nodes()), k
    R, roff = scenes.batch_workload_device(specs, dev)
    total = int(roff[-1])
    d_off = torch.from_numpy(roff).to(dev)
    t = torch.empty(total, device=dev)
    pid = torch.full((total,), -7, dtype=torch.int32, device=dev)
    u = torch.empty(total, device=dev)
    v = torch.empty(total, device=dev)
    b.trace_closest(total, d_off, R, t, pid, u, v, flags=pkg.FLAG_GRAPH)
    b.trace_closest(total, d_off, R, t, pid, u, v, flags=pkg.FLAG_GRAPH)        # a replay
    occ = torch.full((total,), 9, dtype=torch.uint8, device=dev)
    b.trace_any(total, d_off, R, occ, flags=pkg.FLAG_GRAPH)
    torch.cuda.synchronize()
    print("256 scenes: rays", flush=True)
    tg, ig, ug, vg, og = (a.cpu().numpy() for a in (t, pid, u, v, occ))
    hits = 0
    for k, ob in enumerate(orc):
        sl = slice(int(roff[k]), int(roff[k + 1]))
        rays = np.stack([R[key][sl].cpu().numpy() for key in RAY_KEYS])
        ref = oracle.trace_closest(ob.nodes(), ob.leaf_tris(), ob.order(), rays, counts=False)
        assert np.array_equal(ig[sl], ref["id"]), (k, int((ig[sl] != ref["id"]).sum()))
        assert _eq(tg[sl], ref["t"]) and _eq(ug[sl], ref["u"]) and _eq(vg[sl], ref["v"]), k
        assert np.array_equal(og[sl].astype(bool), ref["id"] >= 0), k
        hits += int((ref["id"] >= 0).sum())
    assert 0.2 < hits / total < 0.6
    assert forest.trace_status() == 0
    b.destroy()
