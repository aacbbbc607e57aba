"""GPU parity: the HIP traversal through the C ABI against the CPU oracle on
the same BVH (exported by the library) and the same FP32 rays. IDs, t, u, v
and the node/triangle visit counters must be bit-identical."""
import os

import numpy as np
import pytest

from conftest import hit_buffers, to_dev

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.npz")


def _oracle_inputs(bvh):
    return bvh.node_data(), np.ascontiguousarray(bvh.tri_data().T), bvh.leaf_prim_ids()


def _gpu_closest(bvh, rays, dev, flags=0, counted=False):
    import torch
    n = rays.shape[1]
    R = to_dev(rays, dev)
    H = hit_buffers(n, dev)
    if counted:
        cn = torch.empty(n, dtype=torch.int32, device=dev)
        ct = torch.empty(n, dtype=torch.int32, device=dev)
        bvh.trace_closest_counted(R, H["t"], H["prim_id"], H["u"], H["v"], cn, ct, flags)
    else:
        bvh.trace_closest(R, H["t"], H["prim_id"], H["u"], H["v"], flags)
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in H.items()}
    if counted:
        out["nodes"], out["tris"] = cn.cpu().numpy(), ct.cpu().numpy()
    return out


def _gpu_any(bvh, rays, dev, counted=False, flags=0):
    import torch
    n = rays.shape[1]
    R = to_dev(rays, dev)
    occ = torch.empty(n, dtype=torch.uint8, device=dev)
    if counted:
        cn = torch.empty(n, dtype=torch.int32, device=dev)
        ct = torch.empty(n, dtype=torch.int32, device=dev)
        bvh.trace_any_counted(R, occ, cn, ct, flags)
        torch.cuda.synchronize()
        return occ.cpu().numpy(), cn.cpu().numpy(), ct.cpu().numpy()
    bvh.trace_any(R, occ, flags)
    torch.cuda.synchronize()
    return occ.cpu().numpy()


def _same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def _assert_closest_equal(got, ref):
    assert np.array_equal(got["prim_id"], ref["id"]), f"{(got['prim_id'] != ref['id']).sum()} IDs differ"
    assert _same_bits(got["t"], ref["t"])
    assert _same_bits(got["u"], ref["u"]) and _same_bits(got["v"], ref["v"])


@pytest.fixture(scope="module")
def parity_scene(scenes):
    """The 65,536-triangle parity scene and its 512 x 512 pinhole rays."""
    tris = scenes.random_triangles(65536, seed=0)
    cam = scenes.parity_camera(512, 512)
    return tris, cam


@pytest.fixture(scope="module")
def parity_bvh(pkg, parity_scene, dev):
    tris, _ = parity_scene
    b = pkg.BVH.build(tris)
    yield b
    b.destroy()


def test_builder_matches_oracle(pkg, oracle, parity_scene, parity_bvh):
    tris, _ = parity_scene
    ob = oracle.OracleBVH(tris, max_leaf=4)
    assert np.array_equal(parity_bvh.node_data(), ob.nodes())
    assert np.array_equal(parity_bvh.leaf_prim_ids(), ob.order())
    assert np.array_equal(parity_bvh.stackless_bvh2(), ob.dfs())
    assert np.array_equal(parity_bvh.tri_data().T, tris[ob.order()])
    assert parity_bvh.num_tris == len(tris) and parity_bvh.num_bvh4_nodes == ob.num_bvh4
    assert parity_bvh.depth == ob.depth
    pk = parity_bvh.packed_tris().reshape(-1, 12)
    assert np.array_equal(pk[:, :9], tris[ob.order()])
    assert np.array_equal(pk[:, 9].view(np.int32), ob.order())   # ID bits, as the layer reads them
    cx, cy, cz, ext = parity_bvh.bounds()
    mn, mx = tris.reshape(-1, 3).min(0), tris.reshape(-1, 3).max(0)
    assert np.allclose([cx, cy, cz], (mn + mx) / 2) and np.isclose(ext, np.linalg.norm(mx - mn), rtol=1e-6)


def test_exports_feed_glsl_tracers(oracle, parity_bvh, dev):
    """The layer's GLSL compute tracers (bvh4_trace.comp, bvh2_stackless.comp;
    restated in the oracle) over this library's exports — getNodeData,
    getStacklessBVH2, getPackedTris — see the closest hits that the library's
    own kernel reports for the same rays (their Moller-Trumbore test and
    unsorted order aside, which move no hit on this scene)."""
    pc = np.array([0, 0, -3, 0, 0, 1, 1, 0, 0, 0, 1, 0, np.tan(np.radians(30.0))], np.float32)
    W, H = 256, 192
    rays = oracle.glsl_rays(pc, W, H)
    got = _gpu_closest(parity_bvh, rays, dev)
    assert 0.3 < (got["prim_id"] >= 0).mean() < 0.9
    order = parity_bvh.leaf_prim_ids()
    for kind, nodes in (("bvh4", parity_bvh.node_data()), ("stackless", parity_bvh.stackless_bvh2())):
        t, tri = oracle.glsl_trace(kind, nodes, parity_bvh.packed_tris(), pc, W, H)
        gid = np.where(tri >= 0, order[np.maximum(tri, 0)], -1)
        agree = gid == got["prim_id"]
        assert agree.mean() >= 0.999, (kind, agree.mean())
        both = agree & (gid >= 0)
        assert np.all(np.abs(t[both] - got["t"][both]) <= 1e-4 * got["t"][both]), kind


@pytest.mark.parametrize("isect", [0, 1])
def test_closest_parity_pinhole(oracle, scenes, parity_scene, parity_bvh, dev, isect):
    _, cam = parity_scene
    rays = oracle.gen_pinhole(scenes.cam_array(cam), 512, 512)
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    ref = oracle.trace_closest(nodes, lt, pid, rays, isect=isect)
    got = _gpu_closest(parity_bvh, rays, dev, flags=isect, counted=True)
    _assert_closest_equal(got, ref)
    assert np.array_equal(got["nodes"], ref["nodes"]) and np.array_equal(got["tris"], ref["tris"])
    assert (ref["id"] >= 0).mean() > 0.3
    got2 = _gpu_closest(parity_bvh, rays, dev, flags=isect)      # uncounted kernel, same answers
    _assert_closest_equal(got2, ref)


def test_pinhole_generation_and_fused_trace(pkg, oracle, scenes, parity_scene, parity_bvh, dev):
    import torch
    _, cam = parity_scene
    W = H = 512
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    ref_rays = oracle.gen_pinhole(scenes.cam_array(cam), W, H)
    R = {k: torch.empty(W * H, device=dev) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")}
    pkg.gen_pinhole_rays(W, H, c, R)
    torch.cuda.synchronize()
    gen = np.stack([R[k].cpu().numpy() for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")])
    assert _same_bits(gen, ref_rays)
    hb = hit_buffers(W * H, dev)
    parity_bvh.trace_pinhole_closest(W, H, c, hb["t"], hb["prim_id"], hb["u"], hb["v"])
    torch.cuda.synchronize()
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    ref = oracle.trace_closest(nodes, lt, pid, ref_rays)
    _assert_closest_equal({k: v.cpu().numpy() for k, v in hb.items()}, ref)


def test_any_hit_parity(oracle, parity_scene, parity_bvh, dev):
    rng = np.random.default_rng(8)
    n = 200_000
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = rng.uniform(-1.1, 1.1, (3, n))
    d = rng.normal(size=(3, n))
    rays[3:6] = d / np.linalg.norm(d, axis=0)
    rays[6] = 1e-3
    rays[7] = rng.uniform(0.01, 2.0, n)
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    ref = oracle.trace_any(nodes, lt, pid, rays)
    occ, cn, ct = _gpu_any(parity_bvh, rays, dev, counted=True)
    assert np.array_equal(occ, ref["occ"])
    assert np.array_equal(cn, ref["nodes"]) and np.array_equal(ct, ref["tris"])
    assert np.array_equal(_gpu_any(parity_bvh, rays, dev), ref["occ"])
    refc = oracle.trace_closest(nodes, lt, pid, rays)
    got = _gpu_closest(parity_bvh, rays, dev)
    _assert_closest_equal(got, refc)
    assert np.array_equal(occ.astype(bool), got["prim_id"] >= 0)   # any == (closest exists)
    assert 0.05 < occ.mean() < 0.95


def test_known_answers_on_gpu(pkg, oracle, dev):
    """Exact closed-form answers (by value: the exact answer 0 may come out
    as IEEE -0), and the oracle's answers bit for bit."""
    g = np.load(GOLD)
    with pkg.BVH.build(g["tris"]) as b:
        got = _gpu_closest(b, g["rays"], dev)
        assert np.array_equal(got["prim_id"], g["id"])
        assert np.array_equal(got["t"], g["t"]) and np.array_equal(got["u"], g["u"])
        assert np.array_equal(got["v"], g["v"])
        nodes, lt, pid = _oracle_inputs(b)
        _assert_closest_equal(got, oracle.trace_closest(nodes, lt, pid, g["rays"]))
        assert np.array_equal(_gpu_any(b, g["rays"], dev).astype(bool), g["id"] >= 0)


def test_axis_parallel_rays(pkg, oracle, dev):
    """Rays with zero direction components whose origins lie on box planes
    (the flat-ray slab path), mixed into waves with ordinary rays."""
    g = np.load(GOLD)
    tris = np.concatenate([g["tris"], g["tris"] + np.float32(0.5)])
    rng = np.random.default_rng(9)
    n = 4096
    rays = np.zeros((8, n), np.float32)
    grid = np.float32(0.25) * rng.integers(-8, 40, (2, n)).astype(np.float32)
    rays[0], rays[1], rays[2] = grid[0], grid[1], -1.0
    rays[5] = 1.0
    mixed = rng.random(n) < 0.3                               # some ordinary rays in the same waves
    d = rng.normal(size=(3, n)).astype(np.float32)
    d[2] = np.abs(d[2]) + 0.5
    rays[3:6, mixed] = (d / np.linalg.norm(d, axis=0))[:, mixed]
    rays[7] = 1e30
    with pkg.BVH.build(tris) as b:
        nodes, lt, pid = _oracle_inputs(b)
        ref = oracle.trace_closest(nodes, lt, pid, rays)
        bf = oracle.brute_closest(tris, np.arange(len(tris), dtype=np.int32), rays)
        assert np.array_equal(ref["id"], bf["id"])
        got = _gpu_closest(b, rays, dev, counted=True)
        _assert_closest_equal(got, ref)
        assert np.array_equal(got["nodes"], ref["nodes"])
        assert (ref["id"] >= 0).sum() > 100
        # uncounted launches of every explicit-ray schedule (batch, speculative
        # while-while, refill) on the same mixed waves
        occ_ref = oracle.trace_any(nodes, lt, pid, rays)["occ"]
        for sch in (0, 1, 3):
            b.set_trace_config(pkg.CLASS_INCOHERENT, sch, 8, 0)
            b.set_trace_config(pkg.CLASS_ANY, sch, 8, 0)
            _assert_closest_equal(_gpu_closest(b, rays, dev), ref)
            assert np.array_equal(_gpu_any(b, rays, dev), occ_ref)


def test_edge_cases(pkg, oracle, dev):
    import torch
    one = np.array([[0, 0, 1, 1, 0, 1, 0, 1, 1]], np.float32)
    with pkg.BVH.build(one) as b:
        assert b.num_bvh4_nodes == 1 and b.num_tris == 1
        rays = np.array([[0.25, 2.0], [0.25, 2.0], [0, 0], [0, 0], [0, 0], [1, 1], [0, 0], [1e30, 1e30]], np.float32)
        got = _gpu_closest(b, rays, dev)
        assert got["prim_id"].tolist() == [0, -1] and got["t"][0] == 1.0 and got["t"][1] == np.float32(1e30)
        # zero rays is a no-op
        R = to_dev(np.zeros((8, 0), np.float32), dev)
        H = hit_buffers(0, dev)
        b.trace_closest(R, H["t"], H["prim_id"], H["u"], H["v"])
    # coincident triangles: the lowest original ID wins
    many = np.tile(one, (37, 1))
    with pkg.BVH.build(many) as b:
        got = _gpu_closest(b, np.array([[0.25], [0.25], [0], [0], [0], [1], [0], [1e30]], np.float32), dev)
        assert got["prim_id"][0] == 0
    # invalid input: NULL handle and an empty build, reported through lastError
    with pytest.raises(pkg.CudaBVHError):
        pkg.BVH.build(np.zeros((0, 9), np.float32))
    assert "numTris" in pkg.last_error()


def test_host_pointer_flag(pkg, oracle, scenes, parity_bvh, dev):
    rays = oracle.gen_pinhole(scenes.cam_array(scenes.parity_camera(128, 128)), 128, 128)
    n = rays.shape[1]
    H = dict(t=np.zeros(n, np.float32), prim_id=np.zeros(n, np.int32), u=np.zeros(n, np.float32),
             v=np.zeros(n, np.float32))
    R = {k: np.ascontiguousarray(rays[i]) for i, k in enumerate(("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax"))}
    parity_bvh.trace_closest(R, H["t"], H["prim_id"], H["u"], H["v"], flags=pkg.FLAG_HOST_POINTERS)
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    _assert_closest_equal(H, oracle.trace_closest(nodes, lt, pid, rays))
    occ = np.zeros(n, np.uint8)
    parity_bvh.trace_any(R, occ, flags=pkg.FLAG_HOST_POINTERS)
    assert np.array_equal(occ, oracle.trace_any(nodes, lt, pid, rays)["occ"])


def test_reference_trace_primary(pkg, oracle, parity_scene, parity_bvh):
    """cudaBVH_tracePrimary: side x side rays from camO looking +z, 90 degree
    FOV (layer/cuda_bvh_backend.cu:69-74), hit distances row-major."""
    side = 256
    cam_o = (0.0, 0.0, -3.0)
    out = np.zeros(side * side, np.float32)
    mrs = parity_bvh.trace_primary(side, cam_o, out)
    assert mrs > 0
    cam14 = np.array([0, 0, -3, 0, 0, 1, 1, 0, 0, 0, 1, 0, 1, 1], np.float32)
    rays = oracle.gen_pinhole(cam14, side, side)
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    ref = oracle.trace_closest(nodes, lt, pid, rays)
    assert _same_bits(out, ref["t"])
    assert (out < 1e20).mean() > 0.1
    # later frames run in the cost order of the previous one: same distances
    for _ in range(2):
        again = np.full(side * side, -1.0, np.float32)
        assert parity_bvh.trace_primary(side, cam_o, again) > 0
        assert _same_bits(again, ref["t"])


def test_reference_trace_diffuse(pkg, oracle, parity_bvh):
    """cudaBVH_traceDiffuse: host SoA rays in, host distances out, t in
    (0, 1e30) (layer/cuda_bvh_backend.cu:594). The calls run back to back on
    one handle with growing and shrinking ray counts, so the host staging is
    regrown and then reused with a stale tail; a tracePrimary in between
    shares the staging."""
    rng = np.random.default_rng(12)
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    for n in (3_000, 50_000, 1_000, 120_000, 7):
        o = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
        d = rng.normal(size=(n, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        out = np.full(n, np.nan, np.float32)
        assert parity_bvh.trace_diffuse(o, d, out) > 0
        rays = np.zeros((8, n), np.float32)
        rays[0:3], rays[3:6], rays[6], rays[7] = o.T, d.T, 0.0, 1e30
        assert _same_bits(out, oracle.trace_closest(nodes, lt, pid, rays)["t"]), n
        if n == 50_000:
            side = 600   # 360,000 distances: more than the staging holds after this call
            prim = np.zeros(side * side, np.float32)
            assert parity_bvh.trace_primary(side, (0.0, 0.0, -3.0), prim) > 0
            cam14 = np.array([0, 0, -3, 0, 0, 1, 1, 0, 0, 0, 1, 0, 1, 1], np.float32)
            ref = oracle.trace_closest(nodes, lt, pid, oracle.gen_pinhole(cam14, side, side))
            assert _same_bits(prim, ref["t"])


def test_shading_formats(pkg, oracle, parity_bvh, dev):
    """traceToRGBA / traceToGPUPtr: misses get the sky gradient, hits the
    Lambert+rim shade (layer/cuda_bvh_backend.cu:923-984); the hit mask and
    colours match a numpy restatement on the oracle's hits."""
    import torch
    W, H = 160, 96
    cam_o = (0.3, 0.2, -3.0)
    img = parity_bvh.trace_to_rgba(W, H, cam_o)
    f32 = torch.empty(W * H * 4, device=dev)
    parity_bvh.trace_to_gpu_ptr(W, H, cam_o, f32.data_ptr(), 3)
    f16 = torch.empty(W * H * 4, dtype=torch.float16, device=dev)
    parity_bvh.trace_to_gpu_ptr(W, H, cam_o, f16.data_ptr(), 2)
    bgra = torch.empty(W * H, dtype=torch.int32, device=dev)
    parity_bvh.trace_to_gpu_ptr(W, H, cam_o, bgra.data_ptr(), 1)
    torch.cuda.synchronize()
    rgb = f32.cpu().numpy().reshape(H, W, 4)
    # camera of backend :991-1016
    cx, cy, cz, _ = parity_bvh.bounds()
    fw = np.array([cx, cy, cz], np.float32) - np.array(cam_o, np.float32)
    fw = fw / np.float32(np.sqrt(np.float32(fw @ fw)))
    rt = np.array([fw[2], 0, -fw[0]], np.float32)
    rt = rt / np.float32(np.sqrt(rt[0] ** 2 + rt[2] ** 2))
    up = np.cross(rt, fw).astype(np.float32)
    cam14 = np.concatenate([np.array(cam_o, np.float32), fw, rt, up, np.array([W / H * 0.6, 0.6], np.float32)])
    rays = oracle.gen_pinhole(cam14.astype(np.float32), W, H)
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    ref = oracle.trace_closest(nodes, lt, pid, rays)
    hit = (ref["id"] >= 0).reshape(H, W)
    py = np.arange(H, dtype=np.float32)[:, None] / np.float32(H)
    sky = np.stack([(20 + 10 * (1 - py)) / 255, (30 + 20 * (1 - py)) / 255, (60 + 100 * py) / 255], -1)
    sky = np.broadcast_to(sky, (H, W, 3))
    assert np.allclose(rgb[~hit][:, :3], sky[~hit], atol=1e-6)
    assert hit.mean() > 0.2
    assert (rgb[hit][:, 0] >= 0.08).all() and (rgb[..., 3] == 1).all()
    assert np.allclose(f16.cpu().numpy().reshape(H, W, 4), rgb, atol=2e-3)
    r8 = (img & 0xFF).astype(np.int32)
    assert np.array_equal(r8, (rgb[..., 0] * 255).astype(np.uint8).astype(np.int32))
    b8 = (bgra.cpu().numpy().reshape(H, W).view(np.uint32) & 0xFF).astype(np.int32)
    assert np.array_equal(b8, (rgb[..., 2] * 255).astype(np.uint8).astype(np.int32))
    # the display calls take their tiles in the cost order of the previous
    # frame of the same size: later frames are the same image bit for bit
    for _ in range(2):
        again = torch.full((W * H * 4,), -1.0, device=dev)
        parity_bvh.trace_to_gpu_ptr(W, H, cam_o, again.data_ptr(), 3)
        torch.cuda.synchronize()
        assert np.array_equal(again.cpu().numpy().view(np.uint32), f32.cpu().numpy().view(np.uint32))


def test_determinism_and_stream(pkg, oracle, scenes, parity_bvh, dev):
    import torch
    rays = oracle.gen_pinhole(scenes.cam_array(scenes.parity_camera(256, 256)), 256, 256)
    s = torch.cuda.Stream()
    parity_bvh.set_stream(s.cuda_stream)
    try:
        with torch.cuda.stream(s):
            a = _gpu_closest(parity_bvh, rays, dev)
            b = _gpu_closest(parity_bvh, rays, dev)
    finally:
        parity_bvh.set_stream(None)
    _assert_closest_equal(a, {"id": b["prim_id"], "t": b["t"], "u": b["u"], "v": b["v"]})


@pytest.mark.parametrize("schedule", [0, 1, 2, 3])
@pytest.mark.parametrize("depth", [4, 8, 16])
def test_every_launch_configuration(pkg, oracle, scenes, parity_bvh, dev, schedule, depth):
    """Each wave schedule (batch, while-while, packet, refill) and LDS stack
    depth (4 spills most entries to HBM) gives the oracle's answers for all
    three ray classes."""
    import torch
    W = H = 192
    cam = scenes.parity_camera(W, H)
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    rng = np.random.default_rng(30 + depth)
    n = 60_000
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = rng.uniform(-1.1, 1.1, (3, n))
    d = rng.normal(size=(3, n))
    rays[3:6] = d / np.linalg.norm(d, axis=0)
    rays[6] = 1e-3
    rays[7] = rng.uniform(0.05, 3.0, n)
    try:
        for cls in (pkg.CLASS_COHERENT, pkg.CLASS_INCOHERENT, pkg.CLASS_ANY):
            # the packet schedule exists for the coherent class only
            sch = schedule if (schedule != 2 or cls == pkg.CLASS_COHERENT) else 0
            parity_bvh.set_trace_config(cls, sch, depth, 0 if depth != 8 else 2)
        hb = hit_buffers(W * H, dev)
        parity_bvh.trace_pinhole_closest(W, H, c, hb["t"], hb["prim_id"], hb["u"], hb["v"])
        torch.cuda.synchronize()
        ref = oracle.trace_closest(nodes, lt, pid, oracle.gen_pinhole(scenes.cam_array(cam), W, H))
        _assert_closest_equal({k: v.cpu().numpy() for k, v in hb.items()}, ref)
        _assert_closest_equal(_gpu_closest(parity_bvh, rays, dev), oracle.trace_closest(nodes, lt, pid, rays))
        assert np.array_equal(_gpu_any(parity_bvh, rays, dev), oracle.trace_any(nodes, lt, pid, rays)["occ"])
    finally:
        parity_bvh.set_trace_config(pkg.CLASS_COHERENT, 0, 12, 0)
        parity_bvh.set_trace_config(pkg.CLASS_INCOHERENT, 1, 8, 0)
        parity_bvh.set_trace_config(pkg.CLASS_ANY, 0, 12, 0)


def test_launch_info_reports_configuration(pkg, parity_scene, parity_bvh, dev):
    """The last launch's resources follow the class's launch configuration:
    workgroups per CU as configured (or as many as fit), 256-thread
    workgroups on every CU, no scratch."""
    import torch
    _, cam = parity_scene
    c = pkg.camera(*[np.asarray(cam[k], np.float32) for k in ("o", "f", "r", "u")], cam["scale_x"], cam["scale_y"])
    W = H = 128
    t = torch.empty(W * H, device=dev)
    pid = torch.empty(W * H, dtype=torch.int32, device=dev)
    u = torch.empty(W * H, device=dev)
    v = torch.empty(W * H, device=dev)
    props = torch.cuda.get_device_properties(0)
    saved = parity_bvh.trace_config(pkg.CLASS_COHERENT)
    for bpc in (2, 0):
        parity_bvh.set_trace_config(pkg.CLASS_COHERENT, pkg.SCHEDULE_BATCH, 8, bpc)
        parity_bvh.trace_pinhole_closest(W, H, c, t, pid, u, v)
        torch.cuda.synchronize()
        info = parity_bvh.launch_info()
        assert info["vgprs"] > 0 and info["scratch_bytes"] == 0 and info["threads_per_workgroup"] == 256
        assert 1 <= info["blocks_per_cu"] <= info["max_blocks_per_cu"]
        if bpc:
            assert info["blocks_per_cu"] == bpc
        else:
            assert info["blocks_per_cu"] == info["max_blocks_per_cu"]
        assert info["workgroups"] == info["blocks_per_cu"] * props.multi_processor_count
        assert info["waves_per_simd"] == info["blocks_per_cu"]
    parity_bvh.set_trace_config(pkg.CLASS_COHERENT, *saved)


@pytest.mark.parametrize("W,H", [(1, 1), (9, 3), (17, 9), (67, 45), (130, 72)])
def test_fused_pinhole_odd_sizes(pkg, oracle, scenes, parity_scene, parity_bvh, dev, W, H):
    """Image sizes that leave partial 8x8 tiles, a partial last tile strip,
    fewer tiles than XCD bands or a single pixel: every schedule writes
    every pixel exactly as the oracle traces it, and nothing past W*H; also
    with the cost order (the batch schedule's second cost-ordered launch
    takes the order of the first one's step counts; other schedules ignore
    the flag)."""
    import torch
    _, cam = parity_scene
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    ref_rays = oracle.gen_pinhole(scenes.cam_array(cam), W, H)
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    ref = oracle.trace_closest(nodes, lt, pid, ref_rays)
    saved = parity_bvh.trace_config(pkg.CLASS_COHERENT)
    for sched in (pkg.SCHEDULE_BATCH, pkg.SCHEDULE_WHILE_WHILE, pkg.SCHEDULE_REFILL, pkg.SCHEDULE_PACKET):
        parity_bvh.set_trace_config(pkg.CLASS_COHERENT, sched, 8, 0)
        for flags in (0, pkg.FLAG_COST_ORDER, pkg.FLAG_COST_ORDER):
            hb = {k: torch.full((W * H + 64,), -3, dtype=torch.int32 if k == "prim_id" else torch.float32, device=dev)
                  for k in ("t", "prim_id", "u", "v")}
            parity_bvh.trace_pinhole_closest(W, H, c, hb["t"], hb["prim_id"], hb["u"], hb["v"], flags=flags)
            torch.cuda.synchronize()
            got = {k: v.cpu().numpy() for k, v in hb.items()}
            _assert_closest_equal({k: v[:W * H] for k, v in got.items()}, ref)
            assert (got["prim_id"][W * H:] == -3).all(), (sched, flags)
    parity_bvh.set_trace_config(pkg.CLASS_COHERENT, *saved)


def test_refill_gathered_results_with_evictions(pkg, oracle, dev):
    """The refill kernel gathers closest-hit results per 64-ray work chunk and
    holds two chunks per wave. Here one ray in 64 crosses a dense triangle
    soup while the other 63 miss the scene at the root, so a wave runs through
    many chunks while each chunk's long ray is still walking: older chunks are
    evicted and their long rays write their results directly. Every result
    equals the oracle's, for both stack depths of the refill kernel."""
    rng = np.random.default_rng(44)
    nt = 20000
    c = rng.uniform(-1, 1, (nt, 1, 3)).astype(np.float32)
    tris = (c + rng.normal(scale=0.08, size=(nt, 3, 3))).astype(np.float32).reshape(nt, 9)
    n = 64 * 4096 + 13          # with one workgroup per CU: about four chunks per wave
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = np.array([[0.0], [0.0], [-5.0]], np.float32)      # short rays: look away from the soup
    rays[5] = -1.0
    rays[7] = 1e30
    long_ = np.arange(n) % 64 == 17
    k = int(long_.sum())
    o = rng.uniform(-0.9, 0.9, (k, 3)).astype(np.float32)
    o[:, 2] = -3.0
    d = rng.normal(size=(k, 3)) * 0.02 + np.array([0.0, 0.0, 1.0])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[0:3, long_] = o.T
    rays[3:6, long_] = d.T.astype(np.float32)
    with pkg.BVH.build(tris) as b:
        nodes, lt, pid = _oracle_inputs(b)
        ref = oracle.trace_closest(nodes, lt, pid, rays)
        assert (ref["id"][long_] >= 0).mean() > 0.9 and (ref["id"][~long_] < 0).all()
        for depth in (8, 12):
            b.set_trace_config(pkg.CLASS_INCOHERENT, pkg.SCHEDULE_REFILL, depth, 1)
            _assert_closest_equal(_gpu_closest(b, rays, dev), ref)


@pytest.mark.parametrize("n", [1, 7, 64, 65, 511, 4097])
def test_closest_small_ray_counts(oracle, parity_scene, parity_bvh, dev, n):
    """Ray counts below, at and just past the 64-ray chunk and the eight
    XCD bands: every ray is traced once, as the oracle traces it."""
    _, cam = parity_scene
    rng = np.random.default_rng(n)
    rays = np.zeros((8, n), np.float32)
    d = rng.normal(size=(3, n)); d /= np.linalg.norm(d, axis=0)
    rays[0:3] = rng.uniform(-1.5, 1.5, (3, n)); rays[3:6] = d
    rays[6] = 0.0; rays[7] = 1e30
    nodes, lt, pid = _oracle_inputs(parity_bvh)
    ref = oracle.trace_closest(nodes, lt, pid, rays)
    _assert_closest_equal(_gpu_closest(parity_bvh, rays, dev), ref)
