"""GPU parity of the wavefront path tracer over a two-level scene: image,
first-hit buffers and per-bounce ray counts against the FP32 oracle (bit for
bit) and the FP64 brute-force estimator (relMSE), eager and hipGraph-replayed."""
import numpy as np
import pytest

from test_oracle_pt import rel_mse, tiny_pt_scene

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _render(pkg, gs, cam, p, dev, flags=0):
    import torch
    n = p.width * p.height
    rgb = torch.full((3 * n,), -1.0, device=dev)
    alb = torch.full((3 * n,), -1.0, device=dev)
    nrm = torch.full((3 * n,), -1.0, device=dev)
    dep = torch.full((n,), -1.0, device=dev)
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    cnt = gs.path_trace(c, p, rgb, alb, nrm, dep, counts=True, flags=flags)
    torch.cuda.synchronize()
    H, W = p.height, p.width
    aov = np.concatenate([alb.cpu().numpy().reshape(3, H, W), nrm.cpu().numpy().reshape(3, H, W),
                          dep.cpu().numpy().reshape(1, H, W)])
    return rgb.cpu().numpy().reshape(3, H, W), aov, cnt


def _scene(pkg, sc):
    blas = [pkg.BVH.build(m) for m in sc["meshes"]]
    return pkg.Scene(blas, sc["o2w"], sc["blas_of"], sc["albedo"])


# tail: paths at which one thread per path takes over (0: wavefront only; 200:
# from a later bounce on; 32768: this small image is a tail from the start)
# flags: 1 << 7 hipGraph replay, 1 << 5 octant sort of the extension rays
@pytest.mark.parametrize("spp,depth,rr,flags,tail", [(1, 8, 3, 0, 0), (4, 8, 3, 0, 0), (2, 3, 1, 0, 0),
                                                     (3, 8, 3, 1 << 7, 0), (2, 8, 3, 0, 200), (2, 8, 3, 0, 32768),
                                                     (3, 8, 3, 1 << 7, 200), (2, 8, 3, 1 << 5, 0),
                                                     (3, 8, 3, (1 << 7) | (1 << 5), 200), (1, 2, 1, 1 << 5, 0)])
def test_path_tracer_parity(pkg, oracle, scenes, dev, spp, depth, rr, flags, tail):
    print(f"pt spp={spp} depth={depth} rr={rr} flags={flags} tail={tail}", flush=True)
    sc, cam = tiny_pt_scene(scenes)
    gs = _scene(pkg, sc)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], gs.world_to_object())
    p = pkg.path_params(48, 32, spp=spp, max_depth=depth, rr_start=rr, seed=7, tail_paths=tail)
    rgb, aov, cnt = _render(pkg, gs, cam, p, dev, flags)
    op = oracle.path_params(48, 32, spp=spp, max_depth=depth, rr_start=rr, seed=7)
    ref, raov, rcnt = oracle.path_trace_2l(osc, sc["albedo"], scenes.cam_array(cam), op)
    assert np.array_equal(cnt, rcnt)
    assert np.array_equal(_bits(aov), _bits(raov))
    assert np.array_equal(_bits(rgb), _bits(ref))
    wt, wi, _ = scenes.instance_world_tris(sc)
    r64 = oracle.path_trace_f64(wt, wi, sc["albedo"], scenes.cam_array(cam), op)
    assert rel_mse(rgb, r64) < 1e-6
    gs.destroy()


@pytest.mark.parametrize("W,H", [(50, 37), (7, 5), (130, 9)])
def test_path_tracer_odd_sizes(pkg, oracle, scenes, dev, W, H):
    """Image sizes with partial 8x8 tiles and margins: the camera rays are
    queued in tile order, every pixel once, and the image still equals the
    oracle's bit for bit."""
    print(f"pt odd size {W}x{H}", flush=True)
    sc, cam = tiny_pt_scene(scenes, W, H)
    gs = _scene(pkg, sc)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], gs.world_to_object())
    p = pkg.path_params(W, H, spp=2, max_depth=4, rr_start=2, seed=5)
    rgb, aov, cnt = _render(pkg, gs, cam, p, dev)
    op = oracle.path_params(W, H, spp=2, max_depth=4, rr_start=2, seed=5)
    ref, raov, rcnt = oracle.path_trace_2l(osc, sc["albedo"], scenes.cam_array(cam), op)
    assert np.array_equal(cnt, rcnt)
    assert np.array_equal(_bits(aov), _bits(raov))
    assert np.array_equal(_bits(rgb), _bits(ref))
    gs.destroy()


def test_path_tracer_graph_replays(pkg, oracle, scenes, dev):
    """One captured sample replayed across calls with new seeds, with eager
    work and synchronisation in between; every call equals the oracle."""
    import torch
    sc = scenes.instanced_scene(mesh_res=(12, 20, 28, 36), n_inst=64, seed=9, span=6.0)
    cam = scenes.look_at((0.0, 4.0, -18.0), (0.0, 0.0, 0.0), 96, 64, 60.0)
    gs = _scene(pkg, sc)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], gs.world_to_object())
    for k, seed in enumerate((11, 12, 13)):
        p = pkg.path_params(96, 64, spp=2, seed=seed, tail_paths=(0, 1000, 32768)[k])
        rgb, aov, cnt = _render(pkg, gs, cam, p, dev, flags=pkg.FLAG_GRAPH)
        x = torch.arange(1000, device=dev, dtype=torch.float32).sum()     # eager work between replays
        torch.cuda.synchronize()
        assert float(x) == 499500.0
        op = oracle.path_params(96, 64, spp=2, seed=seed)
        ref, raov, rcnt = oracle.path_trace_2l(osc, sc["albedo"], scenes.cam_array(cam), op)
        assert np.array_equal(cnt, rcnt), k
        assert np.array_equal(_bits(rgb), _bits(ref)), k
        assert np.array_equal(_bits(aov), _bits(raov)), k
    # eager render of the last seed equals the graph render
    e_rgb, _, e_cnt = _render(pkg, gs, cam, pkg.path_params(96, 64, spp=2, seed=13, tail_paths=0), dev)
    assert np.array_equal(_bits(e_rgb), _bits(rgb)) and np.array_equal(e_cnt, cnt)
    gs.destroy()


def test_path_tracer_errors(pkg, scenes, dev):
    import torch
    sc, cam = tiny_pt_scene(scenes)
    gs = _scene(pkg, sc)
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    buf = torch.zeros(3 * 48 * 32, device=dev)
    with pytest.raises(pkg.CudaBVHError):
        gs.path_trace(c, pkg.path_params(48, 32, max_depth=17), buf, buf, buf, buf)
    with pytest.raises(pkg.CudaBVHError):
        gs.path_trace(c, pkg.path_params(48, 32, spp=0), buf, buf, buf, buf)
    with pytest.raises(pkg.CudaBVHError):
        gs.path_trace(c, pkg.path_params(48, 32, sun_dir=(0, 0, 0)), buf, buf, buf, buf)
    gs.destroy()


@pytest.mark.parametrize("configs", [((1, 8, 0), (3, 12, 0), (3, 8, 4)), ((0, 4, 0), (2, 16, 2), (1, 12, 0))])
def test_path_tracer_schedules(pkg, oracle, scenes, dev, configs):
    """The image is the same for every traversal schedule of the three scene
    ray classes (first-bounce closest hit, shadow any hit, later bounces),
    eager and replayed from a hipGraph."""
    c0, c1, c2 = configs
    print(f"pt schedules {configs}", flush=True)
    sc = scenes.instanced_scene(mesh_res=(12, 20, 28, 36), n_inst=64, seed=9, span=6.0)
    cam = scenes.look_at((0.0, 4.0, -18.0), (0.0, 0.0, 0.0), 96, 64, 60.0)
    gs = _scene(pkg, sc)
    gs.set_trace_config(0, *c0)
    gs.set_trace_config(1, *c1)
    gs.set_trace_config(2, *c2)
    assert gs.trace_config(2) == c2
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], gs.world_to_object())
    op = oracle.path_params(96, 64, spp=2, seed=21)
    ref, raov, rcnt = oracle.path_trace_2l(osc, sc["albedo"], scenes.cam_array(cam), op)
    for flags in (0, pkg.FLAG_GRAPH):
        p = pkg.path_params(96, 64, spp=2, seed=21, tail_paths=0)
        rgb, aov, cnt = _render(pkg, gs, cam, p, dev, flags=flags)
        assert np.array_equal(cnt, rcnt), flags
        assert np.array_equal(_bits(rgb), _bits(ref)), flags
        assert np.array_equal(_bits(aov), _bits(raov)), flags
    gs.destroy()


def test_path_tracer_tail_autotune(pkg, oracle, scenes, dev):
    """The tail threshold (queue size below which one thread per path takes
    over) is tuned on the frame; renders with tail_paths = -1 use the tuned
    value and still equal the oracle's image."""
    import torch
    sc = scenes.instanced_scene(mesh_res=(12, 20, 28, 36), n_inst=64, seed=9, span=6.0)
    cam = scenes.look_at((0.0, 4.0, -18.0), (0.0, 0.0, 0.0), 96, 64, 60.0)
    gs = _scene(pkg, sc)
    n = 96 * 64
    bufs = [torch.empty(3 * n, device=dev) for _ in range(3)] + [torch.empty(n, device=dev)]
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    best, table = gs.autotune_path_tracer(c, pkg.path_params(96, 64, spp=1, seed=3), *bufs, reps=2)
    assert [t for t, _ in table] == [0, 4096, 16384, 32768, 65536, 131072]
    assert best in [t for t, _ in table] and all(ms > 0 for _, ms in table)
    osc = oracle.OracleScene(sc["meshes"], sc["o2w"], sc["blas_of"], gs.world_to_object())
    op = oracle.path_params(96, 64, spp=2, seed=21)
    ref, _, rcnt = oracle.path_trace_2l(osc, sc["albedo"], scenes.cam_array(cam), op)
    for flags in (0, pkg.FLAG_GRAPH):
        rgb, _, cnt = _render(pkg, gs, cam, pkg.path_params(96, 64, spp=2, seed=21, tail_paths=-1), dev, flags)
        assert np.array_equal(cnt, rcnt) and np.array_equal(_bits(rgb), _bits(ref)), flags
    assert gs.trace_status() == 0
    gs.destroy()
