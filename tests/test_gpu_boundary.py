"""GPU tests of the reference boundary's shading and background-thread calls
and of the AABB build:

* traceToRGBA / traceToGPUPtr: every pixel's colour against a float32 numpy
  restatement of the reference's shading (layer/cuda_bvh_backend.cu:923-984:
  sky gradient on a miss; face normal turned toward the viewer, Lambert term
  toward the fixed light, ambient and rim) on the oracle's hits;
* traceToHostPtr / traceToGPUAsync (backend :1241-1336): the first call
  blocks until its frame is done; later calls return at once with the last
  finished frame while the background thread renders the new request; every
  frame equals traceToRGBA's / traceToGPUPtr's;
* buildFromAABBs (backend :1586-1610): one degenerate triangle per box, the
  tree byte-identical to the oracle's build of the same triangles."""
import ctypes as C
import time

import numpy as np
import pytest

from test_gpu_trace import _oracle_inputs

pytestmark = pytest.mark.gpu
f32 = np.float32


@pytest.fixture(scope="module")
def scene_bvh(pkg, scenes, dev):
    tris = scenes.random_triangles(65536, seed=0)
    b = pkg.BVH.build(tris)
    yield tris, b
    b.destroy()


def _lookat_rays(oracle, bvh, W, H, cam_o):
    """Rays of the look-at camera toward the bounds centre (backend :991-1016),
    float32 in the library's operation order."""
    cx, cy, cz, _ = bvh.bounds()
    o = np.array(cam_o, f32)
    fx, fy, fz = f32(cx) - o[0], f32(cy) - o[1], f32(cz) - o[2]
    fl = np.sqrt((fx * fx + fy * fy) + fz * fz)
    fx, fy, fz = fx / fl, fy / fl, fz / fl
    rx, rz = fz, -fx
    rl = np.sqrt(rx * rx + rz * rz)
    rx, rz = rx / rl, rz / rl
    ry = f32(0)
    ux, uy, uz = ry * fz - rz * fy, rz * fx - rx * fz, rx * fy - ry * fx
    cam14 = np.array([o[0], o[1], o[2], fx, fy, fz, rx, ry, rz, ux, uy, uz,
                      f32(W) / f32(H) * f32(0.6), f32(0.6)], f32)
    return oracle.gen_pinhole(cam14, W, H)


def _reference_shade(tris, ids, rays, W, H):
    """Float32 restatement of the reference's shading (backend :958-982);
    1/sqrt in IEEE arithmetic and the rim's cube as a product."""
    n = W * H
    py = (np.arange(n) // W).astype(f32)
    vv = py / f32(H)
    rgb = np.stack([(f32(20) + f32(10) * (f32(1) - vv)) / f32(255), (f32(30) + f32(20) * (f32(1) - vv)) / f32(255),
                    (f32(60) + f32(100) * vv) / f32(255)], 1).astype(f32)
    hit = ids >= 0
    T = tris[ids[hit]].astype(f32)
    dx, dy, dz = rays[3][hit], rays[4][hit], rays[5][hit]
    e1 = [T[:, 3 + k] - T[:, k] for k in range(3)]
    e2 = [T[:, 6 + k] - T[:, k] for k in range(3)]
    nx = e1[1] * e2[2] - e1[2] * e2[1]
    ny = e1[2] * e2[0] - e1[0] * e2[2]
    nz = e1[0] * e2[1] - e1[1] * e2[0]
    nl = f32(1) / np.sqrt(((nx * nx + ny * ny) + nz * nz) + f32(1e-20))
    nx, ny, nz = nx * nl, ny * nl, nz * nl
    flip = ((nx * dx + ny * dy) + nz * dz) > 0
    nx, ny, nz = np.where(flip, -nx, nx), np.where(flip, -ny, ny), np.where(flip, -nz, nz)
    lx, ly, lz = f32(0.5), f32(0.7), f32(0.5)
    ll = f32(1) / np.sqrt((lx * lx + ly * ly) + lz * lz)
    lx, ly, lz = lx * ll, ly * ll, lz * ll
    ndotl = np.maximum((nx * lx + ny * ly) + nz * lz, f32(0))
    ndotv = np.abs((nx * -dx + ny * -dy) + nz * -dz)
    k = f32(1) - ndotv
    shade = np.minimum((f32(0.15) + f32(0.65) * ndotl) + f32(0.2) * (k * k * k), f32(1))
    rgb[hit] = np.stack([shade * f32(0.92) + f32(0.08), shade * f32(0.88) + f32(0.07),
                         shade * f32(0.82) + f32(0.06)], 1)
    return rgb.reshape(H, W, 3), hit.reshape(H, W)


def test_shading_matches_reference_formula(pkg, oracle, scene_bvh, dev):
    import torch
    tris, bvh = scene_bvh
    W, H = 160, 96
    cam_o = (0.3, 0.2, -3.0)
    rays = _lookat_rays(oracle, bvh, W, H, cam_o)
    nodes, lt, pid = _oracle_inputs(bvh)
    ref = oracle.trace_closest(nodes, lt, pid, rays)
    want, hit = _reference_shade(tris, ref["id"], rays, W, H)
    assert 0.2 < hit.mean() < 0.95
    f = torch.empty(W * H * 4, device=dev)
    bvh.trace_to_gpu_ptr(W, H, cam_o, f.data_ptr(), 3)
    torch.cuda.synchronize()
    got = f.cpu().numpy().reshape(H, W, 4)
    assert (got[..., 3] == 1).all()
    assert np.array_equal(got[..., :3].view(np.uint32), want.view(np.uint32)), \
        f"{(got[..., :3] != want).any(-1).sum()} pixels differ"
    # 8-bit and half outputs are conversions of those colours
    img = bvh.trace_to_rgba(W, H, cam_o)
    q = (want * f32(255)).astype(np.uint8).astype(np.uint32)
    assert np.array_equal(img.view(np.uint32), (255 << 24) | (q[..., 2] << 16) | (q[..., 1] << 8) | q[..., 0])
    bg = torch.empty(W * H, dtype=torch.int32, device=dev)
    bvh.trace_to_gpu_ptr(W, H, cam_o, bg.data_ptr(), 1)
    h16 = torch.empty(W * H * 4, dtype=torch.float16, device=dev)
    bvh.trace_to_gpu_ptr(W, H, cam_o, h16.data_ptr(), 2)
    torch.cuda.synchronize()
    assert np.array_equal(bg.cpu().numpy().view(np.uint32).reshape(H, W),
                          (255 << 24) | (q[..., 0] << 16) | (q[..., 1] << 8) | q[..., 2])
    assert np.array_equal(h16.cpu().numpy().reshape(H, W, 4)[..., :3], want.astype(np.float16))
    assert bvh.trace_status() == 0


def _wait_for(fn, timeout_s=10.0):
    t0 = time.time()
    while time.time() - t0 < timeout_s:
        if fn():
            return True
        time.sleep(0.01)
    return False


def test_async_host_and_gpu_frames(pkg, scene_bvh, dev):
    import torch
    _, bvh = scene_bvh
    lib = pkg.load_library()
    W, H = 128, 80
    cams = [(0.3, 0.2, -3.0), (-0.5, 0.4, -2.5), (0.0, -0.3, -3.5)]
    refs = [bvh.trace_to_rgba(W, H, c).reshape(-1) for c in cams]
    assert not np.array_equal(refs[0], refs[1]) and not np.array_equal(refs[1], refs[2])
    out = np.zeros(W * H, np.uint32)
    size = W * H * 4

    def host(k):
        rc = lib.cudaBVH_traceToHostPtr(bvh.handle, W, H, *map(float, cams[k]), out.ctypes.data, 0, size)
        assert rc == 0
        return out.copy()

    print("async: host pointer frames", flush=True)
    assert np.array_equal(host(0), refs[0])                       # the first call blocks for its frame
    assert np.array_equal(host(1), refs[0])                       # later calls: the last finished frame
    assert _wait_for(lambda: np.array_equal(host(1), refs[1]))    # ... until the new one is done
    assert np.array_equal(host(2), refs[1])
    assert _wait_for(lambda: np.array_equal(host(2), refs[2]))
    small = np.zeros(10, np.uint32)
    assert lib.cudaBVH_traceToHostPtr(bvh.handle, W, H, 0.0, 0.0, -3.0, small.ctypes.data, 0, 40) == -1

    print("async: device pointer frames", flush=True)
    d = torch.zeros(W * H, dtype=torch.int32, device=dev)
    lib.cudaBVH_traceToGPUAsync(bvh.handle, W, H, *map(float, cams[1]), C.c_void_p(d.data_ptr()), 0)
    assert np.array_equal(d.cpu().numpy().view(np.uint32), refs[1])        # the first call blocks
    lib.cudaBVH_traceToGPUAsync(bvh.handle, W, H, *map(float, cams[0]), C.c_void_p(d.data_ptr()), 0)
    assert _wait_for(lambda: np.array_equal(d.cpu().numpy().view(np.uint32), refs[0]))
    f = torch.zeros(W * H * 4, device=dev)
    lib.cudaBVH_traceToGPUAsync(bvh.handle, W, H, *map(float, cams[2]), C.c_void_p(f.data_ptr()), 3)
    want = torch.empty(W * H * 4, device=dev)
    bvh.trace_to_gpu_ptr(W, H, cams[2], want.data_ptr(), 3)
    torch.cuda.synchronize()
    assert _wait_for(lambda: torch.equal(f, want))


def test_build_from_aabbs_matches_oracle(pkg, oracle, dev):
    rng = np.random.default_rng(17)
    n = 5000
    lo = rng.uniform(-20, 20, (n, 3)).astype(f32)
    boxes = np.concatenate([lo, lo + rng.uniform(0.01, 2.0, (n, 3)).astype(f32)], 1)
    with pkg.BVH.build_from_aabbs(boxes) as b:
        ob = oracle.OracleBVH(oracle.aabb_tris(boxes))
        assert b.num_tris == n and b.num_bvh4_nodes == ob.num_bvh4
        assert np.array_equal(b.node_data(), ob.nodes())
        assert np.array_equal(b.leaf_prim_ids(), ob.order())
        assert np.array_equal(b.stackless_bvh2(), ob.dfs())
        # the degenerate triangle of box i has exactly box i's bounds
        t = np.ascontiguousarray(b.tri_data().T).reshape(-1, 3, 3)
        order = b.leaf_prim_ids()
        assert np.array_equal(t.min(1), boxes[order, :3]) and np.array_equal(t.max(1), boxes[order, 3:])
    with pytest.raises(pkg.CudaBVHError):
        pkg.BVH.build_from_aabbs(np.zeros((0, 6), f32))
