"""cudaBVH_traceClosestNoAlloc through the C ABI: the allocation-free closest
hit (static device scratch) returns exactly what cudaBVH_traceClosest returns,
on two calls in a row (the second with fewer rays, so stale scratch contents
from the first must not leak into it), for a shallow BVH and for a deeper one
whose stack spill caps the resident workgroups; unsupported flags are
rejected with an error."""
import ctypes as C

import numpy as np
import pytest

from conftest import hit_buffers, to_dev

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _trace(pkg, bvh, rays, dev, no_alloc, flags=0):
    import torch
    n = rays.shape[1]
    R = to_dev(rays, dev)
    Hb = hit_buffers(n, dev)
    for v in Hb.values():
        v.fill_(-3)
    rp = pkg.ray_packet(R)
    hp = pkg.hit_packet(Hb["t"], Hb["prim_id"], Hb["u"], Hb["v"])
    lib = pkg.load_library()
    fn = lib.cudaBVH_traceClosestNoAlloc if no_alloc else lib.cudaBVH_traceClosest
    rc = fn(bvh.handle, n, C.byref(rp), C.byref(hp), flags)
    torch.cuda.synchronize()
    assert rc == 0, pkg.last_error()
    return {k: v.cpu().numpy() for k, v in Hb.items()}


def _same(a, b):
    assert np.array_equal(a["prim_id"], b["prim_id"])
    for k in ("t", "u", "v"):
        assert np.array_equal(_bits(a[k]), _bits(b[k])), k


@pytest.mark.parametrize("kind", ["random", "terrain"])
def test_no_alloc_matches_trace_closest(pkg, oracle, scenes, dev, kind):
    print(f"no-alloc closest hit: {kind}", flush=True)
    if kind == "random":
        tris = scenes.random_triangles(65536, seed=0)
        rays = scenes.batch_rays(200000, seed=3)
    else:
        tris = scenes.terrain_scene(res=512, spheres=64, level=3)
        cam = scenes.terrain_camera(640, 360)
        rays = oracle.gen_pinhole(scenes.cam_array(cam), 640, 360)
    with pkg.BVH.build(tris) as bvh:
        for n in (rays.shape[1], rays.shape[1] // 3 + 17):        # second call: smaller input
            r = np.ascontiguousarray(rays[:, :n])
            ref = _trace(pkg, bvh, r, dev, no_alloc=False)
            got = _trace(pkg, bvh, r, dev, no_alloc=True, flags=pkg.FLAG_SYNC)
            _same(got, ref)
            assert (ref["prim_id"] >= 0).mean() > 0.05
        # Moller-Trumbore mode and an asynchronous call
        r = np.ascontiguousarray(rays[:, :5000])
        _same(_trace(pkg, bvh, r, dev, True, pkg.ISECT_REFERENCE_MT),
              _trace(pkg, bvh, r, dev, False, pkg.ISECT_REFERENCE_MT))
        assert bvh.trace_status() == 0


def test_no_alloc_rejects_unsupported_flags(pkg, scenes, dev):
    import torch
    tris = scenes.random_triangles(1000, seed=1)
    rays = scenes.batch_rays(64, seed=2)
    with pkg.BVH.build(tris) as bvh:
        lib = pkg.load_library()
        R = to_dev(rays, dev)
        Hb = hit_buffers(64, dev)
        rp, hp = pkg.ray_packet(R), pkg.hit_packet(Hb["t"], Hb["prim_id"], Hb["u"], Hb["v"])
        for flags in (pkg.FLAG_HOST_POINTERS, pkg.FLAG_SORT_RAYS, pkg.FLAG_CWBVH):
            assert lib.cudaBVH_traceClosestNoAlloc(bvh.handle, 64, C.byref(rp), C.byref(hp), flags) == -1
            assert "device packets only" in pkg.last_error()
        assert lib.cudaBVH_traceClosestNoAlloc(bvh.handle, 64, C.byref(rp), None, 0) == -1
        assert lib.cudaBVH_traceClosestNoAlloc(bvh.handle, 0, C.byref(rp), C.byref(hp), 0) == 0
        torch.cuda.synchronize()
