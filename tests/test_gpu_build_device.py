"""cudaBVH_buildDevice: triangles gathered on the GPU from device vertex and
index buffers build exactly the tree that cudaBVH_build builds from the same
triangles given as a host triangle soup (node bytes, exports and hits)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid_mesh(n, seed):
    """An (n+1) x (n+1) vertex heightfield grid with shared vertices and
    2 n^2 indexed triangles, plus the equivalent triangle soup."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(-1.0, 1.0, n + 1, dtype=np.float32)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    Y = (0.2 * np.sin(3 * X) * np.cos(2 * Z) + 0.02 * rng.standard_normal(X.shape)).astype(np.float32)
    verts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    i = np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]
    a, b, c, d = i, i + 1, i + (n + 1), i + (n + 2)
    idx = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([b, d, c], -1).reshape(-1, 3)])
    rng.shuffle(idx)
    idx = idx.astype(np.int64)
    soup = verts[idx].reshape(-1, 9)
    return verts, idx, soup


def _same_structure(pkg, a, b):
    assert a.num_bvh4_nodes == b.num_bvh4_nodes and a.depth == b.depth
    assert np.array_equal(a.node_data(), b.node_data())
    assert np.array_equal(a.tri_data(), b.tri_data())
    assert np.array_equal(a.packed_tris(), b.packed_tris())


def _pinhole_hits(pkg, bvh, dev):
    import torch
    W = H = 96
    o, f, r, u = (0.0, 1.5, -2.5), (0.0, -0.45, 0.9), (1.0, 0.0, 0.0), (0.0, 0.9, 0.45)
    c = pkg.camera(np.array(o, np.float32), np.array(f, np.float32), np.array(r, np.float32),
                   np.array(u, np.float32), 0.6, 0.6)
    n = W * H
    t = torch.empty(n, device=dev)
    pid = torch.empty(n, dtype=torch.int32, device=dev)
    uu = torch.empty(n, device=dev)
    vv = torch.empty(n, device=dev)
    bvh.trace_pinhole_closest(W, H, c, t, pid, uu, vv)
    torch.cuda.synchronize()
    return t.cpu().numpy(), pid.cpu().numpy()


@pytest.mark.parametrize("index_bytes,stride", [(4, 12), (2, 24), (0, 12), (4, 32)])
def test_build_device_equals_host_build(pkg, dev, index_bytes, stride):
    import torch
    print(f"buildDevice index_bytes={index_bytes} stride={stride}", flush=True)
    verts, idx, soup = _grid_mesh(48, seed=index_bytes + stride)
    n = idx.shape[0]
    if index_bytes == 0:
        vbuf = soup.reshape(-1, 3)          # corners in triangle order
        ind = None
    else:
        vbuf = verts
        ind = torch.from_numpy(idx.astype(np.int32 if index_bytes == 4 else np.int16).reshape(-1)).to(dev)
    floats = stride // 4
    padded = np.zeros((vbuf.shape[0], floats), np.float32)
    padded[:, :3] = vbuf
    padded[:, 3:] = 7.0                      # other vertex attributes, ignored
    V = torch.from_numpy(padded.reshape(-1)).to(dev)
    a = pkg.BVH.build_device(V, n, indices=ind, vertex_stride=stride, num_vertices=vbuf.shape[0])
    b = pkg.BVH.build(soup)
    try:
        _same_structure(pkg, a, b)
        ta, ia = _pinhole_hits(pkg, a, dev)
        tb, ib = _pinhole_hits(pkg, b, dev)
        assert np.array_equal(ia, ib) and np.array_equal(ta.view(np.uint32), tb.view(np.uint32))
        assert (ia >= 0).mean() > 0.1
    finally:
        a.destroy()
        b.destroy()


def test_build_device_rejects_bad_indices(pkg, dev):
    import torch
    verts, idx, soup = _grid_mesh(8, seed=1)
    bad = idx.astype(np.int32).reshape(-1).copy()
    bad[17] = verts.shape[0]                 # one past the last vertex
    V = torch.from_numpy(verts.reshape(-1)).to(dev)
    with pytest.raises(pkg.CudaBVHError, match="numVertices"):
        pkg.BVH.build_device(V, idx.shape[0], indices=torch.from_numpy(bad).to(dev), num_vertices=verts.shape[0])
    with pytest.raises(pkg.CudaBVHError):
        pkg.BVH.build_device(V, idx.shape[0], indices=None, num_vertices=verts.shape[0])   # too few corners


def _geometry(verts, idx, index_bytes, stride, dev, soup=None):
    """Device vertex buffer (position first, `stride` bytes per vertex) and
    index buffer of one geometry."""
    import torch
    vb = soup.reshape(-1, 3) if index_bytes == 0 else verts
    nv = vb.shape[0]
    pad = np.zeros((nv, stride // 4), np.float32)
    pad[:, :3] = vb
    v = torch.from_numpy(pad.reshape(-1)).to(dev)
    ind = None
    if index_bytes == 2:
        ind = torch.from_numpy(idx.astype(np.int16).reshape(-1)).to(dev)
    elif index_bytes == 4:
        ind = torch.from_numpy(idx.astype(np.int32).reshape(-1)).to(dev)
    nt = idx.shape[0]
    return (v, ind, stride, nv, nt)


def test_build_many_device_equals_single_builds(pkg, oracle, dev):
    """cudaBVH_buildManyDevice: BLASes of one or several geometries (16/32-bit
    or no indices, strides 12-32 bytes, and a BLAS below the forest's
    minimum of five triangles) built in one pass; every handle equals
    cudaBVH_buildDevice / cudaBVH_build on the BLAS's triangles (nodes,
    exports, depth, bounds, stackless export) and the oracle's tree."""
    specs = [[(4, 12, 30)], [(2, 16, 20), (0, 12, 12)], [(0, 24, 1)], [(4, 32, 40), (2, 12, 8), (4, 16, 25)],
             [(2, 12, 6)]]
    blases, soups = [], []
    for k, spec in enumerate(specs):
        geoms, soup_k = [], []
        for j, (ib, stride, grid) in enumerate(spec):
            verts, idx, soup = _grid_mesh(grid, seed=100 * k + j)
            if grid == 1:
                idx, soup = idx[:1], soup[:1]                      # a one-triangle BLAS
            geoms.append(_geometry(verts, idx, ib, stride, dev, soup))
            soup_k.append(soup)
        blases.append(geoms)
        soups.append(np.concatenate(soup_k))
    many = pkg.BVH.build_many_device(blases)
    assert len(many) == len(specs)
    for k, (b, soup) in enumerate(zip(many, soups)):
        with pkg.BVH.build(soup) as ref:
            _same_structure(pkg, b, ref)
            assert np.array_equal(b.stackless_bvh2(), ref.stackless_bvh2())
            assert np.array_equal(b.leaf_prim_ids(), ref.leaf_prim_ids())
            assert b.bounds() == ref.bounds() and b.num_tris == soup.shape[0]
            t, pid = _pinhole_hits(pkg, b, dev)
            t2, pid2 = _pinhole_hits(pkg, ref, dev)
            assert np.array_equal(pid, pid2) and np.array_equal(t.view(np.uint32), t2.view(np.uint32))
        ob = oracle.OracleBVH(soup)
        assert np.array_equal(b.node_data(), ob.nodes()), k
        assert b.trace_status() == 0
        b.destroy()


def test_build_many_device_rejects_bad_indices(pkg, dev):
    import torch
    verts, idx, soup = _grid_mesh(8, seed=3)
    good = _geometry(verts, idx, 4, 12, dev)
    bad_idx = idx.copy()
    bad_idx[5, 1] = verts.shape[0]                                # one past the last vertex
    bad = _geometry(verts, bad_idx, 4, 12, dev)
    with pytest.raises(pkg.CudaBVHError):
        pkg.BVH.build_many_device([[good], [bad]])
    with pytest.raises(pkg.CudaBVHError):
        pkg.BVH.build_many_device([[(good[0], good[1], 10, good[3], good[4])]])   # stride below 12
    torch.cuda.synchronize()
