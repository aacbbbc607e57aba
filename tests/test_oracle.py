"""CPU tests that pin the oracle: closed forms, brute force, library routines
(numpy's IEEE float16), invariants. No GPU."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.npz")


# ---------------------------------------------------------------- FP16 --------
def test_f16_rne_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.normal(0, 10, 20000), rng.normal(0, 1e-5, 5000), rng.uniform(-7e4, 7e4, 5000),
                         [0.0, -0.0, 65504.0, 65519.99, 65520.0, -65520.0, 2 ** -24, 2 ** -25, 3 * 2 ** -26,
                          2 ** -14, 1e30, -1e30, np.inf, -np.inf]]).astype(np.float32)
    for x in xs:
        h = oracle.f2h(float(x), "rne")
        assert h == int(np.float16(x).view(np.uint16)), x
        assert oracle.h2f(h) == float(np.float16(x)) or np.isnan(x)


def test_f16_outward_rounding(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.normal(0, 20, 20000), rng.uniform(-1e-3, 1e-3, 2000)]).astype(np.float32)
    for x in xs:
        lo, hi = oracle.f2h(float(x), "down"), oracle.f2h(float(x), "up")
        flo, fhi = oracle.h2f(lo), oracle.h2f(hi)
        assert flo <= x <= fhi
        if float(np.float16(x)) == x:
            assert flo == fhi == x
        else:   # adjacent halves: nothing representable strictly between
            assert np.nextafter(np.float16(flo), np.float16(np.inf)) == np.float16(fhi)


# ---------------------------------------------------------------- ray/tri -----
def test_known_answers_watertight(oracle):
    g = np.load(GOLD)
    r = oracle.brute_closest(g["tris"], np.arange(len(g["tris"]), dtype=np.int32), g["rays"])
    assert np.array_equal(r["id"], g["id"])
    assert np.array_equal(r["t"], g["t"])
    assert np.array_equal(r["u"], g["u"]) and np.array_equal(r["v"], g["v"])


def test_known_answers_through_bvh(oracle):
    g = np.load(GOLD)
    b = oracle.OracleBVH(g["tris"], max_leaf=1)
    r = oracle.trace_closest(b.nodes(), b.leaf_tris(), b.order(), g["rays"])
    assert np.array_equal(r["id"], g["id"]) and np.array_equal(r["t"], g["t"])
    a = oracle.trace_any(b.nodes(), b.leaf_tris(), b.order(), g["rays"])
    assert np.array_equal(a["occ"].astype(bool), g["id"] >= 0)


def test_watertight_closed_mesh_sweep(oracle, scenes):
    """A closed icosphere: every ray from the inside through the surface hits
    (watertightness), including rays aimed exactly at shared vertices."""
    T = scenes.icosphere(2).astype(np.float32)
    tris = T.reshape(-1, 9)
    verts = np.unique(T.reshape(-1, 3), axis=0)
    rng = np.random.default_rng(3)
    d = rng.normal(size=(4000, 3))
    d = np.concatenate([d / np.linalg.norm(d, axis=1, keepdims=True), verts.astype(np.float64)])
    n = d.shape[0]
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = 0.0
    rays[3:6] = d.T
    rays[7] = 1e30
    r = oracle.brute_closest(tris, np.arange(len(tris), dtype=np.int32), rays)
    assert (r["id"] >= 0).all()
    # Moller-Trumbore with the reference's epsilon is not watertight in
    # general; the oracle still reports the hit rate for comparison.
    m = oracle.brute_closest(tris, np.arange(len(tris), dtype=np.int32), rays, isect=1)
    assert (m["id"] >= 0).mean() > 0.99


def test_mt_matches_closed_form(oracle):
    ray = np.array([0.25, 0.125, -1, 0, 0, 1, 0, 1e30], np.float32)
    tri = np.array([0, 0, 2, 1, 0, 2, 1, 1, 2], np.float32)
    hit, t, u, v = oracle.isect(ray, tri, mode=1)
    assert hit and t == 3.0 and u == 0.125 and v == 0.125
    hit, t, u, v = oracle.isect(ray, tri, mode=0)
    assert hit and t == 3.0 and u == 0.125 and v == 0.125


def test_wrong_sign_is_caught(oracle):
    """Planting a mistake (a flipped direction) must change the answer: the
    known-answer set distinguishes it."""
    g = np.load(GOLD)
    rays = g["rays"].copy()
    rays[3:6] *= -1.0
    r = oracle.brute_closest(g["tris"], np.arange(len(g["tris"]), dtype=np.int32), rays)
    assert not np.array_equal(r["id"], g["id"])


# ---------------------------------------------------------------- BVH ---------
@pytest.fixture(scope="module")
def small_scene(scenes):
    return scenes.random_triangles(6000, seed=11)


def test_bvh_invariants(oracle, small_scene):
    tris = small_scene
    b = oracle.OracleBVH(tris, max_leaf=4)
    order = b.order()
    assert sorted(order.tolist()) == list(range(len(tris)))           # every primitive exactly once
    boxes, links = b.bvh2()
    for i in range(len(links)):                                       # children inside parents
        l, r, ts, tc = links[i]
        if tc == 0:
            for c in (l, r):
                assert (boxes[c, :3] >= boxes[i, :3]).all() and (boxes[c, 3:] <= boxes[i, 3:]).all()
                assert c > i                                          # pre-order
        else:
            assert 1 <= tc <= 4
            tb = tris[order[ts:ts + tc]].reshape(-1, 3, 3)
            assert (tb.min(axis=(0, 1)) >= boxes[i, :3]).all() and (tb.max(axis=(0, 1)) <= boxes[i, 3:]).all()
    nodes = b.nodes()
    seen = np.zeros(len(tris), np.int32)
    for nd in nodes:                                                  # FP16 child boxes hold their triangles
        h = nd[:48].view(np.float16).astype(np.float32).reshape(3, 8)
        ch = nd[48:].view(np.int32)
        for c in range(4):
            if ch[c] == -1:
                continue
            lo, hi = h[:, c], h[:, 4 + c]
            if ch[c] <= -2:
                enc = -(ch[c] + 2)
                ts, tc = enc >> 3, (enc & 7) + 1
                tb = tris[order[ts:ts + tc]].reshape(-1, 3)
                seen[order[ts:ts + tc]] += 1
            else:
                cn = nodes[ch[c]]
                ch2 = cn[48:].view(np.int32)
                hh = cn[:48].view(np.float16).astype(np.float32).reshape(3, 8)
                tb = np.concatenate([np.stack([hh[:, k], hh[:, 4 + k]]) for k in range(4) if ch2[k] != -1])
            assert (tb.min(axis=0) >= lo).all() and (tb.max(axis=0) <= hi).all()
    assert (seen == 1).all()


def test_bvh_traversal_matches_brute_force(oracle, small_scene, scenes):
    tris = small_scene
    b = oracle.OracleBVH(tris, max_leaf=4)
    cam = scenes.parity_camera(96, 96)
    rays = oracle.gen_pinhole(scenes.cam_array(cam), 96, 96)
    ids = np.arange(len(tris), dtype=np.int32)
    for mode in (0, 1):
        r = oracle.trace_closest(b.nodes(), b.leaf_tris(), b.order(), rays, isect=mode)
        bf = oracle.brute_closest(tris, ids, rays, isect=mode)
        assert np.array_equal(r["id"], bf["id"])
        assert np.array_equal(r["t"], bf["t"]) and np.array_equal(r["u"], bf["u"])
        assert (r["id"] >= 0).mean() > 0.1
    # FP64 Moller-Trumbore with independent arithmetic: same IDs, t to 1e-5
    f64 = oracle.brute_closest_f64(tris, ids, rays)
    r = oracle.trace_closest(b.nodes(), b.leaf_tris(), b.order(), rays)
    assert (f64["id"] == r["id"]).mean() > 0.999
    both = (f64["id"] >= 0) & (f64["id"] == r["id"])
    assert np.max(np.abs(f64["t"][both] - r["t"][both]) / f64["t"][both]) < 1e-5


def _f64_hit(tris, ids, rays):
    """FP64 Moller-Trumbore t, u, v of the given triangles along the given
    rays, with each quantity's first-order FP32 rounding scale
    (eps * |a| |b| / |det| for a quantity computed as a.b / det)."""
    T = np.asarray(tris, np.float64).reshape(-1, 3, 3)[ids]
    o, d = rays[0:3].T.astype(np.float64), rays[3:6].T.astype(np.float64)
    e1, e2 = T[:, 1] - T[:, 0], T[:, 2] - T[:, 0]
    p = np.cross(d, e2)
    det = (e1 * p).sum(1)
    s = o - T[:, 0]
    q = np.cross(s, e1)
    nrm = lambda a: np.linalg.norm(a, axis=1)
    eps = float(np.finfo(np.float32).eps) / np.abs(det)
    return dict(t=(e2 * q).sum(1) / det, u=(s * p).sum(1) / det, v=(d * q).sum(1) / det,
                st=eps * nrm(e2) * nrm(q), su=eps * nrm(s) * nrm(p), sv=eps * nrm(d) * nrm(q))


def test_hit_attributes_match_fp64(oracle, small_scene, scenes):
    """t, u and v of the oracle's FP32 hits against FP64 Moller-Trumbore on
    the same triangle: each within 8x its rounding scale (measured: at most
    2.4x over these rays), which a swapped u/v, a barycentric taken for the
    wrong vertex or a t measured from the wrong point exceeds by orders of
    magnitude. Watertight t is also within 1e-6 relative on both ray sets."""
    tris = small_scene
    b = oracle.OracleBVH(tris, max_leaf=4)
    cam = scenes.parity_camera(96, 96)
    pin = oracle.gen_pinhole(scenes.cam_array(cam), 96, 96)
    rng = np.random.default_rng(5)
    n = 4000
    rnd = np.zeros((8, n), np.float32)
    rnd[0:3] = rng.uniform(-1.2, 1.2, (3, n))
    d = rng.normal(size=(3, n))
    rnd[3:6] = d / np.linalg.norm(d, axis=0)
    rnd[6], rnd[7] = 1e-3, 3.0
    for rays in (pin, rnd):
        for mode in (0, 1):
            r = oracle.trace_closest(b.nodes(), b.leaf_tris(), b.order(), rays, isect=mode)
            h = r["id"] >= 0
            assert h.sum() > 400
            f = _f64_hit(tris, r["id"][h], rays[:, h])
            for k in "tuv":
                assert (np.abs(r[k][h] - f[k]) <= 8 * f["s" + k]).all(), (mode, k)
            if mode == 0:
                assert np.max(np.abs(r["t"][h] - f["t"]) / f["t"]) < 1e-6
            # the barycentrics name vertices 1 and 2: swapping them fails
            assert not (np.abs(r["v"][h] - f["u"]) <= 8 * f["su"]).all()


def test_incoherent_rays_and_any_hit(oracle, small_scene):
    tris = small_scene
    b = oracle.OracleBVH(tris, max_leaf=3)
    rng = np.random.default_rng(5)
    n = 4000
    rays = np.zeros((8, n), np.float32)
    rays[0:3] = rng.uniform(-1.2, 1.2, (3, n))
    d = rng.normal(size=(3, n))
    rays[3:6] = d / np.linalg.norm(d, axis=0)
    rays[6] = 1e-3
    rays[7] = rng.uniform(0.05, 3.0, n)
    r = oracle.trace_closest(b.nodes(), b.leaf_tris(), b.order(), rays)
    bf = oracle.brute_closest(tris, np.arange(len(tris), dtype=np.int32), rays)
    assert np.array_equal(r["id"], bf["id"]) and np.array_equal(r["t"], bf["t"])
    a = oracle.trace_any(b.nodes(), b.leaf_tris(), b.order(), rays)
    assert np.array_equal(a["occ"].astype(bool), r["id"] >= 0)
    assert 0.05 < a["occ"].mean() < 0.95
    assert (a["nodes"] <= r["nodes"] + 64).all()


def test_degenerate_inputs(oracle):
    one = np.array([[0, 0, 1, 1, 0, 1, 0, 1, 1]], np.float32)
    b = oracle.OracleBVH(one, max_leaf=4)
    assert b.num_bvh4 == 1
    rays = np.array([[0.25], [0.25], [0], [0], [0], [1], [0], [1e30]], np.float32)
    r = oracle.trace_closest(b.nodes(), b.leaf_tris(), b.order(), rays)
    assert r["id"][0] == 0 and r["t"][0] == 1.0
    same = np.tile(one, (50, 1))                       # coincident centroids: median split
    b = oracle.OracleBVH(same, max_leaf=4)
    r = oracle.trace_closest(b.nodes(), b.leaf_tris(), b.order(), rays)
    assert r["id"][0] == 0                             # tie among 50 identical: lowest ID
    with pytest.raises(ValueError):
        oracle.OracleBVH(np.zeros((0, 9), np.float32))


# ---------------------------------------------------------------- exports -----
def _stackless_closest(dfs, leaf_tris, prim_id, ray, oracle):
    """Skip-pointer traversal of the DFS export (layer/shaders/bvh2_stackless.comp
    semantics) with exact slab tests in float64."""
    o, d = ray[:3].astype(np.float64), ray[3:6].astype(np.float64)
    best = (np.inf, -1)
    i = 0
    n = dfs.shape[0]
    while 0 <= i < n:
        bmin = dfs[i, :3].view(np.float32).astype(np.float64)
        bmax = dfs[i, 3:6].view(np.float32).astype(np.float64)
        enc, skip = dfs[i, 6:8].view(np.int32)
        with np.errstate(divide="ignore", invalid="ignore"):
            t1, t2 = (bmin - o) / d, (bmax - o) / d
        t1 = np.where(np.isnan(t1), -np.inf, t1)
        t2 = np.where(np.isnan(t2), np.inf, t2)
        tn, tf = np.max(np.minimum(t1, t2)), np.min(np.maximum(t1, t2))
        if tn <= tf and tf >= 0:
            if enc < 0:
                e = -(enc + 2)
                for k in range(e >> 3, (e >> 3) + (e & 7) + 1):
                    hit, t, _, _ = oracle.isect(ray, leaf_tris[k])
                    if hit and ray[6] < t < ray[7] and (t, prim_id[k]) < best:
                        best = (t, prim_id[k])
                i = skip
            else:
                i += 1
        else:
            i = skip
    return best[1]


def test_dfs_export_semantics(oracle, scenes):
    tris = scenes.random_triangles(800, seed=2)
    b = oracle.OracleBVH(tris, max_leaf=3)
    dfs = b.dfs()
    boxes, links = b.bvh2()
    assert dfs.shape == (len(links), 8)
    assert dfs[0, 7].view(np.int32) == -1
    lt, pid = b.leaf_tris(), b.order()
    cam = scenes.parity_camera(24, 24)
    rays = oracle.gen_pinhole(scenes.cam_array(cam), 24, 24)
    ref = oracle.trace_closest(b.nodes(), lt, pid, rays)
    got = [_stackless_closest(dfs, lt, pid, rays[:, i], oracle) for i in range(rays.shape[1])]
    assert np.array_equal(np.array(got), ref["id"])


def _interleave(x, y, z):
    r = 0
    for i in range(10):
        r |= ((x >> i) & 1) << (3 * i + 2) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i)
    return r


def test_morton_and_sort_keys(oracle):
    rng = np.random.default_rng(4)
    for _ in range(500):
        x, y, z = rng.uniform(-0.1, 1.1, 3).astype(np.float32)
        q = [int(min(max(np.float32(v) * np.float32(1024), 0), 1023)) for v in (x, y, z)]
        assert oracle.lib().orc_morton3d(x, y, z) == _interleave(*q)
    rays = np.zeros((8, 64), np.float32)
    rays[0:3] = rng.uniform(0, 1, (3, 64))
    rays[3:6] = rng.normal(size=(3, 64))
    keys = oracle.sort_keys(rays, [0, 0, 0], [1, 1, 1])
    for i in range(64):
        oct_ = (rays[3, i] < 0) * 4 + (rays[4, i] < 0) * 2 + (rays[5, i] < 0)
        m = oracle.lib().orc_morton3d(*rays[0:3, i])
        assert keys[i] == (oct_ << 29) | (m & 0x1FFFFFFF)


def test_lbvh_tlas(oracle):
    rng = np.random.default_rng(6)
    n = 300
    lo = rng.uniform(-10, 10, (n, 3)).astype(np.float32)
    aabbs = np.concatenate([lo, lo + rng.uniform(0.1, 2, (n, 3)).astype(np.float32)], axis=1)
    nodes, order = oracle.lbvh_tlas(aabbs)
    assert nodes.shape == (2 * n - 1, 8)
    assert sorted(order.tolist()) == list(range(n))
    enc = nodes[:, 6].view(np.int32)
    leaves = enc < 0
    assert leaves.sum() == n
    starts = (-(enc[leaves] + 2)) >> 3
    assert ((-(enc[leaves] + 2)) & 7 == 0).all()                  # one item per leaf
    assert sorted(starts.tolist()) == list(range(n))
    for i in np.where(leaves)[0]:                                  # leaf box = its item's box
        s = (-(enc[i] + 2)) >> 3
        assert np.array_equal(nodes[i, :6].view(np.float32), aabbs[order[s]])
    # Morton order of centroids is non-decreasing
    c = (aabbs[:, :3] + aabbs[:, 3:]) * np.float32(0.5)
    mn, mx = c.min(0), c.max(0)
    inv = (np.float32(1) / (mx - mn)).astype(np.float32)
    codes = [oracle.lib().orc_morton3d(*((c[j] - mn) * inv)) for j in order]
    assert all(a <= b for a, b in zip(codes, codes[1:]))
