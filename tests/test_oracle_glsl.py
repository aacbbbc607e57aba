"""The layer's GLSL compute tracers (layer/shaders/bvh4_trace.comp,
bvh2_stackless.comp), restated in the oracle, consume the exported BVH4 node
array, stackless BVH2 and packed triangles. Fed the oracle's own exports they
must find the same closest hits as the oracle's traversal, which they share
no code with (unsorted or skip-link order, the shaders' Moller-Trumbore test
and fma slabs): a wrong node layout, leaf encoding, skip link or packed-vec4
order would make the images disagree."""
import numpy as np


def _packed(order, leaf_tris):
    pk = np.zeros((len(order), 12), np.float32)
    pk[:, :9] = leaf_tris
    pk[:, 9] = order.astype(np.int32).view(np.float32)   # ID bits
    return pk.reshape(-1, 4)


def _camera():
    # from (0, 0, -3) looking +z, 60 degree vertical field of view
    return np.array([0, 0, -3, 0, 0, 1, 1, 0, 0, 0, 1, 0, np.tan(np.radians(30.0))], np.float32)


def test_glsl_tracers_match_oracle(oracle, scenes):
    tris = scenes.random_triangles(65536, seed=0)
    ob = oracle.OracleBVH(tris, max_leaf=4)
    order, lt = ob.order(), ob.leaf_tris()
    pc, W, H = _camera(), 256, 192
    rays = oracle.glsl_rays(pc, W, H)
    ref = oracle.trace_closest(ob.nodes(), lt, order.astype(np.int32), rays)
    assert 0.3 < (ref["id"] >= 0).mean() < 0.9
    for kind, nodes in (("bvh4", ob.nodes()), ("stackless", ob.dfs())):
        t, tri = oracle.glsl_trace(kind, nodes, _packed(order, lt), pc, W, H)
        gid = np.where(tri >= 0, order[np.maximum(tri, 0)], -1)
        agree = gid == ref["id"]
        assert agree.mean() >= 0.999, (kind, agree.mean())
        both = agree & (gid >= 0)
        # Moller-Trumbore and watertight t of the same hit differ by rounding only
        assert np.all(np.abs(t[both] - ref["t"][both]) <= 1e-4 * ref["t"][both]), kind
        assert np.all(t[gid < 0] == np.float32(1e30))
