"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, where it is the checker. The product library
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I, F = C.c_void_p, C.c_int, C.c_float
        sig = {
            "orc_f2h_rne": (C.c_uint16, [F]), "orc_f2h_down": (C.c_uint16, [F]), "orc_f2h_up": (C.c_uint16, [F]),
            "orc_h2f": (F, [C.c_uint16]),
            "orc_build": (P, [P, I, I]), "orc_free": (None, [P]),
            "orc_num_bvh2": (I, [P]), "orc_num_bvh4": (I, [P]), "orc_num_ordered": (I, [P]),
            "orc_bvh4_depth": (I, [P]), "orc_get_bvh4": (None, [P, P]), "orc_get_order": (None, [P, P]),
            "orc_get_bvh2": (None, [P, P, P]), "orc_num_dfs": (I, [P]), "orc_get_dfs": (None, [P, P]),
            "orc_sah_cost": (C.c_double, [P]),
            "orc_lbvh_tlas": (I, [P, I, P, P]), "orc_release": (None, [P]),
            "orc_morton3d": (C.c_uint32, [F, F, F]),
            "orc_gen_pinhole": (None, [P, I, I, P]),
            "orc_trace_closest": (None, [P, I, P, P, I, P, I, P, P, P, P, P, P, I, I]),
            "orc_trace_any": (None, [P, I, P, P, I, P, I, P, P, P, I, I]),
            "orc_brute_closest": (None, [P, P, I, P, I, P, P, P, P, I, I]),
            "orc_brute_closest_f64": (None, [P, P, I, P, I, P, P, I]),
            "orc_isect_watertight": (I, [P, P, P, P, P]), "orc_isect_mt": (I, [P, P, P, P, P]),
            "orc_sort_keys": (None, [P, I, P, P, P]),
            "orc_num_threads": (I, [I]),
            "orc_trace_closest_2l": (None, [P, I, P, P, P, P, P, I, P, P, P, P, P, P, P, I]),
            "orc_trace_any_2l": (None, [P, I, P, P, P, P, P, I, P, P, P, I]),
            "orc_path_trace_2l": (None, [P, I, P, P, P, P, P, P, P, P, P, P, I]),
            "orc_path_trace_f64": (None, [P, P, I, P, P, P, P, I]),
            "orc_pcg32": (None, [C.c_uint64, C.c_uint64, I, P]),
            "orc_glsl_rays": (None, [P, I, I, P]),
            "orc_scene_hit_normals": (None, [P, P, P, I, I, P, P, P]),
            "orc_glsl_bvh4_trace": (None, [P, P, P, I, I, P, P]),
            "orc_glsl_stackless_trace": (None, [P, P, P, I, I, P, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def num_threads(requested: int = 0) -> int:
    return lib().orc_num_threads(requested)


class OracleBVH:
    """CPU binned-SAH BVH2 -> BVH4 (layer/cuda_bvh_backend.cu:201-251, :391-418)."""

    def __init__(self, tris: np.ndarray, max_leaf: int = 4):
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self.tris = t
        self.h = lib().orc_build(t.ctypes.data, t.shape[0], max_leaf)
        if not self.h:
            raise ValueError("oracle build rejected its input")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    @property
    def num_bvh4(self) -> int:
        return lib().orc_num_bvh4(self.h)

    @property
    def depth(self) -> int:
        return lib().orc_bvh4_depth(self.h)

    def nodes(self) -> np.ndarray:
        out = np.zeros((self.num_bvh4, 64), np.uint8)
        lib().orc_get_bvh4(self.h, out.ctypes.data)
        return out

    def order(self) -> np.ndarray:
        out = np.zeros(lib().orc_num_ordered(self.h), np.int32)
        lib().orc_get_order(self.h, out.ctypes.data)
        return out

    def bvh2(self):
        n = lib().orc_num_bvh2(self.h)
        boxes = np.zeros((n, 6), np.float32)
        links = np.zeros((n, 4), np.int32)
        lib().orc_get_bvh2(self.h, boxes.ctypes.data, links.ctypes.data)
        return boxes, links

    def dfs(self) -> np.ndarray:
        n = lib().orc_num_dfs(self.h)
        out = np.zeros((n, 8), np.uint32)
        lib().orc_get_dfs(self.h, out.ctypes.data)
        return out

    def sah_cost(self) -> float:
        return lib().orc_sah_cost(self.h)

    def leaf_tris(self) -> np.ndarray:
        return np.ascontiguousarray(self.tris[self.order()])


def trace_closest(nodes: np.ndarray, leaf_tris: np.ndarray, prim_id: np.ndarray, rays: np.ndarray,
                  isect: int = 0, threads: int = 0, counts: bool = True):
    """rays: float32 (8, n). Returns dict t, id, u, v (+ nodes, tris counters)."""
    n = rays.shape[1]
    rays = np.ascontiguousarray(rays, np.float32)
    lt = np.ascontiguousarray(leaf_tris, np.float32).reshape(-1, 9)
    pid = np.ascontiguousarray(prim_id, np.int32)
    nd = np.ascontiguousarray(nodes, np.uint8)
    out = dict(t=np.zeros(n, np.float32), id=np.zeros(n, np.int32), u=np.zeros(n, np.float32),
               v=np.zeros(n, np.float32))
    if counts:
        out["nodes"] = np.zeros(n, np.int32)
        out["tris"] = np.zeros(n, np.int32)
    lib().orc_trace_closest(nd.ctypes.data, nd.shape[0], lt.ctypes.data, pid.ctypes.data, lt.shape[0],
                            rays.ctypes.data, n, out["t"].ctypes.data, out["id"].ctypes.data, out["u"].ctypes.data,
                            out["v"].ctypes.data, _p(out.get("nodes")), _p(out.get("tris")), isect, threads)
    return out


def trace_any(nodes, leaf_tris, prim_id, rays, isect: int = 0, threads: int = 0, counts: bool = True):
    n = rays.shape[1]
    rays = np.ascontiguousarray(rays, np.float32)
    lt = np.ascontiguousarray(leaf_tris, np.float32).reshape(-1, 9)
    pid = np.ascontiguousarray(prim_id, np.int32)
    nd = np.ascontiguousarray(nodes, np.uint8)
    out = dict(occ=np.zeros(n, np.uint8))
    if counts:
        out["nodes"] = np.zeros(n, np.int32)
        out["tris"] = np.zeros(n, np.int32)
    lib().orc_trace_any(nd.ctypes.data, nd.shape[0], lt.ctypes.data, pid.ctypes.data, lt.shape[0], rays.ctypes.data,
                        n, out["occ"].ctypes.data, _p(out.get("nodes")), _p(out.get("tris")), isect, threads)
    return out


def brute_closest(tris, prim_id, rays, isect: int = 0, threads: int = 0):
    n = rays.shape[1]
    rays = np.ascontiguousarray(rays, np.float32)
    t9 = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    pid = np.ascontiguousarray(prim_id, np.int32)
    out = dict(t=np.zeros(n, np.float32), id=np.zeros(n, np.int32), u=np.zeros(n, np.float32),
               v=np.zeros(n, np.float32))
    lib().orc_brute_closest(t9.ctypes.data, pid.ctypes.data, t9.shape[0], rays.ctypes.data, n, out["t"].ctypes.data,
                            out["id"].ctypes.data, out["u"].ctypes.data, out["v"].ctypes.data, isect, threads)
    return out


def brute_closest_f64(tris, prim_id, rays, threads: int = 0):
    n = rays.shape[1]
    rays = np.ascontiguousarray(rays, np.float32)
    t9 = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    pid = np.ascontiguousarray(prim_id, np.int32)
    t = np.zeros(n, np.float64)
    ids = np.zeros(n, np.int32)
    lib().orc_brute_closest_f64(t9.ctypes.data, pid.ctypes.data, t9.shape[0], rays.ctypes.data, n, t.ctypes.data,
                                ids.ctypes.data, threads)
    return dict(t=t, id=ids)


def gen_pinhole(cam14: np.ndarray, W: int, H: int) -> np.ndarray:
    out = np.zeros((8, W * H), np.float32)
    c = np.ascontiguousarray(cam14, np.float32)
    lib().orc_gen_pinhole(c.ctypes.data, W, H, out.ctypes.data)
    return out


def isect(ray8, tri9, mode: int = 0):
    r = np.ascontiguousarray(ray8, np.float32)
    t9 = np.ascontiguousarray(tri9, np.float32)
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    f = lib().orc_isect_mt if mode == 1 else lib().orc_isect_watertight
    hit = f(r.ctypes.data, t9.ctypes.data, C.byref(t), C.byref(u), C.byref(v))
    return bool(hit), t.value, u.value, v.value


def lbvh_tlas(aabbs: np.ndarray):
    a = np.ascontiguousarray(aabbs, np.float32).reshape(-1, 6)
    out, ordp = C.c_void_p(), C.c_void_p()
    n = lib().orc_lbvh_tlas(a.ctypes.data, a.shape[0], C.byref(out), C.byref(ordp))
    nodes = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint32)), shape=(n * 8,)).reshape(n, 8).copy()
    order = np.ctypeslib.as_array(C.cast(ordp, C.POINTER(C.c_int32)), shape=(a.shape[0],)).copy()
    lib().orc_release(out)
    lib().orc_release(ordp)
    return nodes, order


def sort_keys(rays: np.ndarray, lo, inv) -> np.ndarray:
    rays = np.ascontiguousarray(rays, np.float32)
    n = rays.shape[1]
    keys = np.zeros(n, np.uint32)
    lo = np.asarray(lo, np.float32)
    inv = np.asarray(inv, np.float32)
    lib().orc_sort_keys(rays.ctypes.data, n, lo.ctypes.data, inv.ctypes.data, keys.ctypes.data)
    return keys


def f2h(x: float, mode: str = "rne") -> int:
    return {"rne": lib().orc_f2h_rne, "down": lib().orc_f2h_down, "up": lib().orc_f2h_up}[mode](x)


def h2f(h: int) -> float:
    return lib().orc_h2f(h)


class _Blas(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("num_nodes", C.c_int), ("tris", C.c_void_p), ("prim_id", C.c_void_p),
                ("num_tris", C.c_int)]


def instance_world_boxes(blas_bounds, o2w, blas_of) -> np.ndarray:
    """World AABB of every instance: the 8 corners of its BLAS box mapped by
    the object-to-world 3x4 map in float64, widened by
    1e-5*extent + 1e-6*(|lo|+|hi|) + 1e-30 per axis, rounded to float32."""
    n = len(blas_of)
    out = np.zeros((n, 6), np.float32)
    for i in range(n):
        lo3, hi3 = blas_bounds[blas_of[i]]
        m = np.asarray(o2w[i], np.float32).astype(np.float64).reshape(3, 4)
        lo = np.full(3, 1e300)
        hi = np.full(3, -1e300)
        for c in range(8):
            p = [float(hi3[0] if c & 1 else lo3[0]), float(hi3[1] if c & 2 else lo3[1]),
                 float(hi3[2] if c & 4 else lo3[2])]
            for r in range(3):
                w = m[r, 0] * p[0] + m[r, 1] * p[1] + m[r, 2] * p[2] + m[r, 3]
                lo[r] = min(lo[r], w)
                hi[r] = max(hi[r], w)
        for r in range(3):
            pad = 1e-5 * (hi[r] - lo[r]) + 1e-6 * (abs(lo[r]) + abs(hi[r])) + 1e-30
            out[i, r] = np.float32(lo[r] - pad)
            out[i, 3 + r] = np.float32(hi[r] + pad)
    return out


def aabb_tris(aabbs: np.ndarray) -> np.ndarray:
    """Degenerate triangles whose bounds are the given boxes: min, max, (max.x, min.y, min.z)."""
    a = np.asarray(aabbs, np.float32).reshape(-1, 6)
    return np.ascontiguousarray(np.stack([a[:, 0], a[:, 1], a[:, 2], a[:, 3], a[:, 4], a[:, 5], a[:, 3], a[:, 1],
                                          a[:, 2]], 1))


class OracleScene:
    """Two-level scene for the oracle: BLAS OracleBVH objects, a TLAS built
    over the instances' world boxes with one instance per leaf, and the
    world-to-object maps (row-major 3x4 float32, one per instance)."""

    def __init__(self, blas_tris: list, o2w: np.ndarray, blas_of, w2o: np.ndarray):
        self.blas = [OracleBVH(t) for t in blas_tris]
        self.blas_nodes = [b.nodes() for b in self.blas]
        self.blas_leaf = [b.leaf_tris() for b in self.blas]
        self.blas_pid = [b.order() for b in self.blas]
        bounds = []
        for t in blas_tris:
            v = np.asarray(t, np.float32).reshape(-1, 3)
            bounds.append((v.min(0), v.max(0)))
        self.blas_of = np.ascontiguousarray(blas_of, np.int32)
        self.world_boxes = instance_world_boxes(bounds, o2w, self.blas_of)
        self.tlas = OracleBVH(aabb_tris(self.world_boxes), max_leaf=1)
        self.tlas_nodes = self.tlas.nodes()
        self.tlas_prim = self.tlas.order()
        self.w2o = np.ascontiguousarray(w2o, np.float32).reshape(-1, 12)
        self._arr = (_Blas * len(self.blas))()
        for i in range(len(self.blas)):
            self._arr[i] = _Blas(self.blas_nodes[i].ctypes.data, self.blas_nodes[i].shape[0],
                                 self.blas_leaf[i].ctypes.data, self.blas_pid[i].ctypes.data,
                                 self.blas_leaf[i].shape[0])

    def trace_closest(self, rays: np.ndarray, threads: int = 0) -> dict:
        n = rays.shape[1]
        rays = np.ascontiguousarray(rays, np.float32)
        out = dict(t=np.zeros(n, np.float32), inst=np.zeros(n, np.int32), id=np.zeros(n, np.int32),
                   u=np.zeros(n, np.float32), v=np.zeros(n, np.float32), nodes=np.zeros(n, np.int32),
                   tris=np.zeros(n, np.int32))
        lib().orc_trace_closest_2l(self.tlas_nodes.ctypes.data, self.tlas_nodes.shape[0], self.tlas_prim.ctypes.data,
                                   self._arr, self.blas_of.ctypes.data, self.w2o.ctypes.data, rays.ctypes.data, n,
                                   *[out[k].ctypes.data for k in ("t", "inst", "id", "u", "v", "nodes", "tris")],
                                   threads)
        return out

    def hit_normals(self, inst: np.ndarray, prim: np.ndarray) -> np.ndarray:
        """World-space unit geometric normals (3, n) of hits (instance, original ID)."""
        inst = np.ascontiguousarray(inst, np.int32)
        prim = np.ascontiguousarray(prim, np.int32)
        n = inst.shape[0]
        out = np.zeros((3, n), np.float32)
        lib().orc_scene_hit_normals(self._arr, self.blas_of.ctypes.data, self.w2o.ctypes.data, self.blas_of.shape[0], n,
                                    inst.ctypes.data, prim.ctypes.data, out.ctypes.data)
        return out

    def trace_any(self, rays: np.ndarray, threads: int = 0, counts: dict | None = None) -> np.ndarray:
        """Occlusion flags; pass a dict as counts to receive per-ray node visits and triangle tests."""
        n = rays.shape[1]
        rays = np.ascontiguousarray(rays, np.float32)
        occ = np.zeros(n, np.uint8)
        nv = np.zeros(n, np.int32) if counts is not None else None
        nt = np.zeros(n, np.int32) if counts is not None else None
        lib().orc_trace_any_2l(self.tlas_nodes.ctypes.data, self.tlas_nodes.shape[0], self.tlas_prim.ctypes.data,
                               self._arr, self.blas_of.ctypes.data, self.w2o.ctypes.data, rays.ctypes.data, n,
                               occ.ctypes.data, _p(nv), _p(nt), threads)
        if counts is not None:
            counts["nodes"], counts["tris"] = nv, nt
        return occ


class PathParams(C.Structure):
    """Layout of orc_pt_params (and of the library's CudaPathParams)."""
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("spp", C.c_int), ("maxDepth", C.c_int),
                ("rrStart", C.c_int), ("seed", C.c_uint32), ("sunDir", C.c_float * 3), ("sunRadiance", C.c_float * 3),
                ("skyZenith", C.c_float * 3), ("skyHorizon", C.c_float * 3), ("clampMax", C.c_float),
                ("tailPaths", C.c_int)]


def path_params(width, height, spp=1, max_depth=8, rr_start=3, seed=1, sun_dir=(0.35, 0.85, -0.40),
                sun_radiance=(3.0, 2.85, 2.6), sky_zenith=(0.35, 0.55, 1.0), sky_horizon=(1.0, 1.0, 1.0),
                clamp_max=20.0) -> PathParams:
    p = PathParams()
    p.width, p.height, p.spp, p.maxDepth, p.rrStart, p.seed = width, height, spp, max_depth, rr_start, seed
    for dst, src in ((p.sunDir, sun_dir), (p.sunRadiance, sun_radiance), (p.skyZenith, sky_zenith),
                     (p.skyHorizon, sky_horizon)):
        for k in range(3):
            dst[k] = float(src[k])
    p.clampMax = float(clamp_max)
    return p


def path_trace_2l(sc: "OracleScene", albedo, cam14, params: PathParams, threads: int = 0):
    """FP32 path tracer: returns (rgb (3, H, W), aov (7, H, W), counts (2*depth,))."""
    W, H = params.width, params.height
    alb = np.ascontiguousarray(albedo, np.float32).reshape(-1, 3)
    cam = np.ascontiguousarray(cam14, np.float32)
    rgb = np.zeros((3, H, W), np.float32)
    aov = np.zeros((7, H, W), np.float32)
    counts = np.zeros(2 * params.maxDepth, np.int64)
    lib().orc_path_trace_2l(sc.tlas_nodes.ctypes.data, sc.tlas_nodes.shape[0], sc.tlas_prim.ctypes.data, sc._arr,
                            sc.blas_of.ctypes.data, sc.w2o.ctypes.data, alb.ctypes.data, cam.ctypes.data,
                            C.byref(params), rgb.ctypes.data, aov.ctypes.data, counts.ctypes.data, threads)
    return rgb, aov, counts


def path_trace_f64(world_tris, tri_inst, albedo, cam14, params: PathParams, threads: int = 0) -> np.ndarray:
    """FP64 brute-force path tracer: rgb (3, H, W) float32."""
    W, H = params.width, params.height
    wt = np.ascontiguousarray(world_tris, np.float64).reshape(-1, 9)
    ti = np.ascontiguousarray(tri_inst, np.int32)
    alb = np.ascontiguousarray(albedo, np.float32).reshape(-1, 3)
    cam = np.ascontiguousarray(cam14, np.float32)
    rgb = np.zeros((3, H, W), np.float32)
    lib().orc_path_trace_f64(wt.ctypes.data, ti.ctypes.data, wt.shape[0], alb.ctypes.data, cam.ctypes.data,
                             C.byref(params), rgb.ctypes.data, threads)
    return rgb


def pcg32(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32(initstate, initseq, n, out.ctypes.data)
    return out


def glsl_rays(pc13, W: int, H: int) -> np.ndarray:
    """Camera rays of the layer's GLSL tracers (8 x W*H SoA)."""
    out = np.zeros((8, W * H), np.float32)
    pc = np.ascontiguousarray(pc13, np.float32)
    lib().orc_glsl_rays(pc.ctypes.data, W, H, out.ctypes.data)
    return out


def glsl_trace(kind: str, nodes: np.ndarray, packed: np.ndarray, pc13, W: int, H: int):
    """The layer's GLSL compute tracer over exported buffers: kind "bvh4"
    (getNodeData bytes) or "stackless" (getStacklessBVH2 words); packed is the
    getPackedTris vec4 array. Returns (t, leaf-order triangle index)."""
    t = np.zeros(W * H, np.float32)
    tri = np.zeros(W * H, np.int32)
    n = np.ascontiguousarray(nodes).view(np.uint32)
    pk = np.ascontiguousarray(packed, np.float32)
    pc = np.ascontiguousarray(pc13, np.float32)
    f = lib().orc_glsl_bvh4_trace if kind == "bvh4" else lib().orc_glsl_stackless_trace
    f(n.ctypes.data, pk.ctypes.data, pc.ctypes.data, W, H, t.ctypes.data, tri.ctypes.data)
    return t, tri
