"""MI355X-native software ray-tracing core: Python host bindings.

The product is ``libcudabvh_hip.so`` (HIP kernels for gfx950 behind the C ABI
of ``include/cudabvh.h``). This module binds every exported symbol with ctypes,
mirroring the reference boundary ``layer/cuda_bvh_backend.h`` function by
function (same names without the ``cudaBVH_`` prefix, same argument meaning,
same return conventions: 0 / -1 for int calls, MRays/s / 0 for float calls,
NULL handles on failure), and adds :class:`BVH`, a small owner of a handle.

Importing succeeds without a GPU; loading the library raises
:class:`LibraryMissing` when it has not been built — there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcudabvh_hip.so")

# Ray classes and wave schedules of the launch configuration / autotuner
CLASS_COHERENT, CLASS_INCOHERENT, CLASS_ANY = 0, 1, 2
SCHEDULE_BATCH, SCHEDULE_WHILE_WHILE, SCHEDULE_PACKET, SCHEDULE_REFILL = 0, 1, 2, 3

# Flags (include/cudabvh.h, part 2)
ISECT_WATERTIGHT = 0
ISECT_REFERENCE_MT = 1
FLAG_HOST_POINTERS = 1 << 4
FLAG_SORT_RAYS = 1 << 5
FLAG_SYNC = 1 << 6
FLAG_GRAPH = 1 << 7
FLAG_CWBVH = 1 << 8
FLAG_TILE_ORDER = 1 << 9
FLAG_COST_ORDER = 1 << 10
PT_MAX_DEPTH = 16
MISS_T = np.float32(1e30)


class LibraryMissing(RuntimeError):
    """libcudabvh_hip.so is not built (run __graft_entry__.build())."""


class CudaBVHError(RuntimeError):
    """A cudaBVH_* call reported failure; the message is cudaBVH_lastError()."""


class CudaTri(C.Structure):
    _fields_ = [("v0", C.c_float * 3), ("v1", C.c_float * 3), ("v2", C.c_float * 3)]


class CudaRayPacket(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class CudaHitPacket(C.Structure):
    _fields_ = [("t", C.c_void_p), ("primId", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p)]


class CudaCamera(C.Structure):
    _fields_ = [("o", C.c_float * 3), ("f", C.c_float * 3), ("r", C.c_float * 3), ("u", C.c_float * 3),
                ("scaleX", C.c_float), ("scaleY", C.c_float)]


class CudaInstance(C.Structure):
    _fields_ = [("blas", C.c_int), ("objectToWorld", C.c_float * 12), ("albedo", C.c_float * 3)]


class CudaGeometry(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("numVertices", C.c_int), ("vertexStride", C.c_int),
                ("indices", C.c_void_p), ("indexBytes", C.c_int), ("numTris", C.c_int)]


class CudaSceneHit(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("t", "instId", "primId", "u", "v")]


class CudaPathParams(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("spp", C.c_int), ("maxDepth", C.c_int),
                ("rrStart", C.c_int), ("seed", C.c_uint32), ("sunDir", C.c_float * 3), ("sunRadiance", C.c_float * 3),
                ("skyZenith", C.c_float * 3), ("skyHorizon", C.c_float * 3), ("clampMax", C.c_float),
                ("tailPaths", C.c_int)]


class CudaPathImage(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("rgb", "albedo", "normal", "depth")]


def path_params(width: int, height: int, spp: int = 1, max_depth: int = 8, rr_start: int = 3, seed: int = 1,
                sun_dir=(0.35, 0.85, -0.40), sun_radiance=(3.0, 2.85, 2.6), sky_zenith=(0.35, 0.55, 1.0),
                sky_horizon=(1.0, 1.0, 1.0), clamp_max: float = 20.0, tail_paths: int = 32768) -> CudaPathParams:
    """Path-tracer parameters; the defaults are the instanced-scene renders'."""
    p = CudaPathParams()
    p.width, p.height, p.spp, p.maxDepth, p.rrStart, p.seed = width, height, spp, max_depth, rr_start, seed
    for dst, src in ((p.sunDir, sun_dir), (p.sunRadiance, sun_radiance), (p.skyZenith, sky_zenith),
                     (p.skyHorizon, sky_horizon)):
        for k in range(3):
            dst[k] = float(src[k])
    p.clampMax = float(clamp_max)
    p.tailPaths = int(tail_paths)
    return p


# name -> (restype, argtypes); the exported C ABI, in header order.
_H = C.c_void_p
_P = C.c_void_p
_F = C.c_float
_I = C.c_int
SIGNATURES = {
    "cudaBVH_build": (_H, [_P, _I]),
    "cudaBVH_buildDevice": (_H, [_P, _I, _I, _P, _I, _I]),
    "cudaBVH_buildManyDevice": (_I, [_P, _P, _I, _P]),
    "cudaBVH_destroy": (None, [_H]),
    "cudaBVH_tracePrimary": (_F, [_H, _I, _F, _F, _F, _P]),
    "cudaBVH_traceDiffuse": (_F, [_H, _I, _P, _P, _P, _P, _P, _P, _P]),
    "cudaBVH_traceToRGBA": (_I, [_H, _I, _I, _F, _F, _F, _P]),
    "cudaBVH_traceToGPUPtr": (_I, [_H, _I, _I, _F, _F, _F, _P, _I]),
    "cudaBVH_traceToGPUAsync": (None, [_H, _I, _I, _F, _F, _F, _P, _I]),
    "cudaBVH_traceToHostPtr": (_I, [_H, _I, _I, _F, _F, _F, _P, _I, C.c_uint64]),
    "cudaBVH_importBufferFd": (_P, [_I, C.c_uint64]),
    "cudaBVH_getBounds": (None, [_H, _P, _P, _P, _P]),
    "cudaBVH_getNumTris": (_I, [_H]),
    "cudaBVH_getNumBVH4Nodes": (_I, [_H]),
    "cudaBVH_getNumCWBVHNodes": (_I, [_H]),
    "cudaBVH_getNodeData": (_P, [_H]),
    "cudaBVH_getTriData": (None, [_H, _P]),
    "cudaBVH_getStacklessBVH2": (_I, [_H, _P]),
    "cudaBVH_getPackedTris": (_I, [_H, _P]),
    "cudaBVH_resetDevice": (None, []),
    "cudaBVH_buildFromAABBs": (_H, [_P, _I]),
    "cudaBVH_buildTLASFast": (_I, [_P, _I, _P, _P]),
    "cudaBVH_readGPUMemoryFd": (_I, [_I, C.c_uint64, C.c_uint64, C.c_uint64, _P]),
    # extensions
    "cudaBVH_traceClosest": (_I, [_H, _I, C.POINTER(CudaRayPacket), C.POINTER(CudaHitPacket), C.c_uint]),
    "cudaBVH_traceAny": (_I, [_H, _I, C.POINTER(CudaRayPacket), _P, C.c_uint]),
    "cudaBVH_traceClosestNoAlloc": (_I, [_H, _I, C.POINTER(CudaRayPacket), C.POINTER(CudaHitPacket), C.c_uint]),
    "cudaBVH_traceClosestCounted": (_I, [_H, _I, C.POINTER(CudaRayPacket), C.POINTER(CudaHitPacket), _P, _P,
                                         C.c_uint]),
    "cudaBVH_traceAnyCounted": (_I, [_H, _I, C.POINTER(CudaRayPacket), _P, _P, _P, C.c_uint]),
    "cudaBVH_genPinholeRays": (_I, [_I, _I, C.POINTER(CudaCamera), C.POINTER(CudaRayPacket), _P]),
    "cudaBVH_tracePinholeClosest": (_I, [_H, _I, _I, C.POINTER(CudaCamera), C.POINTER(CudaHitPacket), C.c_uint]),
    "cudaBVH_genBounceRays": (_I, [_H, _I, _I, C.POINTER(CudaCamera), C.POINTER(CudaHitPacket), C.c_uint32, _P,
                                   C.POINTER(CudaRayPacket), C.POINTER(CudaRayPacket), _P, _P, C.c_uint]),
    "cudaBVH_traceClosestQueued": (_I, [_H, _I, _P, C.POINTER(CudaRayPacket), C.POINTER(CudaHitPacket), C.c_uint]),
    "cudaBVH_traceAnyQueued": (_I, [_H, _I, _P, C.POINTER(CudaRayPacket), _P, C.c_uint]),
    "cudaBVH_setTraceConfig": (_I, [_H, _I, _I, _I, _I]),
    "cudaBVH_getTraceConfig": (_I, [_H, _I, _P, _P, _P]),
    "cudaBVH_autotune": (_I, [_H, _I, _I, _I, _P, _I, _P, _I, _P, _P]),
    "cudaBVH_sortRays": (_I, [_H, _I, C.POINTER(CudaRayPacket), _P, _P]),
    "cudaBVH_setStream": (_I, [_H, _P]),
    "cudaBVH_getStream": (_P, [_H]),
    "cudaBVH_deviceNodes": (_P, [_H]),
    "cudaBVH_deviceTris": (_P, [_H]),
    "cudaBVH_getDepth": (_I, [_H]),
    "cudaBVH_getBuildStats": (_I, [_H, _P]),
    "cudaBVH_buildCWBVH": (_I, [_H]),
    "cudaBVH_getCWBVH": (_I, [_H, _P, _P, _P]),
    "cudaBVH_sceneCreate": (_H, [_P, _I, _P, _I]),
    "cudaBVH_sceneDestroy": (None, [_H]),
    "cudaBVH_sceneSetStream": (_I, [_H, _P]),
    "cudaBVH_sceneTLAS": (_H, [_H]),
    "cudaBVH_sceneWorldToObject": (_I, [_H, _P]),
    "cudaBVH_sceneTraceClosest": (_I, [_H, _I, C.POINTER(CudaRayPacket), C.POINTER(CudaSceneHit), C.c_uint]),
    "cudaBVH_sceneTraceAny": (_I, [_H, _I, C.POINTER(CudaRayPacket), _P, C.c_uint]),
    "cudaBVH_sceneTraceClosestCounted": (_I, [_H, _I, C.POINTER(CudaRayPacket), C.POINTER(CudaSceneHit), _P, _P,
                                              C.c_uint]),
    "cudaBVH_sceneTraceClosestQueued": (_I, [_H, _I, _P, C.POINTER(CudaRayPacket), C.POINTER(CudaSceneHit),
                                             C.c_uint]),
    "cudaBVH_sceneTraceAnyQueued": (_I, [_H, _I, _P, C.POINTER(CudaRayPacket), _P, C.c_uint]),
    "cudaBVH_sceneHitNormals": (_I, [_H, _I, C.POINTER(CudaSceneHit), _P, _P, _P, C.c_uint]),
    "cudaBVH_getLaunchInfo": (_I, [_H, _P]),
    "cudaBVH_sceneGetLaunchInfo": (_I, [_H, _P]),
    "cudaBVH_getTraceStatus": (_I, [_H]),
    "cudaBVH_sceneGetTraceStatus": (_I, [_H]),
    "cudaBVH_sceneSetTraceConfig": (_I, [_H, _I, _I, _I, _I]),
    "cudaBVH_sceneGetTraceConfig": (_I, [_H, _I, _P, _P, _P]),
    "cudaBVH_sceneAutotune": (_I, [_H, _I, _I, C.POINTER(CudaRayPacket), _I, _P, _P]),
    "cudaBVH_scenePathTrace": (_I, [_H, C.POINTER(CudaCamera), C.POINTER(CudaPathParams), C.POINTER(CudaPathImage),
                                    _P, C.c_uint]),
    "cudaBVH_sceneAutotunePathTracer": (_I, [_H, C.POINTER(CudaCamera), C.POINTER(CudaPathParams),
                                             C.POINTER(CudaPathImage), _I, _P, _P]),
    "cudaBVH_buildBatch": (_H, [_P, _P, _I]),
    "cudaBVH_batchDestroy": (None, [_H]),
    "cudaBVH_batchSetStream": (_I, [_H, _P]),
    "cudaBVH_batchForest": (_H, [_H]),
    "cudaBVH_batchRoots": (_I, [_H, _P]),
    "cudaBVH_batchTraceClosest": (_I, [_H, _I, _P, C.POINTER(CudaRayPacket), C.POINTER(CudaHitPacket), C.c_uint]),
    "cudaBVH_batchTraceAny": (_I, [_H, _I, _P, C.POINTER(CudaRayPacket), _P, C.c_uint]),
    "cudaBVH_denoiserCreate": (_H, [_I, C.c_uint64]),
    "cudaBVH_denoiserDestroy": (None, [_H]),
    "cudaBVH_denoiserSetStream": (_I, [_H, _P]),
    "cudaBVH_denoiserNumParams": (C.c_longlong, [_H]),
    "cudaBVH_denoiserGetParams": (_I, [_H, _P, _I]),
    "cudaBVH_denoiserSetParams": (_I, [_H, _P]),
    "cudaBVH_denoiserPack": (_I, [_H, _I, _P, _P, _P, _P, _P]),
    "cudaBVH_denoiserForward": (_I, [_H, _I, _I, _I, _P, _P]),
    "cudaBVH_denoiserTrainStep": (_I, [_H, _I, _I, _I, _P, _P, _F, _P]),
    "cudaBVH_denoise": (_I, [_H, _I, _I, _P, _P, _P, _P, _P]),
    "cudaBVH_lastError": (C.c_char_p, []),
}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libcudabvh_hip.so and declare every signature (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load_library().cudaBVH_lastError()
    return msg.decode() if msg else ""


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise CudaBVHError(f"{what}: {last_error()}")


def _ptr(x) -> int | None:
    """Address of a contiguous numpy array or torch tensor (None passes NULL)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError("array passed to the C ABI must be C-contiguous")
        return x.ctypes.data
    if not x.is_contiguous():  # torch tensor
        raise ValueError("tensor passed to the C ABI must be contiguous")
    return x.data_ptr()


def _out_ptr(x, n: int, dtype=np.float32) -> int | None:
    """Address of an output array the library fills with n elements of dtype:
    checks the element type and the size so the call cannot overrun it."""
    if x is None:
        return None
    dt = np.dtype(dtype)
    if isinstance(x, np.ndarray):
        if x.dtype != dt or x.size < n:
            raise ValueError(f"output must be {dt} with at least {n} elements (got {x.dtype}, {x.size})")
    else:
        import torch
        want = {np.dtype(np.float32): torch.float32, np.dtype(np.int32): torch.int32,
                np.dtype(np.uint8): torch.uint8, np.dtype(np.uint32): torch.uint32}[dt]
        if x.dtype != want or x.numel() < n:
            raise ValueError(f"output must be {want} with at least {n} elements (got {x.dtype}, {x.numel()})")
    return _ptr(x)


def ray_packet(rays) -> CudaRayPacket:
    """Packet from a dict (or object) holding ox oy oz dx dy dz tmin tmax arrays."""
    get = rays.__getitem__ if isinstance(rays, dict) else (lambda k: getattr(rays, k))
    return CudaRayPacket(*[_ptr(get(k)) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")])


def hit_packet(t, prim_id=None, u=None, v=None) -> CudaHitPacket:
    return CudaHitPacket(_ptr(t), _ptr(prim_id), _ptr(u), _ptr(v))


def camera(o, f, r, u, scale_x, scale_y) -> CudaCamera:
    c = CudaCamera()
    for dst, src in ((c.o, o), (c.f, f), (c.r, r), (c.u, u)):
        for i in range(3):
            dst[i] = float(np.float32(src[i]))
    c.scaleX = float(np.float32(scale_x))
    c.scaleY = float(np.float32(scale_y))
    return c


class BVH:
    """Owner of one CudaBVH_t handle; methods mirror the C ABI."""

    def __init__(self, handle: int, num_input: int):
        if not handle:
            raise CudaBVHError(f"build failed: {last_error()}")
        self.lib = load_library()
        self.handle = C.c_void_p(handle)
        self.num_input = num_input

    # --- construction (layer/cuda_bvh_backend.h:21, :123) ---
    @classmethod
    def build(cls, tris: np.ndarray) -> "BVH":
        """tris: float32 array of shape (n, 9) or (n, 3, 3) in original order."""
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        return cls(load_library().cudaBVH_build(t.ctypes.data, t.shape[0]), t.shape[0])

    @classmethod
    def build_device(cls, vertices, num_tris: int, indices=None, vertex_stride: int = 12,
                     num_vertices: int | None = None) -> "BVH":
        """Build from device tensors: vertices (float32, vertex_stride bytes per
        vertex, position first) and optional indices (int32 or int16, three per
        triangle); without indices the vertices are the triangle corners in order."""
        nbytes = vertices.numel() * vertices.element_size()
        nv = num_vertices if num_vertices is not None else nbytes // vertex_stride
        ib = 0 if indices is None else indices.element_size()
        h = load_library().cudaBVH_buildDevice(_ptr(vertices), nv, vertex_stride,
                                               None if indices is None else _ptr(indices), ib, num_tris)
        if not h:
            raise CudaBVHError(f"buildDevice: {last_error()}")
        return cls(h, num_tris)

    @staticmethod
    def build_many_device(blases: list) -> list:
        """One BVH per BLAS from device geometry in one pass. blases: a list
        of BLASes, each a list of geometries (vertices, indices, vertex_stride,
        num_vertices, num_tris) of device tensors (indices may be None)."""
        geoms, offs = [], [0]
        for blas in blases:
            for (vtx, idx, stride, nv, nt) in blas:
                geoms.append(CudaGeometry(_ptr(vtx), int(nv), int(stride), None if idx is None else _ptr(idx),
                                          0 if idx is None else idx.element_size(), int(nt)))
            offs.append(len(geoms))
        arr = (CudaGeometry * len(geoms))(*geoms)
        off = np.asarray(offs, np.int32)
        out = (C.c_void_p * len(blases))()
        lib = load_library()
        rc = lib.cudaBVH_buildManyDevice(arr, off.ctypes.data, len(blases), out)
        if rc != 0:
            raise CudaBVHError(f"buildManyDevice: {last_error()}")
        counts = [sum(int(g[4]) for g in blas) for blas in blases]
        return [BVH(out[k], counts[k]) for k in range(len(blases))]

    @classmethod
    def build_from_aabbs(cls, aabbs: np.ndarray) -> "BVH":
        a = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
        return cls(load_library().cudaBVH_buildFromAABBs(a.ctypes.data, a.shape[0]), a.shape[0])

    def destroy(self) -> None:
        if self.handle and not getattr(self, "_borrowed", False):
            self.lib.cudaBVH_destroy(self.handle)
            self.handle = C.c_void_p(0)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    # --- stats and host views ---
    @property
    def num_tris(self) -> int:
        return self.lib.cudaBVH_getNumTris(self.handle)

    @property
    def num_bvh4_nodes(self) -> int:
        return self.lib.cudaBVH_getNumBVH4Nodes(self.handle)

    @property
    def num_cwbvh_nodes(self) -> int:
        return self.lib.cudaBVH_getNumCWBVHNodes(self.handle)

    @property
    def depth(self) -> int:
        return self.lib.cudaBVH_getDepth(self.handle)

    def build_stats(self) -> dict:
        """upload_ms, device_ms, handle_ms of the build and its BVH2 levels."""
        out = (C.c_float * 4)()
        _check(self.lib.cudaBVH_getBuildStats(self.handle, out), "getBuildStats")
        return {"upload_ms": float(out[0]), "device_ms": float(out[1]), "handle_ms": float(out[2]),
                "levels": int(out[3])}

    def bounds(self):
        v = [C.c_float() for _ in range(4)]
        self.lib.cudaBVH_getBounds(self.handle, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def node_data(self) -> np.ndarray:
        """BVH4 nodes as uint8 (n, 64), copied out of the borrowed pointer."""
        p = self.lib.cudaBVH_getNodeData(self.handle)
        n = self.num_bvh4_nodes
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n * 64,)).reshape(n, 64).copy()

    def tri_data(self) -> np.ndarray:
        """Leaf-order SoA triangles, float32 (9, n)."""
        ptrs = (C.c_void_p * 9)()
        self.lib.cudaBVH_getTriData(self.handle, ptrs)
        n = self.num_tris
        return np.stack([np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n,)).copy() for p in ptrs])

    def stackless_bvh2(self) -> np.ndarray:
        out = C.c_void_p()
        n = self.lib.cudaBVH_getStacklessBVH2(self.handle, C.byref(out))
        if n == 0:
            return np.zeros((0, 8), np.uint32)
        return np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint32)), shape=(n * 8,)).reshape(n, 8).copy()

    def packed_tris(self) -> np.ndarray:
        out = C.c_void_p()
        n = self.lib.cudaBVH_getPackedTris(self.handle, C.byref(out))
        return np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_float)), shape=(n * 4,)).reshape(n, 4).copy()

    def launch_info(self) -> dict:
        """Resources of the last traversal launch on this handle."""
        return _launch_info(self.lib.cudaBVH_getLaunchInfo, self.handle)

    def leaf_prim_ids(self) -> np.ndarray:
        """Original primitive index of each leaf-order triangle (from the packed export)."""
        return self.packed_tris()[2::3, 1].view(np.int32).copy()

    # --- streams ---
    def set_stream(self, stream: int | None) -> None:
        _check(self.lib.cudaBVH_setStream(self.handle, stream), "setStream")

    def build_cwbvh(self) -> int:
        n = self.lib.cudaBVH_buildCWBVH(self.handle)
        if n < 0:
            raise CudaBVHError(f"buildCWBVH: {last_error()}")
        return n

    def cwbvh(self):
        """(nodes uint32 (n, 20), tri_order int32 (m,), depth) host copies."""
        nodes, order, depth = C.c_void_p(), C.c_void_p(), C.c_int()
        n = self.lib.cudaBVH_getCWBVH(self.handle, C.byref(nodes), C.byref(order), C.byref(depth))
        if n <= 0:
            raise CudaBVHError("no CWBVH built")
        nd = np.ctypeslib.as_array(C.cast(nodes, C.POINTER(C.c_uint32)), shape=(n * 20,)).reshape(n, 20).copy()
        od = np.ctypeslib.as_array(C.cast(order, C.POINTER(C.c_int32)), shape=(self.num_tris,)).copy()
        return nd, od, depth.value

    # --- reference traces (layer/cuda_bvh_backend.h:29-73) ---
    def trace_primary(self, side: int, cam_o, out_hit_t: np.ndarray | None = None) -> float:
        return self.lib.cudaBVH_tracePrimary(self.handle, side, *[float(c) for c in cam_o],
                                             _out_ptr(out_hit_t, side * side))

    def trace_diffuse(self, o: np.ndarray, d: np.ndarray, out_hit_t: np.ndarray | None = None) -> float:
        o = [np.ascontiguousarray(o[:, k], np.float32) for k in range(3)]
        d = [np.ascontiguousarray(d[:, k], np.float32) for k in range(3)]
        n = o[0].shape[0]
        return self.lib.cudaBVH_traceDiffuse(self.handle, n, *[x.ctypes.data for x in o + d],
                                             _out_ptr(out_hit_t, n))

    def trace_status(self) -> int:
        """Traversal status word since the last check (waits for the handle's
        stream; bit 0: a stack overflowed); clears it."""
        w = self.lib.cudaBVH_getTraceStatus(self.handle)
        if w < 0:
            raise CudaBVHError(f"getTraceStatus: {last_error()}")
        return w

    def trace_to_rgba(self, width: int, height: int, cam_o) -> np.ndarray:
        out = np.zeros(width * height, np.uint32)
        _check(self.lib.cudaBVH_traceToRGBA(self.handle, width, height, *[float(c) for c in cam_o], out.ctypes.data),
               "traceToRGBA")
        return out.reshape(height, width)

    def trace_to_gpu_ptr(self, width: int, height: int, cam_o, dev_ptr: int, fmt: int) -> None:
        _check(self.lib.cudaBVH_traceToGPUPtr(self.handle, width, height, *[float(c) for c in cam_o], dev_ptr, fmt),
               "traceToGPUPtr")

    # --- extensions (device or host packets) ---
    def trace_closest(self, rays, t, prim_id=None, u=None, v=None, flags: int = 0) -> None:
        n = int(rays["ox"].shape[0]) if isinstance(rays, dict) else int(rays.ox.shape[0])
        rp, hp = ray_packet(rays), hit_packet(t, prim_id, u, v)
        _check(self.lib.cudaBVH_traceClosest(self.handle, n, C.byref(rp), C.byref(hp), flags), "traceClosest")

    def trace_closest_no_alloc(self, rays, t, prim_id=None, u=None, v=None, flags: int = 0) -> None:
        """trace_closest through the library's static device scratch: no
        allocation (device packets only; overlapping calls on a GPU race)."""
        n = int(rays["ox"].shape[0]) if isinstance(rays, dict) else int(rays.ox.shape[0])
        rp, hp = ray_packet(rays), hit_packet(t, prim_id, u, v)
        _check(self.lib.cudaBVH_traceClosestNoAlloc(self.handle, n, C.byref(rp), C.byref(hp), flags),
               "traceClosestNoAlloc")

    def trace_any(self, rays, occluded, flags: int = 0) -> None:
        n = int(rays["ox"].shape[0]) if isinstance(rays, dict) else int(rays.ox.shape[0])
        rp = ray_packet(rays)
        _check(self.lib.cudaBVH_traceAny(self.handle, n, C.byref(rp), _ptr(occluded), flags), "traceAny")

    def trace_closest_counted(self, rays, t, prim_id, u, v, node_visits, tri_tests, flags: int = 0) -> None:
        n = int(rays["ox"].shape[0])
        rp, hp = ray_packet(rays), hit_packet(t, prim_id, u, v)
        _check(self.lib.cudaBVH_traceClosestCounted(self.handle, n, C.byref(rp), C.byref(hp), _ptr(node_visits),
                                                    _ptr(tri_tests), flags), "traceClosestCounted")

    def trace_any_counted(self, rays, occluded, node_visits, tri_tests, flags: int = 0) -> None:
        n = int(rays["ox"].shape[0])
        rp = ray_packet(rays)
        _check(self.lib.cudaBVH_traceAnyCounted(self.handle, n, C.byref(rp), _ptr(occluded), _ptr(node_visits),
                                                _ptr(tri_tests), flags), "traceAnyCounted")

    def trace_pinhole_closest(self, width: int, height: int, cam: CudaCamera, t, prim_id=None, u=None, v=None,
                              flags: int = 0) -> None:
        hp = hit_packet(t, prim_id, u, v)
        _check(self.lib.cudaBVH_tracePinholeClosest(self.handle, width, height, C.byref(cam), C.byref(hp), flags),
               "tracePinholeClosest")


    def gen_bounce_rays(self, width: int, height: int, cam: CudaCamera, hit_t, hit_id, seed: int, light,
                        diffuse, shadow, src_index, d_count, flags: int = 0) -> None:
        hp = hit_packet(hit_t, hit_id)
        lt = (C.c_float * 3)(*[float(x) for x in light])
        dp, sp = ray_packet(diffuse), ray_packet(shadow)
        _check(self.lib.cudaBVH_genBounceRays(self.handle, width, height, C.byref(cam), C.byref(hp), seed, lt,
                                              C.byref(dp), C.byref(sp), _ptr(src_index), _ptr(d_count), flags),
               "genBounceRays")

    def trace_closest_queued(self, capacity: int, d_count, rays, t, prim_id=None, u=None, v=None,
                             flags: int = 0) -> None:
        rp, hp = ray_packet(rays), hit_packet(t, prim_id, u, v)
        _check(self.lib.cudaBVH_traceClosestQueued(self.handle, capacity, _ptr(d_count), C.byref(rp), C.byref(hp),
                                                   flags), "traceClosestQueued")

    def trace_any_queued(self, capacity: int, d_count, rays, occluded, flags: int = 0) -> None:
        rp = ray_packet(rays)
        _check(self.lib.cudaBVH_traceAnyQueued(self.handle, capacity, _ptr(d_count), C.byref(rp), _ptr(occluded),
                                               flags), "traceAnyQueued")


    def sort_rays(self, rays, keys_out, perm_out) -> None:
        n = int(rays["ox"].shape[0])
        rp = ray_packet(rays)
        _check(self.lib.cudaBVH_sortRays(self.handle, n, C.byref(rp), _ptr(keys_out), _ptr(perm_out)), "sortRays")

    def set_trace_config(self, ray_class: int, schedule: int, lds_depth: int, blocks_per_cu: int = 0) -> None:
        _check(self.lib.cudaBVH_setTraceConfig(self.handle, ray_class, schedule, lds_depth, blocks_per_cu),
               "setTraceConfig")

    def trace_config(self, ray_class: int):
        """(schedule, lds_depth, blocks_per_cu) of a ray class."""
        sc, d, b = C.c_int(), C.c_int(), C.c_int()
        _check(self.lib.cudaBVH_getTraceConfig(self.handle, ray_class, C.byref(sc), C.byref(d), C.byref(b)),
               "getTraceConfig")
        return sc.value, d.value, b.value

    def autotune(self, ray_class: int, reps: int = 3, width: int = 0, height: int = 0, cam: CudaCamera | None = None,
                 rays=None):
        """Sweeps the launch configurations of one ray class; returns the
        chosen (schedule, lds_depth, blocks_per_cu) and the table of timings."""
        res = (C.c_float * 192)()
        nres = C.c_int()
        n = 0 if rays is None else int(rays["ox"].shape[0])
        rp = ray_packet(rays) if rays is not None else None
        _check(self.lib.cudaBVH_autotune(self.handle, ray_class, width, height, C.byref(cam) if cam else None, n,
                                         C.byref(rp) if rp is not None else None, reps, res, C.byref(nres)),
               "autotune")
        table = [(int(res[4 * i]), int(res[4 * i + 1]), int(res[4 * i + 2]), float(res[4 * i + 3]))
                 for i in range(nres.value)]
        return self.trace_config(ray_class), table


_LAUNCH_INFO_KEYS = ("vgprs", "lds_bytes", "scratch_bytes", "blocks_per_cu", "max_blocks_per_cu", "workgroups",
                     "threads_per_workgroup", "waves_per_simd")


def _launch_info(fn, handle) -> dict:
    out = (C.c_int * 8)()
    _check(fn(handle, out), "getLaunchInfo")
    return dict(zip(_LAUNCH_INFO_KEYS, [int(x) for x in out]))


class Scene:
    """Owner of a two-level scene: a TLAS over instances of borrowed BLAS
    handles (:class:`BVH`), each with a 3x4 object-to-world map and albedo."""

    def __init__(self, blas: list, object_to_world: np.ndarray, blas_of, albedo=None):
        self.lib = load_library()
        self.blas = list(blas)  # keeps the borrowed handles alive
        m = np.ascontiguousarray(object_to_world, dtype=np.float32).reshape(-1, 12)
        n = m.shape[0]
        alb = np.full((n, 3), 0.5, np.float32) if albedo is None else np.asarray(albedo, np.float32).reshape(n, 3)
        insts = (CudaInstance * n)()
        for i in range(n):
            insts[i].blas = int(blas_of[i])
            for k in range(12):
                insts[i].objectToWorld[k] = float(m[i, k])
            for k in range(3):
                insts[i].albedo[k] = float(alb[i, k])
        hs = (C.c_void_p * len(self.blas))(*[b.handle.value for b in self.blas])
        h = self.lib.cudaBVH_sceneCreate(hs, len(self.blas), insts, n)
        if not h:
            raise CudaBVHError(f"sceneCreate: {last_error()}")
        self.handle = C.c_void_p(h)
        self.num_instances = n

    def destroy(self) -> None:
        if getattr(self, "handle", None):
            self.lib.cudaBVH_sceneDestroy(self.handle)
            self.handle = C.c_void_p(0)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def set_stream(self, stream: int | None) -> None:
        _check(self.lib.cudaBVH_sceneSetStream(self.handle, stream), "sceneSetStream")

    def tlas(self) -> BVH:
        """The TLAS as a borrowed (non-owning) BVH view."""
        v = BVH.__new__(BVH)
        v.lib, v.handle, v.num_input = self.lib, C.c_void_p(self.lib.cudaBVH_sceneTLAS(self.handle)), \
            self.num_instances
        v._borrowed = True
        return v

    def set_trace_config(self, ray_class: int, schedule: int, lds_depth: int, blocks_per_cu: int = 0) -> None:
        _check(self.lib.cudaBVH_sceneSetTraceConfig(self.handle, ray_class, schedule, lds_depth, blocks_per_cu),
               "sceneSetTraceConfig")

    def trace_config(self, ray_class: int):
        sc, d, b = C.c_int(), C.c_int(), C.c_int()
        _check(self.lib.cudaBVH_sceneGetTraceConfig(self.handle, ray_class, C.byref(sc), C.byref(d), C.byref(b)),
               "sceneGetTraceConfig")
        return sc.value, d.value, b.value

    def autotune(self, ray_class: int, n: int, rays, reps: int = 3):
        """Picks the fastest (schedule, LDS depth, blocks/CU) for these rays; returns (config, table)."""
        rp = ray_packet(rays)
        res = np.zeros(4 * 48, np.float32)
        k = C.c_int()
        _check(self.lib.cudaBVH_sceneAutotune(self.handle, ray_class, n, C.byref(rp), reps, res.ctypes.data,
                                              C.byref(k)), "sceneAutotune")
        return self.trace_config(ray_class), [(int(res[4 * i]), int(res[4 * i + 1]), int(res[4 * i + 2]),
                                               float(res[4 * i + 3])) for i in range(k.value)]

    def world_to_object(self) -> np.ndarray:
        out = np.zeros((self.num_instances, 12), np.float32)
        _check(self.lib.cudaBVH_sceneWorldToObject(self.handle, out.ctypes.data), "sceneWorldToObject")
        return out

    def trace_closest(self, rays, t, inst_id=None, prim_id=None, u=None, v=None, flags: int = 0) -> None:
        rp = ray_packet(rays)
        hp = CudaSceneHit(_ptr(t), _ptr(inst_id), _ptr(prim_id), _ptr(u), _ptr(v))
        _check(self.lib.cudaBVH_sceneTraceClosest(self.handle, int(t.shape[0]), C.byref(rp), C.byref(hp), flags),
               "sceneTraceClosest")

    def trace_closest_counted(self, rays, t, inst_id, prim_id, u, v, node_visits, tri_tests, flags: int = 0):
        rp = ray_packet(rays)
        hp = CudaSceneHit(_ptr(t), _ptr(inst_id), _ptr(prim_id), _ptr(u), _ptr(v))
        _check(self.lib.cudaBVH_sceneTraceClosestCounted(self.handle, int(t.shape[0]), C.byref(rp), C.byref(hp),
                                                         _ptr(node_visits), _ptr(tri_tests), flags),
               "sceneTraceClosestCounted")

    def trace_any(self, rays, occluded, flags: int = 0) -> None:
        rp = ray_packet(rays)
        _check(self.lib.cudaBVH_sceneTraceAny(self.handle, int(occluded.shape[0]), C.byref(rp), _ptr(occluded),
                                              flags), "sceneTraceAny")

    def trace_closest_queued(self, capacity: int, d_count, rays, t, inst_id=None, prim_id=None, u=None, v=None,
                             flags: int = 0) -> None:
        rp = ray_packet(rays)
        hp = CudaSceneHit(_ptr(t), _ptr(inst_id), _ptr(prim_id), _ptr(u), _ptr(v))
        _check(self.lib.cudaBVH_sceneTraceClosestQueued(self.handle, capacity, _ptr(d_count), C.byref(rp),
                                                        C.byref(hp), flags), "sceneTraceClosestQueued")

    def trace_status(self) -> int:
        """Traversal status word of every trace on this scene since the last
        check (bit 0: a stack overflowed); clears it."""
        w = self.lib.cudaBVH_sceneGetTraceStatus(self.handle)
        if w < 0:
            raise CudaBVHError(f"sceneGetTraceStatus: {last_error()}")
        return w

    def launch_info(self) -> dict:
        """Resources of the last traversal launch on this scene."""
        return _launch_info(self.lib.cudaBVH_sceneGetLaunchInfo, self.handle)

    def hit_normals(self, inst_id, prim_id, nx, ny, nz, flags: int = 0) -> None:
        """World-space unit geometric normals of closest hits (device arrays)."""
        hp = CudaSceneHit(None, _ptr(inst_id), _ptr(prim_id), None, None)
        _check(self.lib.cudaBVH_sceneHitNormals(self.handle, int(inst_id.shape[0]), C.byref(hp), _ptr(nx), _ptr(ny),
                                                _ptr(nz), flags), "sceneHitNormals")

    def trace_any_queued(self, capacity: int, d_count, rays, occluded, flags: int = 0) -> None:
        rp = ray_packet(rays)
        _check(self.lib.cudaBVH_sceneTraceAnyQueued(self.handle, capacity, _ptr(d_count), C.byref(rp),
                                                    _ptr(occluded), flags), "sceneTraceAnyQueued")


def _scene_path_trace(self, cam: CudaCamera, params: CudaPathParams, rgb, albedo, normal, depth,
                      counts: bool = False, flags: int = 0):
    """Renders into device buffers rgb/albedo/normal (3*W*H) and depth (W*H);
    returns the per-bounce ray counts (extension, then shadow) when asked."""
    img = CudaPathImage(_ptr(rgb), _ptr(albedo), _ptr(normal), _ptr(depth))
    cnt = np.zeros(2 * params.maxDepth, np.int64) if counts else None
    _check(self.lib.cudaBVH_scenePathTrace(self.handle, C.byref(cam), C.byref(params), C.byref(img), _ptr(cnt),
                                           flags), "scenePathTrace")
    return cnt


Scene.path_trace = _scene_path_trace


def _scene_autotune_path_tracer(self, cam: CudaCamera, params: CudaPathParams, rgb, albedo, normal, depth,
                                reps: int = 4):
    """Tunes the tail threshold (the queue size below which one thread per
    path finishes the bounces) on this frame; returns (chosen, [(threshold,
    ms per sample)]). Later renders use it with tail_paths=-1."""
    img = CudaPathImage(_ptr(rgb), _ptr(albedo), _ptr(normal), _ptr(depth))
    res = np.zeros(12, np.float32)
    k = C.c_int()
    _check(self.lib.cudaBVH_sceneAutotunePathTracer(self.handle, C.byref(cam), C.byref(params), C.byref(img), reps,
                                                    res.ctypes.data, C.byref(k)), "sceneAutotunePathTracer")
    table = [(int(res[2 * i]), float(res[2 * i + 1])) for i in range(k.value)]
    return min(table, key=lambda x: x[1])[0], table


Scene.autotune_path_tracer = _scene_autotune_path_tracer


class Batch:
    """Owner of a CudaBVHBatch_t: the BVHs of many scenes built in one pass."""

    def __init__(self, scenes_tris: list):
        self.lib = load_library()
        counts = [len(np.asarray(t).reshape(-1, 9)) for t in scenes_tris]
        self.tri_offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        allt = np.ascontiguousarray(np.concatenate([np.asarray(t, np.float32).reshape(-1, 9) for t in scenes_tris]))
        h = self.lib.cudaBVH_buildBatch(allt.ctypes.data, self.tri_offsets.ctypes.data, len(counts))
        if not h:
            raise CudaBVHError(f"buildBatch: {last_error()}")
        self.handle = C.c_void_p(h)
        self.num_scenes = len(counts)

    def destroy(self) -> None:
        if getattr(self, "handle", None):
            self.lib.cudaBVH_batchDestroy(self.handle)
            self.handle = C.c_void_p(0)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def set_stream(self, stream: int | None) -> None:
        _check(self.lib.cudaBVH_batchSetStream(self.handle, stream), "batchSetStream")

    def forest(self) -> BVH:
        v = BVH.__new__(BVH)
        v.lib, v.handle = self.lib, C.c_void_p(self.lib.cudaBVH_batchForest(self.handle))
        v.num_input = int(self.tri_offsets[-1])
        v._borrowed = True
        return v

    def roots(self) -> np.ndarray:
        out = np.zeros(self.num_scenes, np.int32)
        _check(self.lib.cudaBVH_batchRoots(self.handle, out.ctypes.data), "batchRoots")
        return out

    def trace_closest(self, total: int, d_ray_offsets, rays, t, prim_id=None, u=None, v=None, flags: int = 0) -> None:
        rp, hp = ray_packet(rays), hit_packet(t, prim_id, u, v)
        _check(self.lib.cudaBVH_batchTraceClosest(self.handle, total, _ptr(d_ray_offsets), C.byref(rp), C.byref(hp),
                                                  flags), "batchTraceClosest")

    def trace_any(self, total: int, d_ray_offsets, rays, occluded, flags: int = 0) -> None:
        rp = ray_packet(rays)
        _check(self.lib.cudaBVH_batchTraceAny(self.handle, total, _ptr(d_ray_offsets), C.byref(rp), _ptr(occluded),
                                              flags), "batchTraceAny")


class Denoiser:
    """Owner of a CudaDenoiser_t (the BF16 MFMA U-Net and its Adam state)."""

    def __init__(self, max_pixels: int, seed: int = 0):
        self.lib = load_library()
        h = self.lib.cudaBVH_denoiserCreate(max_pixels, seed)
        if not h:
            raise CudaBVHError(f"denoiserCreate: {last_error()}")
        self.handle = C.c_void_p(h)
        self.num_params = int(self.lib.cudaBVH_denoiserNumParams(self.handle))

    def destroy(self) -> None:
        if getattr(self, "handle", None):
            self.lib.cudaBVH_denoiserDestroy(self.handle)
            self.handle = C.c_void_p(0)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def set_stream(self, stream: int | None) -> None:
        _check(self.lib.cudaBVH_denoiserSetStream(self.handle, stream), "denoiserSetStream")

    def params(self, which: int = 0) -> np.ndarray:
        """0 weights, 1 gradients of the last step, 2/3 Adam moments (FP32, host)."""
        out = np.zeros(self.num_params, np.float32)
        _check(self.lib.cudaBVH_denoiserGetParams(self.handle, out.ctypes.data, which), "denoiserGetParams")
        return out

    def set_params(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.num_params
        _check(self.lib.cudaBVH_denoiserSetParams(self.handle, w.ctypes.data), "denoiserSetParams")

    def pack(self, n: int, rgb, albedo, normal, depth, out) -> None:
        _check(self.lib.cudaBVH_denoiserPack(self.handle, n, _ptr(rgb), _ptr(albedo), _ptr(normal), _ptr(depth),
                                             _ptr(out)), "denoiserPack")

    def forward(self, batch: int, height: int, width: int, x, head_out) -> None:
        _check(self.lib.cudaBVH_denoiserForward(self.handle, batch, height, width, _ptr(x), _ptr(head_out)),
               "denoiserForward")

    def train_step(self, batch: int, height: int, width: int, x, target, lr: float, want_loss: bool = True):
        loss = C.c_float(0.0)
        _check(self.lib.cudaBVH_denoiserTrainStep(self.handle, batch, height, width, _ptr(x), _ptr(target),
                                                  float(lr), C.byref(loss) if want_loss else None),
               "denoiserTrainStep")
        return loss.value if want_loss else None

    def denoise(self, width: int, height: int, rgb, albedo, normal, depth, out) -> None:
        _check(self.lib.cudaBVH_denoise(self.handle, width, height, _ptr(rgb), _ptr(albedo), _ptr(normal),
                                        _ptr(depth), _ptr(out)), "denoise")


def gen_pinhole_rays(width: int, height: int, cam: CudaCamera, rays, stream: int | None = None) -> None:
    rp = ray_packet(rays)
    _check(load_library().cudaBVH_genPinholeRays(width, height, C.byref(cam), C.byref(rp), stream), "genPinholeRays")


def build_tlas_fast(aabbs: np.ndarray):
    """LBVH TLAS: returns (nodes uint32 (n, 8), ordered int32 (m,))."""
    lib = load_library()
    a = np.ascontiguousarray(aabbs, dtype=np.float32).reshape(-1, 6)
    nodes, order = C.c_void_p(), C.c_void_p()
    n = lib.cudaBVH_buildTLASFast(a.ctypes.data, a.shape[0], C.byref(nodes), C.byref(order))
    if n <= 0:
        raise CudaBVHError(f"buildTLASFast: {last_error()}")
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    out = np.ctypeslib.as_array(C.cast(nodes, C.POINTER(C.c_uint32)), shape=(n * 8,)).reshape(n, 8).copy()
    ordv = np.ctypeslib.as_array(C.cast(order, C.POINTER(C.c_int32)), shape=(a.shape[0],)).copy()
    libc.free(nodes)
    libc.free(order)
    return out, ordv
