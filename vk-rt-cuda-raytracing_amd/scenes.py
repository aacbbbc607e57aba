"""Seeded synthetic scenes and cameras for the benchmark workloads.

All generators are deterministic functions of their seed (numpy PCG64 via
``default_rng``) and return float32 triangles of shape (n, 9) in original
order: v0.xyz, v1.xyz, v2.xyz.

* random_triangles  — small random triangles (the 65,536-triangle parity scene
  and the many-scene batches).
* terrain_scene     — a 1024 x 1024-cell sine-octave heightfield over
  [-10, 10]^2 (2,097,152 triangles) plus 256 icospheres of subdivision level 4
  (5,120 triangles each): 3,407,872 triangles in all. The icospheres are added
  on top of the heightfield triangles.
"""
from __future__ import annotations

import math

import numpy as np

TAN30 = math.tan(math.radians(30.0))


def random_triangles(n: int, seed: int = 0, min_len: float = 0.005, max_len: float = 0.05) -> np.ndarray:
    """Centroids uniform in [-1,1]^3; each vertex = centroid + (uniform unit
    direction) * length uniform in [min_len, max_len]."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, size=(n, 1, 3))
    d = rng.normal(size=(n, 3, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    ln = rng.uniform(min_len, max_len, size=(n, 3, 1))
    return (c + d * ln).astype(np.float32).reshape(n, 9)


# ----------------------------------------------------------------------------
# Terrain scene
# ----------------------------------------------------------------------------
def _octaves(seed: int, n: int = 6):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 2.0 * np.pi, n), rng.uniform(0.0, 2.0 * np.pi, n)


def terrain_height(x, z, seed: int = 1, amp0: float = 1.0, freq0: float = 0.35, octaves: int = 6):
    """Sum of `octaves` sine waves; amplitude halves and frequency doubles."""
    theta, phase = _octaves(seed, octaves)
    h = np.zeros(np.broadcast(x, z).shape, np.float64)
    for k in range(octaves):
        a, f = amp0 * 0.5 ** k, freq0 * 2.0 ** k
        h += a * np.sin(f * (x * np.cos(theta[k]) + z * np.sin(theta[k])) + phase[k])
    return h


def heightfield(res: int = 1024, extent: float = 10.0, seed: int = 1) -> np.ndarray:
    xs = np.linspace(-extent, extent, res + 1)
    X, Z = np.meshgrid(xs, xs, indexing="xy")          # X varies along columns
    Y = terrain_height(X, Z, seed)
    V = np.stack([X, Y, Z], axis=-1)                   # (res+1, res+1, 3)
    v00, v10 = V[:-1, :-1], V[:-1, 1:]
    v01, v11 = V[1:, :-1], V[1:, 1:]
    t0 = np.concatenate([v00, v10, v11], axis=-1).reshape(-1, 9)
    t1 = np.concatenate([v00, v11, v01], axis=-1).reshape(-1, 9)
    tris = np.stack([t0, t1], axis=1).reshape(-1, 9)   # cell-interleaved
    return tris.astype(np.float32)


def icosphere(level: int = 4) -> np.ndarray:
    """Unit icosphere as (20*4^level, 3, 3) float64."""
    p = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
         (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = np.array(v, np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    T = V[np.array(f)]                                 # (20, 3, 3)
    for _ in range(level):
        a, b, c = T[:, 0], T[:, 1], T[:, 2]
        ab, bc, ca = a + b, b + c, c + a
        ab /= np.linalg.norm(ab, axis=1, keepdims=True)
        bc /= np.linalg.norm(bc, axis=1, keepdims=True)
        ca /= np.linalg.norm(ca, axis=1, keepdims=True)
        T = np.stack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1), np.stack([ca, bc, c], 1),
                      np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3, 3)
    return T


def terrain_scene(res: int = 1024, spheres: int = 256, level: int = 4, seed: int = 1) -> np.ndarray:
    """Heightfield plus `spheres` icospheres (radii uniform in [0.1, 1.0])
    resting partly above the terrain."""
    hf = heightfield(res, 10.0, seed)
    if spheres <= 0:
        return hf
    rng = np.random.default_rng(seed + 1000)
    cx = rng.uniform(-9.0, 9.0, spheres)
    cz = rng.uniform(-9.0, 9.0, spheres)
    r = rng.uniform(0.1, 1.0, spheres)
    cy = terrain_height(cx, cz, seed) + r * rng.uniform(0.3, 1.2, spheres)
    unit = icosphere(level)                            # (m, 3, 3)
    C = np.stack([cx, cy, cz], 1)[:, None, None, :]
    S = (unit[None] * r[:, None, None, None] + C).reshape(-1, 9).astype(np.float32)
    return np.concatenate([hf, S], axis=0)


# ----------------------------------------------------------------------------
# Cameras: (o, f, r, u, scale_x, scale_y) for the pinhole generator
#   d = f + (sx*scale_x)*r + (sy*scale_y)*u,  sx,sy in (-1,1); row 0 on top.
# ----------------------------------------------------------------------------
def look_at(eye, target, width: int, height: int, vfov_deg: float = 60.0, world_up=(0.0, 1.0, 0.0)):
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(np.asarray(world_up, np.float64), f)
    r /= np.linalg.norm(r)
    up = np.cross(f, r)
    th = math.tan(math.radians(vfov_deg) / 2.0)
    return dict(o=eye.astype(np.float32), f=f.astype(np.float32), r=r.astype(np.float32),
                u=(-up).astype(np.float32), scale_x=np.float32(th * width / height), scale_y=np.float32(th))


def parity_camera(width: int = 512, height: int = 512):
    """(0,0,-3) looking +z, 60 degree vertical FOV."""
    return look_at((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), width, height, 60.0)


# The terrain camera looks down onto the heightfield so that most primary
# rays hit geometry: from (-3,5,-9) toward (0,-1,0), 60 degree vertical FOV,
# about 82% of 1920x1080 rays hit (from (0,8,-18) toward the origin only about
# 32% do, the rest leave over the far edge into the sky).
TERRAIN_EYE = (-3.0, 5.0, -9.0)
TERRAIN_TARGET = (0.0, -1.0, 0.0)


def terrain_camera(width: int = 1920, height: int = 1080):
    return look_at(TERRAIN_EYE, TERRAIN_TARGET, width, height, 60.0)


TERRAIN_LIGHT = (0.0, 20.0, 0.0)


def cam_array(cam) -> np.ndarray:
    """14 floats: o f r u scale_x scale_y (the oracle's camera layout)."""
    return np.concatenate([cam["o"], cam["f"], cam["r"], cam["u"],
                           np.array([cam["scale_x"], cam["scale_y"]], np.float32)]).astype(np.float32)


# ----------------------------------------------------------------------------
# Instanced two-level scenes (TLAS over instances of a few meshes)
# ----------------------------------------------------------------------------
# Four heightfield patches over [-1,1]^2 of 10,082 / 49,928 / 199,712 /
# 500,000 triangles (grid resolutions 71 / 158 / 316 / 500).
INSTANCED_MESH_RES = (71, 158, 316, 500)


def random_rotation(rng) -> np.ndarray:
    """Uniform random rotation (from a normalised Gaussian quaternion), float64 3x3."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def instanced_scene(mesh_res=INSTANCED_MESH_RES, n_inst: int = 64, seed: int = 2, extent: float = 1.0,
                    scale=(0.5, 2.0), span: float = 10.0):
    """Meshes, per-instance object-to-world maps and albedos.

    Instance i uses mesh i % len(mesh_res), a seeded uniform rotation, a
    uniform scale in `scale` and a translation uniform in [-span, span]^3;
    albedo is uniform in [0.2, 0.8] per channel. Returns
    dict(meshes=[(n_k, 9) float32], o2w=(n_inst, 12) float32 row-major 3x4,
    blas_of=(n_inst,) int32, albedo=(n_inst, 3) float32)."""
    rng = np.random.default_rng(seed)
    meshes = [heightfield(res=r, extent=extent, seed=seed + 1 + k) for k, r in enumerate(mesh_res)]
    o2w = np.zeros((n_inst, 12), np.float32)
    for i in range(n_inst):
        R = random_rotation(rng) * rng.uniform(scale[0], scale[1])
        T = rng.uniform(-span, span, 3)
        o2w[i] = np.concatenate([R, T[:, None]], axis=1).reshape(12).astype(np.float32)
    blas_of = (np.arange(n_inst) % len(mesh_res)).astype(np.int32)
    albedo = rng.uniform(0.2, 0.8, (n_inst, 3)).astype(np.float32)
    return dict(meshes=meshes, o2w=o2w, blas_of=blas_of, albedo=albedo)


def instance_world_tris(sc) -> tuple:
    """All instanced triangles mapped to world space in float64 (for brute
    force checks): (tris (m, 9) float64, inst (m,), prim (m,))."""
    out, inst, prim = [], [], []
    for i, b in enumerate(sc["blas_of"]):
        M = sc["o2w"][i].astype(np.float64).reshape(3, 4)
        v = sc["meshes"][b].astype(np.float64).reshape(-1, 3)
        w = v @ M[:, :3].T + M[:, 3]
        out.append(w.reshape(-1, 9))
        inst.append(np.full(len(w) // 3, i, np.int32))
        prim.append(np.arange(len(w) // 3, dtype=np.int32))
    return np.concatenate(out), np.concatenate(inst), np.concatenate(prim)


# Camera of the instanced scene: from (0,2,-12) toward the origin, 60 degree
# vertical FOV, inside the instance volume's near half; about half of the
# 1920x1080 camera rays hit an instance (from (0,6,-26), outside the volume,
# only 16% do).
INSTANCED_EYE = (0.0, 2.0, -12.0)
INSTANCED_TARGET = (0.0, 0.0, 0.0)
SUN_DIR = (0.35, 0.85, -0.40)   # direction towards the sun (normalised on use)


def instanced_camera(width: int = 1920, height: int = 1080):
    return look_at(INSTANCED_EYE, INSTANCED_TARGET, width, height, 60.0)


# ----------------------------------------------------------------------------
# Many independent scenes (batched build and traversal)
# ----------------------------------------------------------------------------
def batch_specs(n: int = 256, seed: int = 3, tris=(1000, 200000), rays=(65536, 1 << 20)):
    """Per scene (triangle count, ray count), both log-uniform."""
    rng = np.random.default_rng(seed)
    t = np.exp(rng.uniform(np.log(tris[0]), np.log(tris[1]), n)).astype(np.int64)
    r = np.exp(rng.uniform(np.log(rays[0]), np.log(rays[1]), n)).astype(np.int64)
    return [(int(a), int(b)) for a, b in zip(t, r)]


def batch_rays(n: int, seed: int) -> np.ndarray:
    """n rays (8, n) for one scene of the random-triangle generator: the first
    half pinhole primaries of the small camera ((0,0,-3) toward +z, 60 degree
    vertical FOV) on a near-square grid, the rest from uniform points of
    [-1,1]^3 in uniform sphere directions; tmin 0, tmax 1e30."""
    rng = np.random.default_rng(seed)
    npri = n // 2
    w = max(1, int(np.ceil(np.sqrt(npri))))
    h = max(1, (npri + w - 1) // w)
    cam = parity_camera(w, h)
    i = np.arange(npri)
    px, py = i % w, i // w
    sx = ((px + 0.5) / w * 2.0 - 1.0) * float(cam["scale_x"])
    sy = ((py + 0.5) / h * 2.0 - 1.0) * float(cam["scale_y"])
    d1 = cam["f"][None].astype(np.float64) + sx[:, None] * cam["r"][None] + sy[:, None] * cam["u"][None]
    d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
    o1 = np.broadcast_to(cam["o"], (npri, 3))
    nr = n - npri
    d2 = rng.normal(size=(nr, 3))
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    o2 = rng.uniform(-1.0, 1.0, (nr, 3))
    out = np.zeros((8, n), np.float32)
    out[0:3] = np.concatenate([o1, o2]).T
    out[3:6] = np.concatenate([d1, d2]).T
    out[7] = 1e30
    return out


def batch_rays_device(n: int, seed: int, dev, cam=None):
    """The many-scenes ray mix generated on a torch device (the generator of
    the 256-scene workload): half pinhole primaries of the small camera on a
    near-square grid, half from uniform points of [-1,1]^3 in uniform sphere
    directions (torch's seeded device generator). Returns (origins (3, n),
    directions (3, n)) float32 device tensors."""
    import torch
    cam = parity_camera(512, 512) if cam is None else cam
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    npri = n // 2
    w = max(1, int(np.ceil(np.sqrt(npri))))
    h = max(1, (npri + w - 1) // w)
    i = torch.arange(npri, device=dev)
    sx = ((i % w).float() + 0.5) / w * 2 - 1
    sy = ((i // w).float() + 0.5) / h * 2 - 1
    f, r, u = (torch.tensor(cam[k], device=dev) for k in ("f", "r", "u"))
    d1 = f[:, None] + (sx * TAN30 * w / h)[None] * r[:, None] + (sy * TAN30)[None] * u[:, None]
    d1 = d1 / d1.norm(dim=0)
    o1 = torch.tensor(cam["o"], device=dev)[:, None].expand(3, npri)
    nr = n - npri
    d2 = torch.randn(3, nr, device=dev, generator=g)
    d2 = d2 / d2.norm(dim=0)
    o2 = torch.rand(3, nr, device=dev, generator=g) * 2 - 1
    return torch.cat([o1, o2], 1), torch.cat([d1, d2], 1)


def batch_workload_device(specs, dev, seed0: int = 7000):
    """Rays of every scene of `specs` (list of (triangles, rays)) concatenated
    on the device: (ray dict of contiguous (total,) tensors, host ray offsets
    (numScenes + 1,) int32). Scene k's rays come from seed seed0 + k."""
    import torch
    roff = np.concatenate([[0], np.cumsum([nr for _, nr in specs])]).astype(np.int32)
    total = int(roff[-1])
    O = torch.empty(3, total, device=dev)
    D = torch.empty(3, total, device=dev)
    for k, (_, nr) in enumerate(specs):
        o, d = batch_rays_device(nr, seed0 + k, dev)
        O[:, roff[k]:roff[k + 1]] = o
        D[:, roff[k]:roff[k + 1]] = d
    R = {"ox": O[0], "oy": O[1], "oz": O[2], "dx": D[0], "dy": D[1], "dz": D[2],
         "tmin": torch.zeros(total, device=dev), "tmax": torch.full((total,), 1e30, device=dev)}
    return {k: v.contiguous() for k, v in R.items()}, roff

