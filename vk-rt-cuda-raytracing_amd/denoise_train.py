"""Training and evaluation of the learned denoiser on path-traced pairs.

Everything runs through the library's C ABI: the path tracer renders the
pairs (1 spp input with first-hit albedo / normal / depth, many-spp target),
the denoiser's own forward/backward kernels and Adam train on them, and the
evaluation compares, at equal GPU time, the 1 spp + denoiser image with plain
sample accumulation against a high-spp reference (PSNR of the tonemapped
image x / (1 + x), peak 1).

Training pairs are 256x256 crops of 1920x1080 views, so the network sees the
feature scale and noise of the 1080p renders it denoises. By default each view
is rendered whole (1 spp input, many-spp target) and cropped; the window mode
renders each crop alone through a camera restricted to the window (same pixel
footprint), which is cheaper for low target spp.
"""
from __future__ import annotations

import importlib
import math
import os
import sys
import time

import numpy as np

if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
pkg = importlib.import_module("vk-rt-cuda-raytracing_amd")
scenes = importlib.import_module("vk-rt-cuda-raytracing_amd.scenes")

# Trained parameters shipped with the package (BF16 bit patterns, .npy),
# written by train_and_save() below.
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "denoiser_instanced.npy")


def load_weights(path: str = WEIGHTS) -> np.ndarray:
    """FP32 parameters from the stored BF16 bit patterns."""
    bits = np.load(path, allow_pickle=False).astype(np.uint16)
    return (bits.astype(np.uint32) << 16).view(np.float32)


def save_weights(params: np.ndarray, path: str = WEIGHTS) -> None:
    """Stores the parameters as BF16 (round to nearest even) bit patterns."""
    u = np.ascontiguousarray(params, np.float32).view(np.uint32).astype(np.uint64)
    bits = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, bits, allow_pickle=False)


def window_camera(cam: dict, W: int, H: int, x0: int, y0: int, w: int, h: int) -> dict:
    """Camera of the w x h window at (x0, y0) of a W x H image of `cam`.

    Pixel (px, py) of the window looks where pixel (x0+px, y0+py) of the full
    image looks: d = f' + sx'*scale_x'*r + sy'*scale_y'*u with
    f' = f + cx*scale_x*r + cy*scale_y*u (cx, cy: window centre in the full
    image's [-1, 1] coordinates), scale' = scale * (w/W, h/H)."""
    cx = (x0 + 0.5 * w) / W * 2.0 - 1.0
    cy = (y0 + 0.5 * h) / H * 2.0 - 1.0
    f = (cam["f"].astype(np.float64) + cx * float(cam["scale_x"]) * cam["r"].astype(np.float64) +
         cy * float(cam["scale_y"]) * cam["u"].astype(np.float64))
    return dict(o=cam["o"], f=f.astype(np.float32), r=cam["r"], u=cam["u"],
                scale_x=np.float32(float(cam["scale_x"]) * w / W), scale_y=np.float32(float(cam["scale_y"]) * h / H))


def cam_struct(cam: dict):
    return pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])


def random_view(rng, W: int = 1920, H: int = 1080, span: float = 10.0) -> dict:
    """A seeded camera inside or near the instance volume looking at a point in it."""
    for _ in range(100):
        eye = rng.uniform(-1.3 * span, 1.3 * span, 3)
        tgt = rng.uniform(-0.5 * span, 0.5 * span, 3)
        if np.linalg.norm(tgt - eye) > 4.0:
            return scenes.look_at(tuple(eye), tuple(tgt), W, H, 60.0)
    return scenes.instanced_camera(W, H)


class Renderer:
    """A two-level scene on the GPU plus image buffers for w x h renders."""

    def __init__(self, sc: dict, dev):
        import torch
        self.torch = torch
        self.dev = dev
        self.blas = [pkg.BVH.build(m) for m in sc["meshes"]]
        self.scene = pkg.Scene(self.blas, sc["o2w"], sc["blas_of"], sc["albedo"])

    def render(self, cam: dict, w: int, h: int, spp: int, seed: int, flags: int = 0):
        t = self.torch
        n = w * h
        out = dict(rgb=t.empty(3 * n, device=self.dev), alb=t.empty(3 * n, device=self.dev),
                   nrm=t.empty(3 * n, device=self.dev), dep=t.empty(n, device=self.dev))
        p = pkg.path_params(w, h, spp=spp, seed=seed)
        cnt = self.scene.path_trace(cam_struct(cam), p, out["rgb"], out["alb"], out["nrm"], out["dep"], counts=True,
                                    flags=flags)
        return out, cnt

    def destroy(self):
        self.scene.destroy()
        for b in self.blas:
            b.destroy()


def tonemap(x):
    return x / (1.0 + x)


def psnr(a, ref) -> float:
    """PSNR of tonemapped images (peak 1)."""
    mse = float(((tonemap(np.maximum(a, 0)) - tonemap(np.maximum(ref, 0))) ** 2).mean())
    return 10.0 * math.log10(1.0 / max(mse, 1e-20))


def rel_mse(a, ref) -> float:
    return float(((a - ref) ** 2 / (ref ** 2 + 1e-2)).mean())


def make_pairs(dev, scene_seeds, views_per_scene: int, windows_per_view: int, target_spp: int, crop: int = 256,
               seed: int = 0, mesh_res=scenes.INSTANCED_MESH_RES, log=print, full_frame: bool = True):
    """Training pairs: returns (inputs int16 [N, crop, crop, 16] bf16 bits on
    the device, targets float32 [N, crop, crop, 3] = log(1 + reference)).
    N = len(scene_seeds) * views_per_scene * windows_per_view; crops start at
    multiples of 8 pixels."""
    t = importlib.import_module("torch")
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    W, H = 1920, 1080
    npx = W * H if full_frame else crop * crop
    dn = pkg.Denoiser((npx + 63) // 64 * 64, seed=0)      # only for packing
    for s in scene_seeds:
        sc = scenes.instanced_scene(mesh_res=mesh_res, seed=int(s))
        r = Renderer(sc, dev)
        for v in range(views_per_scene):
            cam = random_view(rng)
            if full_frame:
                sd = int(rng.integers(1, 2 ** 31))
                noisy, _ = r.render(cam, W, H, 1, sd)
                ref, _ = r.render(cam, W, H, target_spp, sd + 1, flags=pkg.FLAG_GRAPH)
                xf = t.empty((W * H, 16), dtype=t.int16, device=dev)
                dn.pack(W * H, noisy["rgb"], noisy["alb"], noisy["nrm"], noisy["dep"], xf)
                xf = xf.view(H, W, 16)
                yf = t.log1p(ref["rgb"].view(3, H, W).permute(1, 2, 0).clamp_min(0))
                for k in range(windows_per_view):
                    x0 = int(rng.integers(0, (W - crop) // 8 + 1)) * 8
                    y0 = int(rng.integers(0, (H - crop) // 8 + 1)) * 8
                    xs.append(xf[y0:y0 + crop, x0:x0 + crop].contiguous())
                    ys.append(yf[y0:y0 + crop, x0:x0 + crop].contiguous())
                del noisy, ref, xf, yf
                t.cuda.synchronize()
                log(f"pairs: scene {s} view {v}: {target_spp} spp frame done ({len(xs)} pairs)")
                continue
            for k in range(windows_per_view):
                x0 = int(rng.integers(0, (W - crop) // 8 + 1)) * 8
                y0 = int(rng.integers(0, (H - crop) // 8 + 1)) * 8
                wc = window_camera(cam, W, H, x0, y0, crop, crop)
                sd = int(rng.integers(1, 2 ** 31))
                noisy, _ = r.render(wc, crop, crop, 1, sd)
                ref, _ = r.render(wc, crop, crop, target_spp, sd + 1, flags=pkg.FLAG_GRAPH)
                x = t.empty((crop * crop, 16), dtype=t.int16, device=dev)
                dn.pack(crop * crop, noisy["rgb"], noisy["alb"], noisy["nrm"], noisy["dep"], x)
                xs.append(x.view(crop, crop, 16))
                ys.append(t.log1p(ref["rgb"].view(3, crop, crop).permute(1, 2, 0).clamp_min(0)).contiguous())
        t.cuda.synchronize()
        r.destroy()
        log(f"pairs: scene {s} done ({len(xs)} pairs)")
    dn.destroy()
    return t.stack(xs), t.stack(ys)


def train(dn, X, Y, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 0, log=print, log_every: int = 100):
    """Adam on random minibatches (with random flips) of the pairs; cosine
    learning-rate decay. Returns the loss history."""
    t = importlib.import_module("torch")
    rng = np.random.default_rng(seed)
    N, h, w = X.shape[0], X.shape[1], X.shape[2]
    hist = []
    t0 = time.time()
    for it in range(steps):
        idx = t.from_numpy(rng.integers(0, N, batch)).to(X.device)
        xb, yb = X[idx], Y[idx]
        if rng.random() < 0.5:          # horizontal flip (normals keep their world-space values)
            xb, yb = xb.flip(2), yb.flip(2)
        xb = xb.contiguous()
        yb = yb.contiguous()
        cur = lr * 0.5 * (1.0 + math.cos(math.pi * it / steps))
        want = (it % log_every == 0) or it == steps - 1
        loss = dn.train_step(batch, h, w, xb, yb, cur, want_loss=want)
        if want:
            hist.append((it, loss))
            log(f"step {it}: loss {loss:.5f} lr {cur:.2e} ({time.time() - t0:.1f} s)")
    return hist


def train_and_save(dev, scene_seeds=range(100, 108), views: int = 4, windows: int = 16, target_spp: int = 4096,
                   steps: int = 4000, batch: int = 8, lr: float = 1e-3, seed: int = 7, path: str = WEIGHTS, log=print,
                   full_frame: bool = True):
    """Renders the training pairs (default: 8 scenes x 4 views x 16 crops =
    512 pairs of 1 spp / 4096 spp), trains a fresh network and stores it."""
    t0 = time.time()
    X, Y = make_pairs(dev, scene_seeds, views, windows, target_spp, log=log, full_frame=full_frame)
    log(f"{X.shape[0]} pairs in {time.time() - t0:.1f} s")
    dn = pkg.Denoiser(max(batch * X.shape[1] * X.shape[2], 1920 * 1088), seed=seed)
    hist = train(dn, X, Y, steps, batch=batch, lr=lr, log=log)
    save_weights(dn.params(0), path)
    dn.destroy()
    return hist


def time_inference(dev, frames: int = 3, W: int = 1920, H: int = 1080, path: str = WEIGHTS, log=print) -> float:
    """Runs `frames` denoiser passes (pack, U-Net, unpack) on W x H planes of
    uniform noise with the stored weights; returns the mean ms per frame."""
    t = importlib.import_module("torch")
    n = W * H
    dn = pkg.Denoiser(((W + 7) // 8 * 8) * ((H + 7) // 8 * 8), seed=0)     # levels pad to multiples of 8
    dn.set_params(load_weights(path))
    g = t.Generator(device=dev)
    g.manual_seed(0)
    planes = [t.rand(c * n, device=dev, generator=g) for c in (3, 3, 3, 1)]
    out = t.empty(3 * n, device=dev)
    dn.denoise(W, H, *planes, out)
    e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(frames):
        dn.denoise(W, H, *planes, out)
    e1.record()
    t.cuda.synchronize()
    ms = e0.elapsed_time(e1) / frames
    log(f"denoiser {W}x{H}: {ms:.3f} ms/frame over {frames} frames (mean output {float(out.mean()):.4f})")
    dn.destroy()
    return ms


if __name__ == "__main__":
    import argparse
    import torch
    ap = argparse.ArgumentParser(description="render training pairs with the path tracer and train the denoiser")
    ap.add_argument("--time-inference", type=int, default=0, metavar="FRAMES",
                    help="only time FRAMES 1920x1080 denoiser passes with the stored weights")
    ap.add_argument("--scenes", type=int, default=8)
    ap.add_argument("--views", type=int, default=4)
    ap.add_argument("--windows", type=int, default=16)
    ap.add_argument("--target-spp", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--window-mode", action="store_true", help="render each crop alone instead of whole frames")
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args()
    if a.time_inference:
        time_inference(torch.device("cuda:0"), a.time_inference, log=lambda m: print(m, flush=True))
        sys.exit(0)
    train_and_save(torch.device("cuda:0"), range(100, 100 + a.scenes), a.views, a.windows, a.target_spp, a.steps,
                   path=a.out, log=lambda m: print(m, flush=True), full_frame=not a.window_mode)
