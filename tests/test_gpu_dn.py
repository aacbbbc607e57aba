"""GPU checks of the BF16 MFMA denoiser against the FP64 oracle
(oracle/dn_ref.py): forward output, every parameter gradient, the Adam update,
and that training on the GPU lowers the loss."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dn_ref():
    import dn_ref
    return dn_ref


def _bf16_bits(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def _inputs(seed, B, H, W):
    rng = np.random.default_rng(seed)
    x = np.zeros((B, H, W, 16), np.float32)
    x[..., :10] = rng.uniform(-1, 1, (B, H, W, 10))
    bits = _bf16_bits(x)
    xv = (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    t = rng.uniform(-1, 1, (B, H, W, 3)).astype(np.float32)
    return bits, xv, t


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


# BF16 storage makes the GPU and any FP64 restatement drift apart once a
# rounding boundary flips (the GPU's FP32 accumulation order differs from
# FP64) and the flip propagates (through max-pool choices between values that
# BF16 makes equal, ReLU signs, the L1 sign). On these small images most
# (shape, seed) cases show no flip at all (error 0) and a few show one, whose
# effect on the first layers' weight gradients can reach ~10 % — equally for
# every convolution kernel variant with a different accumulation order
# (measured on the MI355X over seeds 3-5). So each shape runs three seeds: every
# layer must stay below 0.15 against the BF16-emulating oracle in every case
# (a wrong tap, channel chunk or transpose gives errors of order 1), and the
# median over the seeds of the worst layer must be below 5e-2 (one unlucky
# flip in two seeds out of three still passes, a systematic error does not);
# the forward output stays below 3e-2 in every case and its median below 1e-2. Against the
# exact FP64 network the BF16 gradients differ by up to ~15 % per layer
# (measured on the CPU between the two oracle modes), hence the loose bound
# there. The gradient check uses targets far below the prediction, so the L1
# loss gradient has the same sign everywhere on both sides.
@pytest.mark.parametrize("B,H,W", [(3, 8, 8), (1, 16, 16), (1, 24, 40), (1, 16, 136)])
def test_forward_and_gradients(pkg, dn_ref, dev, B, H, W):
    import torch
    worst, fwd = [], []
    for seed in (3, 4, 5):
        print(f"dn grads B={B} H={H} W={W} seed={seed}", flush=True)
        d = pkg.Denoiser(4096, seed=1)
        params = dn_ref.init_like_gpu_shapes(seed=2)
        assert params.size == d.num_params
        d.set_params(params)
        bits, xv, t = _inputs(seed, B, H, W)
        X = torch.from_numpy(bits.view(np.int16)).to(dev)
        head = torch.empty((B, H, W, 3), device=dev)
        d.forward(B, H, W, X, head)
        torch.cuda.synchronize()
        hg = head.cpu().numpy()
        h_ref, _ = dn_ref.forward(params, xv)
        h_emu, _ = dn_ref.forward(params, xv, emulate=True)
        assert _rel(hg, h_ref) < 3e-2
        fwd.append(_rel(hg, h_emu))
        assert fwd[-1] < 3e-2
        # loss value on a random target
        T = torch.from_numpy(t).to(dev)
        loss = d.train_step(B, H, W, X, T, 0.0)
        loss_ref, _, _ = dn_ref.loss_and_grads(params, xv, t.astype(np.float64))
        assert abs(loss - loss_ref) < 1e-2 * loss_ref
        # gradients with a one-signed loss gradient
        tgt = (h_ref + xv[..., :3] - 5.0).astype(np.float32)
        loss = d.train_step(B, H, W, X, torch.from_numpy(tgt).to(dev), 0.0)
        g = d.params(1)
        _, g_ref, _ = dn_ref.loss_and_grads(params, xv, tgt.astype(np.float64))
        _, g_emu, _ = dn_ref.loss_and_grads(params, xv, tgt.astype(np.float64), emulate=True)
        offs, _ = dn_ref.offsets()
        errs = []
        for i, (wo, bo) in enumerate(offs):
            cout = dn_ref.LAYERS[i][1]
            errs.append((i, round(_rel(g[wo:bo], g_emu[wo:bo]), 4), round(_rel(g[bo:bo + cout], g_emu[bo:bo + cout]), 4),
                         round(_rel(g[wo:bo], g_ref[wo:bo]), 4)))
        print("per layer (weights vs emulated, bias vs emulated, weights vs exact):", errs, flush=True)
        for i, ew, eb, ex in errs:
            assert ew < 0.15 and eb < 0.15 and ex < 0.25, (seed, i, ew, eb, ex)
        worst.append(max(max(ew, eb) for _, ew, eb, _ in errs))
        cos = float(g @ g_ref / (np.linalg.norm(g) * np.linalg.norm(g_ref)))
        assert cos > 0.98
        d.destroy()
    assert float(np.median(worst)) < 5e-2, worst
    assert float(np.median(fwd)) < 1e-2, fwd


def test_adam_update_and_training(pkg, dn_ref, dev):
    import torch
    d = pkg.Denoiser(8192, seed=4)
    w0 = d.params(0)
    assert np.all(w0[dn_ref.offsets()[0][0][0]:dn_ref.offsets()[0][0][1]].reshape(32, 9, 16)[:, :, 10:] == 0)
    B, H, W = 2, 32, 32
    bits, xv, t = _inputs(5, B, H, W)
    # a learnable target: a smooth function of the input channels
    t = (0.5 * np.tanh(xv[..., :3] + xv[..., 3:6] * xv[..., 6:9])).astype(np.float32)
    X = torch.from_numpy(bits.view(np.int16)).to(dev)
    T = torch.from_numpy(t).to(dev)
    lr = 1e-3
    l0 = d.train_step(B, H, W, X, T, lr)
    g = d.params(1).astype(np.float64)
    # first Adam step from zero moments: w1 = w0 - lr * m/(1-b1) / (sqrt(v/(1-b2)) + eps)
    m = 0.1 * g
    v = 0.001 * g * g
    want = w0 - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    w1 = d.params(0)
    assert np.allclose(w1, want, rtol=1e-5, atol=1e-7)
    losses = [l0]
    for _ in range(100):
        losses.append(d.train_step(B, H, W, X, T, lr))
    assert losses[-1] < 0.5 * losses[0], losses[::10]
    # denoise() = exp(head + log(1+rgb)) - 1 on packed planes
    n = H * W
    rng = np.random.default_rng(6)
    planes = {k: torch.from_numpy(rng.uniform(0, 1, (c * n,)).astype(np.float32)).to(dev)
              for k, c in (("rgb", 3), ("alb", 3), ("nrm", 3), ("dep", 1))}
    out = torch.empty(3 * n, device=dev)
    d.denoise(W, H, planes["rgb"], planes["alb"], planes["nrm"], planes["dep"], out)
    packed = torch.empty((n, 16), dtype=torch.int16, device=dev)
    d.pack(n, planes["rgb"], planes["alb"], planes["nrm"], planes["dep"], packed)
    head = torch.empty((n, 3), device=dev)
    d.forward(1, H, W, packed, head)
    torch.cuda.synchronize()
    pk = packed.cpu().numpy().view(np.uint16)
    logrgb = (pk[:, :3].astype(np.uint32) << 16).view(np.float32)
    want = np.maximum(np.expm1(head.cpu().numpy() + logrgb), 0.0).T.reshape(-1)
    assert np.allclose(out.cpu().numpy(), want, rtol=1e-5, atol=1e-6)
    # packing: log1p(rgb), albedo, normal, log1p(depth)/4 rounded to bf16, zeros after
    p = {k: v.cpu().numpy() for k, v in planes.items()}
    ref = np.zeros((n, 16), np.float32)
    ref[:, 0:3] = np.log1p(p["rgb"].reshape(3, n).T)
    ref[:, 3:6] = p["alb"].reshape(3, n).T
    ref[:, 6:9] = p["nrm"].reshape(3, n).T
    ref[:, 9] = 0.25 * np.log1p(p["dep"])
    got = (pk.astype(np.uint32) << 16).view(np.float32)
    assert np.allclose(got, ref, rtol=2 ** -7, atol=1e-6)
    assert np.all(pk[:, 10:] == 0)
    d.destroy()


def test_forward_large_tiles(pkg, dn_ref, dev):
    """Images of 65536 pixels and more use 256-pixel convolution tiles: the
    output stays within BF16 drift of the oracle (a tiling error would not)."""
    import torch
    d = pkg.Denoiser(1 << 17, seed=1)
    params = dn_ref.init_like_gpu_shapes(seed=4)
    d.set_params(params)
    B, H, W = 1, 256, 256
    bits, xv, _ = _inputs(6, B, H, W)
    X = torch.from_numpy(bits.view(np.int16)).to(dev)
    head = torch.empty((B, H, W, 3), device=dev)
    d.forward(B, H, W, X, head)
    torch.cuda.synchronize()
    h_emu, _ = dn_ref.forward(params, xv, emulate=True)
    err = _rel(head.cpu().numpy(), h_emu)
    print("256x256 head vs emulated", err, flush=True)
    assert err < 3e-2
    d.destroy()
