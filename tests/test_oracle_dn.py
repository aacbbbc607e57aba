"""CPU tests that pin the denoiser oracle (oracle/dn_ref.py): its forward and
every parameter gradient against torch's own float64 convolution / pooling /
interpolation routines with autograd (library routines, an independent
implementation), finite differences, and the BF16 rounding rule. No GPU."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def dn_ref():
    import dn_ref
    return dn_ref


def _torch_loss(dn_ref, params, x, target):
    import torch
    import torch.nn.functional as F
    W, Bs = dn_ref.unpack(params)
    # w[co][t][ci] with t = 3*ky + kx  ->  conv2d weight [co][ci][ky][kx]
    tw = [torch.tensor(np.ascontiguousarray(w.transpose(0, 2, 1)).reshape(w.shape[0], w.shape[2],
                                                                         *((3, 3) if w.shape[1] == 9 else (1, 1))),
                       requires_grad=True) for w in W]
    tb = [torch.tensor(b, requires_grad=True) for b in Bs]
    X = torch.tensor(x.transpose(0, 3, 1, 2))
    cv = lambda i, v: F.conv2d(v, tw[i], tb[i], padding=1 if W[i].shape[1] == 9 else 0)
    r = F.relu
    up = lambda v: F.interpolate(v, scale_factor=2, mode="nearest")
    e1 = r(cv(1, r(cv(0, X))))
    e2 = r(cv(3, r(cv(2, F.max_pool2d(e1, 2)))))
    e3 = r(cv(5, r(cv(4, F.max_pool2d(e2, 2)))))
    e4 = r(cv(7, r(cv(6, F.max_pool2d(e3, 2)))))
    d3 = r(cv(9, r(cv(8, torch.cat([up(e4), e3], 1)))))
    d2 = r(cv(11, r(cv(10, torch.cat([up(d3), e2], 1)))))
    d1 = r(cv(13, r(cv(12, torch.cat([up(d2), e1], 1)))))
    head = cv(14, d1)
    pred = head + X[:, :3]
    loss = (pred - torch.tensor(target.transpose(0, 3, 1, 2))).abs().mean()
    loss.backward()
    grads = []
    for w, b in zip(tw, tb):
        g = w.grad.numpy()
        grads.append(g.reshape(g.shape[0], g.shape[1], -1).transpose(0, 2, 1).reshape(-1))
        grads.append(b.grad.numpy())
    return float(loss.detach()), np.concatenate(grads), head.detach().numpy().transpose(0, 2, 3, 1)


def _inputs(seed, B=2, H=16, W=16):
    rng = np.random.default_rng(seed)
    x = np.zeros((B, H, W, 16))
    x[..., :10] = rng.uniform(-1, 1, (B, H, W, 10))
    t = rng.uniform(-1, 1, (B, H, W, 3))
    return x, t


def test_bf16_rounding(dn_ref):
    """Nearest bfloat16 with ties to even, by brute force over both neighbours."""
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.normal(0, 10, 20000), rng.normal(0, 1e-30, 2000),
                        [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 3e38, 1e-40]]).astype(np.float32)
    r = dn_ref.bf16(v)
    u = v.view(np.uint32)
    lo = (u & 0xFFFF0000).view(np.float32).astype(np.float64)
    hi = ((u & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32).astype(np.float64)
    x = v.astype(np.float64)
    dlo, dhi = np.abs(x - lo), np.abs(hi - x)
    even_lo = ((u >> 16) & 1) == 0
    want = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(even_lo, lo, hi)))
    assert np.array_equal(r, want)
    assert r[-5] == 1.0 and r[-4] == 1.0 + 2.0 ** -6 and r[-3] == -1.0


def test_oracle_matches_torch_autograd(dn_ref):
    pytest.importorskip("torch")
    params = dn_ref.init_like_gpu_shapes(seed=1)
    x, t = _inputs(2)
    loss, grads, head = dn_ref.loss_and_grads(params, x, t)
    tl, tg, th = _torch_loss(dn_ref, params, x, t)
    assert abs(loss - tl) < 1e-12
    assert np.allclose(head, th, rtol=1e-10, atol=1e-12)
    offs, total = dn_ref.offsets()
    assert grads.shape == tg.shape == (total,)
    assert np.allclose(grads, tg, rtol=1e-8, atol=1e-14)
    # every layer receives gradient
    for (wo, bo) in offs:
        assert np.abs(grads[wo:bo]).max() > 0


def test_oracle_finite_differences(dn_ref):
    """Central differences of the loss (exact FP64 weights) on parameters of
    several layers, including both ends of the skip connections and the head."""
    params = dn_ref.init_like_gpu_shapes(seed=3).astype(np.float64)
    x, t = _inputs(4, B=1, H=8, W=8)
    loss, grads, _ = dn_ref.loss_and_grads(params, x, t, round_bf16=False)
    offs, _ = dn_ref.offsets()
    rng = np.random.default_rng(5)
    checked = 0
    for li in (0, 5, 7, 8, 12, 14):
        wo, bo = offs[li]
        ks = [k for k in rng.integers(wo, bo, 12) if not (li == 0 and (k - wo) % 16 >= 10)][:3] + [bo]
        for k in ks:
            h = 1e-6
            p1 = params.copy(); p1[k] += h
            p0 = params.copy(); p0[k] -= h
            fd = (dn_ref.loss_and_grads(p1, x, t, False)[0] - dn_ref.loss_and_grads(p0, x, t, False)[0]) / (2 * h)
            assert abs(fd - grads[k]) <= 1e-4 * abs(grads[k]) + 1e-8, (li, k, fd, grads[k])
            checked += 1
    assert checked >= 20
