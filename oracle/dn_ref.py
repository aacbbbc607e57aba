"""FP64 NumPy restatement of the denoiser network, its loss and gradients.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and the benchmark's quality
report) as the checker of the HIP denoiser. The product never imports it.

Network (the shape of the reference's U-Net, train_denoiser.py:13-54 and
train_denoiser_v2.py:60-110, at 32/64/128/256 channels): two 3x3 convolutions
(+ReLU) per level, 2x2 max pooling, 2x nearest upsampling, skip concatenation
(upsampled channels first, then the encoder's), a 1x1 head whose 3 outputs are
added to the log-radiance input channels, L1 loss in log space. Activations
are NHWC. Parameters: per layer weights [cout][taps][cin] then biases [cout],
in the layer order of LAYERS. The weights the GPU multiplies are the FP32
masters rounded to BF16 (round to nearest even); the oracle uses those values
and exact FP64 arithmetic everywhere else.
"""
from __future__ import annotations

import numpy as np

# (cin, cout, taps): c1a c1b c2a c2b c3a c3b c4a c4b c5a c5b c6a c6b c7a c7b head
LAYERS = [(16, 32, 9), (32, 32, 9), (32, 64, 9), (64, 64, 9), (64, 128, 9), (128, 128, 9), (128, 256, 9),
          (256, 256, 9), (384, 128, 9), (128, 128, 9), (192, 64, 9), (64, 64, 9), (96, 32, 9), (32, 32, 9),
          (32, 3, 1)]


def offsets():
    out, off = [], 0
    for cin, cout, taps in LAYERS:
        w = off
        off += cout * taps * cin
        b = off
        off += cout
        out.append((w, b))
    return out, off


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, round_bf16: bool = True):
    offs, _ = offsets()
    W, Bs = [], []
    for (cin, cout, taps), (wo, bo) in zip(LAYERS, offs):
        w = params[wo:wo + cout * taps * cin]
        W.append((bf16(w) if round_bf16 else w.astype(np.float64)).reshape(cout, taps, cin))
        Bs.append(params[bo:bo + cout].astype(np.float64))
    return W, Bs


def _shifted(x, t, taps):
    """x [B,H,W,C] -> x[b, y+dy, x+dx] with zero padding (tap t of 3x3)."""
    if taps == 1:
        return x
    dy, dx = t // 3 - 1, t % 3 - 1
    B, H, W, C = x.shape
    out = np.zeros_like(x)
    ys, yd = slice(max(0, dy), H + min(0, dy)), slice(max(0, -dy), H - max(0, dy))
    xs, xd = slice(max(0, dx), W + min(0, dx)), slice(max(0, -dx), W - max(0, dx))
    out[:, yd, xd] = x[:, ys, xs]
    return out


def _unshift(g, t, taps):
    """Adjoint of _shifted."""
    if taps == 1:
        return g
    dy, dx = t // 3 - 1, t % 3 - 1
    B, H, W, C = g.shape
    out = np.zeros_like(g)
    ys, yd = slice(max(0, dy), H + min(0, dy)), slice(max(0, -dy), H - max(0, dy))
    xs, xd = slice(max(0, dx), W + min(0, dx)), slice(max(0, -dx), W - max(0, dx))
    out[:, ys, xs] = g[:, yd, xd]
    return out


def conv(x, w, b):
    cout, taps, cin = w.shape
    y = np.zeros(x.shape[:3] + (cout,))
    for t in range(taps):
        y += _shifted(x, t, taps) @ w[:, t, :].T
    return y + b


def conv_bwd(x, w, g):
    cout, taps, cin = w.shape
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    gf = g.reshape(-1, cout)
    for t in range(taps):
        xs = _shifted(x, t, taps)
        dw[:, t, :] = gf.T @ xs.reshape(-1, cin)
        dx += _unshift(g @ w[:, t, :], t, taps)
    return dx, dw, gf.sum(0)


def pool(x):
    B, H, W, C = x.shape
    win = np.stack([x[:, 0::2, 0::2], x[:, 0::2, 1::2], x[:, 1::2, 0::2], x[:, 1::2, 1::2]])
    return win.max(0), win.argmax(0)          # argmax: first maximum in window order


def pool_bwd(arg, g, shape):
    dx = np.zeros(shape)
    for s, (oy, ox) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
        dx[:, oy::2, ox::2] = np.where(arg == s, g, 0.0)
    return dx


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def up_bwd(g):
    return g[:, 0::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 0::2] + g[:, 1::2, 1::2]


def forward(params, x, round_bf16: bool = True, emulate: bool = False):
    """x: [B,H,W,16] float64 (the BF16 input values). Returns (head [B,H,W,3], cache).

    emulate=True also rounds every stored activation to BF16 (through FP32),
    where the HIP kernels store BF16, so that the remaining differences are
    only the FP32 accumulation order."""
    W, Bs = unpack(params, round_bf16)
    c = {}
    R = (lambda v: bf16(v.astype(np.float32))) if emulate else (lambda v: v)
    relu = lambda v: R(np.maximum(v, 0.0))
    c["A0"] = relu(conv(x, W[0], Bs[0]))
    c["E1"] = relu(conv(c["A0"], W[1], Bs[1]))
    c["P1"], c["arg1"] = pool(c["E1"])
    c["A2"] = relu(conv(c["P1"], W[2], Bs[2]))
    c["E2"] = relu(conv(c["A2"], W[3], Bs[3]))
    c["P2"], c["arg2"] = pool(c["E2"])
    c["A4"] = relu(conv(c["P2"], W[4], Bs[4]))
    c["E3"] = relu(conv(c["A4"], W[5], Bs[5]))
    c["P3"], c["arg3"] = pool(c["E3"])
    c["A6"] = relu(conv(c["P3"], W[6], Bs[6]))
    c["E4"] = relu(conv(c["A6"], W[7], Bs[7]))
    c["CAT3"] = np.concatenate([up(c["E4"]), c["E3"]], -1)
    c["A8"] = relu(conv(c["CAT3"], W[8], Bs[8]))
    c["D3"] = relu(conv(c["A8"], W[9], Bs[9]))
    c["CAT2"] = np.concatenate([up(c["D3"]), c["E2"]], -1)
    c["A10"] = relu(conv(c["CAT2"], W[10], Bs[10]))
    c["D2"] = relu(conv(c["A10"], W[11], Bs[11]))
    c["CAT1"] = np.concatenate([up(c["D2"]), c["E1"]], -1)
    c["A12"] = relu(conv(c["CAT1"], W[12], Bs[12]))
    c["D1"] = relu(conv(c["A12"], W[13], Bs[13]))
    head = conv(c["D1"], W[14], Bs[14])
    c["x"] = x
    return head, c


def loss_and_grads(params, x, target, round_bf16: bool = True, emulate: bool = False):
    """L1 between head + x[..., :3] and target (mean over pixels and channels).
    Returns (loss, grads flat in the parameter layout, head).

    emulate=True rounds stored activations and stored gradients to BF16 where
    the HIP kernels do (see forward)."""
    W, Bs = unpack(params, round_bf16)
    head, c = forward(params, x, round_bf16, emulate)
    R = (lambda v: bf16(v.astype(np.float32))) if emulate else (lambda v: v)
    if emulate:
        head = head.astype(np.float32).astype(np.float64)
    d = head + x[..., :3] - target
    n = d.size
    loss = np.abs(d).sum() / n
    g = R(np.sign(d) / n)
    offs, total = offsets()
    grads = np.zeros(total)

    def put(i, dw, db):
        wo, bo = offs[i]
        grads[wo:wo + dw.size] = dw.reshape(-1)
        grads[bo:bo + db.size] = db

    def back(i, inp, out, gy, relu=True):
        if relu:
            gy = gy * (out > 0)
        dx, dw, db = conv_bwd(inp, W[i], gy)
        put(i, dw, db)
        return R(dx)

    def pool_add(gE, arg, gP, shape):
        pb = pool_bwd(arg, gP, shape)
        hit = np.zeros(shape, bool)
        for s_, (oy, ox) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
            hit[:, oy::2, ox::2] = arg == s_
        return np.where(hit, R(gE + pb), gE)

    gD1 = back(14, c["D1"], None, g, relu=False)
    gA12 = back(13, c["A12"], c["D1"], gD1)
    gCAT1 = back(12, c["CAT1"], c["A12"], gA12)
    gD2 = R(up_bwd(gCAT1[..., :64]))
    gE1 = gCAT1[..., 64:]
    gA10 = back(11, c["A10"], c["D2"], gD2)
    gCAT2 = back(10, c["CAT2"], c["A10"], gA10)
    gD3 = R(up_bwd(gCAT2[..., :128]))
    gE2 = gCAT2[..., 128:]
    gA8 = back(9, c["A8"], c["D3"], gD3)
    gCAT3 = back(8, c["CAT3"], c["A8"], gA8)
    gE4 = R(up_bwd(gCAT3[..., :256]))
    gE3 = gCAT3[..., 256:]
    gA6 = back(7, c["A6"], c["E4"], gE4)
    gP3 = back(6, c["P3"], c["A6"], gA6)
    gE3 = pool_add(gE3, c["arg3"], gP3, c["E3"].shape)
    gA4 = back(5, c["A4"], c["E3"], gE3)
    gP2 = back(4, c["P2"], c["A4"], gA4)
    gE2 = pool_add(gE2, c["arg2"], gP2, c["E2"].shape)
    gA2 = back(3, c["A2"], c["E2"], gE2)
    gP1 = back(2, c["P1"], c["A2"], gA2)
    gE1 = pool_add(gE1, c["arg1"], gP1, c["E1"].shape)
    gA0 = back(1, c["A0"], c["E1"], gE1)
    back(0, x, c["A0"], gA0)
    return loss, grads, head


def init_like_gpu_shapes(seed=0, scale=1.0):
    """Random FP32 parameters in the layer layout (He-uniform; zero weights on
    the first layer's input-padding channels), for tests."""
    rng = np.random.default_rng(seed)
    offs, total = offsets()
    p = np.zeros(total, np.float32)
    for i, ((cin, cout, taps), (wo, bo)) in enumerate(zip(LAYERS, offs)):
        fan = taps * (10 if i == 0 else cin)
        w = rng.uniform(-1, 1, (cout, taps, cin)) * np.sqrt(6.0 / fan) * scale
        if i == 0:
            w[:, :, 10:] = 0.0
        p[wo:wo + w.size] = w.reshape(-1)
        p[bo:bo + cout] = rng.uniform(-0.05, 0.05, cout)
    return p
