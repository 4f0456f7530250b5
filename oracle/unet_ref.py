"""CPU oracle of the warm-start U-Net (TEST INFRASTRUCTURE ONLY).

Plain numpy forward and backward passes of the network that the HIP kernels of
simple-hpc_amd/csrc/sg_unet.hip implement, used as the checker of those kernels: its own
gradients are pinned by central finite differences in float64 (tests/test_unet_oracle.py), and the
GPU's gradients and outputs are compared with it on the same inputs (tests/test_gpu_unet.py).
With bf16=True it rounds to bfloat16 (round to nearest even) at exactly the points where the GPU
stores bfloat16: weights, every activation written by a convolution, pooling or concatenation,
and every activation gradient; accumulations are float32 (the MFMA accumulators), so GPU and CPU
then differ only by summation order.

The reference (arfazeli/simple-hpc) has no learned component; nothing here restates it.

Network: 4 levels with 16, 32, 64, 128 channels; two 3x3 convolutions (zero padding, bias, ReLU)
per level; 2x2 average pooling down; nearest-neighbour 2x upsampling and channel concatenation
with the encoder features on the way up; a 1x1 head to one output channel (no activation).
Tensors are NHWC. The input has 3 channels (mass imbalance, a_P, Re), zero-padded to 8.
"""
import numpy as np

LEVELS = (16, 32, 64, 128)
IN_CH = 3
IN_PAD = 8


def conv_specs():
    """(C_in, C_out) of the 14 3x3 convolutions in execution order; C_in as stored (padded)."""
    c0, c1, c2, c3 = LEVELS
    return [(IN_PAD, c0), (c0, c0),          # enc0
            (c0, c1), (c1, c1),              # enc1
            (c1, c2), (c2, c2),              # enc2
            (c2, c3), (c3, c3),              # bottleneck
            (c3 + c2, c2), (c2, c2),         # dec2
            (c2 + c1, c1), (c1, c1),         # dec1
            (c1 + c0, c0), (c0, c0)]         # dec0


def n_params():
    n = 0
    for k, (ci, co) in enumerate(conv_specs()):
        ci_real = IN_CH if k == 0 else ci
        n += co * ci_real * 9 + co
    return n + LEVELS[0] + 1


def unflatten(theta):
    """Flat float32 parameter vector -> list of (W[co, ci, 3, 3], b[co]) + (head w[c0], head b).
    Layout: per conv W (co, ci_real, ky, kx) row-major then b; then head w, head b."""
    out, o = [], 0
    for k, (ci, co) in enumerate(conv_specs()):
        ci_real = IN_CH if k == 0 else ci
        w = theta[o:o + co * ci_real * 9].reshape(co, ci_real, 3, 3)
        o += co * ci_real * 9
        b = theta[o:o + co]
        o += co
        if k == 0:  # padded input channels carry zero weights
            wp = np.zeros((co, ci, 3, 3), dtype=theta.dtype)
            wp[:, :ci_real] = w
            w = wp
        out.append((w, b))
    hw = theta[o:o + LEVELS[0]]
    hb = theta[o + LEVELS[0]]
    return out, (hw, hb)


def init_params(seed=0):
    """He-normal weights, zero biases (float32), from numpy's PCG64(seed)."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, (ci, co) in enumerate(conv_specs()):
        ci_real = IN_CH if k == 0 else ci
        std = np.sqrt(2.0 / (ci_real * 9))
        parts.append((rng.standard_normal(co * ci_real * 9) * std).astype(np.float32))
        parts.append(np.zeros(co, np.float32))
    parts.append((rng.standard_normal(LEVELS[0]) * np.sqrt(1.0 / LEVELS[0])).astype(np.float32))
    parts.append(np.zeros(1, np.float32))
    return np.concatenate(parts)


def bf16(a):
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def _q(a, on):
    return bf16(a) if on else a


def conv3x3(x, w, b):
    """x [N,H,W,Ci], w [Co,Ci,3,3], b [Co] -> [N,H,W,Co]; zero padding; accumulation in x's
    dtype (float32 or float64)."""
    n, h, wd, ci = x.shape
    xp = np.zeros((n, h + 2, wd + 2, ci), dtype=x.dtype)
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((n, h, wd, w.shape[0]), dtype=x.dtype)
    for ky in range(3):
        for kx in range(3):
            y += xp[:, ky:ky + h, kx:kx + wd] @ w[:, :, ky, kx].T.astype(x.dtype)
    return y + b.astype(x.dtype)


def conv3x3_dx(dy, w, h, wd):
    """Gradient of conv3x3 w.r.t. its input."""
    n = dy.shape[0]
    ci = w.shape[1]
    dxp = np.zeros((n, h + 2, wd + 2, ci), dtype=dy.dtype)
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + h, kx:kx + wd] += dy @ w[:, :, ky, kx].astype(dy.dtype)
    return dxp[:, 1:-1, 1:-1]


def conv3x3_dw(x, dy):
    """Gradients of conv3x3 w.r.t. its weights [Co,Ci,3,3] and bias [Co]."""
    n, h, wd, ci = x.shape
    co = dy.shape[-1]
    xp = np.zeros((n, h + 2, wd + 2, ci), dtype=x.dtype)
    xp[:, 1:-1, 1:-1] = x
    dw = np.zeros((co, ci, 3, 3), dtype=x.dtype)
    dyf = dy.reshape(-1, co)
    for ky in range(3):
        for kx in range(3):
            dw[:, :, ky, kx] = dyf.T @ xp[:, ky:ky + h, kx:kx + wd].reshape(-1, ci)
    return dw, dyf.sum(axis=0)


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_back(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) / 4


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_back(d):
    n, h, w, c = d.shape
    return d.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def forward(theta, x, bf=False, keep=False):
    """x [N,H,W,8] (float32 or float64) -> y [N,H,W] (H, W multiples of 8)."""
    convs, (hw, hb) = unflatten(theta)
    dt = x.dtype
    if bf:
        convs = [(bf16(w), b) for w, b in convs]
    cache = []
    a = _q(x, bf)

    def cv(k, a_in):
        w, b = convs[k]
        z = conv3x3(a_in, w.astype(dt), b.astype(dt))
        out = _q(np.maximum(z, 0), bf)
        cache.append((k, a_in, out))
        return out

    e0 = cv(1, cv(0, a))
    e1 = cv(3, cv(2, _q(pool(e0), bf)))
    e2 = cv(5, cv(4, _q(pool(e1), bf)))
    bt = cv(7, cv(6, _q(pool(e2), bf)))
    d2 = cv(9, cv(8, np.concatenate([up(bt), e2], axis=-1)))
    d1 = cv(11, cv(10, np.concatenate([up(d2), e1], axis=-1)))
    d0 = cv(13, cv(12, np.concatenate([up(d1), e0], axis=-1)))
    y = d0 @ hw.astype(dt) + dt.type(hb)
    if keep:
        return y, (cache, d0, convs, (hw, hb))
    return y


def loss_and_grad(theta, x, t, mask=None, bf=False):
    """Mean squared error over the masked pixels and its gradient w.r.t. theta (flat, same
    layout). mask [N,H,W] of 0/1 (default: all)."""
    y, (cache, d0, convs, (hw, hb)) = forward(theta, x, bf=bf, keep=True)
    dt = x.dtype
    m = np.ones_like(y) if mask is None else mask.astype(dt)
    cnt = m.sum()
    r = (y - t) * m
    loss = float((r * r).sum() / cnt)
    dy = (2.0 / cnt) * r                       # d loss / d y (float, not rounded)
    g_hw = (d0 * dy[..., None]).sum(axis=(0, 1, 2))
    g_hb = dy.sum()
    dd0 = _q(dy[..., None] * hw.astype(dt)[None, None, None, :], bf)
    grads = {}
    by_k = {k: (a_in, out) for k, a_in, out in cache}

    def cv_back(k, dout):
        a_in, out = by_k[k]
        dz = dout * (out > 0)
        w, _ = convs[k]
        dw, db = conv3x3_dw(a_in, dz)
        grads[k] = (dw, db)
        h, wd = a_in.shape[1], a_in.shape[2]
        return _q(conv3x3_dx(dz, w.astype(dt), h, wd), bf)

    def split(d, c_up):
        return d[..., :c_up], d[..., c_up:]

    c0, c1, c2, c3 = LEVELS
    d = cv_back(12, cv_back(13, dd0))
    d_up1, d_e0 = split(d, c1)
    d = cv_back(10, cv_back(11, _q(up_back(d_up1), bf)))
    d_up2, d_e1 = split(d, c2)
    d = cv_back(8, cv_back(9, _q(up_back(d_up2), bf)))
    d_bt, d_e2 = split(d, c3)
    d = cv_back(6, cv_back(7, _q(up_back(d_bt), bf)))
    d = _q(pool_back(d) + d_e2, bf)
    d = cv_back(4, cv_back(5, d))
    d = _q(pool_back(d) + d_e1, bf)
    d = cv_back(2, cv_back(3, d))
    d = _q(pool_back(d) + d_e0, bf)
    cv_back(0, cv_back(1, d))
    parts = []
    for k, (ci, co) in enumerate(conv_specs()):
        dw, db = grads[k]
        if k == 0:
            dw = dw[:, :IN_CH]
        parts += [dw.reshape(-1), db.reshape(-1)]
    parts += [g_hw.reshape(-1), np.array([g_hb], dtype=dt)]
    return loss, np.concatenate(parts).astype(dt)


def features(b, ap, re):
    """Network input of one p' system on the coarse grid: [H, W, 8] float32 with channels
    b / rms(b), a_P / mean(a_P) - 1, log10(Re / 1000), zero padding. Returns (x, scale) where the
    prediction times scale is p' (scale = rms(b) / mean(a_P))."""
    b = np.asarray(b, np.float64)
    ap = np.asarray(ap, np.float64)
    sb = np.sqrt(np.mean(b * b)) or 1.0
    ma = np.mean(ap) or 1.0
    x = np.zeros(b.shape + (IN_PAD,), np.float32)
    x[..., 0] = b / sb
    x[..., 1] = ap / ma - 1.0
    x[..., 2] = np.log10(re / 1000.0)
    return x, sb / ma
