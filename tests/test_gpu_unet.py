"""The warm-start U-Net's bfloat16 MFMA kernels against the CPU oracle (oracle/unet_ref.py, whose
own gradients are pinned by finite differences), on the same inputs and weights.

The oracle in bf16 mode rounds where the GPU stores bfloat16, so the two differ only by float32
summation order (and the rare bfloat16 rounding flip that order causes): outputs are compared to
within about 10x the measured error per image size, gradients to 1e-6 relative L2 per parameter
block (measured: at most 8.9e-8). Also: training
lowers the loss; an all-zero network's guess is discarded and the run is bit-identical to the
plain mixed-precision run (the fallback is exact).
"""
import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu

import simple_gpu as sg  # noqa: E402
import unet_ref as U  # noqa: E402


def _inputs(seed, n, h, w):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, h, w, U.IN_PAD)).astype(np.float32)
    x[..., U.IN_CH:] = 0
    t = rng.standard_normal((n, h, w)).astype(np.float32)
    return x, t


def _theta(seed):
    rng = np.random.default_rng(seed + 100)
    th = U.init_params(seed)
    return (th + rng.standard_normal(th.size).astype(np.float32) * 0.01).astype(np.float32)


def test_param_count_matches_oracle():
    net = sg.UNet()
    assert net.nparams == U.n_params()
    th = _theta(0)
    net.set_params(th)
    assert np.array_equal(net.get_params(), th)
    net.close()


# Bounds about 10x the measured errors (7.7e-8, 4.7e-4, 1.9e-3, 3.9e-3): beyond the smallest
# image a bfloat16 rounding flip in an intermediate activation (summation order) moves outputs by
# a bfloat16 ulp of that activation.
@pytest.mark.parametrize("n,h,w,bound", [(1, 8, 8, 1e-6), (2, 16, 24, 5e-3), (1, 32, 32, 2e-2),
                                         (1, 256, 256, 4e-2)])
def test_forward_matches_oracle(n, h, w, bound):
    x, _ = _inputs(1, n, h, w)
    th = _theta(1)
    net = sg.UNet(th)
    y = net.forward(x)
    ref = U.forward(th, x, bf=True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print(f"forward {n}x{h}x{w}: max rel err {err:.2e}", flush=True)
    assert err < bound
    assert np.corrcoef(y.ravel(), ref.ravel())[0, 1] > 0.9995
    net.close()


@pytest.mark.parametrize("n,h,w", [(2, 8, 8), (1, 16, 16)])
def test_gradients_match_oracle(n, h, w):
    x, t = _inputs(2, n, h, w)
    th = _theta(2)
    net = sg.UNet(th)
    loss, g = net.loss_grad(x, t)
    lref, gref = U.loss_and_grad(th, x, t, bf=True)
    print(f"loss gpu {loss:.6f} cpu {lref:.6f}", flush=True)
    assert abs(loss - lref) <= 1e-5 * lref  # equal to 7 digits as measured
    o = 0
    for k, (ci, co) in enumerate(U.conv_specs()):
        ci_real = U.IN_CH if k == 0 else ci
        for name, sz in (("W", co * ci_real * 9), ("b", co)):
            a, b = g[o:o + sz].astype(np.float64), gref[o:o + sz].astype(np.float64)
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = np.dot(a, b) / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
            print(f"conv {k:2d} {name}: rel L2 {rel:.2e} cos {cos:.6f}", flush=True)
            # measured: rel L2 at most 8.9e-8 in any block, cosine 1.000000
            assert rel < 1e-6 and cos > 0.999999, (k, name, rel, cos)
            o += sz
    a, b = g[o:], gref[o:]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    net.close()


def test_training_reduces_loss():
    # A learnable target: a smooth function of the inputs (channel 0 blurred).
    rng = np.random.default_rng(3)
    x = rng.standard_normal((8, 32, 32, U.IN_PAD)).astype(np.float32)
    x[..., U.IN_CH:] = 0
    k = np.ones((5, 5)) / 25.0
    t = np.zeros((8, 32, 32), np.float32)
    for i in range(8):
        xp = np.pad(x[i, :, :, 0], 2)
        for a in range(5):
            for b in range(5):
                t[i] += k[a, b] * xp[a:a + 32, b:b + 32]
    net = sg.UNet(U.init_params(4))
    first = None
    for step in range(200):
        loss = net.train_step(x, t, 2e-3)
        first = loss if first is None else first
    final, _ = net.loss_grad(x, t)
    print(f"loss {first:.4f} -> {final:.4f}", flush=True)
    assert final < 0.2 * first
    net.close()


def _mixed_params(n, re, learned, ny=None):
    text = (f"N_cells_x = {n}\nN_cells_y = {ny or n}\nRe = {re}\nconvection_scheme = QUICK\n"
            f"diffusion_order = 2\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
            f"mg_precision = mixed\nerror_tolerance = 0\nmaxit = 100000\nbackend = gpu\n"
            f"write_fields = False\n")
    if learned:
        text += "pressure_warm_start = learned\n"
    return sg.params_from_config(text)[0]


def test_automatic_warm_start_level():
    """Level -1 resolves to the coarsest multigrid level whose grid is still at least 64 x 64 (the
    coarse-grid size the network was trained on); an explicit level is kept."""
    net = sg.UNet(np.zeros(U.n_params(), np.float32))
    for n, want in ((64, 0), (128, 1), (256, 2), (1024, 4)):
        s = sg.Solver(_mixed_params(n, 400.0, True))
        assert sg.warm_level(s) == -1  # none attached yet
        sg.attach_warm_start(s, net, -1)
        assert sg.warm_level(s) == want, (n, sg.warm_level(s))
        sg.attach_warm_start(s, net, 1)
        assert sg.warm_level(s) == 1
        with pytest.raises(sg.GpuError, match="warm-start level"):
            sg.attach_warm_start(s, net, 9)
        s.close()
    net.close()


def test_rejected_guess_is_exactly_the_plain_solve():
    net = sg.UNet(np.zeros(U.n_params(), np.float32))  # predicts 0: a degenerate guess
    plain = sg.Solver(_mixed_params(64, 400.0, False))
    warm = sg.Solver(_mixed_params(64, 400.0, True))
    sg.attach_warm_start(warm, net, level=2)
    plain.advance(25)
    warm.advance(25)
    for f in ("u", "v", "p"):
        assert bits_equal(warm.get(f), plain.get(f)), f
    acc, rej = sg.warm_stats(warm)
    assert acc == 0 and rej == 25, (acc, rej)
    plain.close()
    warm.close()
    net.close()


def test_rejected_guess_wide_grid_is_exactly_the_plain_solve():
    """Level 0 on a 64 x 2048 grid: the prediction's rows are 2048 wide, more than the guess
    kernels copy to LDS, so they interpolate from global memory instead; the zero network's guess
    is rejected every time and the run equals the plain mixed one bit for bit."""
    net = sg.UNet(np.zeros(U.n_params(), np.float32))
    plain = sg.Solver(_mixed_params(64, 400.0, False, ny=2048))
    warm = sg.Solver(_mixed_params(64, 400.0, True, ny=2048))
    sg.attach_warm_start(warm, net, level=0)
    assert sg.warm_level(warm) == 0
    plain.advance(12)
    warm.advance(12)
    for f in ("u", "v", "p"):
        assert bits_equal(warm.get(f), plain.get(f)), f
    acc, rej = sg.warm_stats(warm)
    assert acc == 0 and rej == 12, (acc, rej)
    plain.close()
    warm.close()
    net.close()


def test_features_shape_and_scale():
    s = sg.Solver(_mixed_params(64, 400.0, False))
    s.advance(5)
    x, scale = sg.warm_features(s, level=2)
    assert x.shape == (16, 16, 8) and scale > 0
    assert abs(np.sqrt(np.mean(x[..., 0] ** 2)) - 1.0) < 1e-3  # b channel normalised to rms 1
    assert abs(np.mean(x[..., 1])) < 1e-3                       # a_P channel: mean 0
    assert np.allclose(x[..., 2], np.log10(0.4)) and np.all(x[..., 3:] == 0)
    s.close()


def test_cli_learned_warm_start(tmp_path):
    import json
    import os
    import subprocess
    import warmstart as ws
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    wpath = tmp_path / "w.bin"
    ws.save(str(wpath), ws.init_params(0))
    assert np.array_equal(ws.load(str(wpath)), ws.init_params(0))
    out = tmp_path / "run"
    cfg = tmp_path / "c.cfg"
    cfg.write_text(f"N_cells_x = 64\nN_cells_y = 64\nRe = 400\nconvection_scheme = QUICK\n"
                   f"diffusion_order = 2\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
                   f"mg_precision = mixed\npressure_warm_start = learned\n"
                   f"warm_start_weights = {wpath}\nerror_tolerance = 0\nmaxit = 60\n"
                   f"backend = gpu\noutput_dir = {out}\n")
    r = subprocess.run([os.path.join(root, "build", "simple"), str(cfg)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads((out / "run_info.json").read_text())
    w = info["warm_start"]
    # one guess per p' solve; iteration 1 runs eagerly too
    # the automatic level: the coarsest grid of at least 64 x 64 cells, here level 0
    assert w["kept"] + w["discarded"] == 60 and w["level"] == 0, w
    cfg.write_text(cfg.read_text().replace(f"warm_start_weights = {wpath}\n", ""))
    r = subprocess.run([os.path.join(root, "build", "simple"), str(cfg)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 1 and "needs warm_start_weights" in r.stderr
