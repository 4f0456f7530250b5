"""The U-Net's CPU oracle pinned by mathematics: central finite differences in float64 against
its analytic gradients (every parameter block, a tiny input), the bfloat16 rounding against
known values, and the convolution against a direct loop."""
import numpy as np

import unet_ref as U  # noqa: E402


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3],
                 np.float32)
    got = U.bf16(x)
    # 1+2^-8 is a tie between 1 and 1+2^-7: to even (1.0); 1+3*2^-8 ties up to 1+2^-6;
    # 3e-3 = 1.536 * 2^-9 and 1.536 * 128 = 196.6, so 197/128 * 2^-9.
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 197 / 128 * 2 ** -9],
                    np.float32)
    assert np.array_equal(got, want), got


def test_conv3x3_equals_direct_loop():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    y = U.conv3x3(x, w, b)
    ref = np.zeros_like(y)
    for n in range(2):
        for i in range(5):
            for j in range(6):
                for co in range(4):
                    s = b[co]
                    for ky in range(3):
                        for kx in range(3):
                            ii, jj = i + ky - 1, j + kx - 1
                            if 0 <= ii < 5 and 0 <= jj < 6:
                                s += np.dot(x[n, ii, jj], w[co, :, ky, kx])
                    ref[n, i, j, co] = s
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_parameter_count():
    assert U.init_params(0).size == U.n_params()


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(7)
    theta = U.init_params(3).astype(np.float64)
    theta += rng.standard_normal(theta.size) * 0.01  # nonzero biases too
    x = rng.standard_normal((2, 8, 8, U.IN_PAD))
    x[..., U.IN_CH:] = 0
    t = rng.standard_normal((2, 8, 8))
    mask = (rng.random((2, 8, 8)) > 0.2).astype(np.float64)
    loss, g = U.loss_and_grad(theta, x, t, mask)
    # sample parameters from every block (each conv's W and b, the head)
    idx, o = [], 0
    for k, (ci, co) in enumerate(U.conv_specs()):
        ci_real = U.IN_CH if k == 0 else ci
        nw = co * ci_real * 9
        idx += list(rng.choice(nw, 3, replace=False) + o) + [o + nw + rng.integers(co)]
        o += nw + co
    idx += [o, o + U.LEVELS[0]]
    eps = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (U.loss_and_grad(tp, x, t, mask)[0] - U.loss_and_grad(tm, x, t, mask)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bf16_mode_stays_close_to_fp64():
    rng = np.random.default_rng(2)
    theta = U.init_params(5)
    x = rng.standard_normal((1, 16, 16, U.IN_PAD)).astype(np.float32)
    x[..., U.IN_CH:] = 0
    t = rng.standard_normal((1, 16, 16)).astype(np.float32)
    l64, g64 = U.loss_and_grad(theta.astype(np.float64), x.astype(np.float64),
                               t.astype(np.float64))
    l16, g16 = U.loss_and_grad(theta, x, t, bf=True)
    assert abs(l16 - l64) < 2e-2 * l64
    cos = float(np.dot(g16, g64) / np.linalg.norm(g16) / np.linalg.norm(g64))
    assert cos > 0.99, cos


def test_product_initialisation_matches_layout():
    import warmstart as ws
    assert [(ci, co) for ci, co, _ in ws.CONVS] == U.conv_specs()
    assert np.array_equal(ws.init_params(0), U.init_params(0))
