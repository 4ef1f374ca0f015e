"""CPU: the U-Net oracle (oracle/unet_ref.py) pinned by central finite differences of its
own loss (float64) and by closed forms of its pieces. No reference counterpart exists (the
reference solver has no learned part), so these checks are what pins it."""
import numpy as np

from oracle import unet_ref as U


def _params(seed, scale=0.3):
    lay, n = U.param_layout()
    rng = np.random.default_rng(seed)
    th = np.zeros(n)
    for cin, cout, relu, wo, bo in lay:
        th[wo:wo + cout * 9 * cin] = rng.uniform(-1, 1, cout * 9 * cin) * scale / np.sqrt(9 * cin) * 3
        th[bo:bo + cout] = rng.uniform(0.0, 0.2, cout)        # positive biases keep units active
    return th


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    th = _params(2)
    x = rng.standard_normal((1, 8, 8, 5))
    t = rng.standard_normal((1, 8, 8, 4))
    loss, g, _ = U.forward_backward(th, x, t, bf=False)
    lay, n = U.param_layout()
    # one weight and one bias of every convolution
    idx = []
    for q, (cin, cout, relu, wo, bo) in enumerate(lay):
        idx.append(wo + int(rng.integers(cout * 9 * cin)))
        idx.append(bo + int(rng.integers(cout)))
    h = 1e-6
    for k in idx:
        tp = th.copy(); tp[k] += h
        tm = th.copy(); tm[k] -= h
        fd = (U.loss_only(tp, x, t) - U.loss_only(tm, x, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])


def test_conv_identity_kernel_and_input_grad_adjoint():
    """A centre-tap identity kernel reproduces its input; <conv(x), g> = <x, dX(g)> and
    <conv(x), g> = <W, dW(x, g)> (the two gradients are the adjoints of the forward)."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 6, 5, 3))
    W = np.zeros((3, 3, 3, 3))
    for c in range(3):
        W[c, 1, 1, c] = 1.0
    assert np.allclose(U.conv(x, W, None), x)
    W = rng.standard_normal((4, 3, 3, 3))
    g = rng.standard_normal((2, 6, 5, 4))
    y = U.conv(x, W, None)
    assert np.isclose(np.sum(y * g), np.sum(x * U.conv_input_grad(g, W)))
    assert np.isclose(np.sum(y * g), np.sum(W * U.conv_weight_grad(x, g)))


def test_pool_upsample_and_bf16_closed_forms():
    x = np.array([[[[1.0], [3.0]], [[3.0], [2.0]]]])        # one 2x2 window, two maxima
    y = U.maxpool(x)
    assert y[0, 0, 0, 0] == 3.0
    gx = U.maxpool_bwd(x, y, np.full(y.shape, 5.0))
    assert gx[0, 0, 1, 0] == 5.0 and gx[0, 1, 0, 0] == 0.0     # first maximum in row-major order
    g = np.arange(16.0).reshape(1, 4, 4, 1)
    assert np.array_equal(U.upsample_bwd(g)[0, :, :, 0], [[10.0, 18.0], [42.0, 50.0]])
    assert np.array_equal(U.upsample(np.array([[[[7.0]]]]))[0, :, :, 0], [[7.0, 7.0], [7.0, 7.0]])
    assert U.bf16(1.0 + 2 ** -9) == 1.0                        # tie to even
    assert U.bf16(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7
    assert U.bf16(-2.5) == -2.5


def test_library_layer_table_matches_oracle():
    """The package's layer table (used for FLOP counts) is the oracle's network."""
    from cfd_amd import lib as L
    assert [tuple(s) for s in L.UNET_LAYERS] == [tuple(s) for s in U.SPEC]


def test_adam_closed_forms():
    """Adam's bias correction makes the first step lr * g / (|g| + eps), i.e. lr in the
    direction of -g whatever |g| is; under a constant gradient every later step is the same
    size; after g then -g the moments are m = -0.01 g, v = 0.001999 g^2, so the second step
    is lr * (0.01 / 0.19) against the first direction. A missing or off-by-one bias correction fails these."""
    rng = np.random.default_rng(3)
    th = rng.standard_normal(257)
    g = rng.standard_normal(257) * np.logspace(-6, 3, 257)
    lr = 1e-3
    z = np.zeros_like(th)
    th1, m1, v1 = U.adam(th, g, z, z, 1, lr)
    assert np.allclose(th1, th - lr * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    # constant gradient: each step moves lr * g / (|g| + eps), steps 2..5
    t, m, v = th1, m1, v1
    for k in range(2, 6):
        t2, m, v = U.adam(t, g, m, v, k, lr)
        assert np.allclose(t2 - t, -lr * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-17), k
        t = t2
    # reversal: m2 = -0.01 g, v2 = 0.001999 g^2; corrected by 0.19 and 0.001999
    th2, _, _ = U.adam(th1, -g, m1, v1, 2, lr)
    exp = lr * (0.01 / 0.19) * g / (np.abs(g) + 1e-8)
    assert np.allclose(th2 - th1, exp, rtol=1e-9, atol=1e-18)
