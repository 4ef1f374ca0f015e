"""GPU: the U-Net kernels (BF16 MFMA convolutions, pooling, upsampling) and the whole
network's loss and gradients against the CPU oracle (oracle/unet_ref.py).

Operands are BF16-representable, so each convolution differs from the float64 oracle only by
FP32 accumulation (bar 1e-5 normwise). The network's gradients are compared with the oracle
run with the same BF16 rounding points; the remaining differences come from FP32 sums
flipping a BF16 rounding now and then (bar 3e-2 normwise per parameter tensor, loss 1e-3).
"""
import sys

import numpy as np
import pytest

from cfd_amd import lib as L
from oracle import unet_ref as U
from cfd_amd.hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    """float array (BF16-representable) -> uint16 bit patterns."""
    f = np.asarray(U.bf16(a), dtype=np.float32)
    return (f.view(np.uint32) >> 16).astype(np.uint16)


def unbits(u):
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


# channel counts off the 32 grid (the element-wise loading kernels) and on it, partial tiles at
# the image edges, images smaller than one tile, and (2, 48, 80, 256, 256): more pixel tiles
# (30) than weight-gradient blocks per channel block (24), so blocks loop over tiles
CONV_SHAPES = [(2, 8, 12, 5, 7), (1, 16, 16, 40, 70), (3, 6, 10, 96, 32), (1, 8, 8, 256, 130),
               (2, 20, 36, 64, 128), (1, 33, 17, 32, 64), (2, 8, 8, 128, 256), (1, 16, 48, 192, 64),
               (2, 48, 80, 256, 256), (1, 40, 40, 32, 4)]


@pytest.mark.parametrize("B,H,W,cin,cout", CONV_SHAPES)
def test_conv3x3_forward_and_gradients_match_oracle(B, H, W, cin, cout):
    say("case: conv3x3", B, H, W, cin, cout)
    rng = np.random.default_rng(B * 1000 + cin)
    x = U.bf16(rng.standard_normal((B, H, W, cin)))
    w = U.bf16(rng.standard_normal((cout, 3, 3, cin)) / np.sqrt(9 * cin))
    b = rng.standard_normal(cout).astype(np.float32).astype(np.float64)
    g = U.bf16(rng.standard_normal((B, H, W, cout)))
    lib = L.load()
    dx = DeviceArray(bits(x)); dw = DeviceArray(bits(w).reshape(cout, -1)); db = DeviceArray(b.astype(np.float32))
    dy = DeviceArray(np.zeros((B, H, W, cout), dtype=np.uint16)); dy32 = DeviceArray(np.zeros((B, H, W, cout), np.float32))
    for relu in (0, 1):
        assert lib.cfd_nn_conv3x3(dx.ptr, B, H, W, cin, dw.ptr, db.ptr, cout, relu, dy.ptr, dy32.ptr) == 0
        ref = U.conv(x, w, b)
        if relu:
            ref = np.maximum(ref, 0.0)
        got = dy32.get().astype(np.float64)
        assert rel(got, ref) <= 1e-5, rel(got, ref)
        assert np.array_equal(dy.get(), bits(got))                 # BF16 store = rounding of the FP32 value
    # input gradient through the flipped weights
    dw32 = DeviceArray(w.astype(np.float32).reshape(cout, -1))
    dwf = DeviceArray(np.zeros((cin, 9 * cout), dtype=np.uint16))
    assert lib.cfd_nn_flip_weights(dw32.ptr, cin, cout, dwf.ptr) == 0
    dg = DeviceArray(bits(g)); dgx32 = DeviceArray(np.zeros((B, H, W, cin), np.float32))
    assert lib.cfd_nn_conv3x3(dg.ptr, B, H, W, cout, dwf.ptr, None, cin, 0, None, dgx32.ptr) == 0
    ref = U.conv_input_grad(g, w)
    assert rel(dgx32.get(), ref) <= 1e-5
    # weight gradient, alone and with the bias gradient (both accumulate into what is there)
    dgw = DeviceArray(np.zeros((cout, 9 * cin), np.float32))
    assert lib.cfd_nn_conv3x3_wgrad(dx.ptr, dg.ptr, B, H, W, cin, cout, dgw.ptr, None) == 0
    ref = U.conv_weight_grad(x, g).reshape(cout, -1)
    assert rel(dgw.get(), ref) <= 1e-5
    dgb = DeviceArray(np.full(cout, 0.5, np.float32))
    assert lib.cfd_nn_conv3x3_wgrad(dx.ptr, dg.ptr, B, H, W, cin, cout, dgw.ptr, dgb.ptr) == 0
    assert rel(dgw.get(), 2 * ref) <= 1e-5
    refb = g.sum(axis=(0, 1, 2)) + 0.5
    assert rel(dgb.get(), refb) <= 1e-5, rel(dgb.get(), refb)
    say("ok")


def test_pool_and_upsample_match_oracle():
    rng = np.random.default_rng(7)
    B, H, W, C = 2, 8, 6, 5
    x = U.bf16(np.round(rng.standard_normal((B, H, W, C)) * 4) / 4)      # ties inside windows
    lib = L.load()
    dx = DeviceArray(bits(x)); dy = DeviceArray(np.zeros((B, H // 2, W // 2, C), np.uint16))
    assert lib.cfd_nn_maxpool2(dx.ptr, B, H, W, C, dy.ptr) == 0
    y = unbits(dy.get())
    assert np.array_equal(y, U.maxpool(x))
    gy = U.bf16(rng.standard_normal(y.shape))
    dgy = DeviceArray(bits(gy)); dgx = DeviceArray(np.zeros((B, H, W, C), np.float32))
    assert lib.cfd_nn_maxpool2_bwd(dx.ptr, dy.ptr, dgy.ptr, B, H, W, C, dgx.ptr) == 0
    assert np.array_equal(dgx.get().astype(np.float64), U.maxpool_bwd(x, y, gy))
    # upsample into channels 3..7 of an 8-channel tensor, and back
    Cy = 8
    dup = DeviceArray(np.zeros((B, 2 * H, 2 * W, Cy), np.uint16))
    assert lib.cfd_nn_upsample2(dx.ptr, B, H, W, C, dup.ptr, Cy, 3) == 0
    up = unbits(dup.get())
    assert np.array_equal(up[..., 3:], U.upsample(x)) and not up[..., :3].any()
    gu = U.bf16(rng.standard_normal((B, 2 * H, 2 * W, Cy)))
    dgu = DeviceArray(bits(gu)); dgx2 = DeviceArray(np.zeros((B, H, W, C), np.float32))
    assert lib.cfd_nn_upsample2_bwd(dgu.ptr, B, H, W, C, Cy, 3, dgx2.ptr) == 0
    assert np.allclose(dgx2.get(), U.upsample_bwd(gu[..., 3:]), rtol=1e-6, atol=1e-6)


def test_unet_loss_and_gradients_match_oracle():
    say("case: U-Net gradients 16x16, batch 2")
    net = L.UNet(16, 16, 2, seed=11)
    th = net.get_params().astype(np.float64)
    rng = np.random.default_rng(12)
    x = rng.standard_normal((2, 16, 16, 5)).astype(np.float32)
    t = rng.standard_normal((2, 16, 16, 4)).astype(np.float32)
    loss, g = net.loss_grad(x, t)
    loss_ref, g_ref, _ = U.forward_backward(th, x.astype(np.float64), t.astype(np.float64), bf=True)
    say("loss", loss, loss_ref)
    assert abs(loss - loss_ref) <= 1e-3 * abs(loss_ref)
    lay, n = U.param_layout()
    worst = 0.0
    for q, (cin, cout, relu, wo, bo) in enumerate(lay):
        ew = rel(g[wo:wo + cout * 9 * cin], g_ref[wo:wo + cout * 9 * cin])
        eb = rel(g[bo:bo + cout], g_ref[bo:bo + cout])
        worst = max(worst, ew, eb)
        say("conv", q, "dW", ew, "db", eb)
        assert ew <= 3e-2 and eb <= 3e-2, q
    net.close()


def test_unet_training_fits_a_batch():
    """Adam on a fixed batch of smooth targets: the loss falls by at least 10x in 60 steps."""
    say("case: U-Net training")
    net = L.UNet(32, 32, 4, seed=3)
    ii, jj = np.meshgrid(np.linspace(0, 1, 32), np.linspace(0, 1, 32), indexing="ij")
    x = np.zeros((4, 32, 32, 5), np.float32)
    t = np.zeros((4, 32, 32, 4), np.float32)
    for b in range(4):
        x[b, ..., 0] = (ii > 0.3 + 0.1 * b) & (jj < 0.6)
        x[b, ..., 1] = 0.5 + 0.1 * b
        x[b, ..., 2] = 0.2 * b
        t[b, ..., 0] = np.sin(np.pi * jj) * (1 - 0.5 * x[b, ..., 0])
        t[b, ..., 1] = 0.1 * np.cos(np.pi * ii) * b
        t[b, ..., 2] = 1 - ii
        t[b, ..., 3] = jj * x[b, ..., 1]
    losses = [net.train_step(x, t, 2e-3) for _ in range(60)]
    say("loss", losses[0], losses[-1])
    assert losses[-1] <= 0.1 * losses[0]
    p = net.predict(x)
    assert np.all(np.isfinite(p)) and np.mean((p - t) ** 2) <= 0.2 * losses[0]
    net.close()


def test_adam_steps_match_oracle():
    """Two training steps: the FP32 master weights after each equal the oracle's Adam
    (float64, Kingma & Ba's algorithm with bias correction) applied to the gradient the GPU
    computed at the same weights. Bar: the FP32 rounding of the weights and of the
    constants 0.9f, 0.999f (1e-7 |w| + 1e-5 lr per entry)."""
    say("case: Adam, two steps, 16x16 batch 2")
    net = L.UNet(16, 16, 2, seed=21)
    rng = np.random.default_rng(22)
    x = rng.standard_normal((2, 16, 16, 5)).astype(np.float32)
    t = rng.standard_normal((2, 16, 16, 4)).astype(np.float32)
    lr = 1e-3
    th = net.get_params().astype(np.float64)
    m = np.zeros_like(th); v = np.zeros_like(th)
    for step in (1, 2):
        _, g = net.loss_grad(x, t)
        _, g2 = net.loss_grad(x, t)
        assert np.array_equal(g, g2)             # the gradient is deterministic
        net.train_step(x, t, lr)
        got = net.get_params().astype(np.float64)
        exp, m, v = U.adam(th, g.astype(np.float64), m, v, step, lr)
        err = np.abs(got - exp) / (1e-7 * np.abs(exp) + 1e-5 * lr)
        moved = np.abs(got - th)
        say("step", step, "worst error / bar", float(err.max()), "median move / lr", float(np.median(moved) / lr))
        assert err.max() <= 1.0
        th = got
    net.close()


def test_unet_full_size_loss_and_gradients_match_oracle():
    """The warm-start network at its production size: 256 x 256 inputs (mask, porosity,
    log Re, log Da, Ri), batch 1, loss and every parameter tensor's gradient against the
    oracle with the same BF16 rounding points (bars as at 16 x 16)."""
    say("case: U-Net gradients 256x256, batch 1")
    H = W = 256
    net = L.UNet(H, W, 1, seed=31)
    th = net.get_params().astype(np.float64)
    rng = np.random.default_rng(32)
    ii, jj = np.meshgrid(np.linspace(0, 1, H), np.linspace(0, 1, W), indexing="ij")
    x = np.zeros((1, H, W, 5), np.float32)
    blk = (ii > 0.3) & (ii < 0.62) & (jj > 0.25) & (jj < 0.7)
    x[0, ..., 0] = blk
    x[0, ..., 1] = np.where(blk, rng.uniform(0.3, 0.95, (H, W)), 1.0)
    x[0, ..., 2] = np.log(120.0) / np.log(500.0)
    x[0, ..., 3] = np.log10(3e-3) / 4 + 1
    x[0, ..., 4] = 2.5 / 5
    t = np.stack([np.sin(np.pi * jj) * (1 - 0.5 * blk), 0.1 * np.cos(np.pi * ii), 1 - ii, jj * blk], -1)
    t = t[None].astype(np.float32)
    loss, g = net.loss_grad(x, t)
    loss_ref, g_ref, _ = U.forward_backward(th, x.astype(np.float64), t.astype(np.float64), bf=True)
    say("loss", loss, loss_ref)
    assert abs(loss - loss_ref) <= 1e-3 * abs(loss_ref)
    lay, n = U.param_layout()
    for q, (cin, cout, relu, wo, bo) in enumerate(lay):
        ew = rel(g[wo:wo + cout * 9 * cin], g_ref[wo:wo + cout * 9 * cin])
        eb = rel(g[bo:bo + cout], g_ref[bo:bo + cout])
        say("conv", q, "dW", ew, "db", eb)
        assert ew <= 3e-2 and eb <= 3e-2, q
    net.close()
