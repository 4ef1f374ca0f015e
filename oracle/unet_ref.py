"""CPU oracle of the warm-start U-Net (TEST INFRASTRUCTURE ONLY).

numpy restatement of the network the library trains on the GPU (solver-cfd_amd/csrc/
cfd_unet.hip): NHWC tensors, 3x3 convolutions with zero padding, ReLU, 2x2 max pooling,
nearest x2 upsampling with channel concatenation, mean-squared-error loss, and the exact
backward pass. Parameters are one flat FP32 vector in the library's order (for each of the
15 convolutions: weights [Cout][3][3][Cin], then biases [Cout]).

`bf16=True` rounds to BF16 at the points where the GPU stores BF16 (network input, every
convolution output, the loss gradient, every input gradient, and the FP32 gradient
accumulators where the skip and the upsampled paths meet), so that GPU gradients can be
compared with a tight tolerance; `bf16=False` is plain float64 (used to pin the oracle
against central finite differences). The reference solver has no learned component; this
oracle is pinned by finite differences and by closed forms of its pieces only.
"""
import numpy as np

SPEC = [(5, 32, 1), (32, 32, 1), (32, 64, 1), (64, 64, 1), (64, 128, 1), (128, 128, 1), (128, 256, 1),
        (256, 256, 1), (384, 128, 1), (128, 128, 1), (192, 64, 1), (64, 64, 1), (96, 32, 1), (32, 32, 1),
        (32, 4, 0)]


def bf16(a):
    """Round float64/float32 values to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def param_layout():
    out, off = [], 0
    for cin, cout, relu in SPEC:
        wo = off
        off += cout * 9 * cin
        bo = off
        off += cout
        out.append((cin, cout, relu, wo, bo))
    return out, off


def unpack(theta):
    lay, n = param_layout()
    assert theta.size == n
    Ws, bs = [], []
    for cin, cout, relu, wo, bo in lay:
        Ws.append(theta[wo:wo + cout * 9 * cin].reshape(cout, 3, 3, cin).astype(np.float64))
        bs.append(theta[bo:bo + cout].astype(np.float64))
    return Ws, bs


def im2col(x):
    """x [B,H,W,C] -> [B,H,W,9C], k = (dy*3 + dx)*C + c, zero padding."""
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2, W + 2, C))
    xp[:, 1:-1, 1:-1] = x
    cols = [xp[:, dy:dy + H, dx:dx + W, :] for dy in range(3) for dx in range(3)]
    return np.concatenate(cols, axis=3)


def conv(x, W, b):
    """y[b,h,w,n] = sum_k im2col(x)[...,k] W[n,k] + b[n]."""
    cout = W.shape[0]
    A = im2col(x)
    return A @ W.reshape(cout, -1).T + (b if b is not None else 0.0)


def conv_input_grad(g, W):
    """dX of conv for output gradient g: convolution of g with the flipped, transposed weights."""
    Wf = W[:, ::-1, ::-1, :].transpose(3, 1, 2, 0)        # [cin][dy'][dx'][cout]
    return conv(g, Wf, None)


def conv_weight_grad(x, g):
    """dW[n, dy, dx, c] = sum_m g[m, n] im2col(x)[m, k]."""
    A = im2col(x).reshape(-1, 9 * x.shape[3])
    G = g.reshape(-1, g.shape[3])
    return (G.T @ A).reshape(g.shape[3], 3, 3, x.shape[3])


def maxpool(x):
    B, H, W, C = x.shape
    r = x.reshape(B, H // 2, 2, W // 2, 2, C)
    return r.max(axis=(2, 4))


def maxpool_bwd(x, y, gy):
    """Gradient to the first input of each 2x2 window (row-major) that equals the maximum."""
    B, H, W, C = x.shape
    gx = np.zeros_like(x)
    taken = np.zeros(y.shape, dtype=bool)
    for dy in range(2):
        for dx in range(2):
            xv = x[:, dy::2, dx::2, :]
            hit = (xv == y) & ~taken
            gx[:, dy::2, dx::2, :] = np.where(hit, gy, 0.0)
            taken |= hit
    return gx


def upsample(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def upsample_bwd(g):
    B, H, W, C = g.shape
    return g.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def forward_backward(theta, x, t, bf=True):
    """Loss and flat gradient (float64) of the U-Net for inputs x [B,H,W,5], targets t [B,H,W,4]."""
    R = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs = unpack(theta)
    Wq = [R(w) for w in Ws]
    acts = {}

    def cv(q, a):
        y = conv(a, Wq[q], bs[q])
        if SPEC[q][2]:
            y = np.maximum(y, 0.0)
        return y

    a_in = R(x)
    e1a = R(cv(0, a_in)); e1 = R(cv(1, e1a)); p1 = maxpool(e1)
    e2a = R(cv(2, p1)); e2 = R(cv(3, e2a)); p2 = maxpool(e2)
    e3a = R(cv(4, p2)); e3 = R(cv(5, e3a)); p3 = maxpool(e3)
    ba = R(cv(6, p3)); bo = R(cv(7, ba))
    c3 = np.concatenate([upsample(bo), e3], axis=3)
    d3a = R(cv(8, c3)); d3 = R(cv(9, d3a))
    c2 = np.concatenate([upsample(d3), e2], axis=3)
    d2a = R(cv(10, c2)); d2 = R(cv(11, d2a))
    c1 = np.concatenate([upsample(d2), e1], axis=3)
    d1a = R(cv(12, c1)); d1 = R(cv(13, d1a))
    out = cv(14, d1)
    if bf:
        out = out.astype(np.float32).astype(np.float64)
    n = out.size
    loss = float(np.mean((out - t) ** 2))
    g_out = R(2.0 * (out - t) / n)

    dW = [None] * 15
    db = [None] * 15

    def bwd(q, a_in_, gy, need_dx=True):
        dW[q] = conv_weight_grad(a_in_, gy)
        db[q] = gy.reshape(-1, gy.shape[3]).sum(axis=0)
        return R(conv_input_grad(gy, Wq[q])) if need_dx else None

    def mask(g, y):
        return np.where(y > 0, g, 0.0)

    g = bwd(14, d1, g_out)
    g = bwd(13, d1a, mask(g, d1))
    gc1 = bwd(12, c1, mask(g, d1a))
    g_d2 = upsample_bwd(gc1[..., :64]); g_e1 = gc1[..., 64:].copy()
    g = bwd(11, d2a, mask(R(g_d2), d2))
    gc2 = bwd(10, c2, mask(g, d2a))
    g_d3 = upsample_bwd(gc2[..., :128]); g_e2 = gc2[..., 128:].copy()
    g = bwd(9, d3a, mask(R(g_d3), d3))
    gc3 = bwd(8, c3, mask(g, d3a))
    g_b = upsample_bwd(gc3[..., :256]); g_e3 = gc3[..., 256:].copy()
    g = bwd(7, ba, mask(R(g_b), bo))
    gp3 = bwd(6, p3, mask(g, ba))
    g_e3 = g_e3 + maxpool_bwd(e3, p3, gp3)
    g = bwd(5, e3a, mask(R(g_e3), e3))
    gp2 = bwd(4, p2, mask(g, e3a))
    g_e2 = g_e2 + maxpool_bwd(e2, p2, gp2)
    g = bwd(3, e2a, mask(R(g_e2), e2))
    gp1 = bwd(2, p1, mask(g, e2a))
    g_e1 = g_e1 + maxpool_bwd(e1, p1, gp1)
    g = bwd(1, e1a, mask(R(g_e1), e1))
    bwd(0, a_in, mask(g, e1a), need_dx=False)

    lay, ntot = param_layout()
    grad = np.zeros(ntot)
    for q, (cin, cout, relu, wo, bo_) in enumerate(lay):
        grad[wo:wo + cout * 9 * cin] = dW[q].reshape(-1)
        grad[bo_:bo_ + cout] = db[q]
    return loss, grad, out


def loss_only(theta, x, t):
    return forward_backward(theta, x, t, bf=False)[0]


def adam(theta, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) in float64: moments m, v updated from
    the gradient g, bias-corrected by 1 - b1^step and 1 - b2^step, and the step
    lr * m_hat / (sqrt(v_hat) + eps) taken. Returns (theta', m', v')."""
    g = np.asarray(g, dtype=np.float64)
    m = b1 * np.asarray(m, dtype=np.float64) + (1.0 - b1) * g
    v = b2 * np.asarray(v, dtype=np.float64) + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return np.asarray(theta, dtype=np.float64) - lr * mh / (np.sqrt(vh) + eps), m, v
