"""FP64 numpy restatement of the learned preconditioner's U-net (forward and
backward), used only by the tests to check the GPU kernels' outputs and
parameter gradients.  TEST INFRASTRUCTURE ONLY.

Same architecture and flat parameter layout as ifem_amd/csrc/unet.hip:
  enc l: conv(in_l -> c_l) relu, conv(c_l -> c_l) relu, 2x2x2 average pool
  dec l (l = L-2..0): nearest up x2, concat [up, enc_l],
         conv(c_{l+1}+c_l -> c_l) relu, conv(c_l -> c_l) relu
  out: conv(c_0 -> Co)
Convolutions are 3x3x3, zero padding 1; per conv the weights [co][ci][27]
(tap index k = kz*9 + ky*3 + kx) are followed by the biases [co].
Loss: 0.5*sum_n w_n*sum((y_n - t_n)^2)/N (w_n = 1 without weights).

With quant (a rounding function, e.g. to bf16) the tensors that the GPU
kernels store between layers are rounded as they are there: every hidden
convolution's output, the pooled inputs, the output gradient, every data
gradient, the summed up-sampling adjoint and the skip + pooling-adjoint sum.
The arithmetic itself stays FP64.
"""
import numpy as np


def layout(levels, channels, cin, cout):
    convs = []
    off = 0

    def add(ci, co, relu):
        nonlocal off
        w = off
        off += co * ci * 27
        b = off
        off += co
        convs.append((ci, co, relu, w, b))

    prev = cin
    for l in range(levels):
        add(prev, channels[l], True)
        add(channels[l], channels[l], True)
        prev = channels[l]
    for l in range(levels - 2, -1, -1):
        add(channels[l + 1] + channels[l], channels[l], True)
        add(channels[l], channels[l], True)
    add(channels[0], cout, False)
    return convs, off


def _shift(xp, k, D, H, W):
    kz, ky, kx = k // 9, (k // 3) % 3, k % 3
    return xp[..., kz:kz + D, ky:ky + H, kx:kx + W]


def conv(x, w, b):
    """x (N,ci,D,H,W), w (co,ci,27), b (co,)"""
    N, ci, D, H, W = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1), (1, 1)))
    out = np.zeros((N, w.shape[0], D, H, W)) + b[None, :, None, None, None]
    for k in range(27):
        out += np.einsum("oc,ncdhw->nodhw", w[:, :, k], _shift(xp, k, D, H, W))
    return out


def conv_bwd(x, w, g):
    """returns (gx, gw, gb) for out = conv(x, w, b), g = dL/dout"""
    N, ci, D, H, W = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1), (1, 1)))
    gw = np.zeros_like(w)
    gxp = np.zeros_like(xp)
    for k in range(27):
        gw[:, :, k] = np.einsum("nodhw,ncdhw->oc", g, _shift(xp, k, D, H, W))
        kz, ky, kx = k // 9, (k // 3) % 3, k % 3
        gxp[..., kz:kz + D, ky:ky + H, kx:kx + W] += np.einsum("oc,nodhw->ncdhw", w[:, :, k], g)
    gb = g.sum(axis=(0, 2, 3, 4))
    return gxp[..., 1:-1, 1:-1, 1:-1], gw, gb


def pool(x):
    N, C, D, H, W = x.shape
    return x.reshape(N, C, D // 2, 2, H // 2, 2, W // 2, 2).mean(axis=(3, 5, 7))


def pool_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 2), 2, 3), 2, 4) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 2), 2, 3), 2, 4)


def up_bwd(g):
    N, C, D, H, W = g.shape
    return g.reshape(N, C, D // 2, 2, H // 2, 2, W // 2, 2).sum(axis=(3, 5, 7))


def forward_backward(params, x, t, levels, channels, cin, cout, weights=None, linear=False,
                     quant=None):
    """Returns (y, loss, grad) in FP64; weights: optional per-sample loss weights."""
    q = quant if quant is not None else (lambda v: v)
    convs, n = layout(levels, channels, cin, cout)
    P = np.asarray(params, dtype=np.float64)
    act = (lambda v: v) if linear else (lambda v: np.maximum(v, 0.0))
    mask = (lambda v: 1.0) if linear else (lambda v: v > 0)

    def wb(i):
        ci, co, relu, w, b = convs[i]
        return P[w:w + co * ci * 27].reshape(co, ci, 27), P[b:b + co]

    N = x.shape[0]
    acts = {}
    ic = 0
    h = x.astype(np.float64)
    enc_in, enc_a, enc_b = [], [], []
    for l in range(levels):
        inp = h if l == 0 else q(pool(enc_b[l - 1]))
        enc_in.append(inp)
        w, b = wb(ic); ic += 1
        a = q(act(conv(inp, w, b)))
        enc_a.append(a)
        w, b = wb(ic); ic += 1
        bb = q(act(conv(a, w, b)))
        enc_b.append(bb)
    cur = enc_b[-1]
    dec = {}
    for l in range(levels - 2, -1, -1):
        u = up(cur)
        cat = np.concatenate([u, enc_b[l]], axis=1)
        w, b = wb(ic); ic += 1
        a = q(act(conv(cat, w, b)))
        w, b = wb(ic); ic += 1
        bb = q(act(conv(a, w, b)))
        dec[l] = (cat, a, bb)
        cur = bb
    w, b = wb(ic)
    y = conv(cur, w, b)
    d = y - t
    wn = np.ones(N) if weights is None else np.asarray(weights, dtype=np.float64)
    d = d * wn[:, None, None, None, None]
    loss = 0.5 * np.sum((y - t) * d) / N
    grad = np.zeros(n)

    def put(i, gw, gb):
        ci, co, relu, wo, bo = convs[i]
        grad[wo:wo + co * ci * 27] = gw.ravel()
        grad[bo:bo + co] = gb

    g = q(d / N)
    gx, gw, gb = conv_bwd(cur, wb(ic)[0], g)
    put(ic, gw, gb)
    ic -= 1
    g = q(gx)
    skip = {}
    for l in range(0, levels - 1):
        cat, a, bb = dec[l]
        g = g * mask(bb)
        gx, gw, gb = conv_bwd(a, wb(ic)[0], g)
        put(ic, gw, gb); ic -= 1
        g = q(gx) * mask(a)
        gx, gw, gb = conv_bwd(cat, wb(ic)[0], g)
        put(ic, gw, gb); ic -= 1
        gx = q(gx)
        cu = channels[l + 1]
        skip[l] = gx[:, cu:]
        g = q(up_bwd(gx[:, :cu]))
    for l in range(levels - 1, -1, -1):
        if l < levels - 1:
            g = q(g + skip[l])
        g = g * mask(enc_b[l])
        gx, gw, gb = conv_bwd(enc_a[l], wb(2 * l + 1)[0], g)
        put(2 * l + 1, gw, gb)
        g = q(gx) * mask(enc_a[l])
        gx, gw, gb = conv_bwd(enc_in[l], wb(2 * l)[0], g)
        put(2 * l, gw, gb)
        if l > 0:
            g = pool_bwd(q(gx))
    return y, loss, grad
