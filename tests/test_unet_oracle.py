"""Pins the U-net oracle (oracle/unet_ref.py, the FP64 numpy restatement the
GPU kernels are checked against) independently of it: the same network
written with torch's own layers (conv3d, avg_pool3d, nearest interpolation,
concatenation, ReLU) in float64 on the CPU gives the same output, loss and
parameter gradient (autograd), and central finite differences of the loss
agree with the oracle's gradient.  A dropped term, a wrong tap order, a
transposed weight or a wrong pooling/up-sampling adjoint fails these."""
import numpy as np
import pytest

from oracle import unet_ref


def _torch_net(params, x, levels, channels, cin, cout, linear):
    import torch
    import torch.nn.functional as F
    convs, n = unet_ref.layout(levels, channels, cin, cout)
    P = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    act = (lambda v: v) if linear else F.relu

    def conv(i, h):
        ci, co, relu, w, b = convs[i]
        W = P[w:w + co * ci * 27].reshape(co, ci, 3, 3, 3)   # tap k = kz*9 + ky*3 + kx
        return F.conv3d(h, W, P[b:b + co], padding=1)

    h = torch.tensor(x, dtype=torch.float64)
    ic = 0
    enc = []
    for l in range(levels):
        inp = h if l == 0 else F.avg_pool3d(enc[-1], 2)
        a = act(conv(ic, inp)); ic += 1
        bb = act(conv(ic, a)); ic += 1
        enc.append(bb)
    cur = enc[-1]
    for l in range(levels - 2, -1, -1):
        u = F.interpolate(cur, scale_factor=2, mode="nearest")
        a = act(conv(ic, torch.cat([u, enc[l]], 1))); ic += 1
        cur = act(conv(ic, a)); ic += 1
    y = conv(ic, cur)
    return P, y


@pytest.mark.parametrize("levels,channels,dims,linear", [
    (3, [3, 4, 5], (8, 8, 4), False), (2, [2, 3], (4, 6, 8), False), (1, [3], (4, 4, 4), False),
    (3, [2, 3, 4], (8, 4, 8), True)])
def test_unet_oracle_matches_torch_autograd(levels, channels, dims, linear):
    import torch
    rng = np.random.default_rng(11)
    cin, cout, N = 2, 2, 2
    _, n = unet_ref.layout(levels, channels, cin, cout)
    p = 0.3 * rng.standard_normal(n)
    x = rng.standard_normal((N, cin) + dims)
    t = rng.standard_normal((N, cout) + dims)
    wts = np.array([0.5, 1.5])
    y, loss, g = unet_ref.forward_backward(p, x, t, levels, channels, cin, cout, weights=wts,
                                           linear=linear)
    P, yt = _torch_net(p, x, levels, channels, cin, cout, linear)
    tt = torch.tensor(t)
    lt = 0.5 * (torch.tensor(wts)[:, None, None, None, None] * (yt - tt) ** 2).sum() / N
    lt.backward()
    assert np.max(np.abs(y - yt.detach().numpy())) <= 1e-12 * np.max(np.abs(y))
    assert abs(loss - float(lt.detach())) <= 1e-12 * abs(loss)
    gt = P.grad.numpy()
    assert np.max(np.abs(g - gt)) <= 1e-11 * np.max(np.abs(gt))


def test_unet_oracle_gradient_by_finite_differences():
    rng = np.random.default_rng(5)
    levels, channels, cin, cout, dims = 2, [2, 3], 1, 1, (4, 4, 4)
    _, n = unet_ref.layout(levels, channels, cin, cout)
    p = 0.4 * rng.standard_normal(n)
    x = rng.standard_normal((1, cin) + dims)
    t = rng.standard_normal((1, cout) + dims)
    _, _, g = unet_ref.forward_backward(p, x, t, levels, channels, cin, cout)
    h = 1e-6
    for i in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[i] = h
        lp = unet_ref.forward_backward(p + e, x, t, levels, channels, cin, cout)[1]
        lm = unet_ref.forward_backward(p - e, x, t, levels, channels, cin, cout)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
