"""Learned preconditioner: the U-net's GPU forward and parameter gradients
(BF16 operands on matrix cores, FP32 accumulation, bf16 activations) against
the FP64 restatement oracle/unet_ref.py (itself pinned by torch autograd and
finite differences in test_unet_oracle.py), training, and flexible CG with the
learned correction.

Tolerance for the BF16 kernels: the oracle is fed the same bf16-rounded
weights and inputs and rounds every tensor the kernels store between layers
to bf16 as well (its quant option), so the ReLU masks agree and what remains
is the FP32 accumulation order and values that round to neighbouring bf16
numbers, which spread from layer to layer (most in the deepest, widest
network, 12/24/48 channels: 5e-3): relative L2 error <= 1e-2 for the output
and the parameter gradient, <= 2e-3 for the loss.  A wrong tap, channel or
adjoint gives errors of order one."""
import numpy as np
import pytest

from oracle import unet_ref

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def _rel2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def _bf16(a):
    """Round float32 values to bf16 (nearest even), as the kernels store them."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _weights_bf16(p, levels, channels, cin, cout):
    """Parameters with the convolution weights rounded to bf16 (the biases
    stay FP32, as in the kernels)."""
    q = p.astype(np.float32).copy()
    convs, _ = unet_ref.layout(levels, list(channels), cin, cout)
    for ci, co, relu, w, b in convs:
        q[w:w + co * ci * 27] = _bf16(q[w:w + co * ci * 27])
    return q


@pytest.mark.parametrize("levels,channels,dims,seed,linear", [
    (3, (4, 6, 8), (8, 8, 4), 3, False), (2, (3, 5), (8, 8, 4), 2, False),
    (1, (4,), (8, 8, 4), 1, False), (4, (3, 4, 5, 6), (16, 16, 8), 4, False),
    (3, (4, 5, 6), (12, 20, 12), 4, False), (3, (4, 5, 6), (12, 20, 12), 3, True),
    (3, (12, 24, 48), (16, 16, 16), 5, False), (2, (20, 36), (8, 24, 40), 6, False)])
def test_unet_forward_and_gradient_match_oracle(levels, channels, dims, seed, linear):
    import torch
    from ifem_amd import Session
    print(f"[gpu] unet gradient check levels={levels} channels={channels} dims={dims} "
          f"linear={linear}", flush=True)
    (D, H, W), N, cin, cout = dims, 2, 2, 1
    rng = np.random.default_rng(seed)
    x = _bf16(rng.standard_normal((N, cin, D, H, W)).astype(np.float32))
    t = rng.standard_normal((N, cout, D, H, W)).astype(np.float32)
    with Session() as s:
        n = s.unet_create(levels, channels, cin, cout, seed=7, linear=linear)
        assert n == unet_ref.layout(levels, list(channels), cin, cout)[1]
        p = s.unet_params()
        # non-zero biases so that their gradients and ReLU masks are exercised
        p = p + 0.05 * rng.standard_normal(p.shape).astype(np.float32)
        s.unet_set_params(p)
        xd = torch.from_numpy(x).cuda()
        td = torch.from_numpy(t).cuda()
        yd = torch.empty((N, cout, D, H, W), dtype=torch.float32, device="cuda")
        s.unet_forward(xd.data_ptr(), yd.data_ptr(), N, D, H, W)
        torch.cuda.synchronize()
        wts = np.array([0.7, 1.6], dtype=np.float32)
        wd = torch.from_numpy(wts).cuda()
        loss, g = s.unet_gradient(xd.data_ptr(), td.data_ptr(), N, D, H, W, w_ptr=wd.data_ptr())
        y = yd.cpu().numpy()
    pq = _weights_bf16(p, levels, channels, cin, cout).astype(np.float64)
    yr, lr_, gr = unet_ref.forward_backward(pq, x, t, levels, list(channels), cin, cout,
                                            weights=wts, linear=linear,
                                            quant=lambda v: _bf16(v).astype(np.float64))
    ey, eg = _rel2(y, yr), _rel2(g, gr)
    print(f"  rel. L2 error: output {ey:.2e}, loss {abs(loss - lr_) / abs(lr_):.2e}, "
          f"gradient {eg:.2e}", flush=True)
    assert ey <= 1e-2
    assert abs(loss - lr_) <= 2e-3 * abs(lr_)
    assert eg <= 1e-2


def test_unet_backward_with_output_gradient():
    """The gradient of sum(dy * U(x)) equals the regression gradient for the
    target t = U(x) - N*dy (unit weights), whose dL/dy is dy."""
    import torch
    from ifem_amd import Session
    print("[gpu] unet backward with output gradient", flush=True)
    N, D, H, W, cin, cout = 2, 8, 8, 8, 3, 3
    rng = np.random.default_rng(11)
    x = torch.from_numpy(rng.standard_normal((N, cin, D, H, W)).astype(np.float32)).cuda()
    dy = torch.from_numpy(rng.standard_normal((N, cout, D, H, W)).astype(np.float32)).cuda()
    with Session() as s:
        s.unet_create(3, (4, 6, 8), cin, cout, seed=5, linear=True)
        y = torch.empty((N, cout, D, H, W), dtype=torch.float32, device="cuda")
        s.unet_forward(x.data_ptr(), y.data_ptr(), N, D, H, W)
        torch.cuda.synchronize()
        t = (y - N * dy).contiguous()
        _, g_ref = s.unet_gradient(x.data_ptr(), t.data_ptr(), N, D, H, W)
        g = s.unet_backward(x.data_ptr(), dy.data_ptr(), N, D, H, W, want_grad=True)
    # both paths round the same output gradient dy (up to FP32 rounding of
    # (y - t)/N) to bf16
    assert _rel2(g, g_ref) <= 1e-3


def test_unet_training_reduces_loss():
    import torch
    from ifem_amd import Session
    print("[gpu] unet training", flush=True)
    N, D, H, W = 16, 8, 8, 8
    rng = np.random.default_rng(3)
    x = rng.standard_normal((N, 1, D, H, W)).astype(np.float32)
    # target: a smoothed copy of the input (learnable by a conv net)
    t = x.copy()
    for ax in (2, 3, 4):
        t = 0.5 * t + 0.25 * (np.roll(t, 1, ax) + np.roll(t, -1, ax))
    with Session() as s:
        s.unet_create(2, (6, 8), 1, 1, seed=3)
        xd = torch.from_numpy(x).cuda()
        td = torch.from_numpy(t.astype(np.float32)).cuda()
        l0 = s.unet_train_step(xd.data_ptr(), td.data_ptr(), N, D, H, W, 0.0, want_loss=True)
        for _ in range(150):
            l = s.unet_train_step(xd.data_ptr(), td.data_ptr(), N, D, H, W, 3e-3, want_loss=True)
    assert l < 0.2 * l0, (l0, l)


def test_harmful_correction_is_dropped_by_the_safeguard_without_fallback():
    """A learned correction that points against the residual (an identity
    network with a large negative scale: zn = -1e6 |r|_rms r^) is not usable:
    the safeguard drops it at the start of every flexible-CG pass and in every
    iteration, and the solve runs on its base (Jacobi or multigrid) to the
    tolerance without falling back to Jacobi-PCG."""
    from ifem_amd import PC_JACOBI, PC_MG, learned, problems, setup_problem
    print("[gpu] learned preconditioner safeguard", flush=True)
    pb = problems.poisson_cube(8, 2)
    with setup_problem(pb) as s:
        s.unet_create(2, (2, 2), 1, 1, linear=True)
        s.unet_set_params(_identity_unet_params(2, 2, 1))
        learned.enable(s, pb, scale=-1e6)
        for pc in (PC_JACOBI, PC_MG):
            st = s.solve(rtol=1e-8, maxits=5000, check_every=2, pc=pc)
            assert st["converged"] and st["rel_residual"] <= 1e-8
            assert not st["fallback"] and st["pc_iterations"] == st["iterations"]
            assert st["pc_rejected"] == st["iterations"] + st["outer"], st


def _identity_unet_params(levels, c, ncomp):
    """Flat parameters of a linear U-net (channels c at every level) whose
    output equals its input: identity centre taps along the encoder level 0,
    the skip half of the decoder's first convolution and the output
    convolution; every other weight and all biases zero."""
    convs, n = unet_ref.layout(levels, [c] * levels, ncomp, ncomp)
    p = np.zeros(n, dtype=np.float32)

    def eye(i, ci_off=0):
        ci, co, relu, w, b = convs[i]
        W = p[w:w + co * ci * 27].reshape(co, ci, 27)
        for k in range(min(co, ncomp)):
            W[k, ci_off + k, 13] = 1.0

    eye(0)
    eye(1)
    if levels > 1:
        nenc = 2 * levels
        eye(nenc + 2 * (levels - 2), ci_off=c)  # decoder level 0, first conv: skip part
        eye(nenc + 2 * (levels - 2) + 1)
    eye(len(convs) - 1)
    return p


def test_learned_preconditioner_on_padded_lattice_matches_numpy_pcg():
    """Elasticity 4^3 (free box 6x6x5 nodes, clamped face) on a 6x6x6 lattice
    (levels 2): with an identity network the learned preconditioner is
    z = D^-1 r + s r, an SPD operator, so 10 flexible-CG iterations equal 10
    PCG iterations with M^-1 = D^-1 + s I computed in numpy.  A wrong box <->
    lattice mapping would make M^-1 a different, non-symmetric operator."""
    import scipy.sparse as sp
    from ifem_amd import learned, problems, setup_problem
    print("[gpu] learned preconditioner on a padded lattice", flush=True)
    pb = problems.elasticity_cube(4, 2, seed=3)
    levels, c = 2, 3
    dims = learned.lattice_dims(pb, levels)
    assert learned.lattice(pb)[:3] == (6, 6, 5) and dims == (6, 6, 6)
    with setup_problem(pb) as s:
        rp, ci, A, b = s.system()
        dinv = learned.jacobi_inverse(s)
        scale = float(np.mean(dinv))
        s.unet_create(levels, (c, c), pb.ncomp, pb.ncomp, linear=True)
        s.unet_set_params(_identity_unet_params(levels, c, pb.ncomp))
        learned.enable(s, pb, scale=scale, dims=dims)
        st = s.solve(rtol=1e-30, maxits=10, check_every=2)
        x = s.solution()
    assert st["iterations"] == 10
    K = sp.csr_matrix((A, ci, rp), shape=(pb.neq, pb.neq))
    xr = np.zeros(pb.neq)
    r = b.copy()
    z = dinv * r + scale * r
    p = z.copy()
    rz = r @ z
    for _ in range(10):
        q = K @ p
        a = rz / (p @ q)
        xr += a * p
        r -= a * q
        z = dinv * r + scale * r
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    # the network output is r rounded to bf16 (relative 2^-9): a slightly
    # perturbed M^-1; a wrong box <-> lattice map would differ by O(1)
    assert np.max(np.abs(x - xr)) <= 1e-2 * np.max(np.abs(xr))
