"""Learned preconditioner: data from the method's own solves, training of
the U-net on the GPU, and evaluation inside flexible CG.

The free-DOF box can be placed in a zero-padded lattice whose extents divide
by 2^(levels-1) (``lattice_dims``, ``to_lattice``, ``from_lattice``).

The network sees the residual on the lattice of free DOFs (normalised to unit
RMS) plus auxiliary channels describing the problem (for diffusion: the
log-conductivity at the nodes), and predicts the part of the error that a
base preconditioner M0 misses: target = (e - M0 r) / |r|_rms with e = x* - x_k
and r = b - A x_k taken at the base solver's own iterates x_k.  M0 is the
Jacobi step D^-1 (base "jacobi", iterates of Jacobi-PCG) or one V-cycle of the
spline multigrid (base "mg", iterates of multigrid flexible CG); the second is
the multigrid-augmented arrangement of Azulay and Treister (SIAM J. Sci.
Comput. 2022), where the network learns what the cycle leaves.  Inside
flexible CG the preconditioner is z = M0 r + s |r|_rms U(r / |r|_rms, aux),
and an iteration drops the learned term when it would make <r, z> smaller
than a tenth of <r, M0 r> (counted in the solver's statistics).

All network arithmetic runs in the library's own HIP kernels
(ifem_gpu_unet_*); torch only holds device buffers here.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import List, Optional, Sequence

import numpy as np

from . import FP64, Session, model as M, problems, setup_problem


def lattice(pb) -> tuple:
    """(D, H, W, ncomp) of the free-DOF lattice, or raise if it is not a box."""
    n1, n2, n3 = pb.patch.n
    nc = pb.ncomp
    free = (pb.meqn.reshape(n3, n2, n1, nc) > 0)
    ks, js, is_ = np.nonzero(free.any(axis=3))
    if ks.size == 0:
        raise ValueError("no free DOFs")
    D, H, W = ks.max() - ks.min() + 1, js.max() - js.min() + 1, is_.max() - is_.min() + 1
    if D * H * W * nc != pb.neq or not free[ks.min():ks.max() + 1, js.min():js.max() + 1,
                                            is_.min():is_.max() + 1].all():
        raise ValueError("the free DOFs do not form a box lattice")
    return int(D), int(H), int(W), nc


def lattice_dims(pb, levels: int) -> tuple:
    """(D, H, W): the free-DOF box extents rounded up to multiples of
    2^(levels-1) (the U-net pools levels-1 times)."""
    D, H, W, _ = lattice(pb)
    m = 1 << (levels - 1)
    return tuple(int(-(-v // m) * m) for v in (D, H, W))


def to_lattice(v, pb, dims):
    """Equation-order vector(s) (..., neq) -> (..., nc, D, H, W) float32 with
    the free box at the origin and zeros beyond it."""
    import torch
    bd, bh, bw, nc = lattice(pb)
    D, H, W = dims
    lead = v.shape[:-1]
    t = v.reshape(*lead, bd, bh, bw, nc).float()
    t = t.movedim(-1, -4)
    out = torch.zeros(*lead, nc, D, H, W, dtype=torch.float32, device=v.device)
    out[..., :bd, :bh, :bw] = t
    return out


def from_lattice(t, pb):
    """Inverse of to_lattice: (..., nc, D, H, W) -> (..., neq) float64."""
    bd, bh, bw, nc = lattice(pb)
    lead = t.shape[:-4]
    box = t[..., :bd, :bh, :bw].movedim(-4, -1)
    return box.reshape(*lead, bd * bh * bw * nc).double()


def aux_channels(pb) -> np.ndarray:
    """Auxiliary input channels on the lattice, (naux, D, H, W) float32.

    Diffusion: log-conductivity at the free control points, scaled by the
    field's sigma.  Problems with constant coefficients have no channel."""
    D, H, W, nc = lattice(pb)
    f = pb.integrand.kappa
    if pb.integrand.kind != M.POISSON or f.type not in (M.FN_FOURIER, M.FN_EXP_FOURIER):
        return np.zeros((0, D, H, W), dtype=np.float32)
    n1, n2, n3 = pb.patch.n
    X = pb.patch.coefs.reshape(n3, n2, n1, 3)
    free = pb.meqn.reshape(n3, n2, n1) > 0
    P = X[free].reshape(-1, 3)
    md = f.modes
    g = f.mean + f.scale * (md[:, 4][None, :] * np.cos(P @ md[:, :3].T + md[:, 3][None, :])).sum(1)
    g = g / max(f.scale, 1e-30)
    return g.reshape(1, D, H, W).astype(np.float32)


@dataclasses.dataclass
class NetSpec:
    levels: int = 3
    channels: Sequence[int] = (8, 16, 48)
    seed: int = 1

    def params(self, cin, cout):
        n = 0
        prev = cin
        c = list(self.channels)
        for l in range(self.levels):
            n += prev * c[l] * 27 + c[l] + c[l] * c[l] * 27 + c[l]
            prev = c[l]
        for l in range(self.levels - 2, -1, -1):
            n += (c[l + 1] + c[l]) * c[l] * 27 + c[l] + c[l] * c[l] * 27 + c[l]
        n += c[0] * cout * 27 + cout
        return n

    def flops(self, cin, cout, dims):
        """Forward multiply-add work of one sample, 2*Cin*Cout*27 per voxel
        and 3^3 convolution (the same layers as params(), level l at 1/8^l of
        the D*H*W voxels)."""
        vox = [int(np.prod(dims)) // 8 ** l for l in range(self.levels)]
        c = list(self.channels)
        f, prev = 0, cin
        for l in range(self.levels):
            f += 2 * 27 * vox[l] * (prev * c[l] + c[l] * c[l])
            prev = c[l]
        for l in range(self.levels - 2, -1, -1):
            f += 2 * 27 * vox[l] * ((c[l + 1] + c[l]) * c[l] + c[l] * c[l])
        return f + 2 * 27 * vox[0] * c[0] * cout


def jacobi_inverse(s: Session) -> np.ndarray:
    rp, ci, A, _ = s.system()
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    d = np.zeros(len(rp) - 1)
    m = ci == rows
    d[rows[m]] = A[m]
    return np.where(d != 0, 1.0 / d, 1.0)


def pairs(s: Session, pb, base: str = "jacobi", rhs=None, rtol: float = 1e-10, every: int = 1,
          kmax: int = 100000, torch_device="cuda", features=1, target: str = "defect",
          dims=None):
    """(input, target) tensors from the iterates of one flexible-CG solve of
    an assembled system (right-hand side replaced by rhs when given), driven
    from the host with the device SpMV and base preconditioner M0 = D^-1
    (base "jacobi") or one V-cycle (base "mg"): every `every`-th iterate
    x_k before convergence to rtol gives the pair (r_k / |r_k|_rms [, with
    aux], (x* - x_k - M0 r_k) / |r_k|_rms), x* the converged iterate
    (target "defect"), or (x* - x_k) / |r_k|_rms (target "error"), on the
    free-DOF box zero-padded to dims (default: the box itself).
    features as in ifem_gpu_set_learned_preconditioner."""
    import torch
    D, H, W, nc = lattice(pb)
    aux = aux_channels(pb)
    naux = aux.shape[0]
    if rhs is not None:
        s.set_rhs(rhs)
    b = s.rhs()
    if base == "mg":
        s.mg_setup()
        prec = s.mg_apply
    else:
        dinv = jacobi_inverse(s)
        prec = lambda r: dinv * r
    x = np.zeros_like(b)
    r = b.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    hist = []
    for k in range(kmax):
        if np.linalg.norm(r) <= rtol * bn:
            break
        if k % every == 0:
            hist.append((x.copy(), r.copy(), z.copy()))
        q = s.spmv(p)
        a = rz / (p @ q)
        x += a * p
        ro = r
        r = r - a * q
        z = prec(r)
        rzn = r @ z
        p = z + ((rzn - ro @ z) / rz) * p
        rz = rzn
    xs = x
    ins, tgs = [], []
    pdims = tuple(dims) if dims is not None else (D, H, W)
    auxt = torch.from_numpy(padded_aux(pb, pdims)).to(torch_device)

    def pad(t):
        if pdims == (D, H, W):
            return t
        out = torch.zeros((t.shape[0],) + pdims, dtype=t.dtype, device=t.device)
        out[:, :D, :H, :W] = t
        return out

    for xk, rk, zk in hist:
        rms = float(np.sqrt(np.mean(rk * rk)))
        if not rms > 0:
            continue
        rin = torch.from_numpy((rk / rms).astype(np.float32)).to(torch_device)
        rin = pad(rin.reshape(D, H, W, nc).permute(3, 0, 1, 2))
        e = xs - xk if target == "error" else xs - xk - zk
        tgt = torch.from_numpy((e / rms).astype(np.float32)).to(torch_device)
        tgt = pad(tgt.reshape(D, H, W, nc).permute(3, 0, 1, 2))
        if naux and features:
            mod = [rin * auxt[j:j + 1] for j in range(naux)]
            ins.append(torch.cat([rin] + mod, 0).contiguous())
        else:
            ins.append((torch.cat([rin, auxt], 0) if naux else rin).contiguous())
        tgs.append(tgt.contiguous())
    return ins, tgs


def pcg_pairs(s: Session, pb, every: int = 4, torch_device="cuda", features=1):
    """pairs() with the Jacobi base."""
    return pairs(s, pb, "jacobi", None, every=every, torch_device=torch_device, features=features)


def random_rhs(pb, seed: int, nmodes: int = 8, kmax: int = 3) -> np.ndarray:
    """A seeded smooth right-hand side in equation order: per component a sum
    of nmodes low-frequency products of sines at the free control points."""
    n1, n2, n3 = pb.patch.n
    nc = pb.ncomp
    X = pb.patch.coefs
    u = M.uniform(seed, nmodes * 4 * nc).reshape(nc, nmodes, 4)
    vals = np.zeros((len(X), nc))
    for c in range(nc):
        for m in range(nmodes):
            k = 1 + np.floor(u[c, m, :3] * kmax)
            vals[:, c] += (2 * u[c, m, 3] - 1) * np.prod(np.sin(np.pi * k[None, :] * X), axis=1)
    v = vals.reshape(-1)
    out = np.zeros(pb.neq)
    free = pb.meqn > 0
    out[pb.meqn[free] - 1] = v[free]
    return out * (1.0 / max(1.0, float(pb.patch.nel[0]) ** 3))


def default_iters(n: int, kmax: int) -> List[int]:
    """n iteration counts spread geometrically over [1, kmax]."""
    ks = np.unique(np.round(np.geomspace(1, kmax, n)).astype(int))
    return [int(k) for k in ks]


def target_scale(T) -> float:
    """RMS of the training targets: the network learns T / scale, and the
    preconditioner multiplies its output by the same scale."""
    import torch
    return float(torch.sqrt(torch.mean(T.double() ** 2)))


def sample_weights(T):
    """Per-sample loss weights 1/mean(t_n^2), normalised to mean 1: every pair
    counts by its relative error, whatever the amplification of its residual."""
    import torch
    m = torch.mean(T.double().reshape(T.shape[0], -1) ** 2, dim=1)
    w = 1.0 / torch.clamp(m, min=1e-30)
    return (w / w.mean()).float().contiguous()


def train(s: Session, X, T, epochs=20, batch=8, lr=2e-3, log=None, seed=0, weights=None):
    """Adam on the pairs (X[i], T[i]) with a cosine learning-rate schedule and
    optional per-sample loss weights.  T must already be divided by
    target_scale(T)."""
    import torch
    n = X.shape[0]
    D, H, W = X.shape[2:]
    rng = np.random.default_rng(seed)
    steps = epochs * max(1, n // batch)
    step = 0
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot, cnt = 0.0, 0
        for i in range(0, n - batch + 1, batch):
            idx = torch.from_numpy(perm[i:i + batch]).to(X.device)
            xb = X.index_select(0, idx).contiguous()
            tb = T.index_select(0, idx).contiguous()
            wb = weights.index_select(0, idx).contiguous() if weights is not None else None
            torch.cuda.synchronize()
            a = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(steps, 1)))
            l = s.unet_train_step(xb.data_ptr(), tb.data_ptr(), batch, D, H, W, a, want_loss=True,
                                  w_ptr=wb.data_ptr() if wb is not None else 0)
            step += 1
            tot += l
            cnt = (i // batch) + 1
        losses.append(tot / max(cnt, 1))
        if log:
            log(f"[learned] epoch {ep} loss {tot:.4e}")
    torch.cuda.synchronize()
    return losses


def enable(s: Session, pb, scale=1.0, features=1, dims=None):
    """Turn the trained network into the preconditioner of session s (on the
    lattice dims, default the free box itself)."""
    import torch
    D, H, W, nc = lattice(pb)
    dims = dims or (D, H, W)
    aux = padded_aux(pb, dims)
    auxt = torch.from_numpy(np.ascontiguousarray(aux)).to("cuda") if aux.shape[0] else None
    s.set_learned(True, dims[0], dims[1], dims[2], auxt.data_ptr() if auxt is not None else 0,
                  aux.shape[0], scale, features)
    torch.cuda.synchronize()
    return auxt


def padded_aux(pb, dims) -> np.ndarray:
    aux = aux_channels(pb)
    D, H, W = dims
    out = np.zeros((aux.shape[0], D, H, W), dtype=np.float32)
    out[:, :aux.shape[1], :aux.shape[2], :aux.shape[3]] = aux
    return out


def calibrate_scale(pb, params, spec, cin, cout, tscale, candidates=(1.0, 0.5, 0.25, 0.125),
                    pc=None, rtol=1e-8, precision=None):
    """Output scale of a trained network, chosen on a validation problem pb
    (not a test problem): the factor in candidates (times the training
    target scale) with the fewest flexible-CG iterations to rtol, ties to the
    larger factor.  Returns (scale, {factor: iterations})."""
    from ._native import FP64, PC_MG
    pc = PC_MG if pc is None else pc
    precision = FP64 if precision is None else precision
    its = {}
    with setup_problem(pb, precision=precision) as s:
        s.unet_create(spec.levels, spec.channels, cin, cout)
        s.unet_set_params(params)
        for f in candidates:
            enable(s, pb, scale=tscale * f, dims=lattice_dims(pb, spec.levels))
            st = s.solve(rtol=rtol, maxits=5000, pc=pc, precision=precision, check_every=8)
            its[f] = st["iterations"] if st["converged"] else 10 ** 9
    best = min(candidates, key=lambda f: (its[f], -f))
    return tscale * best, its
