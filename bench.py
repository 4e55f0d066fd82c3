#!/usr/bin/env python3
"""Benchmark of the isogeometric assembly + Jacobi-PCG hot path on one MI355X.

Workload: 3D isotropic linear elasticity (E=1, nu=0.3) on one p=2 B-spline
patch of 64x64x64 elements (66^3 x 3 = 862,488 DOFs), 3x3x3 Gauss points,
unit cube with interior control points jittered by U(-0.15h, 0.15h)
(splitmix64 seed 2024), clamped at x=0, traction (0,0,-1) on x=1.

One step = numeric FP64 assembly of K and b (volume + traction) followed by
the plain FP64 Jacobi-PCG solve to a true relative residual of 1e-8.  The
metric value is assembled DOF/s (all DOFs incl. eliminated ones / numeric
assembly device time).  Inputs (patch, SAM, sparsity pattern) are resident in
HBM before the timed region; the symbolic pattern build is reported apart.

Also reported: the FP32-inner / FP64-refinement solve (time, iterations,
energy-norm and max-nodal error against the FP64 solve), flexible CG with the
FP32 spline multigrid preconditioner (setup included in its time), the roofline of the
assembly kernel and of the SpMV, the oracle (CPU restatement) timed on a
bounded sample on the host cores, and the batch of 256 uneven diffusion
patches (one persistent assembly launch per precision, per-patch PCG).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12       # MI355X HBM3E bytes/s
PEAK_FP64 = 78.6e12     # MI355X FP64 (vector and matrix) FLOP/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a 1 GiB copy), or None."""
    fn = os.path.join(ROOT, "profiles", "hbm_traffic_elasticity64.md")
    col = None
    try:
        with open(fn) as f:
            for line in f:
                cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
                if cells and cells[0] == "kernel":
                    col = next((i for i, c in enumerate(cells) if c.startswith("traffic GB")), None)
                elif col is not None and len(cells) > col and cells[0] == kernel:
                    return round(float(cells[col]) * 1e9)
    except (OSError, ValueError):
        pass
    return None


def _train_net(LN, Session, spec, X, T, epochs, lr=1e-3):
    """Adam in three stages of halving learning rate; returns (params,
    target scale, loss first/last, seconds)."""
    import time as _t
    tscale = LN.target_scale(T)
    Tn = (T / tscale).contiguous()
    w = LN.sample_weights(Tn)
    tr = Session()
    tr.unet_create(spec.levels, spec.channels, X.shape[1], T.shape[1], seed=spec.seed)
    t0 = _t.perf_counter()
    first, last = None, None
    for stage in range(3):
        losses = LN.train(tr, X, Tn, epochs=max(1, epochs // 3), batch=8, lr=lr * 0.5 ** stage,
                          weights=w, seed=stage)
        first = losses[0] if first is None else first
        last = losses[-1]
    import torch
    torch.cuda.synchronize()
    secs = _t.perf_counter() - t0
    params = tr.unet_params()
    tr.close()
    return params, tscale, (first, last), secs


def _net_forward(Session, stream, spec, params, cin, cout, dims, reps=20):
    """Device time of one network application (one sample, the solver's
    use) on a D x H x W lattice: HIP events on the library's stream around
    reps forward calls; with its forward work in TFLOP/s."""
    import torch
    D, H, W = dims
    x = torch.randn((1, cin, D, H, W), device="cuda", dtype=torch.float32)
    y = torch.empty((1, cout, D, H, W), device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    with Session(stream=stream.cuda_stream) as t:
        t.unet_create(spec.levels, spec.channels, cin, cout)
        t.unet_set_params(params)
        t.unet_forward(x.data_ptr(), y.data_ptr(), 1, D, H, W)  # workspace, warm-up
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            t.unet_forward(x.data_ptr(), y.data_ptr(), 1, D, H, W)
        e1.record(stream)
        e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"lattice": f"{D}x{H}x{W}", "ms": round(ms, 4),
            "tflops": round(spec.flops(cin, cout, dims) / ms / 1e9, 2)}


def learned_legs(args, stream, s, pb, x64, solve64_ms):
    """The learned correction inside flexible CG (library kernels for
    training and inference):
    * diffusion 48^3 p=2, log-normal kappa (sigma 1.5, length 0.1):
      a 3-level U-net trained on (residual, multigrid defect) pairs from the
      method's own multigrid flexible-CG iterates over seeded coefficient
      fields, evaluated on held-out seeds against Jacobi-PCG and multigrid
      without the network;
    * elasticity: the network trained the same way at 48^3, applied
      at 64^3 (learned + FP32: FP32 Krylov, FP32 multigrid and the network
      inside FP64 refinement), timed against plain FP64 Jacobi-PCG."""
    import numpy as np
    import torch

    from ifem_amd import FP32, FP64, PC_JACOBI, PC_MG, Session, learned as LN, problems, \
        setup_problem
    spec = LN.NetSpec(levels=3, channels=(12, 24, 48), seed=1)
    out = {}

    def gen(mk, seed0, npairs, dims=None, rhs_per_seed=2):
        X, T = [], []
        seed = seed0
        while len(X) < npairs:
            pbq = mk(seed)
            with setup_problem(pbq) as sq:
                for j in range(rhs_per_seed):
                    rhs = None if j == 0 else LN.random_rhs(pbq, 7919 * seed + j)
                    ins, tgs = LN.pairs(sq, pbq, "mg", rhs, every=1, dims=dims)
                    X += ins
                    T += tgs
            seed += 1
        return (torch.stack(X[:npairs]).contiguous(), torch.stack(T[:npairs]).contiguous(),
                seed - seed0)

    # ---- diffusion 48^3, sigma 1.5, length 0.1 -----------------------------------
    mkd = lambda seed: problems.diffusion_cube(48, 2, sigma=1.5, ell=0.1, seed=seed)
    t0 = time.perf_counter()
    X, T, nseeds = gen(mkd, 100, args.learned_pairs)
    gen_s = time.perf_counter() - t0
    params, tscale, loss, train_s = _train_net(LN, Session, spec, X, T, args.learned_epochs)
    cin, cout = int(X.shape[1]), int(T.shape[1])
    del X, T
    torch.cuda.empty_cache()
    # output scale chosen on a validation seed (neither a training nor a test seed)
    scale, cal = LN.calibrate_scale(mkd(99999), params, spec, cin, cout, tscale)
    fwd48 = _net_forward(Session, stream, spec, params, cin, cout, LN.lattice_dims(mkd(0), 3))
    rows = []
    for seed in range(100000, 100000 + args.learned_eval_seeds):
        pbd = mkd(seed)
        with setup_problem(pbd, precision=FP32) as sd:
            j = sd.solve(rtol=args.rtol, maxits=200000, check_every=64)
            m = sd.solve(rtol=args.rtol, maxits=5000, pc=PC_MG, check_every=8)
            sd.unet_create(spec.levels, spec.channels, cin, cout)
            sd.unet_set_params(params)
            LN.enable(sd, pbd, scale=scale)
            l64 = sd.solve(rtol=args.rtol, maxits=5000, pc=PC_MG, check_every=8)
            l32 = sd.solve(rtol=args.rtol, maxits=5000, pc=PC_MG, precision=FP32, check_every=8)
        rows.append((j["iterations"], j["solve_ms"], m["iterations"], m["solve_ms"],
                     l64["iterations"], l64["solve_ms"], l32["iterations"], l32["solve_ms"],
                     l64["pc_rejected"] + l32["pc_rejected"],
                     int(l64["fallback"]) + int(l32["fallback"]),
                     bool(j["converged"] and m["converged"] and l64["converged"] and l32["converged"]
                          and max(l64["rel_residual"], l32["rel_residual"]) <= args.rtol)))
    R = np.array([r[:10] for r in rows], dtype=float)
    med = lambda k: float(np.median(R[:, k]))
    out["diffusion48"] = {
        "workload": "diffusion 48^3 p=2, log-normal kappa (sigma 1.5, length 0.1), smooth source, "
                    "u=0; U-net 3 levels (12,24,48 channels), BF16 operands, FP32 accumulation",
        "parameters": int(params.size), "train_pairs": args.learned_pairs, "train_seeds": nseeds,
        "pairs_generation_s": round(gen_s, 1), "epochs": args.learned_epochs,
        "train_s": round(train_s, 1), "loss_first_last": [round(loss[0], 4), round(loss[1], 4)],
        "output_scale_factor": round(scale / tscale, 3),
        "validation_iterations_by_factor": {str(k): v for k, v in cal.items()},
        "network_forward": fwd48,
        "eval_seeds": len(rows),
        "iterations_median": {"jacobi_fp64": med(0), "multigrid_fp64": med(2),
                              "learned_mg_fp64": med(4), "learned_mg_fp32": med(6)},
        "time_to_tol_ms_median": {"jacobi_fp64": round(med(1), 3), "multigrid_fp64": round(med(3), 3),
                                  "learned_mg_fp64": round(med(5), 3),
                                  "learned_mg_fp32": round(med(7), 3)},
        "learned_gain_iterations_vs_multigrid": round(float(np.median(R[:, 2] / R[:, 4])), 3),
        "learned_gain_time_vs_multigrid": round(float(np.median(R[:, 3] / R[:, 5])), 3),
        "learned_fp32_speedup_vs_jacobi_fp64": round(float(np.median(R[:, 1] / R[:, 7])), 3),
        "safeguard_rejections_total": int(R[:, 8].sum()), "fallbacks_total": int(R[:, 9].sum()),
        "all_converged": bool(all(r[10] for r in rows)),
    }
    log(f"[bench] learned diffusion48: {out['diffusion48']}")

    if s is None:
        return out
    # ---- elasticity network trained at 48^3, applied at 64^3 --------------------
    mke = lambda seed: problems.elasticity_cube(48, 2, seed=seed)
    dims48 = LN.lattice_dims(mke(0), spec.levels)
    t0 = time.perf_counter()
    X, T, nseeds = gen(mke, 300, args.learned_elast_pairs, dims=dims48)
    gen_s = time.perf_counter() - t0
    params, tscale, loss, train_s = _train_net(LN, Session, spec, X, T, args.learned_elast_epochs)
    cin, cout = int(X.shape[1]), int(T.shape[1])
    del X, T
    torch.cuda.empty_cache()
    scale, cal = LN.calibrate_scale(mke(999), params, spec, cin, cout, tscale, precision=FP32)
    dims64 = LN.lattice_dims(pb, spec.levels)
    fwd64 = _net_forward(Session, stream, spec, params, cin, cout, dims64)
    with torch.cuda.stream(stream):
        s.assemble(pb.integrand, pb.ng, FP32)  # FP32 K for the inner Krylov loop
    stream.synchronize()
    s.unet_create(spec.levels, spec.channels, cin, cout)
    s.unet_set_params(params)
    LN.enable(s, pb, scale=scale, dims=dims64)
    runs = [s.solve(rtol=args.rtol, maxits=5000, pc=PC_MG, precision=FP32, check_every=8)
            for _ in range(3)]
    x = s.solution()
    # the same network and FP32 V-cycle inside FP64 Krylov (no refinement restarts)
    runs64 = [s.solve(rtol=args.rtol, maxits=5000, pc=PC_MG, check_every=8) for _ in range(3)]
    s.set_learned(False)
    base = [s.solve(rtol=args.rtol, maxits=5000, pc=PC_MG, precision=FP32, check_every=8)
            for _ in range(3)]  # the same solve without the network
    e = x - x64
    en = float(np.sqrt(abs(e @ s.spmv(e))) / np.sqrt(x64 @ s.spmv(x64)))
    mn = float(np.max(np.abs(e)) / np.max(np.abs(x64)))
    ms = float(np.median([r["solve_ms"] for r in runs]))
    out["elasticity64"] = {
        "workload": "elasticity 64^3 p=2 (the benchmark patch); network trained on 48^3 "
                    "elasticity (seeded jitter, random smooth loads), applied on the padded "
                    f"{dims64[0]}x{dims64[1]}x{dims64[2]} lattice",
        "train_pairs": args.learned_elast_pairs, "train_seeds": nseeds,
        "pairs_generation_s": round(gen_s, 1), "epochs": args.learned_elast_epochs,
        "train_s": round(train_s, 1), "loss_first_last": [round(loss[0], 4), round(loss[1], 4)],
        "output_scale_factor": round(scale / tscale, 3),
        "validation_iterations_by_factor": {str(k): v for k, v in cal.items()},
        "network_forward": fwd64,
        "learned_fp32_solve_ms": round(ms, 3), "iterations": runs[-1]["iterations"],
        "outer": runs[-1]["outer"], "rel_residual": runs[-1]["rel_residual"],
        "converged": all(r["converged"] for r in runs),
        "safeguard_rejections": runs[-1]["pc_rejected"], "fallback": runs[-1]["fallback"],
        "energy_norm_err_vs_fp64": en, "max_nodal_err_vs_fp64": mn,
        "speedup_vs_fp64_jacobi": round(solve64_ms / ms, 3),
        "multigrid_fp32_without_network": {
            "solve_ms": round(float(np.median([r["solve_ms"] for r in base])), 3),
            "iterations": base[-1]["iterations"]},
        "learned_fp64_krylov": {
            "solve_ms": round(float(np.median([r["solve_ms"] for r in runs64])), 3),
            "iterations": runs64[-1]["iterations"], "rel_residual": runs64[-1]["rel_residual"],
            "converged": all(r["converged"] for r in runs64)},
        "learned_gain_time_vs_multigrid": round(float(np.median([r["solve_ms"] for r in base]))
                                                / ms, 3),
    }
    return out


def large_leg(args, stream):
    """Largest single-GPU case: elasticity 96^3 p=3, 4x4x4 Gauss points."""
    import numpy as np
    import torch
    from ifem_amd import FP32, FP64, Session, problems
    from ifem_amd import PC_MG
    pbL = problems.elasticity_cube(96, 3)
    log(f"[bench] large: elasticity 96^3 p=3, ndof={pbL.ndof}")
    with Session(stream=stream.cuda_stream) as sL:
        sL.bind(sL.add_patch(pbL.patch), 3, pbL.meqn, pbL.neq)
        t0 = time.perf_counter()
        nnzL = sL.init_assembly()
        symL = (time.perf_counter() - t0) * 1e3
        faceL, tracL, ngL = pbL.traction
        runs = []
        for _ in range(3):
            with torch.cuda.stream(stream):
                sL.assemble(pbL.integrand, pbL.ng, FP64)
                sL.add_traction(faceL, tracL, ngL)
            stream.synchronize()
            runs.append((sL.timings()[0], *sL.assembly_parts()))
        kL = float(np.median([r[0] for r in runs]))
        gL = float(np.median([r[1] for r in runs]))
        lL = float(np.median([r[2] for r in runs]))
        nenL = pbL.patch.nen
        flopsL = (36 * nenL * nenL + 6 * nenL) * int(np.prod(pbL.ng)) * int(np.prod(pbL.patch.nel))
        if args.assembly_only:
            return {"assembly_kernels_ms": round(kL, 3), "line_kernel_ms": round(lL, 3),
                    "geometry_pass_ms": round(gL, 3), "nnz": nnzL}
        st64 = sL.solve(rtol=args.rtol, maxits=200000, check_every=64)
        xL = sL.solution()
        KxL = sL.spmv(xL)

        def errL(x):
            e = x - xL
            return (float(np.sqrt(abs(e @ sL.spmv(e))) / np.sqrt(xL @ KxL)),
                    float(np.max(np.abs(e)) / np.max(np.abs(xL))))
        with torch.cuda.stream(stream):
            sL.assemble(pbL.integrand, pbL.ng, FP32)
        stream.synchronize()
        a32 = sL.timings()[0]
        stm = sL.solve(rtol=args.rtol, maxits=200000, precision=FP32, inner_rtol=1e-4,
                       check_every=64)
        em = errL(sL.solution())
        stg = [sL.solve(rtol=args.rtol, maxits=5000, pc=PC_MG, check_every=8) for _ in range(2)]
        eg = errL(sL.solution())
        large = {
            "workload": "elasticity 96^3 p=3 jittered cube, 4x4x4 Gauss, clamped x=0, "
                        "traction (0,0,-1) on x=1",
            "dofs": pbL.ndof, "neq": pbL.neq, "nnz": nnzL, "symbolic_ms": round(symL, 1),
            "assembly_kernels_ms": round(kL, 3),
            "assembled_dof_per_s": round(pbL.ndof / (kL * 1e-3), 1),
            "line_kernel_ms": round(lL, 3), "geometry_pass_ms": round(gL, 3),
            "line_kernel_tflops": round(flopsL / (lL * 1e-3) / 1e12, 3),
            "line_kernel_frac_fp64_peak": round(flopsL / (lL * 1e-3) / PEAK_FP64, 4),
            "fp64": {"solve_ms": round(st64["solve_ms"], 3), "iterations": st64["iterations"],
                     "rel_residual": st64["rel_residual"], "converged": st64["converged"]},
            "mixed": {"fp32_assembly_kernels_ms": round(a32, 3),
                      "solve_ms": round(stm["solve_ms"], 3), "iterations": stm["iterations"],
                      "outer": stm["outer"], "rel_residual": stm["rel_residual"],
                      "converged": stm["converged"], "energy_norm_err_vs_fp64": em[0],
                      "max_nodal_err_vs_fp64": em[1]},
            "multigrid": {"solve_ms": round(stg[1]["solve_ms"], 3),
                          "first_solve_ms": round(stg[0]["solve_ms"], 3),
                          "iterations": stg[1]["iterations"],
                          "rel_residual": stg[1]["rel_residual"],
                          "converged": all(r["converged"] for r in stg),
                          "energy_norm_err_vs_fp64": eg[0], "max_nodal_err_vs_fp64": eg[1]},
            "speedup_mg_vs_fp64": round(st64["solve_ms"] / stg[1]["solve_ms"], 3),
        }
    return large



def batch_leg(args):
    """256 uneven diffusion patches: one persistent assembly launch per precision."""
    import numpy as np
    from ifem_amd import FP32, FP64, problems
    from ifem_amd import setup_batch
    bpbs = problems.diffusion_batch()
    bdofs = int(sum(q.ndof for q in bpbs))
    bs = setup_batch(bpbs, precisions=())
    tune = bs.batch_autotune(FP32)
    bt = {}
    for prec in (FP64, FP32):
        ts = []
        for _ in range(3):
            bs.batch_assemble(prec)
            ts.append(bs.batch_timings()["assemble_ms"])
        bt[prec] = float(np.median(ts))
    if args.assembly_only:
        bs.close()
        return {"assembly_fp64_ms": round(bt[FP64], 2), "assembly_fp32_ms": round(bt[FP32], 2)}
    b64 = bs.batch_solve(rtol=args.rtol, maxits=20000, check_every=32)
    bmx = bs.batch_solve(rtol=args.rtol, maxits=20000, precision=FP32, check_every=32)
    degs = [q.patch.p[0] for q in bpbs]
    batch = {
        "workload": "256 diffusion patches, nel per direction U{8..48}, p in {2,3}, "
                    "log-normal kappa (sigma 1, length 0.2), smooth source, u=0",
        "patches": len(bpbs), "p2": degs.count(2), "p3": degs.count(3),
        "dofs": bdofs, "neq": int(sum(q.neq for q in bpbs)),
        "symbolic_ms": round(bs.batch_timings()["symbolic_ms"], 2),
        "assembly_fp64_ms": round(bt[FP64], 2), "assembly_fp32_ms": round(bt[FP32], 2),
        "assembly_fp32_note": "FP32 values after an FP64 assembly of the same batch: the "
                              "Gauss-point geometry records of the FP64 pass are reused",
        "assembled_dof_per_s_fp64": round(bdofs / (bt[FP64] * 1e-3), 1),
        "fp64_solve_ms": round(b64["solve_ms"], 2), "fp64_max_iterations": b64["iterations"],
        "mixed_solve_ms": round(bmx["solve_ms"], 2), "mixed_max_iterations": bmx["iterations"],
        "mixed_outer": bmx["outer"],
        "all_converged": bool(b64["converged"] and bmx["converged"]),
        "max_rel_residual": float(max(b64["rel_residual"], bmx["rel_residual"])),
        "tuning": {f"p{p}": tune[p] for p in (2, 3)},
    }
    bs.close()
    return batch



def copy_calibration(stream):
    """A 1 GiB device-to-device copy: the reference kernel whose FETCH_SIZE /
    WRITE_SIZE counters scale the per-kernel HBM traffic."""
    import torch
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    with torch.cuda.stream(stream):
        dst.copy_(src)
    stream.synchronize()
    log("[bench] copy calibration: 1 GiB read + 1 GiB written")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nel", type=int, default=64)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-mg", action="store_true", help="skip the multigrid-preconditioned leg")
    ap.add_argument("--no-multipatch", action="store_true",
                    help="skip the two-patch assembly leg")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-preconditioner leg")
    ap.add_argument("--learned-pairs", type=int, default=2000,
                    help="training pairs of the 48^3 diffusion network")
    ap.add_argument("--learned-epochs", type=int, default=30)
    ap.add_argument("--learned-eval-seeds", type=int, default=200)
    ap.add_argument("--learned-elast-pairs", type=int, default=1500,
                    help="training pairs of the elasticity network (48^3)")
    ap.add_argument("--learned-elast-epochs", type=int, default=30)
    ap.add_argument("--only", choices=["large", "batch", "learned"], default=None,
                    help="run only that leg and print its record (for kernel profiles of one "
                         "workload)")
    ap.add_argument("--assembly-only", action="store_true",
                    help="with --only large/batch: assemble, skip the solves (counter passes)")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 96^3 p=3 elasticity leg (largest single-GPU case)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched uneven-patch leg (256 diffusion patches)")
    ap.add_argument("--cpu-layers", type=int, default=16,
                    help="element layers (in w) of the patch assembled by the CPU baseline")
    ap.add_argument("--spmv-reps", type=int, default=20)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (reference kernel for HBM counters)")
    return ap.parse_args()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE {world}); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from ifem_amd import FP32, FP64, Session, problems

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()  # default priority
    if args.only and args.copy_calibration:
        copy_calibration(stream)
    if args.only == "large":
        print(json.dumps({"large": large_leg(args, stream)}), flush=True)
        return
    if args.only == "batch":
        print(json.dumps({"batch": batch_leg(args)}), flush=True)
        return
    if args.only == "learned":  # the 48^3 diffusion part alone
        print(json.dumps({"learned": learned_legs(args, stream, None, None, None, None)}),
              flush=True)
        return
    pb = problems.elasticity_cube(args.nel, 2)
    ndof, neq = pb.ndof, pb.neq
    nel_tot = int(np.prod(pb.patch.nel))
    nen = pb.patch.nen
    ngp = int(np.prod(pb.ng))
    log(f"[bench] workload elasticity {args.nel}^3 p=2: ndof={ndof} neq={neq}")

    s = Session(stream=stream.cuda_stream)
    pid = s.add_patch(pb.patch)
    s.bind(pid, 3, pb.meqn, neq)
    t0 = time.perf_counter()
    nnz = s.init_assembly()
    symbolic_ms = (time.perf_counter() - t0) * 1e3
    log(f"[bench] pattern nnz={nnz} in {symbolic_ms:.1f} ms")

    face, trac, ngf = pb.traction
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]

    def step():
        with torch.cuda.stream(stream):
            ev[0].record(stream)
            s.assemble(pb.integrand, pb.ng, FP64)
            s.add_traction(face, trac, ngf)
            ev[1].record(stream)
            st = s.solve(rtol=args.rtol, maxits=200000, check_every=64)
            ev[2].record(stream)
        stream.synchronize()
        kern_ms, solve_ms = s.timings()
        geo_ms, lines_ms = s.assembly_parts()
        return (ev[0].elapsed_time(ev[1]), kern_ms, ev[1].elapsed_time(ev[2]), st, geo_ms,
                lines_ms)

    for w in range(args.warmup):
        r = step()
        log(f"[bench] warmup {w}: assemble {r[0]:.2f} ms, solve {r[2]:.1f} ms, iters {r[3]['iterations']}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for i, r in enumerate(res):
        log(f"[bench] step {i}: assemble {r[0]:.2f} ms (kernel {r[1]:.2f}), solve {r[2]:.1f} ms, "
            f"iters {r[3]['iterations']}, rel res {r[3]['rel_residual']:.2e}")
    if not all(r[3]["converged"] for r in res):
        raise SystemExit("FP64 solve did not converge")
    asm_ms = float(np.median([r[0] for r in res]))
    kern_ms = float(np.median([r[1] for r in res]))
    geo_ms = float(np.median([r[4] for r in res]))
    lines_ms = float(np.median([r[5] for r in res]))
    solve64_ms = float(np.median([r[2] for r in res]))
    iters64 = res[-1][3]["iterations"]
    x64 = s.solution()

    # ---- SpMV roofline (device vectors, ctx stream) ---------------------------
    xs = torch.randn(neq, dtype=torch.float64, device="cuda")
    ys = torch.empty_like(xs)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        s.spmv_device(xs.data_ptr(), ys.data_ptr(), FP64)
        e0.record(stream)
        for _ in range(args.spmv_reps):
            s.spmv_device(xs.data_ptr(), ys.data_ptr(), FP64)
        e1.record(stream)
    stream.synchronize()
    spmv_ms = e0.elapsed_time(e1) / args.spmv_reps
    # node-triple index: values + one int32 per 9 values + 2 row pointers per
    # node + x read + y written (3-component node-uniform system)
    spmv_bytes = nnz * 8 + (nnz // 9) * 4 + (neq // 3) * 2 * 8 + 2 * neq * 8

    if args.copy_calibration:
        copy_calibration(stream)

    # ---- mixed precision: FP32 inner PCG + FP64 refinement ---------------------
    mixed = None
    if not args.no_mixed:
        with torch.cuda.stream(stream):
            s.assemble(pb.integrand, pb.ng, FP32)
        fp32_asm_ms = s.timings()[0]
        runs = []
        for _ in range(2):
            st = s.solve(rtol=args.rtol, maxits=200000, precision=FP32, inner_rtol=1e-4,
                         check_every=64)
            runs.append(st)
        xm = s.solution()
        e = xm - x64
        Ke = s.spmv(e)
        Kx = s.spmv(x64)
        energy = float(np.sqrt(abs(e @ Ke)) / np.sqrt(x64 @ Kx))
        maxnodal = float(np.max(np.abs(e)) / np.max(np.abs(x64)))
        mixed = {
            "solve_ms": round(float(np.median([r["solve_ms"] for r in runs])), 3),
            "iterations": runs[-1]["iterations"], "outer": runs[-1]["outer"],
            "rel_residual": runs[-1]["rel_residual"], "converged": runs[-1]["converged"],
            "energy_norm_err_vs_fp64": energy, "max_nodal_err_vs_fp64": maxnodal,
            "fp32_assembly_kernel_ms": round(fp32_asm_ms, 3),
        }
        log(f"[bench] mixed: {mixed}")

    # ---- spline multigrid (FP32) inside FP64 flexible CG ----------------------
    mg = None
    if not args.no_mg:
        from ifem_amd import PC_MG
        def err_vs_fp64(x):
            e = x - x64
            return (float(np.sqrt(abs(e @ s.spmv(e))) / np.sqrt(x64 @ s.spmv(x64))),
                    float(np.max(np.abs(e)) / np.max(np.abs(x64))))
        legs = {}
        for name, prec in (("fp64_krylov", FP64), ("fp32_krylov", FP32)):
            if prec == FP32 and mixed is None:
                continue
            runs = [s.solve(rtol=args.rtol, maxits=5000, pc=PC_MG, precision=prec, mg_smooth=2,
                            check_every=8) for _ in range(3)]
            en, mn = err_vs_fp64(s.solution())
            legs[name] = {"solve_ms": round(float(np.median([r["solve_ms"] for r in runs])), 3),
                          # the first multigrid solve of a system also builds the hierarchy
                          # (levels, transfers, masks: once per sparsity pattern)
                          "first_solve_ms": round(runs[0]["solve_ms"], 3),
                          "iterations": runs[-1]["iterations"], "outer": runs[-1]["outer"],
                          "rel_residual": runs[-1]["rel_residual"],
                          "converged": all(r["converged"] for r in runs),
                          "energy_norm_err_vs_fp64": en, "max_nodal_err_vs_fp64": mn}
        nlev, setup_ms = s.mg_setup(2)
        mg = {"levels": nlev, "setup_ms_included": round(setup_ms, 3), **legs}
        log(f"[bench] multigrid: {mg}")

    # ---- several patches sharing faces: assembly into shared rows -------------
    multipatch = None
    if not args.no_multipatch:
        import types
        from ifem_amd import model as M
        mp = M.x_split_patches(2, [16, 16], 32, jitter=0.15, seed=5)
        mmsc = M.face_status(types.SimpleNamespace(n=mp.dims), 3, [7, 0, 0, 0, 0, 0])
        mmeqn, mneq = M.equation_numbers(mmsc)
        itg_e = pb.integrand
        ms = Session(stream=stream.cuda_stream)
        pids = [ms.add_patch(P) for P in mp.patches]
        ms.bind_patches(pids, mp.mlgn, 3, mmeqn, mneq)
        t0 = time.perf_counter()
        mnnz = ms.init_assembly()
        msym = (time.perf_counter() - t0) * 1e3
        mt = []
        for _ in range(4):
            ms.assemble(itg_e, pb.ng, FP64)
            mt.append(ms.timings()[0])
        sp_ms = float(np.median(mt[1:]))
        # one patch of 32 x 32 x 32 elements for comparison
        one = problems.elasticity_cube(32, 2)
        os_ = Session(stream=stream.cuda_stream)
        opid = os_.add_patch(one.patch)
        os_.bind(opid, 3, one.meqn, one.neq)
        os_.init_assembly()
        ot = []
        for _ in range(4):
            os_.assemble(one.integrand, one.ng, FP64)
            ot.append(os_.timings()[0])
        multipatch = {
            "workload": "elasticity p=2, two 16x32x32-element patches sharing the x=0.5 face, "
                        "clamped x=0",
            "dofs": int(len(mmeqn)), "neq": int(mneq), "nnz": int(mnnz),
            "pattern_ms": round(msym, 1), "pattern_built_on": "device", "assembly_kernels_ms": round(sp_ms, 3),
            "assembled_dof_per_s": round(len(mmeqn) / (sp_ms * 1e-3), 1),
            "one_patch_32cubed_assembly_kernels_ms": round(float(np.median(ot[1:])), 3),
        }
        ms.close()
        os_.close()
        log(f"[bench] multipatch: {multipatch}")

    # ---- learned preconditioner: 48^3 diffusion and 64^3 elasticity -------------
    learned_leg = None
    if not args.no_learned:
        learned_leg = learned_legs(args, stream, s, pb, x64, solve64_ms)
        log(f"[bench] learned: {learned_leg}")

    # ---- largest single-GPU case: 96^3 p=3 elasticity, 4x4x4 Gauss points -----
    large = None if args.no_large else large_leg(args, stream)
    log(f"[bench] large: {large}")

    # ---- batched uneven patches: one persistent assembly launch ----------------
    batch = None if args.no_batch else batch_leg(args)
    log(f"[bench] batch: {batch}")

    # ---- CPU baseline: the oracle on a bounded sample ----------------------------
    cpu = None
    if not args.no_cpu:
        import oracle
        rp, ci, A, b = s.system(values=True)
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        threads = max(1, min(16, avail or 1))
        layers = min(args.cpu_layers, pb.patch.nel[2])
        # warm-up on one layer, then the median of 3 timed runs
        oracle.assemble(pb.patch, pb.integrand, pb.ng, pb.meqn, neq, rp, ci, nthreads=threads,
                        e3_range=(0, 1))
        t_runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.assemble(pb.patch, pb.integrand, pb.ng, pb.meqn, neq, rp, ci,
                            nthreads=threads, e3_range=(0, layers))
            t_runs.append(time.perf_counter() - t0)
        t_asm = float(np.median(t_runs))
        frac = layers / pb.patch.nel[2]
        cpu_dofs = ndof / (t_asm / frac)
        # one thread on a smaller sample (2 element layers)
        l1 = min(2, pb.patch.nel[2])
        t0 = time.perf_counter()
        oracle.assemble(pb.patch, pb.integrand, pb.ng, pb.meqn, neq, rp, ci, nthreads=1,
                        e3_range=(0, l1))
        t_1 = time.perf_counter() - t0
        cpu_dofs_1 = ndof / (t_1 / (l1 / pb.patch.nel[2]))
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        its = 10
        t_its = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.pcg_fixed(rp, ci, A, b, its)
            t_its.append((time.perf_counter() - t0) / its)
        t_it = float(np.median(t_its))
        cpu = {"value": round(cpu_dofs, 1), "unit": "assembled DOF/s", "cores": threads,
               "kind": "port",
               "sample": (f"oracle assembly of element layers w<{layers} of the same patch "
                          f"({layers * pb.patch.nel[0] * pb.patch.nel[1]} of {nel_tot} elements; "
                          f"median of 3 runs after a warm-up: {t_asm:.2f} s, runs "
                          f"{', '.join(f'{t:.2f}' for t in t_runs)} s) scaled to the whole patch; "
                          f"Jacobi-PCG {its} iterations on the full FP64 system, median of 3"),
               "pcg_ms_per_iter": round(t_it * 1e3, 2),
               "time_to_tol_est_s": round(t_it * iters64, 1),
               "value_1thread": round(cpu_dofs_1, 1),
               "sample_1thread": f"element layers w<{l1}, one thread, {t_1:.1f} s",
               "host_cpus_available": avail, "cpu_model": model}
        del A
        log(f"[bench] cpu baseline: {cpu}")

    # ---- roofline of the dominant assembly kernel -------------------------------
    flops = (36 * nen * nen + 6 * nen) * ngp * nel_tot
    asm_bytes = nnz * 8 + neq * 8 + pb.patch.nnod * 3 * 8
    achieved = flops / (lines_ms * 1e-3)
    traffic = measured_traffic("ifem::k_assemble_lines<2, 3, 3, double, 1, false, 0>") \
        if args.nel == 64 else None
    # compute-bound: FP64 FMAs on the vector ALUs (MI355X's FP64 vector and
    # FP64 matrix peaks are the same 78.6 TFLOP/s)
    roofline = {"kernel": "k_assemble_lines<2,3,3,double,1,false,0>", "bound": "mfma",
                "engine": "valu_fp64",
                "achieved": round(achieved / 1e12, 3), "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64, 4), "traffic": traffic,
                "algorithmic_flops": flops, "algorithmic_bytes": asm_bytes,
                "kernel_ms": round(lines_ms, 3),
                "whole_assembly": {"ms": round(kern_ms, 3), "geometry_pass_ms": round(geo_ms, 3),
                                   "max_roofline_frac": round(max(flops / PEAK_FP64,
                                                                  asm_bytes / PEAK_HBM)
                                                              / (kern_ms * 1e-3), 4)}}
    spmv_roof = {"kernel": "k_spmv_node3<double,double,false,false,double>", "bound": "hbm",
                 "achieved": round(spmv_bytes / (spmv_ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                 "unit": "GB/s", "frac": round(spmv_bytes / (spmv_ms * 1e-3) / PEAK_HBM, 4),
                 "traffic": (measured_traffic("ifem::k_spmv_node3<double, double, false, false, double>")
                             if args.nel == 64 else None), "bytes": spmv_bytes, "kernel_ms": round(spmv_ms, 4)}

    value = ndof / (asm_ms * 1e-3)
    out = {
        "metric": "assembled DOF/s; time-to-1e-8 solve, learned+FP32 vs plain FP64 PCG, 1 GPU",
        "value": round(value, 1), "unit": "assembled DOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded jittered cube patch, splitmix64)",
        "config": {"workload": f"elasticity {args.nel}^3 p=2 jittered cube, 3x3x3 Gauss, "
                               "clamped x=0, traction (0,0,-1) on x=1",
                   "dofs": ndof, "neq": neq, "nnz": nnz, "elements": nel_tot,
                   "parallelism": "single GPU"},
        "assembly_ms": round(asm_ms, 3), "assembly_kernels_ms": round(kern_ms, 3),
        "symbolic_ms": round(symbolic_ms, 1),
        "fp64_time_to_tol_ms": round(solve64_ms, 2), "fp64_iterations": iters64,
        "mixed": mixed,
        "speedup_mixed_vs_fp64": (round(solve64_ms / mixed["solve_ms"], 3) if mixed else None),
        "multigrid": mg,
        "speedup_mg_vs_fp64": (round(solve64_ms / min(v["solve_ms"] for k, v in mg.items()
                                                      if isinstance(v, dict)), 3)
                               if mg else None),
        "learned_fp32_time_to_tol_ms": (learned_leg["elasticity64"]["learned_fp32_solve_ms"]
                                        if learned_leg else None),
        "speedup_learned_fp32_vs_fp64": (learned_leg["elasticity64"]["speedup_vs_fp64_jacobi"]
                                         if learned_leg else None),
        "roofline": roofline, "roofline_spmv": spmv_roof, "cpu_baseline": cpu,
        "batch": batch,
        "multipatch": multipatch,
        "large": large,
        "learned": learned_leg,
    }
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
