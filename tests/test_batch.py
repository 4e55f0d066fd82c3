"""Batch of independent patches (one persistent assembly launch, per-patch
PCG from a captured graph) against the oracle, patch by patch.

Tolerances as for the single-patch path: FP64 blocks 1e-12 relative to
max|.|, FP32 blocks 1e-5, FP64 iterates 1e-12, solutions 1e-12 relative to the
oracle's PCG solution and to a direct solve; mixed precision: true residual <= rtol and energy-norm
error vs the FP64 solution <= 1e-6 per patch.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from ifem_amd import FP32, FP64, problems, setup_batch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _small_batch(seed=5, npatch=7, nel=(3, 6)):
    # degrees 1..3, 3..6 elements per direction: every block is small
    return problems.diffusion_batch(npatch, nel, (1, 2, 3), sigma=1.0, ell=0.2, seed=seed,
                                    nmodes=32)


def _oracle(pb):
    return oracle.system(pb.patch, pb.integrand, pb.ng, pb.meqn, pb.neq, None)


def test_batch_assembly_matches_oracle():
    pbs = _small_batch()
    _say(f"[gpu] batch assembly parity, {len(pbs)} patches, p={[p.patch.p[0] for p in pbs]}")
    with setup_batch(pbs, precisions=(FP64, FP32)) as s:
        for k, pb in enumerate(pbs):
            rp, ci, A, b = _oracle(pb)
            grp, gci, gA, gA32, gb = s.batch_system(k, values32=True)
            assert np.array_equal(grp, rp), k
            assert np.array_equal(gci, ci), k
            assert np.max(np.abs(gA - A)) <= 1e-12 * np.max(np.abs(A)), k
            assert np.max(np.abs(gb - b)) <= 1e-12 * np.max(np.abs(b)), k
            assert np.max(np.abs(gA32 - A)) <= 1e-5 * np.max(np.abs(A)), k


def test_batch_iterates_match_oracle_across_graph_replays():
    """20 Jacobi-PCG iterations from x = 0 with a host check every 4
    iterations (5 replays of the captured loop): every patch's iterate equals
    the oracle's 20th iterate."""
    pbs = _small_batch(seed=6, nel=(5, 8))
    _say("[gpu] batch iterate parity over graph replays")
    with setup_batch(pbs) as s:
        st = s.batch_solve(rtol=1e-30, maxits=20, check_every=4)
        assert st["rc"] == 4
        for k, pb in enumerate(pbs):
            rp, ci, A, _, b = s.batch_system(k)
            x = s.batch_solution(k)
            it = int(st["patch_iterations"][k])
            assert it == 20, (k, it)
            xo = oracle.pcg_fixed(rp, ci, A, b, it)
            assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo)), k


def test_batch_solution_matches_oracle_and_is_reproducible():
    pbs = _small_batch(seed=7)
    _say("[gpu] batch FP64 solve vs oracle")
    with setup_batch(pbs) as s:
        st = s.batch_solve(rtol=1e-14, maxits=5000, check_every=8)
        xs = [s.batch_solution(k) for k in range(len(pbs))]
        st2 = s.batch_solve(rtol=1e-14, maxits=5000, check_every=8)
        xs2 = [s.batch_solution(k) for k in range(len(pbs))]
        assert st["converged"] and st2["converged"]
        assert np.array_equal(st["patch_iterations"], st2["patch_iterations"])
        for k, pb in enumerate(pbs):
            assert np.array_equal(xs[k], xs2[k]), k
            rp, ci, A, _, b = s.batch_system(k)
            xo, it, _ = oracle.pcg(rp, ci, A, b, rtol=1e-14)
            assert abs(int(st["patch_iterations"][k]) - it) <= 2, (k, it)
            # at a relative residual of 1e-14 both solutions lie within 1e-13
            # of the exact one for these patches (direct solve below)
            assert np.max(np.abs(xs[k] - xo)) <= 1e-12 * np.max(np.abs(xo)), k
            r = b - oracle.spmv(rp, ci, A, xs[k])
            assert np.linalg.norm(r) <= 1e-14 * np.linalg.norm(b) * 1.01, k
            K = sp.csr_matrix((A, ci, rp), shape=(len(b), len(b))).tocsc()
            xd = spla.spsolve(K, b)
            assert np.max(np.abs(xs[k] - xd)) <= 1e-12 * np.max(np.abs(xd)), k


def test_batch_mixed_precision_and_autotune():
    pbs = _small_batch(seed=8)
    _say("[gpu] batch mixed precision + autotune")
    with setup_batch(pbs, precisions=(FP64, FP32)) as s:
        table = s.batch_autotune(FP32)
        for p in (1, 2, 3):
            assert table[p]["half"] in (0, 1) and table[p]["lanes"] in (16, 32, 64)
        # the re-assembly with the tuned variants keeps parity
        for k, pb in enumerate(pbs):
            rp, ci, A, b = _oracle(pb)
            _, _, gA, gA32, gb = s.batch_system(k, values32=True)
            assert np.max(np.abs(gA - A)) <= 1e-12 * np.max(np.abs(A)), k
            assert np.max(np.abs(gA32 - A)) <= 1e-5 * np.max(np.abs(A)), k
        st64 = s.batch_solve(rtol=1e-11, maxits=5000)
        x64 = [s.batch_solution(k) for k in range(len(pbs))]
        st = s.batch_solve(rtol=1e-8, maxits=5000, precision=FP32, inner_rtol=1e-4)
        assert st64["converged"] and st["converged"]
        assert np.all(st["patch_rel_residual"] <= 1e-8)
        for k, pb in enumerate(pbs):
            rp, ci, A, _, b = s.batch_system(k)
            x = s.batch_solution(k)
            K = sp.csr_matrix((A, ci, rp), shape=(pb.neq, pb.neq))
            e = x - x64[k]
            en = np.sqrt(e @ (K @ e)) / np.sqrt(x64[k] @ (K @ x64[k]))
            assert en <= 1e-6, (k, en)
            assert np.linalg.norm(b - K @ x) <= 1e-8 * np.linalg.norm(b) * 1.01, k


def test_batch_errors():
    from ifem_amd import GpuError, Session
    from ifem_amd import model as M
    pb = problems.poisson_cube(4, 2)
    with Session() as s:
        pid = s.add_patch(pb.patch)
        with pytest.raises(GpuError):
            s.batch_init()  # empty batch
        with pytest.raises(GpuError):
            s.batch_add(pid, pb.meqn, pb.neq, pb.integrand, 4)  # ng != p+1
        with pytest.raises(GpuError):
            s.batch_add(pid, pb.meqn, pb.neq, M.Integrand(M.ELASTICITY), 3)
        s.batch_add(pid, pb.meqn, pb.neq, pb.integrand, 3)
        with pytest.raises(GpuError):
            s.batch_solve()  # before init
        s.batch_init()
        with pytest.raises(GpuError):
            s.batch_solve()  # before assembly
        s.batch_assemble(FP64)
        st = s.batch_solve(rtol=1e-10)
        assert st["converged"]


def test_batch_full_size_properties():
    """The 256-patch benchmark batch (p in {2,3}, 8..48 elements per direction,
    about 5.6M equations, 1.1e9 nonzeros): closed-form nonzero counts per patch,
    the smallest block and a mid-size p = 3 block against the oracle, every
    block's FP32 values against its FP64 values, FP64 and mixed solves to 1e-8
    on every patch's true residual."""
    pbs = problems.diffusion_batch()
    _say("[gpu] full-size batch 256 patches")

    def nnz_closed(pb):
        q = pb.patch.p[0]
        r = 1
        for d in range(3):
            m = pb.patch.n[d] - 2
            r *= sum(min(i + q, m - 1) - max(i - q, 0) + 1 for i in range(m))
        return r

    with setup_batch(pbs, precisions=(FP64, FP32)) as s:
        for k, pb in enumerate(pbs):
            neq, nnz, _ = s.batch_patch_info(k)
            assert neq == pb.neq and nnz == nnz_closed(pb), k
        k = int(np.argmin([pb.neq for pb in pbs]))
        rp, ci, A, b = _oracle(pbs[k])
        _, _, gA, _, gb = s.batch_system(k)
        assert np.max(np.abs(gA - A)) <= 1e-12 * np.max(np.abs(A))
        assert np.max(np.abs(gb - b)) <= 1e-12 * np.max(np.abs(b))
        # a mid-size p = 3 block against the oracle (full bundles, many work
        # items per workgroup), and every block's FP32 values against its FP64
        # ones (the two precisions run different kernels)
        mid = [(pb.neq, q) for q, pb in enumerate(pbs)
               if pb.patch.p[0] == 3 and min(np.array(pb.patch.n) - 3) >= 12]
        k = min(mid)[1]
        rp, ci, A, b = _oracle(pbs[k])
        _, _, gA, gA32, gb = s.batch_system(k, values32=True)
        assert np.max(np.abs(gA - A)) <= 1e-12 * np.max(np.abs(A)), k
        assert np.max(np.abs(gb - b)) <= 1e-12 * np.max(np.abs(b)), k
        for q in range(len(pbs)):
            _, _, gA, gA32, _ = s.batch_system(q, values32=True)
            assert np.max(np.abs(gA32 - gA)) <= 1e-5 * np.max(np.abs(gA)), q
        st = s.batch_solve(rtol=1e-8, maxits=20000)
        assert st["converged"] and np.all(st["patch_rel_residual"] <= 1e-8)
        _say(f"  fp64: max iterations {st['iterations']}, {st['solve_ms']:.1f} ms")
        st = s.batch_solve(rtol=1e-8, maxits=20000, precision=FP32)
        assert st["converged"] and np.all(st["patch_rel_residual"] <= 1e-8)
        _say(f"  mixed: max iterations {st['iterations']}, outer {st['outer']}, "
             f"{st['solve_ms']:.1f} ms")
