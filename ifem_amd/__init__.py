"""ifem_amd -- MI355X-native isogeometric assembly + PCG hot path of IFEM.

Package layout:
  csrc/        HIP kernels for gfx950 and the C ABI (libifem_gpu.so)
  facade/      C++ classes with the reference's interface names
               (ASMs3D::integrate, SAM, SystemMatrix, LinSolParams) on top
               of the C ABI
  _native.py   ctypes binding of the C ABI (Session)
  model.py     patch / SAM / integrand inputs (model generator)
  problems.py  the benchmark problems as model inputs
  build.py     hipcc / g++ build recipe
"""
from . import model, problems
from ._native import FP32, FP64, PC_JACOBI, PC_MG, GpuError, Session, lib

__all__ = ["model", "problems", "Session", "GpuError", "FP32", "FP64", "PC_JACOBI", "PC_MG", "lib",
           "setup_problem", "setup_batch"]


def setup_problem(pb, session=None, precision=FP64):
    """Upload a Problem, build the pattern and assemble K, b (and the traction)."""
    s = session or Session()
    pid = s.add_patch(pb.patch)
    if pb.sam is not None:
        s.bind_constraints(pid, pb.ncomp, pb.sam)
    else:
        s.bind(pid, pb.ncomp, pb.meqn, pb.neq)
    s.init_assembly()
    s.assemble(pb.integrand, pb.ng, FP64)
    if pb.traction is not None:
        face, t, ng = pb.traction
        s.add_traction(face, t, ng)
    if precision == FP32:
        s.assemble(pb.integrand, pb.ng, FP32)
    return s


def setup_batch(problems, session=None, precisions=(FP64,)):
    """Upload independent scalar Problems as one batch, build the global
    pattern and assemble (one persistent launch per precision)."""
    s = session or Session()
    for pb in problems:
        pid = s.add_patch(pb.patch)
        s.batch_add(pid, pb.meqn, pb.neq, pb.integrand, pb.ng[0])
    s.batch_init()
    for prec in precisions:
        s.batch_assemble(prec)
    return s
