"""ctypes binding of libifem_gpu.so (the C ABI in include/ifem_gpu.h).

The library is loaded from the package directory only; if it is missing or a
call fails, an exception is raised -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libifem_gpu.so")

# every symbol declared in include/ifem_gpu.h
EXPORTS = [
    "ifem_gpu_last_error", "ifem_gpu_abi_version", "ifem_gpu_ctx_create", "ifem_gpu_ctx_destroy",
    "ifem_gpu_ctx_stream", "ifem_gpu_synchronize", "ifem_gpu_patch_create", "ifem_gpu_sam_bind",
    "ifem_gpu_init_assembly", "ifem_gpu_assemble", "ifem_gpu_assemble_neumann",
    "ifem_gpu_solver_defaults", "ifem_gpu_solve", "ifem_gpu_system_info", "ifem_gpu_get_system",
    "ifem_gpu_get_values_fp32", "ifem_gpu_get_solution", "ifem_gpu_expand_solution",
    "ifem_gpu_spmv", "ifem_gpu_spmv_noalloc", "ifem_gpu_spmv_device", "ifem_gpu_set_rhs", "ifem_gpu_device_pointers",
    "ifem_gpu_last_timings", "ifem_gpu_unet_create", "ifem_gpu_unet_get_params",
    "ifem_gpu_unet_set_params", "ifem_gpu_unet_forward", "ifem_gpu_unet_gradient",
    "ifem_gpu_unet_train_step", "ifem_gpu_set_learned_preconditioner",
    "ifem_gpu_batch_clear", "ifem_gpu_batch_add", "ifem_gpu_batch_init", "ifem_gpu_batch_assemble",
    "ifem_gpu_batch_autotune", "ifem_gpu_batch_solve", "ifem_gpu_batch_results",
    "ifem_gpu_batch_patch_info", "ifem_gpu_batch_get_system", "ifem_gpu_batch_get_solution",
    "ifem_gpu_batch_timings", "ifem_gpu_autotune_assembly", "ifem_gpu_unet_backward",
    "ifem_gpu_last_assembly_parts", "ifem_gpu_integrate_norms", "ifem_gpu_sam_bind_constraints",
    "ifem_gpu_mg_setup", "ifem_gpu_mg_level", "ifem_gpu_mg_apply", "ifem_gpu_sam_bind_patches",
    "ifem_gpu_assemble_neumann_patch", "ifem_gpu_set_assembly_variant",
]

FP64, FP32 = 64, 32
PC_JACOBI, PC_MG = 0, 1


class GpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ifem_gpu error {code}: {msg}")
        self.code = code


class Function(C.Structure):
    _fields_ = [("type", C.c_int), ("nmodes", C.c_int), ("scale", C.c_double),
                ("mean", C.c_double), ("modes", C.POINTER(C.c_double))]


class Integrand(C.Structure):
    _fields_ = [("kind", C.c_int), ("E", C.c_double), ("nu", C.c_double),
                ("kappa", Function), ("source", Function), ("body", C.c_double * 3)]


class SolverParams(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("dtol", C.c_double),
                ("maxits", C.c_int), ("precision", C.c_int), ("inner_rtol", C.c_double),
                ("check_every", C.c_int), ("verbosity", C.c_int), ("pc", C.c_int),
                ("mg_smooth", C.c_int)]


class UNetConfig(C.Structure):
    _fields_ = [("levels", C.c_int), ("channels", C.c_int * 4), ("in_channels", C.c_int),
                ("out_channels", C.c_int), ("seed", C.c_uint64), ("activation", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_iterations", C.c_int), ("converged", C.c_int),
                ("rel_residual", C.c_double), ("solve_ms", C.c_double),
                ("pc_iterations", C.c_int), ("fallback", C.c_int), ("pc_rejected", C.c_int)]


_lib = None


def lib():
    """Load libifem_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(python -m ifem_amd.build); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.c_int, C.c_double
    L.ifem_gpu_last_error.restype = C.c_char_p
    L.ifem_gpu_ctx_create.argtypes = [ip, vp, C.POINTER(vp)]
    L.ifem_gpu_ctx_destroy.argtypes = [vp]
    L.ifem_gpu_ctx_stream.argtypes = [vp]
    L.ifem_gpu_ctx_stream.restype = vp
    L.ifem_gpu_synchronize.argtypes = [vp]
    L.ifem_gpu_patch_create.argtypes = [vp, C.POINTER(ip), C.POINTER(ip),
                                        C.POINTER(C.POINTER(dp)), C.POINTER(dp), ip,
                                        C.POINTER(ip)]
    L.ifem_gpu_sam_bind.argtypes = [vp, ip, ip, ip, C.POINTER(ip), ip]
    L.ifem_gpu_sam_bind_constraints.argtypes = [vp, ip, ip, ip, C.POINTER(ip), ip, ip,
                                                C.POINTER(ip), C.POINTER(ip), C.POINTER(dp)]
    L.ifem_gpu_mg_setup.argtypes = [vp, ip, C.POINTER(ip), C.POINTER(dp)]
    L.ifem_gpu_mg_level.argtypes = [vp, ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(dp), vp]
    L.ifem_gpu_mg_apply.argtypes = [vp, vp, vp]
    L.ifem_gpu_sam_bind_patches.argtypes = [vp, ip, C.POINTER(ip), C.POINTER(C.POINTER(ip)), ip,
                                            ip, C.POINTER(ip), ip, ip, C.POINTER(ip),
                                            C.POINTER(ip), C.POINTER(dp)]
    L.ifem_gpu_assemble_neumann_patch.argtypes = [vp, ip, ip, C.POINTER(dp), ip]
    L.ifem_gpu_init_assembly.argtypes = [vp, C.POINTER(C.c_int64)]
    L.ifem_gpu_assemble.argtypes = [vp, C.POINTER(Integrand), C.POINTER(ip), ip]
    L.ifem_gpu_assemble_neumann.argtypes = [vp, ip, C.POINTER(dp), ip]
    L.ifem_gpu_autotune_assembly.argtypes = [vp, C.POINTER(Integrand), C.POINTER(ip), ip,
                                             C.POINTER(ip)]
    L.ifem_gpu_set_assembly_variant.argtypes = [vp, ip, ip, ip, ip, ip]
    L.ifem_gpu_solver_defaults.argtypes = [C.POINTER(SolverParams)]
    L.ifem_gpu_solver_defaults.restype = None
    L.ifem_gpu_solve.argtypes = [vp, C.POINTER(SolverParams), C.POINTER(Stats)]
    L.ifem_gpu_system_info.argtypes = [vp, C.POINTER(ip), C.POINTER(C.c_int64), C.POINTER(ip)]
    L.ifem_gpu_get_system.argtypes = [vp, vp, vp, vp, vp]
    L.ifem_gpu_get_values_fp32.argtypes = [vp, vp]
    L.ifem_gpu_get_solution.argtypes = [vp, vp]
    L.ifem_gpu_expand_solution.argtypes = [vp, vp]
    L.ifem_gpu_spmv.argtypes = [vp, vp, vp]
    L.ifem_gpu_spmv_noalloc.argtypes = [vp, vp, vp]
    L.ifem_gpu_set_rhs.argtypes = [vp, vp]
    L.ifem_gpu_spmv_device.argtypes = [vp, ip, vp, vp]
    L.ifem_gpu_device_pointers.argtypes = [vp] + [C.POINTER(vp)] * 5
    L.ifem_gpu_last_timings.argtypes = [vp, C.POINTER(dp), C.POINTER(dp)]
    L.ifem_gpu_last_assembly_parts.argtypes = [vp, C.POINTER(dp), C.POINTER(dp)]
    L.ifem_gpu_integrate_norms.argtypes = [vp, C.POINTER(Integrand), C.POINTER(ip),
                                           C.POINTER(Function), C.POINTER(dp)]
    L.ifem_gpu_unet_create.argtypes = [vp, C.POINTER(UNetConfig), C.POINTER(C.c_int64)]
    L.ifem_gpu_unet_get_params.argtypes = [vp, vp]
    L.ifem_gpu_unet_set_params.argtypes = [vp, vp]
    L.ifem_gpu_unet_forward.argtypes = [vp, vp, vp, ip, ip, ip, ip]
    L.ifem_gpu_unet_gradient.argtypes = [vp, vp, vp, vp, ip, ip, ip, ip, vp, C.POINTER(C.c_float)]
    L.ifem_gpu_unet_train_step.argtypes = [vp, vp, vp, vp, ip, ip, ip, ip, C.c_float,
                                           C.POINTER(C.c_float)]
    L.ifem_gpu_set_learned_preconditioner.argtypes = [vp, ip, ip, ip, ip, vp, ip, ip, dp]
    L.ifem_gpu_unet_backward.argtypes = [vp, vp, vp, ip, ip, ip, ip, C.c_float, vp]
    L.ifem_gpu_batch_clear.argtypes = [vp]
    L.ifem_gpu_batch_add.argtypes = [vp, ip, C.POINTER(ip), ip, C.POINTER(Integrand), ip,
                                     C.POINTER(ip)]
    L.ifem_gpu_batch_init.argtypes = [vp, C.POINTER(ip), C.POINTER(C.c_int64)]
    L.ifem_gpu_batch_assemble.argtypes = [vp, ip]
    L.ifem_gpu_batch_autotune.argtypes = [vp, ip, C.POINTER(ip)]
    L.ifem_gpu_batch_solve.argtypes = [vp, C.POINTER(SolverParams), C.POINTER(Stats)]
    L.ifem_gpu_batch_results.argtypes = [vp, vp, vp, vp]
    L.ifem_gpu_batch_patch_info.argtypes = [vp, ip, C.POINTER(ip), C.POINTER(C.c_int64),
                                            C.POINTER(ip)]
    L.ifem_gpu_batch_get_system.argtypes = [vp, ip, vp, vp, vp, vp, vp]
    L.ifem_gpu_batch_get_solution.argtypes = [vp, ip, vp]
    L.ifem_gpu_batch_timings.argtypes = [vp, C.POINTER(dp), C.POINTER(dp), C.POINTER(dp)]
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        raise GpuError(rc, lib().ifem_gpu_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Session:
    """One context holding one patch and its assembled system (AlgEqSystem analogue)."""

    def __init__(self, device: int = 0, stream=None):
        L = lib()
        self._keep = []
        h = C.c_void_p()
        check(L.ifem_gpu_ctx_create(device, C.c_void_p(stream) if stream else None, C.byref(h)))
        self.h = h
        self.neq = 0
        self.ncomp = 0
        self.nnz = 0

    def close(self):
        if self.h:
            lib().ifem_gpu_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return lib().ifem_gpu_ctx_stream(self.h) or 0

    def synchronize(self):
        check(lib().ifem_gpu_synchronize(self.h))

    def add_patch(self, patch) -> int:
        order = (C.c_int * 3)(*[q + 1 for q in patch.p])
        ncoef = (C.c_int * 3)(*patch.n)
        ks = [np.ascontiguousarray(k, dtype=np.float64) for k in patch.knots]
        kp = (C.POINTER(C.c_double) * 3)(*[k.ctypes.data_as(C.POINTER(C.c_double)) for k in ks])
        co = np.ascontiguousarray(patch.coefs.reshape(-1), dtype=np.float64)
        pid = C.c_int(-1)
        check(lib().ifem_gpu_patch_create(self.h, order, ncoef, kp,
                                          co.ctypes.data_as(C.POINTER(C.c_double)), 0,
                                          C.byref(pid)))
        self._mg_p = patch.p[0]
        return pid.value

    def bind(self, patch_id: int, ncomp: int, meqn: np.ndarray, neq: int):
        m = np.ascontiguousarray(meqn, dtype=np.int32)
        check(lib().ifem_gpu_sam_bind(self.h, patch_id, ncomp, len(m),
                                      m.ctypes.data_as(C.POINTER(C.c_int)), neq))
        self.neq = neq
        self.ncomp = ncomp

    def bind_constraints(self, patch_id: int, ncomp: int, sam):
        """Binds SAM arrays with constraint equations: ``sam`` has meqn, neq,
        mpmceq, mmceq, ttcc (1-based, SAM layout; prescribed values and
        constraints u_s = c0 + sum c_m u_m with free masters)."""
        m = np.ascontiguousarray(sam.meqn, dtype=np.int32)
        mp = np.ascontiguousarray(sam.mpmceq, dtype=np.int32)
        mm = np.ascontiguousarray(sam.mmceq, dtype=np.int32)
        tt = np.ascontiguousarray(sam.ttcc, dtype=np.float64)
        nceq = len(mp) - 1
        ipp = C.POINTER(C.c_int)
        check(lib().ifem_gpu_sam_bind_constraints(
            self.h, patch_id, ncomp, len(m), m.ctypes.data_as(ipp), sam.neq, nceq,
            mp.ctypes.data_as(ipp), mm.ctypes.data_as(ipp),
            tt.ctypes.data_as(C.POINTER(C.c_double))))
        self.neq = sam.neq
        self.ncomp = ncomp

    def bind_patches(self, patch_ids, mlgns, ncomp: int, meqn, neq: int, sam=None):
        """Several patches sharing nodes: mlgns[k] maps patch k's local nodes
        to global nodes (1-based); meqn/neq per global DOF, or sam (SAM arrays
        with prescribed values)."""
        ipp = C.POINTER(C.c_int)
        ids = (C.c_int * len(patch_ids))(*patch_ids)
        keep = [np.ascontiguousarray(m, dtype=np.int32) for m in mlgns]
        arr = (ipp * len(keep))(*[k.ctypes.data_as(ipp) for k in keep])
        if sam is not None:
            meqn, neq = sam.meqn, sam.neq
        m = np.ascontiguousarray(meqn, dtype=np.int32)
        nnod = len(m) // ncomp
        if sam is not None:
            mp = np.ascontiguousarray(sam.mpmceq, dtype=np.int32)
            mm = np.ascontiguousarray(sam.mmceq, dtype=np.int32)
            tt = np.ascontiguousarray(sam.ttcc, dtype=np.float64)
            args = (len(mp) - 1, mp.ctypes.data_as(ipp), mm.ctypes.data_as(ipp),
                    tt.ctypes.data_as(C.POINTER(C.c_double)))
        else:
            args = (0, None, None, None)
        check(lib().ifem_gpu_sam_bind_patches(self.h, len(patch_ids), ids, arr, nnod, ncomp,
                                              m.ctypes.data_as(ipp), neq, *args))
        self.neq = neq
        self.ncomp = ncomp

    def add_traction_patch(self, index: int, face: int, t, ng: int):
        """Traction on face 0..5 of patch `index` of a multi-patch system."""
        tv = (C.c_double * 3)(*[float(x) for x in t] + [0.0] * (3 - len(t)))
        check(lib().ifem_gpu_assemble_neumann_patch(self.h, index, face + 1, tv, ng))

    def init_assembly(self) -> int:
        nnz = C.c_int64(0)
        check(lib().ifem_gpu_init_assembly(self.h, C.byref(nnz)))
        self.nnz = nnz.value
        return self.nnz

    @staticmethod
    def _function(f, keep) -> Function:
        s = Function()
        s.type = int(f.type)
        s.scale = float(f.scale)
        s.mean = float(f.mean)
        s.nmodes = f.nmodes
        if f.nmodes:
            m = np.ascontiguousarray(f.modes.reshape(-1), dtype=np.float64)
            keep.append(m)
            s.modes = m.ctypes.data_as(C.POINTER(C.c_double))
        return s

    def _integrand(self, itg, keep) -> Integrand:
        s = Integrand()
        s.kind = int(itg.kind)
        s.E = float(itg.E)
        s.nu = float(itg.nu)
        s.kappa = self._function(itg.kappa, keep)
        s.source = self._function(itg.source, keep)
        for i in range(3):
            s.body[i] = float(itg.body[i])
        return s

    def assemble(self, itg, ng, precision: int = FP64):
        keep = []
        s = Integrand()
        s.kind = int(itg.kind)
        s.E = float(itg.E)
        s.nu = float(itg.nu)
        s.kappa = self._function(itg.kappa, keep)
        s.source = self._function(itg.source, keep)
        for i in range(3):
            s.body[i] = float(itg.body[i])
        ngv = (C.c_int * 3)(*ng)
        check(lib().ifem_gpu_assemble(self.h, C.byref(s), ngv, precision))

    def autotune_assembly(self, itg, ng, precision: int = FP64) -> int:
        """Pick the fastest element-chunk variant for this system; returns 0
        for full chunks, 1 for half chunks, 2 for full chunks in two
        Gauss-point slabs.  K and b are left as after assemble()."""
        keep = []
        s = self._integrand(itg, keep)
        ngv = (C.c_int * 3)(*ng)
        h = C.c_int(-1)
        check(lib().ifem_gpu_autotune_assembly(self.h, C.byref(s), ngv, precision, C.byref(h)))
        return h.value

    def set_assembly_variant(self, p, ng, ncomp, precision, variant):
        """Fix the element-chunk variant for (p, ng, ncomp, precision); -1
        restores the built-in default."""
        check(lib().ifem_gpu_set_assembly_variant(self.h, p, ng, ncomp, precision, variant))

    def norms(self, itg, ng, exact=None):
        """[||u_h||_L2, ||u_h||_a, ||u||_L2, ||u||_a, ||u-u_h||_L2, ||u-u_h||_a] of the
        current solution (exact: optional analytic scalar solution)."""
        keep = []
        s = self._integrand(itg, keep)
        ngv = (C.c_int * 3)(*ng)
        out = (C.c_double * 6)()
        ex = self._function(exact, keep) if exact is not None else None
        check(lib().ifem_gpu_integrate_norms(self.h, C.byref(s), ngv,
                                             C.byref(ex) if ex is not None else None, out))
        return np.array(list(out))

    def add_traction(self, face: int, t, ng: int):
        """face: 0..5 (u0,u1,v0,v1,w0,w1) -> IFEM face index face+1."""
        tv = (C.c_double * 3)(*[float(x) for x in t] + [0.0] * (3 - len(t)))
        check(lib().ifem_gpu_assemble_neumann(self.h, face + 1, tv, ng))

    def solve(self, rtol=1e-8, atol=1e-300, maxits=100000, precision=FP64, inner_rtol=1e-4,
              check_every=32, dtol=1e30, pc=PC_JACOBI, mg_smooth=2):
        p = SolverParams()
        lib().ifem_gpu_solver_defaults(C.byref(p))
        p.rtol, p.atol, p.dtol, p.maxits = rtol, atol, dtol, maxits
        p.precision, p.inner_rtol, p.check_every = precision, inner_rtol, check_every
        p.pc, p.mg_smooth = pc, mg_smooth
        st = Stats()
        rc = lib().ifem_gpu_solve(self.h, C.byref(p), C.byref(st))
        if rc not in (0, 4):
            check(rc)
        return {"iterations": st.iterations, "outer": st.outer_iterations,
                "converged": bool(st.converged), "rel_residual": st.rel_residual,
                "solve_ms": st.solve_ms, "rc": rc, "pc_iterations": st.pc_iterations,
                "fallback": bool(st.fallback), "pc_rejected": st.pc_rejected}

    def system(self, values=True):
        rowptr = np.zeros(self.neq + 1, dtype=np.int64)
        colind = np.zeros(max(self.nnz, 1), dtype=np.int32)
        vals = np.zeros(max(self.nnz, 1), dtype=np.float64) if values else None
        rhs = np.zeros(max(self.neq, 1), dtype=np.float64)
        check(lib().ifem_gpu_get_system(self.h, _ptr(rowptr), _ptr(colind),
                                        _ptr(vals) if values else None, _ptr(rhs)))
        return rowptr, colind[:self.nnz], (vals[:self.nnz] if values else None), rhs[:self.neq]

    def values_fp32(self):
        v = np.zeros(max(self.nnz, 1), dtype=np.float32)
        check(lib().ifem_gpu_get_values_fp32(self.h, _ptr(v)))
        return v[:self.nnz]

    def solution(self):
        x = np.zeros(max(self.neq, 1))
        check(lib().ifem_gpu_get_solution(self.h, _ptr(x)))
        return x[:self.neq]

    def mg_setup(self, smooth=2):
        """Builds the multigrid hierarchy; returns (levels, setup ms)."""
        nl, ms = C.c_int(0), C.c_double(0.0)
        check(lib().ifem_gpu_mg_setup(self.h, smooth, C.byref(nl), C.byref(ms)))
        return nl.value, ms.value

    def mg_level(self, level: int, operator=False):
        """(n[3], free DOFs, lambda_max, stencil or None) of a multigrid level."""
        n = (C.c_int * 3)()
        nf, lam = C.c_int(0), C.c_double(0.0)
        check(lib().ifem_gpu_mg_level(self.h, level, n, C.byref(nf), C.byref(lam), None))
        S = None
        if operator:
            nnode = n[0] * n[1] * n[2]
            p = self._mg_p
            S = np.zeros((2 * p + 1) ** 3 * self.ncomp ** 2 * nnode)
            check(lib().ifem_gpu_mg_level(self.h, level, n, C.byref(nf), C.byref(lam), _ptr(S)))
        return list(n), nf.value, lam.value, S

    def mg_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.neq)
        check(lib().ifem_gpu_mg_apply(self.h, _ptr(r), _ptr(z)))
        return z

    def expanded_solution(self, ndof: int):
        """Solution in DOF order: eliminated DOFs 0 or their prescribed value."""
        x = np.zeros(max(ndof, 1))
        check(lib().ifem_gpu_expand_solution(self.h, _ptr(x)))
        return x[:ndof]

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.neq)
        check(lib().ifem_gpu_spmv(self.h, _ptr(x), _ptr(y)))
        return y

    def spmv_device(self, x_ptr: int, y_ptr: int, precision: int = FP64):
        """y = K x on device pointers (no synchronisation)."""
        check(lib().ifem_gpu_spmv_device(self.h, precision, C.c_void_p(x_ptr), C.c_void_p(y_ptr)))

    def rhs(self):
        """The device right-hand side b (equation order)."""
        b = np.zeros(max(self.neq, 1))
        check(lib().ifem_gpu_get_system(self.h, None, None, None, _ptr(b)))
        return b[:self.neq]

    def set_rhs(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        check(lib().ifem_gpu_set_rhs(self.h, _ptr(b)))

    # ---- learned preconditioner ------------------------------------------------
    def unet_create(self, levels, channels, in_channels, out_channels, seed=1, linear=False):
        cfg = UNetConfig()
        cfg.levels = levels
        for i, c in enumerate(channels):
            cfg.channels[i] = int(c)
        cfg.in_channels, cfg.out_channels, cfg.seed = in_channels, out_channels, seed
        cfg.activation = 1 if linear else 0
        n = C.c_int64(0)
        check(lib().ifem_gpu_unet_create(self.h, C.byref(cfg), C.byref(n)))
        self.unet_cfg = (levels, list(channels)[:levels], in_channels, out_channels)
        self.unet_nparams = n.value
        return n.value

    def unet_params(self):
        p = np.zeros(self.unet_nparams, dtype=np.float32)
        check(lib().ifem_gpu_unet_get_params(self.h, _ptr(p)))
        return p

    def unet_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.unet_nparams
        check(lib().ifem_gpu_unet_set_params(self.h, _ptr(p)))

    def unet_forward(self, x_ptr: int, y_ptr: int, N, D, H, W):
        check(lib().ifem_gpu_unet_forward(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), N, D, H, W))

    def unet_gradient(self, x_ptr: int, t_ptr: int, N, D, H, W, w_ptr: int = 0):
        g = np.zeros(self.unet_nparams, dtype=np.float32)
        loss = C.c_float(0.0)
        check(lib().ifem_gpu_unet_gradient(self.h, C.c_void_p(x_ptr), C.c_void_p(t_ptr),
                                           C.c_void_p(w_ptr) if w_ptr else None, N, D, H, W,
                                           _ptr(g), C.byref(loss)))
        return loss.value, g

    def unet_train_step(self, x_ptr: int, t_ptr: int, N, D, H, W, lr, want_loss=False,
                        w_ptr: int = 0):
        loss = C.c_float(0.0)
        check(lib().ifem_gpu_unet_train_step(self.h, C.c_void_p(x_ptr), C.c_void_p(t_ptr),
                                             C.c_void_p(w_ptr) if w_ptr else None, N, D, H, W, lr,
                                             C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def unet_backward(self, x_ptr: int, dy_ptr: int, N, D, H, W, lr=0.0, want_grad=False):
        """Parameter gradient of sum(dy * U(x)); lr > 0 also takes an Adam step."""
        g = np.zeros(self.unet_nparams, dtype=np.float32) if want_grad else None
        check(lib().ifem_gpu_unet_backward(self.h, C.c_void_p(x_ptr), C.c_void_p(dy_ptr), N, D, H,
                                           W, float(lr), _ptr(g) if want_grad else None))
        return g

    def set_learned(self, enable, D=0, H=0, W=0, aux_ptr: int = 0, naux=0, scale=1.0,
                    features=0):
        check(lib().ifem_gpu_set_learned_preconditioner(self.h, int(enable), D, H, W,
                                                        C.c_void_p(aux_ptr) if aux_ptr else None,
                                                        naux, features, scale))

    def timings(self):
        a = C.c_double(0.0); s = C.c_double(0.0)
        check(lib().ifem_gpu_last_timings(self.h, C.byref(a), C.byref(s)))
        return a.value, s.value

    def assembly_parts(self):
        """(geometry pass ms, line kernel ms) of the last assembly."""
        g = C.c_double(0.0); l = C.c_double(0.0)
        check(lib().ifem_gpu_last_assembly_parts(self.h, C.byref(g), C.byref(l)))
        return g.value, l.value

    # ---- batch of independent patches -----------------------------------------
    def batch_clear(self):
        check(lib().ifem_gpu_batch_clear(self.h))
        self.batch_size = 0

    def batch_add(self, patch_id: int, meqn: np.ndarray, neq: int, itg, ng: int) -> int:
        m = np.ascontiguousarray(meqn, dtype=np.int32)
        keep = []
        s = self._integrand(itg, keep)
        idx = C.c_int(-1)
        check(lib().ifem_gpu_batch_add(self.h, patch_id, m.ctypes.data_as(C.POINTER(C.c_int)), neq,
                                       C.byref(s), ng, C.byref(idx)))
        self.batch_size = idx.value + 1
        return idx.value

    def batch_init(self):
        n = C.c_int(0)
        nnz = C.c_int64(0)
        check(lib().ifem_gpu_batch_init(self.h, C.byref(n), C.byref(nnz)))
        return n.value, nnz.value

    def batch_assemble(self, precision: int = FP64):
        check(lib().ifem_gpu_batch_assemble(self.h, precision))

    def batch_autotune(self, precision: int = FP64):
        t = (C.c_int * 8)()
        check(lib().ifem_gpu_batch_autotune(self.h, precision, t))
        return {p: {"half": t[2 * p], "lanes": t[2 * p + 1]} for p in range(1, 4)}

    def batch_solve(self, rtol=1e-8, atol=1e-300, maxits=100000, precision=FP64, inner_rtol=1e-4,
                    check_every=32, dtol=1e30):
        p = SolverParams()
        lib().ifem_gpu_solver_defaults(C.byref(p))
        p.rtol, p.atol, p.dtol, p.maxits = rtol, atol, dtol, maxits
        p.precision, p.inner_rtol, p.check_every = precision, inner_rtol, check_every
        st = Stats()
        rc = lib().ifem_gpu_batch_solve(self.h, C.byref(p), C.byref(st))
        if rc not in (0, 4):
            check(rc)
        n = self.batch_size
        it = np.zeros(max(n, 1), dtype=np.int32)
        cv = np.zeros(max(n, 1), dtype=np.int32)
        rr = np.zeros(max(n, 1))
        check(lib().ifem_gpu_batch_results(self.h, _ptr(it), _ptr(cv), _ptr(rr)))
        return {"iterations": st.iterations, "outer": st.outer_iterations,
                "converged": bool(st.converged), "rel_residual": st.rel_residual,
                "solve_ms": st.solve_ms, "rc": rc, "patch_iterations": it[:n],
                "patch_converged": cv[:n].astype(bool), "patch_rel_residual": rr[:n]}

    def batch_patch_info(self, k: int):
        n = C.c_int(0); nnz = C.c_int64(0); e0 = C.c_int(0)
        check(lib().ifem_gpu_batch_patch_info(self.h, k, C.byref(n), C.byref(nnz), C.byref(e0)))
        return n.value, nnz.value, e0.value

    def batch_system(self, k: int, values=True, values32=False, rhs=True):
        neq, nnz, _ = self.batch_patch_info(k)
        rowptr = np.zeros(neq + 1, dtype=np.int64)
        colind = np.zeros(max(nnz, 1), dtype=np.int32)
        v = np.zeros(max(nnz, 1)) if values else None
        v32 = np.zeros(max(nnz, 1), dtype=np.float32) if values32 else None
        b = np.zeros(max(neq, 1)) if rhs else None
        check(lib().ifem_gpu_batch_get_system(self.h, k, _ptr(rowptr), _ptr(colind),
                                              _ptr(v) if values else None,
                                              _ptr(v32) if values32 else None,
                                              _ptr(b) if rhs else None))
        return (rowptr, colind[:nnz], v[:nnz] if values else None,
                v32[:nnz] if values32 else None, b[:neq] if rhs else None)

    def batch_solution(self, k: int):
        neq, _, _ = self.batch_patch_info(k)
        x = np.zeros(max(neq, 1))
        check(lib().ifem_gpu_batch_get_solution(self.h, k, _ptr(x)))
        return x[:neq]

    def batch_timings(self):
        a = C.c_double(0.0); b = C.c_double(0.0); s = C.c_double(0.0)
        check(lib().ifem_gpu_batch_timings(self.h, C.byref(a), C.byref(b), C.byref(s)))
        return {"symbolic_ms": a.value, "assemble_ms": b.value, "solve_ms": s.value}
