"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and fails cleanly (error code + message) without a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "ifem_gpu.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(ifem_gpu_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    from ifem_amd import _native
    assert sorted(_native.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    from ifem_amd import _native
    L = C.CDLL(_native.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name


def test_abi_version_and_defaults():
    from ifem_amd import _native
    L = _native.lib()
    assert L.ifem_gpu_abi_version() == 2
    p = _native.SolverParams()
    L.ifem_gpu_solver_defaults(C.byref(p))
    # LinSolParams defaults (src/LinAlg/LinSolParams.C:70-76)
    assert p.rtol == 1e-6 and p.atol == 1e-20 and p.dtol == 1e6 and p.maxits == 1000
    assert p.precision == 64


def test_null_arguments_are_rejected():
    from ifem_amd import _native
    L = _native.lib()
    assert L.ifem_gpu_ctx_create(0, None, None) == 1
    assert b"null" in L.ifem_gpu_last_error()
    assert L.ifem_gpu_system_info(None, None, None, None) == 1


def test_no_device_is_an_error_not_a_fallback():
    """Without a GPU the context cannot be created: EDEVICE, no CPU path."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from ifem_amd import _native
    L = _native.lib()
    h = C.c_void_p()
    rc = L.ifem_gpu_ctx_create(0, None, C.byref(h))
    assert rc == 3 and not h.value
    assert len(L.ifem_gpu_last_error()) > 0


def test_product_does_not_import_oracle():
    """The product package never imports or loads the oracle."""
    pkg = os.path.join(ROOT, "ifem_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".hpp", ".C", ".h", ".cpp")):
                with open(os.path.join(dirpath, fn)) as f:
                    txt = f.read()
                assert "import oracle" not in txt and "liboracle" not in txt, fn
