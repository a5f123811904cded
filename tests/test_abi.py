"""The C-ABI library: loads, exports every entry point include/hfx.h declares,
and fails loudly (status + message, no fallback) when no GPU is present."""
import os
import re

import numpy as np
import pytest

from hfsolver_amd import engine as E
from hfsolver_amd.basis import Basis

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "hfx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hfx_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_abi():
    names = declared()
    for n in ["hfx_create", "hfx_fock_2e", "hfx_fock_2e_device", "hfx_grad_2e", "hfx_grad_1e", "hfx_one_electron",
              "hfx_set_geometry", "hfx_eri_shells", "hfx_last_error", "hfx_destroy", "hfx_default_opts"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    L = E.lib()
    for n in declared():
        assert hasattr(L, n), n
        assert n in E.SIGNATURES, n


def test_default_opts():
    import ctypes
    o = E.HfxOpts()
    assert E.lib().hfx_default_opts(ctypes.byref(o)) == 0
    assert o.schwarz_threshold == 1e-12 and o.fp32_threshold == 0.0 and o.device == 0 and o.kernel_mode == 0


def test_create_without_gpu_fails_loudly():
    b = Basis([8, 1, 1], [[0, 0, 0], [1.809, 0, 0], [-0.453549, 1.751221, 0]])
    try:
        eng = E.Engine(b)
    except E.HfxError as e:
        assert "device" in str(e) or "hip" in str(e).lower()
        return
    eng.close()
    pytest.skip("a HIP device is present; covered by the gpu tests")


def test_bad_basis_is_rejected_before_device_work():
    # incomplete Cartesian d shell: the ABI must reject, not guess
    b = Basis([8, 1, 1], [[0, 0, 0], [1.809, 0, 0], [-0.453549, 1.751221, 0]])
    b.l = b.l.copy()
    b.l[9] = 0  # first d component of O no longer x^2
    with pytest.raises(E.HfxError):
        E.Engine(b)


def test_cython_declarations_cover_the_abi():
    """c_hfx.pxd (the Cython layer, as hfsolver/c_hfsolver.pxd) declares every
    entry point of include/hfx.h."""
    src = open(os.path.join(ROOT, "hfsolver_amd", "c_hfx.pxd")).read()
    pxd = sorted(set(re.findall(r"\b(hfx_[a-z0-9_]+)\s*\(", src)))
    assert pxd == declared()


def test_cython_module_loads_and_fails_loudly_without_gpu():
    from hfsolver_amd import hfx
    names = ["fock_2e", "fock_2e_device", "one_electron", "grad_2e", "grad_2e_stats", "grad_1e", "set_geometry",
             "eri_shells", "int2", "shells", "pair_bounds", "set_density_screening", "screening", "close"]
    for n in names:
        assert hasattr(hfx.Engine, n) and hasattr(E.Engine, n), n
    b = Basis([8, 1, 1], [[0, 0, 0], [1.809, 0, 0], [-0.453549, 1.751221, 0]])
    try:
        eng = hfx.Engine(b)
    except hfx.HfxError as e:
        assert "device" in str(e)
        return
    eng.close()
    pytest.skip("a HIP device is present; covered by the gpu tests")


def test_empty_basis_is_rejected():
    """nbf = 0 (or a null array) fails hfx_create with a message, before any device work."""
    import ctypes
    L = E.lib()
    o = E.HfxOpts()
    L.hfx_default_opts(ctypes.byref(o))
    h = ctypes.c_void_p()
    z = np.zeros(1, dtype=np.int32)
    d = np.zeros(1)
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    rc = L.hfx_create(0, p(z), p(z), p(d), p(d), p(d), p(z), p(z), p(z), 1, p(d), p(d), 1, p(z), p(d),
                      ctypes.byref(o), ctypes.byref(h))
    assert rc != 0
    assert b"invalid arguments" in L.hfx_last_error(h)
    L.hfx_destroy(h)
