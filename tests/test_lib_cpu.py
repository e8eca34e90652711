"""CPU-side checks of libhexat.so: the library loads, exports every symbol of
include/hexat.h, and its host-side element matrices match the oracle.  No
kernel is launched here."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported(hx):
    L = hx._lib.load()
    declared = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            hdr = open(os.path.join(ROOT, "include", h)).read()
            declared |= set(re.findall(r"^(?:int|void|const char\*)\s+(hx_\w+)\(", hdr, flags=re.M))
    assert declared == set(hx._lib.EXPORTS) | set(hx._lib.TEST_EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_version_and_device_count_without_gpu(hx):
    L = hx._lib.load()
    assert L.hx_version() >= 10
    assert hx.device_count() >= 0


@pytest.mark.parametrize("dims,nu", [((1, 1, 1), 0.3), ((1.0, 2.0, 0.5), 0.3), ((0.6186, 0.625, 0.625), 0.3),
                                     ((0.3, 0.7, 1.1), 0.45)])
def test_host_element_stiffness_matches_oracle(hx, ox, dims, nu):
    K = hx.element_stiffness(*dims, nu, E=1.7)
    Kx = ox.ke_hex(*dims, nu, E=1.7)
    assert np.allclose(K, Kx, rtol=1e-13, atol=1e-14 * np.abs(Kx).max())


@pytest.mark.parametrize("quirk", [0, 1])
def test_host_scalar_matrices_match_oracle(hx, ox, quirk):
    Ks, Ms = hx.scalar_matrices(0.9, 1.1, 0.7, quirk)
    Kx, Mx = ox.scalar_ke_me(0.9, 1.1, 0.7, quirk)
    assert np.allclose(Ks, Kx, rtol=1e-13, atol=1e-15)
    assert np.allclose(Ms, Mx, rtol=1e-13, atol=1e-15)


def test_bad_arguments_rejected(hx):
    with pytest.raises(hx.HexatError):
        hx.element_stiffness(-1.0, 1.0, 1.0, 0.3)
    with pytest.raises(hx.HexatError):
        hx.element_stiffness(1.0, 1.0, 1.0, 0.5)


def test_context_without_gpu_fails_loudly(hx):
    if hx.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(hx.HexatError) as ei:
        hx.Context(0)
    assert ei.value.code == hx._lib.HX_ERR_NO_DEVICE


def test_solver_config_preconditioner_mapping(hx):
    """solver_parameters.preconditioner selects the multilevel or V-cycle
    form; "amg" and "jacobi_amg" (the reference's CPU AMG, not built) run
    Jacobi; anything else is rejected before any device work."""
    import importlib
    solver = importlib.import_module(hx.__name__ + ".solver")
    L = hx._lib
    assert solver._precond_of(None) == (L.HX_PC_JACOBI, "Jacobi")
    assert solver._precond_of({"solver_parameters": {"preconditioner": "Multilevel"}}) == (L.HX_PC_MULTILEVEL,
                                                                                           "Multilevel")
    assert solver._precond_of({"solver_parameters": {"preconditioner": "vcycle"}}) == (L.HX_PC_VCYCLE, "V-cycle")
    assert solver._precond_of({"solver_parameters": {"preconditioner": "jacobi_amg"}}) == (L.HX_PC_JACOBI, "Jacobi")
    assert solver._precond_of({"solver_parameters": {"preconditioner": "amg"}}) == (L.HX_PC_JACOBI, "Jacobi")
    with pytest.raises(ValueError):
        solver._precond_of({"solver_parameters": {"preconditioner": "ilu"}})
