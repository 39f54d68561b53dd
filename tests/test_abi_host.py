"""CPU tests: the C ABI library loads and exports every declared entry point; host-side
logic (shape bookkeeping, polarisability, chunk plan, direction sets) matches the
reference rules. No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pdda.h")
LIB = os.path.join(ROOT, "precondda_amd", "libpdda.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdda_\w+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libpdda.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from precondda_amd import _lib
    assert set(names) == set(_lib.EXPORTED_SYMBOLS), set(names) ^ set(_lib.EXPORTED_SYMBOLS)
    L = _lib.load_library()
    assert L.pdda_abi_version() == 1
    assert L.pdda_fft_length_supported(96) == 1 and L.pdda_fft_length_supported(58) == 0


@pytest.mark.skipif(not os.path.exists(LIB), reason="libpdda.so not built")
def test_abi_rejects_bad_arguments_without_gpu():
    from precondda_amd import _lib
    L = _lib.load_library()
    h = ctypes.c_void_p()
    # unsupported FFT length (2*29) and bad precision are caught before any device call
    rc = L.pdda_conv_create(ctypes.byref(h), 5, 4, 4, 4, 4, 4, 4, 1)
    assert rc == -1 and b"prec" in L.pdda_last_error()
    rc = L.pdda_conv_create(ctypes.byref(h), 0, 29, 4, 4, 29, 4, 4, 1)
    assert rc == -2 and b"FFT length 58" in L.pdda_last_error()
    assert L.pdda_solver_iterate(None, 1, None) == -1


def test_calculate_alpha_host_matches_oracle():
    from oracle import dda_oracle as O
    from precondda_amd.targets import calculate_alpha
    for m in (1.5 + 0.01j, 1.33 + 0.001j, 1.7 + 0.1j, 2.0):
        for meth in ("CM", "RR", "LDR"):
            a = calculate_alpha(m, 1.0, 0.31, [0, 0, 1], [1, 0, 0], meth)
            b = O.calculate_alpha(m, 1.0, 0.31, [0, 0, 1], [1, 0, 0], meth)
            assert abs(a - b) <= 1e-15 * abs(b)
    with pytest.raises(ValueError):
        calculate_alpha(1.5, 1.0, 0.3)


def test_shape_manager_rules():
    from precondda_amd.targets import ShapeManager
    sm = ShapeManager()
    a = sm.add_shape("sphere", center=[7.2, 8.0, 8.5], radius=8, n=1.5)
    b = sm.add_shape("rectangle", center=[3.0, 3.0, 3.0], dimensions=[2, 2, 2], n=1.5)
    assert (a, b) == ("sphere_1", "rectangle_1")
    sm.process_shapes(0.3, [0, 0, 1], [1, 0, 0], 1.0)
    assert sm.shapes[a].center == [8.0, 8.0, 9.0]
    assert sm.shapes[a].value == sm.shapes[b].value
    assert sm.calculate_grid_extents(1.0) == (17, 17, 18)
    with pytest.raises(ValueError):
        sm.add_shape("torus", center=[0, 0, 0], n=1.2)


def test_chunk_plan_matches_reference_rule():
    from precondda_amd.operators import setup_chunked_values
    for nx, (nch, cs) in {16: (1, 32), 32: (2, 32), 64: (4, 32), 256: (16, 32), 24: (1, 48), 48: (2, 48),
                          96: (4, 48)}.items():
        plan = setup_chunked_values(nx, 8, 8)
        assert len(plan) == nch and plan[0]["chunk_size"] == cs


def test_batch_directions_are_orthonormal_triads():
    from precondda_amd.workloads import batch_directions
    d, p = batch_directions(64, 0)
    assert d.shape == (128, 3)
    assert np.allclose(np.linalg.norm(d, axis=1), 1) and np.allclose(np.linalg.norm(p, axis=1), 1)
    assert np.abs(np.sum(d * p, axis=1)).max() < 1e-12
    # the two polarisations of one direction are orthogonal
    assert np.abs(np.sum(p[0::2] * p[1::2], axis=1)).max() < 1e-12
    v = np.random.default_rng(0).standard_normal((64, 3))
    assert np.allclose(d[0::2], v / np.linalg.norm(v, axis=1, keepdims=True))


def test_excited_field_size_mismatch_is_value_error():
    """E_field_solver rejects a polarisation of the wrong size with ValueError and the
    reference's message (ExcitedField.py:141-152), before any device work."""
    from precondda_amd import operators
    with pytest.raises(ValueError, match="^Polarization size 10 doesn't match expected size 24$"):
        operators.E_field_solver(np.zeros(10, complex), None, (2, 2, 2))
