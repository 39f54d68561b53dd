"""Results files in the reference's layout: text sections and the polarisation array."""
import numpy as np

from precondda_amd.results import convergence_status, load_polarization, save_results_to_file


def _problem():
    return {"grid_size": (16, 16, 16), "lattice_spacing": 1.0, "wavelength": 2 * np.pi / 0.25,
            "k_direction": [0, 0, 1], "E_polarization": [1, 0, 0], "k": 0.25}


def test_save_and_reload(tmp_path):
    P = (np.arange(16 * 16 * 16 * 3) * (1 + 0.5j)).reshape(16, 16, 16, 3).astype(np.complex128)
    shapes = [{"type": "sphere", "center": [8, 8, 8], "radius": 8, "n": 1.5 + 0.01j},
              {"type": "rectangle", "center": [4, 4, 4], "dimensions": [2, 2, 2], "n": 2.0}]
    fn = str(tmp_path / "run.txt")
    npy = save_results_to_file(fn, _problem(), [1.0, -2.0], shapes, (24, P, 3.2e-9), extra={"C_ext": 363.86})
    assert npy == str(tmp_path / "run_polarization.npy")
    txt = open(fn).read().splitlines()
    assert txt[0] == "DDA Problem Analysis Results" and txt[1] == "=" * 50
    order = [txt.index(h) for h in ("PROBLEM PARAMETERS:", "SHAPES:", "AI PREDICTION:", "SOLVER RESULTS:")]
    assert order == sorted(order)
    assert "Grid size: (16, 16, 16)" in txt and "k (wave number): 0.250000" in txt
    assert "Shape 1: sphere" in txt and "  Radius: 8" in txt and "  Dimensions: [2, 2, 2]" in txt
    assert "Optimal multiplier: 1.000000 + -2.000000j" in txt
    assert "Iterations: 24" in txt and "Final norm: 3.200000e-09" in txt and "Status: Converged" in txt
    assert "C_ext: 363.86" in txt
    assert "Polarization array shape: (16, 16, 16, 3)" in txt
    Q = load_polarization(npy)
    assert Q.dtype == np.complex128 and np.array_equal(Q, P)


def test_status_and_missing_parts(tmp_path):
    assert convergence_status(10000) == "Did not converge (max iterations reached)"
    assert convergence_status(-5) == "Failed (NaN or error condition)"
    assert convergence_status(0) == "Unknown"
    fn = str(tmp_path / "a.txt")
    assert save_results_to_file(fn, _problem(), [np.nan, np.nan], [], None) is None
    txt = open(fn).read()
    assert "AI PREDICTION: Failed" in txt and "SOLVER RESULTS: Not run" in txt
