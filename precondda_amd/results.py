"""Results on disk in the reference's layout (ddaforgit.py:1653-1749): a text summary of
the problem, the shapes, the preconditioner multiplier and the solve, plus the full
polarisation array as `<name>_polarization.npy` beside it.

The sections, their order and the line formats follow the reference so that files
written here read the same as its files. Additions (cross sections, the FP64
residual) come as extra lines in the solver section.
"""
from __future__ import annotations

import datetime
import math
from typing import Optional, Sequence

import numpy as np


def _host(a):
    if a is None:
        return None
    if hasattr(a, "detach"):  # torch tensor (device or host)
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _shape_lines(i: int, shape: dict) -> list:
    out = [f"Shape {i}: {shape['type']}", f"  Center: {shape['center']}", f"  Refractive index: {shape['n']}"]
    t = shape["type"]
    if t in ("sphere",):
        out.append(f"  Radius: {shape['radius']}")
    elif t == "cylinder":
        out += [f"  Radius: {shape['radius']}", f"  Height: {shape['height']}", f"  Axis: {shape['axis']}"]
    elif t == "rectangle":
        out.append(f"  Dimensions: {shape['dimensions']}")
    elif t == "ellipsoid":
        out.append(f"  Semi-axes: {shape['semi_axes']}")
    elif t == "prism":
        out += [f"  Radius: {shape['radius']}", f"  Height: {shape['height']}", f"  Sides: {shape['sides']}",
                f"  Axis: {shape['axis']}"]
    return out


def convergence_status(iterations: Optional[int], max_iter: int = 10000) -> str:
    """The reference's reading of run_solver's iteration count (ddaforgit.py:1708-1716)."""
    if iterations is None:
        return "Unknown"
    if 0 < iterations < max_iter:
        return "Converged"
    if iterations == max_iter:
        return "Did not converge (max iterations reached)"
    if iterations < 0:
        return "Failed (NaN or error condition)"
    return "Unknown"


def save_results_to_file(filename: str, problem: dict, prediction, shapes: Sequence[dict], solver_result=None,
                         max_iter: int = 10000, extra: Optional[dict] = None) -> Optional[str]:
    """Writes the summary to `filename` and, when a polarisation array is given, the
    array to `filename` with `.txt` replaced by `_polarization.npy` (complex, the
    solver's (nx, ny, nz, 3) layout). `solver_result` is run_solver's (iterations, P,
    final_norm). `prediction` is the multiplier as (re, im) or NaNs. `extra` adds
    "key: value" lines to the solver section. Returns the .npy path (or None)."""
    lines = ["DDA Problem Analysis Results", "=" * 50, "", "PROBLEM PARAMETERS:",
             f"Grid size: {tuple(problem['grid_size'])}",
             f"Lattice spacing: {float(problem['lattice_spacing']):.6f}",
             f"Wavelength: {float(problem['wavelength']):.6f}",
             f"Wave direction: {problem['k_direction']}",
             f"Polarization: {problem['E_polarization']}",
             f"k (wave number): {float(problem['k']):.6f}", "", "SHAPES:"]
    for i, sh in enumerate(shapes, 1):
        lines += _shape_lines(i, sh) + [""]
    pred = np.asarray(prediction if prediction is not None else [np.nan, np.nan], dtype=np.float64)
    if not np.any(np.isnan(pred)):
        mult = complex(pred[0], pred[1])
        lines += ["AI PREDICTION:", f"Optimal multiplier: {mult.real:.6f} + {mult.imag:.6f}j",
                  f"Magnitude: {abs(mult):.6f}",
                  f"Phase: {math.atan2(mult.imag, mult.real):.6f} radians "
                  f"({math.degrees(math.atan2(mult.imag, mult.real)):.2f} degrees)", ""]
    else:
        lines += ["AI PREDICTION: Failed", ""]
    npy = None
    if solver_result is not None:
        iterations, pol, final_norm = solver_result
        lines += ["SOLVER RESULTS:", f"Iterations: {iterations}", f"Final norm: {float(final_norm):.6e}",
                  "First norm: Not available from run_solver",
                  f"Status: {convergence_status(iterations, max_iter)}"]
        for k, v in (extra or {}).items():
            lines.append(f"{k}: {v}")
        p = _host(pol)
        if p is not None:
            lines += [f"Polarization array shape: {p.shape}", f"Polarization array dtype: {p.dtype}"]
            npy = filename.replace(".txt", "_polarization.npy")
            np.save(npy, p)
            lines.append(f"Full polarization array saved to: {npy}")
        else:
            lines.append("Polarization array: Not available")
    else:
        lines.append("SOLVER RESULTS: Not run")
    lines += ["", f"Results saved on: {datetime.datetime.now().replace(microsecond=0).isoformat()}"]
    with open(filename, "w") as f:
        f.write("\n".join(lines) + "\n")
    return npy


def load_polarization(npy_path: str) -> np.ndarray:
    """Reads a saved polarisation array (no pickled objects are accepted)."""
    return np.load(npy_path, allow_pickle=False)
