"""Writes the golden fixtures used by the tests.

The reference repository ships no fixtures and cannot run here (its modules import
cupy, which is absent), so the fixtures come from the complex128 oracle and are pinned
independently: the cfg1 polarisation is cross-checked against a dense direct solve of
the explicitly assembled DDA matrix on the occupied sites.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import dda_oracle as O  # noqa: E402


def cfg1():
    pb, nd = O.sphere_problem(8, 1.5 + 0.01j, 2.0)
    grid, k = pb["grid_size"], pb["k"]
    ia, mask, G, M = O.prepare_interaction_matrices(grid, k, 1.0, pb["alpha_array"])
    it, P, lo, fi = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 2000, 1e-8)
    D, rows, _ = O.dense_matrix(grid, k, 1.0, ia, occupied_only=True)
    b = pb["E_inc"].reshape(-1, 3)[rows].ravel()
    Pd = np.linalg.solve(D, b)
    rel = np.linalg.norm(P.reshape(-1, 3)[rows].ravel() - Pd) / np.linalg.norm(Pd)
    cext, cabs = O.cross_sections(P, pb["E_inc"], ia, k)
    np.savez_compressed(os.path.join(HERE, "cfg1_sphere16.npz"), P=P, iters=it, lowest=lo, first=fi,
                        c_ext=cext, c_abs=cabs, k=k, n_dipoles=nd, mask=mask, ghat_corner=G[:3, :3, :3],
                        dense_vs_krylov_rel=rel)
    print("cfg1", grid, nd, "iters", it, "C_ext", cext, "C_abs", cabs, "dense rel", rel)


if __name__ == "__main__":
    cfg1()
