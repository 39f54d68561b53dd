"""Pins the CPU oracle against closed forms, brute force, independent formulations
and the committed golden fixtures (CPU only)."""
import itertools
import math
import os

import numpy as np
import pytest

from oracle import dda_oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ polarisability
def test_alpha_cm_closed_form():
    for m in (1.5 + 0.01j, 1.33 + 0.001j, 2.0 + 0.1j):
        eps = m * m
        assert abs(O.calculate_alpha(m, 1.0, 0.3, method="CM") - 3 / (4 * np.pi) * (eps - 1) / (eps + 2)) < 1e-15
        # d^3 scaling of CM
        assert abs(O.calculate_alpha(m, 2.0, 0.3, method="CM") / O.calculate_alpha(m, 1.0, 0.3, method="CM") - 8) < 1e-13


@pytest.mark.parametrize("method", ["RR", "LDR"])
def test_alpha_real_m_radiative_reaction(method):
    """For real m, Im(1/alpha) = -(2/3) k^3 exactly (optical theorem for one dipole)."""
    for m, k, d in [(1.5, 0.25, 1.0), (1.33, 0.9, 0.5), (2.0, 0.1, 2.0)]:
        a = O.calculate_alpha(m, d, k, [0, 0, 1], [1, 0, 0], method=method)
        assert abs((1 / a).imag + (2.0 / 3.0) * k ** 3) < 1e-12 * max(1.0, abs(1 / a))


def test_alpha_small_k_limit_and_transverse_S():
    m = 1.6 + 0.05j
    cm = O.calculate_alpha(m, 1.0, 0.0, method="CM")
    assert abs(O.calculate_alpha(m, 1.0, 1e-9, [0, 0, 1], [1, 0, 0]) - cm) < 1e-14
    # reference S = (a.e)^2 vanishes for any transverse pair: alpha independent of direction
    a1 = O.calculate_alpha(m, 1.0, 0.3, [0, 0, 1], [1, 0, 0])
    a2 = O.calculate_alpha(m, 1.0, 0.3, [0.6, 0.0, 0.8], [0.0, 1.0, 0.0])
    assert abs(a1 - a2) < 1e-15
    # a longitudinal pair (S = 1) changes the LDR correction by b3 m^2 (kd)^2
    a3 = O.calculate_alpha(m, 1.0, 0.3, [0, 0, 1], [0, 0, 1])
    lhs = 1 / a3 - 1 / a1
    assert abs(lhs - O.LDR_B3 * m * m * 0.09) < 1e-12
    with pytest.raises(ValueError):
        O.calculate_alpha(m, 1.0, 0.3, method="XX")


# ------------------------------------------------------------------ lattice
def test_dipole_counts_of_benchmark_targets():
    _, n1 = O.sphere_problem(8, 1.5 + 0.01j, 2.0)
    assert n1 == 2176
    _, n2 = O.sphere_problem(32, 1.33 + 0.001j, 10.0)
    assert n2 == 137376
    sh = O.snap_centers([{"type": "ellipsoid", "center": [24, 24, 48], "semi_axes": [24, 24, 48], "n": 1.6}], 1.0)
    g = O.trim_zero_faces(O.mark_shapes(O.grid_extents(sh, 1.0), 1.0, sh, [1.0]))
    assert g.shape == (48, 48, 96) and np.count_nonzero(g) == 115776
    occ = np.random.default_rng(0).random((256, 256, 128)) < 0.7
    assert int(occ.sum()) == 5873665


def test_marking_brute_force():
    shapes = [{"type": "sphere", "center": [3.2, 2.7, 4.1], "radius": 2.6, "n": 1.5},
              {"type": "rectangle", "center": [5.0, 5.0, 5.0], "dimensions": [3.0, 2.0, 4.0], "n": 1.7},
              {"type": "ellipsoid", "center": [4.0, 4.0, 3.0], "semi_axes": [1.5, 3.0, 2.0], "n": 2.0}]
    sh = O.snap_centers(shapes, 1.0)
    grid = O.grid_extents(sh, 1.0)
    vals = [1.0, 2.0, 3.0]
    got = O.mark_shapes(grid, 1.0, sh, vals)
    exp = np.zeros(grid, dtype=complex)
    for (i, j, l) in itertools.product(*(range(n) for n in grid)):
        p = (i + 0.5, j + 0.5, l + 0.5)
        for s, v in zip(sh, vals):
            c = s["center"]
            if s["type"] == "sphere":
                r = math.dist(p, c)
                inside = r < s["radius"] or abs(r - s["radius"]) < 1e-7
            elif s["type"] == "rectangle":
                inside = all(abs(p[a] - c[a]) <= s["dimensions"][a] / 2 + 1e-12 for a in range(3))
            else:
                inside = math.sqrt(sum(((p[a] - c[a]) / s["semi_axes"][a]) ** 2 for a in range(3))) <= 1 + 1e-7
            if inside:
                exp[i, j, l] = v
    assert np.array_equal(got, exp)


def test_trim_pads_odd_dims():
    a = np.zeros((9, 9, 9), complex)
    a[2:5, 3:7, 1:4] = 1.0
    t = O.trim_zero_faces(a)
    assert t.shape == (4, 4, 4)
    assert np.count_nonzero(t) == 3 * 4 * 3 and np.all(t[3] == 0) and np.all(t[:, :, 3] == 0)


# ------------------------------------------------------------------ Green's tensor
def _scalar_green(r_vec, k):
    r = np.linalg.norm(r_vec)
    return np.exp(1j * k * r) / r


def test_green_tensor_equals_dyadic_finite_differences():
    """A_ref = -(k^2 I + grad grad) e^{ikr}/r, checked by central differences."""
    k, h = 0.7, 1e-3
    nx, ny, nz = 4, 5, 6
    A = O.greens_real(nx, ny, nz, k, 1.0)
    for idx in [(1, 0, 0), (1, 2, 3), (6, 1, 2), (3, 9, 11), (2, 2, 2)]:
        pos = np.array([idx[0] - (2 * nx if idx[0] >= nx else 0), idx[1] - (2 * ny if idx[1] >= ny else 0),
                        idx[2] - (2 * nz if idx[2] >= nz else 0)], dtype=float)
        H = np.zeros((3, 3), complex)
        for a in range(3):
            for b in range(3):
                ea, eb = np.eye(3)[a] * h, np.eye(3)[b] * h
                H[a, b] = (_scalar_green(pos + ea + eb, k) - _scalar_green(pos + ea - eb, k)
                           - _scalar_green(pos - ea + eb, k) + _scalar_green(pos - ea - eb, k)) / (4 * h * h)
        G = k * k * _scalar_green(pos, k) * np.eye(3) + H
        for c, (a, b) in enumerate(O.COMP_AB):
            assert abs(A[idx + (c,)] + G[a, b]) < 2e-5 * abs(G).max()
    # zeros: origin and the i == n planes
    assert np.all(A[0, 0, 0] == 0) and np.all(A[nx] == 0) and np.all(A[:, ny] == 0) and np.all(A[:, :, nz] == 0)


def test_reduced_table_mirror_rule():
    G = O.generate_interaction_row(5, 4, 6, 0.4, 1.0)
    F = O.generate_interaction_row(5, 4, 6, 0.4, 1.0, reduced=False)
    T = O.expand_reduced(G)
    for c, (a, b) in enumerate(O.COMP_AB):
        assert np.abs(T[..., a, b] - F[..., c]).max() < 1e-13 * np.abs(F).max()


def test_fft_matches_brute_force_dft():
    rng = np.random.default_rng(0)
    for n in (6, 10, 14):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        W = np.exp(-2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
        assert np.abs(O._fft(x) - W @ x).max() < 1e-12


# ------------------------------------------------------------------ matvec and solve
def _small(R=3, m=1.5 + 0.1j, x=1.0):
    pb, _ = O.sphere_problem(R, m, x)
    grid, k = pb["grid_size"], pb["k"]
    ia, mask, G, M = O.prepare_interaction_matrices(grid, k, 1.0, pb["alpha_array"])
    return pb, grid, k, ia, mask, G, M


def test_fft_matvec_equals_dense_matrix():
    pb, grid, k, ia, mask, G, M = _small(3)
    D, _, _ = O.dense_matrix(grid, k, 1.0, ia)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(D.shape[0]) + 1j * rng.standard_normal(D.shape[0])
    v = O.interaction_mult(x, G, np.repeat(ia.ravel(), 3), mask, grid)
    assert np.linalg.norm(v - D @ x) / np.linalg.norm(D @ x) < 1e-13


def test_solver_matches_dense_solve_and_quirks():
    pb, grid, k, ia, mask, G, M = _small(3)
    trace = []
    it, P, lo, fi = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 500, 1e-12, trace=trace)
    assert it > 0 and lo < 1e-12 and fi > 0
    D, rows, _ = O.dense_matrix(grid, k, 1.0, ia, occupied_only=True)
    b = pb["E_inc"].reshape(-1, 3)[rows].ravel()
    Pd = np.linalg.solve(D, b)
    assert np.linalg.norm(P.reshape(-1, 3)[rows].ravel() - Pd) / np.linalg.norm(Pd) < 1e-10
    assert np.all(P[~mask] == 0)
    # preconditioner on/off reach the same solution
    it2, P2, _, _ = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 500, 1e-12, precon=False)
    assert np.linalg.norm(P2 - P) / np.linalg.norm(P) < 1e-10


def test_solver_failure_codes():
    pb, grid, k, ia, mask, G, M = _small(3)
    # max_iter too small -> returns max_iter and the best iterate
    it, P, lo, fi = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 2, 1e-12)
    assert it == 2 and lo < 1
    # NaN in E_inc -> -(i+1) at the first iteration
    E = pb["E_inc"].copy()
    E[mask][0, 0] = np.nan
    E2 = E.copy()
    idx = np.argwhere(mask)[0]
    E2[tuple(idx) + (0,)] = np.nan
    it, P, lo, fi = O.solve_gpbicgstab(grid, ia, G, M, E2, mask, 100, 50, 1e-12)
    assert it == -1
    # tiny stagnation window
    it, P, lo, fi = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 0, 200, 0.0, stagnation=1)
    assert it < 0


def test_cross_sections_energy_conservation():
    for m in (1.5 + 0.05j, 1.5 + 0.0j):
        pb, grid, k, ia, mask, G, M = _small(4, m, 1.5)
        it, P, _, _ = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 1000, 1e-13)
        cext, cabs = O.cross_sections(P, pb["E_inc"], ia, k)
        csca = O.c_sca_farfield(P, k)
        assert abs(cext - cabs - csca) < 1e-8 * cext
        if m.imag == 0:
            assert abs(cabs) < 1e-10 * cext
        else:
            assert cabs > 0


def test_excited_field_invariant():
    pb, grid, k, ia, mask, G, M = _small(3)
    it, P, _, _ = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 500, 1e-13)
    Es = O.E_field_solver(P, G, grid)
    lhs = (pb["E_inc"] + Es)[mask]
    rhs = (ia[..., None] * P)[mask]
    assert np.linalg.norm(lhs - rhs) / np.linalg.norm(rhs) < 1e-11
    with pytest.raises(ValueError):
        O.E_field_solver(P.ravel()[:-3], G, grid)


def test_mixed_refinement_reaches_fp64_tolerance():
    pb, grid, k, ia, mask, G, M = _small(4, 1.33 + 0.001j, 2.0)
    P, total, outer, rel = O.solve_mixed(grid, ia, G, M, pb["E_inc"], mask, tol=1e-9, inner_target=1e-3)
    assert rel < 1e-9 and outer >= 2
    assert O.true_residual(P, pb["E_inc"], ia, mask, G, grid) < 1e-9


# ------------------------------------------------------------------ golden fixtures
def test_golden_cfg1_fixture():
    path = os.path.join(GOLDEN, "cfg1_sphere16.npz")
    z = np.load(path, allow_pickle=False)
    pb, nd = O.sphere_problem(8, 1.5 + 0.01j, 2.0)
    assert nd == int(z["n_dipoles"])
    assert np.array_equal(pb["alpha_array"] != 0, z["mask"])
    grid, k = pb["grid_size"], pb["k"]
    assert abs(k - float(z["k"])) < 1e-15
    # the dense direct solve recorded in the fixture pins P independent of Krylov
    assert float(z["dense_vs_krylov_rel"]) < 1e-7  # Krylov stopped at relative residual 1e-8
    ia, mask, G, M = O.prepare_interaction_matrices(grid, k, 1.0, pb["alpha_array"])
    assert np.abs(G[:3, :3, :3] - z["ghat_corner"]).max() < 1e-13 * np.abs(z["ghat_corner"]).max()
    cext, cabs = O.cross_sections(z["P"], pb["E_inc"], ia, k)
    assert abs(cext - float(z["c_ext"])) < 1e-12 * abs(cext)
    assert abs(cabs - float(z["c_abs"])) < 1e-12 * abs(cabs)
