"""Device GPBiCGSTAB against the step-by-step complex128 oracle."""
import numpy as np
import pytest
import torch

from oracle import dda_oracle as O

pytestmark = pytest.mark.gpu


def _setup(R, m, x, precon=True):
    pb, _ = O.sphere_problem(R, m, x)
    grid, k = pb["grid_size"], pb["k"]
    ia, mask, G, M = O.prepare_interaction_matrices(grid, k, 1.0, pb["alpha_array"])
    return pb, grid, k, ia, mask, G, M


@pytest.mark.parametrize("p3", ["auto", "register", "stockham"])
@pytest.mark.parametrize("precon", [True, False])
def test_fp64_solve_matches_oracle(monkeypatch, precon, p3):
    """p3 = register: the x-convolution reads the x-planar copies the solver makes of its
    tables at bind time."""
    import precondda_amd as P
    if p3 != "auto":
        monkeypatch.setenv("PDDA_P3_KERNEL", p3)
    pb, grid, k, ia, mask, G, M = _setup(4, 1.5 + 0.01j, 1.0)
    ref = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 300, 1e-10, precon=precon)
    out = P.solve_gpbicgstab(grid, ia, torch.as_tensor(G).cuda(), torch.as_tensor(M).cuda(), pb["E_inc"], mask,
                             100, 300, 1e-10, precon=precon)
    print("iters", ref[0], out[0])
    assert out[0] == ref[0]
    Pd = out[1].cpu().numpy()
    err = np.linalg.norm(Pd - ref[1]) / np.linalg.norm(ref[1])
    print("P rel err", err)
    assert err < 1e-10
    assert abs(out[2] - ref[2]) <= 1e-6 * abs(ref[2]) + 1e-14
    ce, ca = O.cross_sections(ref[1], pb["E_inc"], ia, k)
    de, da = P.cross_sections(out[1], pb["E_inc"], ia, k)
    assert abs(de - ce) < 1e-10 * abs(ce) and abs(da - ca) < 1e-10 * abs(ca)


@pytest.mark.parametrize("precon", [True, False])
def test_fp64_solve_register_z_passes(monkeypatch, precon):
    """The register z-inverse in every epilogue the solver launches: interaction with
    the alpha and zeta/eta dot products, the residual with its norms (replacement at
    iteration 1) and without (reset recompute), and the preconditioner apply. Without
    the first vector update (s = r - alpha v, y = s_prev - s - alpha w) is computed
    by the register z-forward that first reads s: the preconditioner apply's, or the
    second matvec's without the preconditioner."""
    import precondda_amd as P
    monkeypatch.setenv("PDDA_Z_KERNEL", "register")
    pb, grid, k, ia, mask, G, M = _setup(4, 1.5 + 0.01j, 1.0)
    ref = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 300, 1e-10, precon=precon)
    out = P.solve_gpbicgstab(grid, ia, torch.as_tensor(G).cuda(), torch.as_tensor(M).cuda(), pb["E_inc"], mask,
                             100, 300, 1e-10, precon=precon)
    print("iters", ref[0], out[0])
    assert out[0] == ref[0]
    err = np.linalg.norm(out[1].cpu().numpy() - ref[1]) / np.linalg.norm(ref[1])
    print("P rel err", err)
    assert err < 1e-10


@pytest.mark.parametrize("precon", [True, False])
def test_fp64_solve_stockham_z_passes(monkeypatch, precon):
    """LDS-tiled z passes: the first vector update runs as its own kernel (k_upd1)."""
    import precondda_amd as P
    monkeypatch.setenv("PDDA_Z_KERNEL", "stockham")
    pb, grid, k, ia, mask, G, M = _setup(4, 1.5 + 0.01j, 1.0)
    ref = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 300, 1e-10, precon=precon)
    out = P.solve_gpbicgstab(grid, ia, torch.as_tensor(G).cuda(), torch.as_tensor(M).cuda(), pb["E_inc"], mask,
                             100, 300, 1e-10, precon=precon)
    print("iters", ref[0], out[0])
    assert out[0] == ref[0]
    err = np.linalg.norm(out[1].cpu().numpy() - ref[1]) / np.linalg.norm(ref[1])
    print("P rel err", err)
    assert err < 1e-10


def test_fp64_solve_cfg1_sphere():
    """16^3 sphere, 2176 dipoles, m = 1.5+0.01i, x = 2, target 1e-8."""
    import precondda_amd as P
    pb, grid, k, ia, mask, G, M = _setup(8, 1.5 + 0.01j, 2.0)
    ref = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 1000, 1e-8)
    out = P.solve_gpbicgstab(grid, ia, torch.as_tensor(G).cuda(), torch.as_tensor(M).cuda(), pb["E_inc"], mask,
                             100, 1000, 1e-8)
    print("iters", ref[0], out[0])
    assert out[0] == ref[0]
    err = np.linalg.norm(out[1].cpu().numpy() - ref[1]) / np.linalg.norm(ref[1])
    print("P rel err", err)
    assert err < 1e-10


@pytest.mark.parametrize("p3", ["auto", "register"])
def test_mixed_precision_small(monkeypatch, p3):
    import precondda_amd as P
    if p3 != "auto":
        monkeypatch.setenv("PDDA_P3_KERNEL", p3)
    pb, grid, k, ia, mask, G, M = _setup(6, 1.33 + 0.001j, 2.0)
    ref = O.solve_gpbicgstab(grid, ia, G, M, pb["E_inc"], mask, 100, 1000, 1e-12)
    Gd = torch.as_tensor(G).cuda()
    Md = torch.as_tensor(M).cuda()
    Pm, info = P.solve_mixed(grid, ia, Gd, Gd.to(torch.complex64), Md.to(torch.complex64), pb["E_inc"], mask,
                             tol=1e-6, inner_target=1e-3)
    print(info)
    Pm = Pm.cpu().numpy()
    rel = O.true_residual(Pm, pb["E_inc"], ia, mask, G, grid)
    print("true residual", rel)
    assert rel < 1e-6
    ce, _ = O.cross_sections(ref[1], pb["E_inc"], ia, k)
    me, _ = O.cross_sections(Pm, pb["E_inc"], ia, k)
    assert abs(me - ce) < 1e-5 * abs(ce)


@pytest.mark.parametrize("graph", [False, True])
def test_iterate_before_bind_is_an_error(graph):
    """Iterating a solver whose operands were never bound returns PDDA_ERR_STATE with
    the same message as loading one, from a host check before any launch."""
    import ctypes
    from precondda_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.pdda_solver_create(ctypes.byref(h), 1, 8, 8, 8, 8, 8, 8, 2, 0) == 0
    try:
        it = L.pdda_solver_iterate_graph if graph else L.pdda_solver_iterate
        assert it(h, 2, None) == -5  # PDDA_ERR_STATE
        assert b"not bound" in L.pdda_last_error()
    finally:
        L.pdda_solver_destroy(h)
    m = ctypes.c_void_p()
    assert L.pdda_mixed_create(ctypes.byref(m), 8, 8, 8, 8, 8, 8, 2, 0) == 0
    try:
        assert L.pdda_mixed_iterate(m, 2, 1 if graph else 0, None) == -5
    finally:
        L.pdda_mixed_destroy(m)


def test_solver_failure_codes_match_oracle():
    """The reference's return conventions (Gpbicgstab.py:1173-1201), on the device and in
    the oracle with the same inputs: an exhausted iteration limit returns max_iter; a
    NaN in E_inc returns -(i+1) at the first iteration; a one-iteration stagnation
    window returns -(i+1) at the first step that does not lower the residual. That step
    is taken on a target (x = 3, no preconditioner) whose residual first rises at
    iteration 6, far above rounding; near the rounding floor the first rise moves with
    the last bits."""
    import precondda_amd as P
    pb, grid, k, ia, mask, G, M = _setup(3, 1.5 + 0.1j, 1.0)
    Gd, Md = torch.as_tensor(G).cuda(), torch.as_tensor(M).cuda()

    def both(E, ratio, max_iter, target, **kw):
        ref = O.solve_gpbicgstab(grid, ia, G, M, E, mask, ratio, max_iter, target, **kw)
        out = P.solve_gpbicgstab(grid, ia, Gd, Md, E, mask, ratio, max_iter, target, **kw)
        print("oracle", ref[0], "device", out[0], flush=True)
        return ref, out

    ref, out = both(pb["E_inc"], 100, 2, 1e-12)
    assert ref[0] == 2 and out[0] == 2
    assert abs(out[2] - ref[2]) <= 1e-10 * abs(ref[2])
    E2 = pb["E_inc"].copy()
    E2[tuple(np.argwhere(mask)[0]) + (0,)] = np.nan
    ref, out = both(E2, 100, 50, 1e-12)
    assert ref[0] == -1 and out[0] == -1
    pb, grid, k, ia, mask, G, M = _setup(3, 1.5 + 0.1j, 3.0)
    Gd, Md = torch.as_tensor(G).cuda(), torch.as_tensor(M).cuda()
    ref, out = both(pb["E_inc"], 0, 200, 0.0, precon=False, stagnation=1)
    assert ref[0] == -7 and out[0] == -7
    assert abs(out[2] - ref[2]) <= 1e-10 * abs(ref[2])


def test_run_solver_matches_oracle_and_wraps_failures():
    """run_solver (ddaforgit.py:1175-1210): ratio forced to 100 and the default target
    1e-5, preconditioned, FP64 path; same iterations and P as the oracle's run_solver,
    and a failure inside surfaces as RuntimeError("Solver failed: ...")."""
    import precondda_amd as P
    pb, _ = O.sphere_problem(4, 1.5 + 0.01j, 1.0)
    grid, k = pb["grid_size"], pb["k"]
    nx, ny, nz = grid
    rm = 1 - 2j
    ref = O.run_solver(grid, k, pb["alpha_array"], 1.0, refract_mult=rm, E_inc=pb["E_inc"], precon=True,
                       max_iter=500, expansion=(nx, ny, nz))
    out = P.run_solver(grid, k, pb["alpha_array"], 1.0, nx, ny, nz, rm, pb["E_inc"], precon=True, max_iter=500,
                       double=True, ratio=7)
    print("iters", ref[0], out[0])
    assert out[0] == ref[0] and out[0] > 0
    err = np.linalg.norm(out[1].cpu().numpy() - ref[1]) / np.linalg.norm(ref[1])
    print("P rel err", err)
    assert err < 1e-10
    with pytest.raises(RuntimeError, match="^Solver failed: "):
        P.run_solver(grid, k, pb["alpha_array"], 1.0, nx, ny, nz, rm, pb["E_inc"].reshape(-1)[:-3], double=True)
