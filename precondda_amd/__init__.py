"""MI355X-native DDA scattering solver.

Drop-in for the solve path of Toastlovesjam/preconDDA: polarisability, target and
lattice setup, incident and excited fields, the interaction-tensor build, the FFT
matvec and the preconditioned GPBiCGSTAB solve, computed by hand-written gfx950 HIP
kernels in libpdda.so (C ABI in include/pdda.h) driven from Python over ctypes.
"""
from .operators import (E_field_solver, build_preconditioner, calculate_zeta_eta_twostep, cross_sections,
                        fft_matvec, fft_matvec3, generate_interaction_row, inner_product, interaction_mult,
                        prepare_interaction_matrices, setup_chunked_values)
from .solver import DeviceSolver, MixedSolver, run_solver, solve_gpbicgstab, solve_mixed

__all__ = [
    "E_field_solver", "build_preconditioner", "calculate_zeta_eta_twostep", "cross_sections", "fft_matvec",
    "fft_matvec3", "generate_interaction_row", "inner_product", "interaction_mult", "prepare_interaction_matrices",
    "setup_chunked_values", "DeviceSolver", "MixedSolver", "run_solver", "solve_gpbicgstab", "solve_mixed",
]
