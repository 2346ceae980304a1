"""MI355X (gfx950) Bethe-Salpeter hot path of fdDGAsolver.jl: host API over libfdga.so."""
from ._lib import build, load, LIB_PATH, HEADER  # noqa: F401
from .solver import (  # noqa: F401
    FdgaError, Solver, SolverGroup, pCh, tCh, aCh,
    siam_bare_Green, hubbard_bare_Green, hubbard_band, compute_occupation,
    ParquetSolver, NL_ParquetSolver,
    parquet_solver_siam_parquet_approximation, parquet_solver_hubbard_parquet_approximation,
    Dyson, bubbles, build_K3_cache, build_K3_cache_mfRG, BSE_L_K2, BSE_L_K3, BSE_K1, BSE_K2, BSE_K3,
    BSE_K1_new, BSE_K2_new, update_F, SDE, mfRGLinearMap,
    tangent_solver, BSE_K3_jvp, BSE_K3_vjp, k3_gather_map,
    iterate_solver, flatten, flatten_vertex, unflatten, fixed_point, solve, init_sym_grp, reset_sym_grp,
    susceptibilities, chi_from_k1,
)
