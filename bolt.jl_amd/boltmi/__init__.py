"""boltmi -- Bolt.jl's perturbation hot path on MI355X (libbolt_hip.so)."""
from .api import (MASS_NATURAL, Background, autotune, backgrounds, BasicNewtonian, Context, CosmoParams, HIPNewtonian, Hierarchy,
                  IonizationHistory, PerturbationIntegrator, RECFAST, Server, Solution, SourceGrid, boltsolve,
                  boltsolve_batch, boltsolve_rsa, cl_on_grid, clee, clee_on_grid, clte, clte_on_grid, cltt,
                  cltt_on_grid, default_context, FP32_SOLVE, hier_cfg, precision_flags,
                  initial_conditions_batch, log10_k, plin, quadratic_k, recfast_xe, rhs_batch, source_grid,
                  source_grid_P, source_grid_batch, source_grids_TP,
                  state_length)
from .dual import PARAM_FIELDS, Dual, derivative, tables_jacobian
from .bessel import (MomentTable, filon_sum, integrate_sph_bessel_filon, sph_bessel_interpolator,
                     sph_j_moment_asymp, sph_j_moment_maclaurin_1F2, sph_j_moment_series)

__all__ = [n for n in dir() if not n.startswith("_")]
