"""boltmi -- Bolt.jl's perturbation hot path on MI355X (libbolt_hip.so)."""
from .api import (MASS_NATURAL, Background, BasicNewtonian, Context, CosmoParams, HIPNewtonian, Hierarchy,
                  IonizationHistory, PerturbationIntegrator, RECFAST, Solution, SourceGrid, boltsolve,
                  boltsolve_batch, boltsolve_rsa, cltt, cltt_on_grid, default_context, hier_cfg,
                  initial_conditions_batch, log10_k, plin, quadratic_k, recfast_xe, rhs_batch, source_grid,
                  state_length)

__all__ = [n for n in dir() if not n.startswith("_")]
