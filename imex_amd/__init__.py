"""imex_amd — MI355X-native IMEX-MHD per-step hot path (HIP kernels behind a C ABI).

The compute path is ``libimex_amd.so`` (include/imex.h); this package only binds it
with ctypes for tests and benchmarks.
"""
from .capi import (  # noqa: F401
    Solver, Batch, ImexError, Params, Scalars, Stats, SolveInfo, lib, init_case, new_stage,
    default_params, vortex_params, orszag_tang_params, blast_params, case_of, grid_shape, array_shape,
    BC_PERIODIC, BC_TRANSMISSIVE, RECON_MINMOD, RECON_VANLEER, RECON_WENO5, EOS_IDEAL, EOS_RK,
    CT_CONSISTENT, CT_PLAIN, PREC_FP64, PREC_MIXED, OP_MAG, OP_P_CARRIER, OP_P_JACOBIAN, OP_ENERGY,
    FORMULA_ENERGY, FORMULA_PRESSURE, MAGFIELD_STAGGER, MAGFIELD_UNSTAGGER, MAGFIELD_POTENTIAL_A,
    MAGFIELD_NODIVFREE, OP_POTENTIAL, init_potential, RECON_NONE, RECON_ENO3, RECONSTATE_PRIM, RECONSTATE_CONS, write_vtk_host, write_fields_host, write_history, output_path,
    CASE_VORTEX, CASE_ORSZAG_TANG, CASE_BLAST, CASE_JET, BC_JET_INFLOW, BC_REFLECTIVE, BC_DIRICHLET, jet_params,
    SCHED_GRAPH, SCHED_BUDGET, SCHED_HOST, SCHED_AUTO, GC, NVAR,
    TUNE_EXPLICIT, TUNE_KRYLOV, ERR_ARG, ERR_HIP, ERR_NONFINITE, ERR_UNSUPPORTED, ERR_NOMEM,
)
