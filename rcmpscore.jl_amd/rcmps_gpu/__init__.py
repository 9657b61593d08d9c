"""rcmps_gpu on MI355X: fg_phi4 (phi^4 energy and Riemannian gradient of a relativistic
continuous MPS), its state preparation and optimizer, running on hand-written HIP kernels
(librcmps_hip.so, C ABI in include/rcmps.h).  Importing this package does not touch the GPU.
"""
from ._lib import DeviceControlFailure, SolverFailure, gemm_tuning, gemm_variant, nojump_threshold
from .autograd import Phi4Energy, lbfgs_rk, phi4_energy
from .kinetic import KineticKrylovOptions, KineticOptions, kinetic1, kinetic2
from .excitation import (DavidsonOptions, ExcitationContext, ExcitationModel, ExcitationOperator,
                         TangentCoordinates, action_dual_poisson, davidson, dispersion_relation, excitation_eigsolve,
                         excitation_hamiltonian_action, excitation_parity, excitation_poisson, field_seed,
                         iterative_dispersion, kinetic_preconditioner, relativistic_diagnostics, excitation_element,
                         tangent_metric, excitation_matrix_element)
from .jld2 import read_jld2
from .io import StateRecord, rcmps_load, rcmps_store, read_portable, write_portable
from .observables import (boundary_environment, fg_phi4, fg_phi4_batch, fg_phi4_raw, fg_phi4_rk,
                          trajectory_samples)
from .linesearch import BackwardLineSearch, HagerZhangLineSearch, SafeguardedLineSearch
from . import controller
from .controller import (OptimizationControls, RobustnessControls, gpu_objective, gpu_stationary,
                         load_resume_state, save_resume_state)
from .optimization import metric, optimize_state, precondition, retraction
from .state import (IntegratorParameters, ManifoldParameters, ModelParameters, RandomState, RandomStateZ2,
                    State, random_state_arrays, random_state_z2_arrays)

__version__ = "0.1.0"

__all__ = [
    "BackwardLineSearch", "HagerZhangLineSearch", "SafeguardedLineSearch", "DeviceControlFailure", "IntegratorParameters", "ManifoldParameters",
    "ModelParameters", "RandomState", "RandomStateZ2", "SolverFailure", "State", "boundary_environment",
    "fg_phi4", "fg_phi4_raw", "metric", "optimize_state", "precondition", "random_state_arrays",
    "random_state_z2_arrays", "retraction", "trajectory_samples", "StateRecord", "rcmps_load", "rcmps_store",
    "read_jld2", "read_portable", "write_portable", "KineticKrylovOptions", "KineticOptions", "kinetic1", "kinetic2", "Phi4Energy", "fg_phi4_batch", "fg_phi4_rk", "lbfgs_rk", "phi4_energy",
    "DavidsonOptions", "ExcitationContext", "ExcitationModel", "ExcitationOperator", "TangentCoordinates",
    "action_dual_poisson", "davidson", "dispersion_relation", "excitation_eigsolve", "excitation_hamiltonian_action",
    "excitation_parity", "excitation_poisson", "field_seed", "iterative_dispersion", "kinetic_preconditioner",
    "relativistic_diagnostics", "excitation_element", "tangent_metric", "excitation_matrix_element",
    "gemm_tuning", "gemm_variant", "nojump_threshold", "controller", "OptimizationControls", "RobustnessControls", "gpu_objective",
    "gpu_stationary", "load_resume_state", "save_resume_state",
]
