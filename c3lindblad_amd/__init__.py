"""MI355X-native coupled-cluster hot path of C3L (haoen2021/C3Lindblad).

Public names mirror the reference API (``README.md:62-83``):

    from c3lindblad_amd import (
        FermionicLindbladModel, factorized_cc_problem, ComponentStaticProblem,
        solve_component_ccsd, eom_ccsd_problem, ComponentEOMProblem,
        component_eom_context, component_eom_sources, ComponentSigmaOperator,
        component_retarded_green,
    )

Host code lowers the model in FP64; residuals, JVPs, sigma actions, Krylov
orthogonalisation and shifted solves run in the HIP library ``libc3l_hip.so``
(``include/c3l_hip.h``).  There is no CPU fallback: without the library or a
GPU the device entry points raise.
"""

from .model import FermionicLindbladModel, SparseTwoBodyInteraction, TolerancePolicy, density_density_interaction
from .lowering import factorized_cc_problem, FactorizedCCProblem
from .static import (
    ComponentStaticProblem,
    component_static_jvp,
    component_static_linearization,
    component_static_linearized_jvp,
    component_static_residual,
)
from .solver import ComponentCCSolution, NewtonKrylovPolicy, RapidityPreconditionerDiagnostics, solve_component_ccsd
from .eom import (
    ComponentEOMProblem,
    ComponentSigmaOperator,
    component_eom_context,
    component_eom_sigma,
    component_eom_sources,
    eom_ccsd_problem,
)
from .response import (
    ArnoldiResolvent,
    arnoldi_resolvent,
    component_retarded_green,
    component_retarded_matrix_green,
    component_retarded_matrix_green_arnoldi,
    component_retarded_matrix_green_shifts,
    component_retarded_matrix_time,
    component_retarded_time,
    solve_component_shifted,
)
from .matsubara import MatsubaraGrid, component_matsubara_green, matsubara_grid
from .checkpoint import CheckpointedArnoldiResult, checkpointed_arnoldi_resolvent, eom_checkpoint_identity
from .s135 import CompiledFullS135Action, FullS135ResponseOperator, full_s135_problem, full_s135_sources
from .observables import (
    ComponentObservableProblem,
    contact_currents,
    factorized_ccsd_covariance,
    two_body_expectations,
)

__all__ = [
    "ArnoldiResolvent",
    "CheckpointedArnoldiResult",
    "CompiledFullS135Action",
    "FullS135ResponseOperator",
    "full_s135_problem",
    "full_s135_sources",
    "checkpointed_arnoldi_resolvent",
    "eom_checkpoint_identity",
    "ComponentCCSolution",
    "MatsubaraGrid",
    "RapidityPreconditionerDiagnostics",
    "component_matsubara_green",
    "component_retarded_matrix_green",
    "component_retarded_matrix_green_arnoldi",
    "component_retarded_matrix_time",
    "matsubara_grid",
    "ComponentEOMProblem",
    "ComponentObservableProblem",
    "ComponentSigmaOperator",
    "ComponentStaticProblem",
    "FactorizedCCProblem",
    "FermionicLindbladModel",
    "NewtonKrylovPolicy",
    "SparseTwoBodyInteraction",
    "TolerancePolicy",
    "arnoldi_resolvent",
    "component_eom_context",
    "component_eom_sigma",
    "component_eom_sources",
    "contact_currents",
    "factorized_ccsd_covariance",
    "component_retarded_green",
    "component_retarded_matrix_green_shifts",
    "component_retarded_time",
    "component_static_jvp",
    "component_static_linearization",
    "component_static_linearized_jvp",
    "component_static_residual",
    "density_density_interaction",
    "eom_ccsd_problem",
    "factorized_cc_problem",
    "solve_component_ccsd",
    "solve_component_shifted",
    "two_body_expectations",
]
