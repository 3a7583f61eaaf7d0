"""Full S1 + S3 + S5 response: the EOM space with the whole rank-five shell.

Drop-in for ``CompiledFullS135Action`` and ``FullS135ResponseOperator``
(``src/c3l/cc/s135_full.py:101-361``, used as in ``examples/shared/response.py:105-138``): the
exact similarity-transformed response operator among the rank-one, rank-three and rank-five
creator strings of one charge block, with the CCSD amplitudes fixed.

The term compiler expands ``[K̄, X]|0>`` with the rank-five input blocks and keeps every output
term of rank one, three or five (the reference's retained plans plus its rank-five input plans,
``s135_full.py:51-98,109-170``); the plan engine evaluates it with the same kernels as the S13
sigma, extended to rank-five output blocks, rank-five input blocks and rank-five creation
tensors.  Vectors hold the S13 coordinates first, in the S13 order, and then the rank-five
shell, as the reference's operator does (``s135_full.py:299-328``), so S13 sources embed with
zeros (the sources have no rank-five part).
"""

from __future__ import annotations

from time import perf_counter

import numpy as np

from .lowering import reference_response_order
from .eom import (
    ComponentEOMContext,
    ComponentEOMProblem,
    ComponentSigmaOperator,
    component_eom_context,
    component_eom_sources,
    eom_ccsd_problem,
)


def full_s135_problem(problem: ComponentEOMProblem) -> ComponentEOMProblem:
    """The S135 problem of the charge block of an S13 problem (compiled once per problem)."""

    if problem.oracle.space == "S135":
        return problem
    cached = problem.__dict__.get("_s135")
    if cached is None:
        cached = ComponentEOMProblem.compile(
            eom_ccsd_problem(problem.oracle.cc, response_spin=problem.oracle.response_spin, space="S135"),
            problem.static, problem.charge)
        problem.__dict__["_s135"] = cached
    return cached


class CompiledFullS135Action:
    """Compiled S135 action on the amplitudes of an S13 (or S135) EOM context.

    ``materialization_batch_size`` is the reference's host batching knob
    (``s135_full.py:109-118``): it is validated and has no effect here (the plan engine
    materialises every creation tensor in one launch).
    """

    def __init__(self, context: ComponentEOMContext, *, materialization_batch_size: int = 16):
        if materialization_batch_size <= 0 or materialization_batch_size > 16:
            raise ValueError("materialization_batch_size must be in [1,16]")
        started = perf_counter()
        s13 = context.problem
        problem = full_s135_problem(s13)
        shared = context.__dict__.get("_device_t")
        self.context = component_eom_context(problem, context.amplitudes, device_amplitudes=shared)
        self.p_dimension = int(problem.dimension - _shell_dimension(problem))
        self.compile_wall_seconds = float(perf_counter() - started)


def _shell_dimension(problem: ComponentEOMProblem) -> int:
    return int(sum(b.positions.size for b in problem.coordinates.blocks if sum(b.key) == 5))


class FullS135ResponseOperator(ComponentSigmaOperator):
    """Exact right action on the complete ``S1 + S3 + S5`` space (``s135_full.py:296-361``).

    A ``ComponentSigmaOperator`` of the S135 context, so every resolvent driver of this package
    (``component_retarded_green``, ``arnoldi_resolvent``, the matrix and Matsubara drivers)
    accepts it.  ``s5_tile_size`` is the reference's host tiling knob: validated, no effect.
    """

    def __init__(self, action: CompiledFullS135Action, *, s5_tile_size: int = 16384, slot: int = 0):
        if s5_tile_size <= 0:
            raise ValueError("s5_tile_size must be positive")
        super().__init__(action.context, slot=slot)
        self.action = action
        self.s5_tile_size = int(s5_tile_size)
        self.p_dimension = action.p_dimension
        self.s5_dimension = self.dimension - self.p_dimension
        self.coordinates = action.context.problem.coordinates
        # the reference orders the rank-five shell by key and index combination, this build by
        # mask value: vector[reference_order] is a vector of this operator in the reference's order
        self.reference_order = reference_response_order(self.coordinates)

    def on_slot(self, slot: int) -> "FullS135ResponseOperator":
        return FullS135ResponseOperator(self.action, s5_tile_size=self.s5_tile_size, slot=slot)

    def matvec(self, vector) -> np.ndarray:
        value = np.asarray(vector, dtype=np.complex128)
        if value.shape != (self.dimension,) or not np.all(np.isfinite(value)):
            raise ValueError("the full-S135 input has the wrong shape or is nonfinite")
        return super().matvec(value)


def full_s135_sources(operator: FullS135ResponseOperator, *, site: int | None = None):
    """Addition and removal sources and the sink of ``site`` on the S135 space: the S13 sources
    embedded with zeros in the rank-five shell (``examples/shared/response.py:128-136``), made
    on the GPU by the S135 problem's source plans."""

    return component_eom_sources(operator.context, site=site)


__all__ = ["CompiledFullS135Action", "FullS135ResponseOperator", "full_s135_problem", "full_s135_sources"]
