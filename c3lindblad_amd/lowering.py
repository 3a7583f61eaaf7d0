"""Lower a Lindblad model to the coupled-cluster problem data the device path consumes.

Host-side restatement of the reference set-up chain
(``src/c3l/cc/factorized.py:248-465``):

1. full Hartree--Fock fixed point ``solve_ghf`` (``factorized.py:303-318``);
2. per-spin normal modes of the Fock blocks (``ci/reference.py:422-500``);
3. the particle-hole mode layout: for spin ``s`` the ``4n`` modes are ordered
   ``[s*2n, s*2n+n)`` (unoccupied, charge +1) then ``[s*2n+n, s*2n+2n)``
   (occupied, charge -1) -- four *groups* up+, up-, down+, down-;
4. linear particle-hole forms of physical and tilde operators
   (``ci/matrix_free.py:95-139``) and the factorized interaction products plus
   the one-body counterterms ``E - F`` (``ci/matrix_free.py:142-183``,
   ``factorized.py:374-402``);
5. the retained CCSD creator strings in the reference's canonical order
   (PH rank, then integer mask; ``ci/basis.py:149-178``) and the induced
   amplitude tilde map with its real orbit coordinates (``cc/symmetry.py``).

Masks of ``4n`` modes overflow 64 bits at ``n >= 17``, so coordinates are kept
as sorted mode tuples; the order key ``sum_k m_k 256^k`` over the
ascending modes sorts exactly like the integer mask within one rank.
"""

from __future__ import annotations

from functools import lru_cache
from dataclasses import dataclass, field
from itertools import combinations
from math import comb

import numpy as np

from .meanfield import solve_ghf
from .model import FermionicLindbladModel, TolerancePolicy
from .normal_modes import ModeTilde, NormalModeReference, normal_mode_reference, tilde_map_from_references

# Block keys count creators per group (up+, up-, down+, down-).
BlockKey = tuple[int, int, int, int]
STATIC_KEYS: tuple[BlockKey, ...] = ((1, 1, 0, 0), (0, 0, 1, 1), (2, 2, 0, 0), (0, 0, 2, 2), (1, 1, 1, 1))


@dataclass(frozen=True)
class HartreeFockReference:
    model: FermionicLindbladModel
    fock: np.ndarray
    gamma: np.ndarray
    E_up: np.ndarray
    E_down: np.ndarray
    up: NormalModeReference
    down: NormalModeReference

    @property
    def n_sites(self) -> int:
        return self.model.n_sites

    @property
    def mode_references(self) -> tuple[NormalModeReference, NormalModeReference]:
        return (self.up, self.down)

    @property
    def one_body_blocks(self) -> tuple[np.ndarray, np.ndarray]:
        return (self.E_up, self.E_down)


def full_hartree_fock_reference(model: FermionicLindbladModel, fock, gamma, policy=None) -> HartreeFockReference:
    """Normal modes of the spin-diagonal Fock blocks (``ci/reference.py:422-500``)."""

    F = np.asarray(fock, dtype=np.complex128)
    E_up, E_down = np.asarray(F[0::2, 0::2]), np.asarray(F[1::2, 1::2])
    spin_mixing = max(np.linalg.norm(F[0::2, 1::2]), np.linalg.norm(F[1::2, 0::2]))
    if spin_mixing > 1e-10:
        raise ValueError("the component coordinates require a spin-conserving HF reference")
    up = normal_mode_reference(E_up, model.gamma_loss, model.gamma_gain, policy)
    down = normal_mode_reference(E_down, model.gamma_loss, model.gamma_gain, policy)
    return HartreeFockReference(model, F, np.asarray(gamma), E_up, E_down, up, down)


@dataclass(frozen=True)
class LinearModeOperator:
    """``sum (coefficient * bar_P_mode or P_mode)`` over ``(kind, mode, coefficient)``."""

    label: str
    components: tuple[tuple[str, int, complex], ...]


@dataclass(frozen=True)
class FactorizedModeProduct:
    """Coefficient times a left-to-right product of linear mode operators."""

    label: str
    coefficient: complex
    factors: tuple[LinearModeOperator, ...]


def physical_mode_operator(reference: HartreeFockReference, site: int, spin: int, channel: str) -> LinearModeOperator:
    """Expand ``c``, ``c†``, ``c̃``, ``c̃†`` of one site/spin in PH modes.

    ``c`` and ``c̃†`` are dual rows of ``V`` (occupied modes create), ``c†``
    and ``c̃`` are sharp columns of ``V_inv`` (occupied modes annihilate),
    exactly as ``ci/matrix_free.py:95-139``.
    """

    if channel not in ("c", "c_dagger", "tilde_c", "tilde_c_dagger"):
        raise ValueError("unknown physical/tilde operator channel")
    ref = reference.up if spin == 0 else reference.down
    n = reference.n_sites
    offset = spin * 2 * n
    if channel in ("c", "tilde_c_dagger"):
        sharp = False
        coefficients = ref.V[site if channel == "c" else n + site, :]
    else:
        sharp = True
        coefficients = ref.V_inv[:, site if channel == "c_dagger" else n + site]
    components = []
    for p, coefficient in enumerate(coefficients):
        if coefficient == 0.0:
            continue
        occupied = bool(ref.vacuum_mask[p])
        kind = ("annihilate" if occupied else "create") if sharp else ("create" if occupied else "annihilate")
        components.append((kind, offset + p, complex(coefficient)))
    prefix = "up" if spin == 0 else "down"
    return LinearModeOperator(f"{prefix}_site_{site}_{channel}", tuple(components))


def interaction_products(reference: HartreeFockReference) -> list[FactorizedModeProduct]:
    """Physical and tilde copies of every canonical two-body term (``matrix_free.py:142-183``)."""

    cache: dict[tuple[int, str], LinearModeOperator] = {}

    def op(orbital: int, channel: str) -> LinearModeOperator:
        key = (orbital, channel)
        if key not in cache:
            site, spin = reference.model.orbital_site_spin(orbital)
            cache[key] = physical_mode_operator(reference, site, spin, channel)
        return cache[key]

    products = []
    for index, (p, q, r, s, value) in enumerate(reference.model.interaction.terms):
        products.append(
            FactorizedModeProduct(
                f"two_body_{index}:physical", value, (op(p, "c_dagger"), op(q, "c_dagger"), op(s, "c"), op(r, "c"))
            )
        )
        products.append(
            FactorizedModeProduct(
                f"two_body_{index}:tilde",
                -value.conjugate(),
                (op(p, "tilde_c_dagger"), op(q, "tilde_c_dagger"), op(s, "tilde_c"), op(r, "tilde_c")),
            )
        )
    return products


def counterterm_products(model: FermionicLindbladModel, reference: HartreeFockReference) -> list[FactorizedModeProduct]:
    """One-body ``E - E_spin`` corrections, physical and tilde (``factorized.py:374-402``)."""

    products = []
    delta = tuple(np.asarray(model.E) - block for block in reference.one_body_blocks)
    n = model.n_sites
    for i in range(n):
        for j in range(n):
            for spin in range(2):
                value = delta[spin][i, j]
                if value == 0.0:
                    continue
                normal = (
                    physical_mode_operator(reference, i, spin, "c_dagger"),
                    physical_mode_operator(reference, j, spin, "c"),
                )
                tilde = (
                    physical_mode_operator(reference, i, spin, "tilde_c_dagger"),
                    physical_mode_operator(reference, j, spin, "tilde_c"),
                )
                products.append(FactorizedModeProduct(f"quadratic_counter:c_{i}_{spin}_dagger_c_{j}_{spin}", complex(value), normal))
                products.append(
                    FactorizedModeProduct(
                        f"quadratic_counter:-tilde_c_{i}_{spin}_dagger_tilde_c_{j}_{spin}", -complex(delta[spin][j, i]), tilde
                    )
                )
    return products


# ----------------------------------------------------------------------------- coordinates


def _order_key(modes: np.ndarray) -> np.ndarray:
    """Integer that sorts sorted-mode rows like their bit masks (modes < 256)."""

    key = np.zeros(modes.shape[0], dtype=np.int64)
    for k in range(modes.shape[1]):
        key += modes[:, k].astype(np.int64) << (8 * k)
    return key


def _block_modes(n: int, key: BlockKey) -> np.ndarray:
    """All canonical coordinates of one block as ascending global mode rows."""

    parts = []
    for g, count in enumerate(key):
        if count == 0:
            continue
        combos = np.array(list(combinations(range(n), count)), dtype=np.int64).reshape(-1, count) + g * n
        parts.append(combos)
    if not parts:
        return np.zeros((1, 0), dtype=np.int64)
    grids = np.meshgrid(*[np.arange(p.shape[0]) for p in parts], indexing="ij")
    rows = [part[grid.ravel()] for part, grid in zip(parts, grids)]
    return np.concatenate(rows, axis=1)


@dataclass
class BlockCoordinates:
    """One block of a coordinate set: vector positions and per-slot local indices."""

    key: BlockKey
    positions: np.ndarray  # int64 [count] index into the ordered vector
    indices: np.ndarray  # int64 [count, degree] local (within-group) indices, slot order

    @property
    def degree(self) -> int:
        return int(sum(self.key))


@dataclass
class CoordinateSet:
    """Ordered creator strings of one vector space with its block decomposition."""

    n: int
    modes: np.ndarray  # int64 [dim, max_degree] ascending global modes, padded with -1
    degrees: np.ndarray  # int64 [dim]
    blocks: list[BlockCoordinates] = field(default_factory=list)

    @property
    def dimension(self) -> int:
        return int(self.degrees.size)

    def block(self, key: BlockKey) -> BlockCoordinates | None:
        for b in self.blocks:
            if b.key == key:
                return b
        return None


def coordinate_set(n: int, keys_by_rank: list[list[BlockKey]]) -> CoordinateSet:
    """Enumerate blocks rank by rank; within a rank sort all strings by mask value."""

    all_modes = []
    degrees = []
    block_rows: dict[BlockKey, tuple[np.ndarray, np.ndarray]] = {}
    cursor = 0
    width = max(sum(k) for keys in keys_by_rank for k in keys)
    for keys in keys_by_rank:
        rank_modes = []
        rank_tags = []
        for tag, key in enumerate(keys):
            modes = _block_modes(n, key)
            rank_modes.append(modes)
            rank_tags.append(np.full(modes.shape[0], tag, dtype=np.int64))
        deg = sum(keys[0])
        modes = np.concatenate(rank_modes)
        tags = np.concatenate(rank_tags)
        local = np.arange(modes.shape[0])
        order = np.argsort(_order_key(modes), kind="stable")
        modes, tags, local = modes[order], tags[order], local[order]
        offsets = np.cumsum([0] + [m.shape[0] for m in rank_modes])
        for tag, key in enumerate(keys):
            sel = np.flatnonzero(tags == tag)
            # block-local canonical order = the block enumeration order
            block_local = local[sel] - offsets[tag]
            positions = np.empty(sel.size, dtype=np.int64)
            positions[block_local] = cursor + sel
            bm = _block_modes(n, key)
            block_rows[key] = (positions, bm % n)
        padded = np.full((modes.shape[0], width), -1, dtype=np.int64)
        padded[:, :deg] = modes
        all_modes.append(padded)
        degrees.append(np.full(modes.shape[0], deg, dtype=np.int64))
        cursor += modes.shape[0]
    cs = CoordinateSet(n, np.concatenate(all_modes), np.concatenate(degrees))
    for keys in keys_by_rank:
        for key in keys:
            positions, indices = block_rows[key]
            cs.blocks.append(BlockCoordinates(key, positions, indices))
    return cs


def _frozen_set(cs: CoordinateSet) -> CoordinateSet:
    for a in (cs.modes, cs.degrees, *[x for b in cs.blocks for x in (b.positions, b.indices)]):
        a.setflags(write=False)
    return cs


@lru_cache(maxsize=64)
def static_coordinates(n: int) -> CoordinateSet:
    """CCSD strings: PH rank 1 ((1,1,0,0),(0,0,1,1)) then rank 2 (``ci/basis.py:149-178``).

    The set depends on ``n`` only: it is built once per site count and shared (read-only) by
    every problem of that size."""

    return _frozen_set(coordinate_set(n, [[(1, 1, 0, 0), (0, 0, 1, 1)], [(2, 2, 0, 0), (1, 1, 1, 1), (0, 0, 2, 2)]]))


RESPONSE_SPACES = ("S13", "S135")


def response_keys(charge: int, space: str = "S13") -> list[list[BlockKey]]:
    """Keys of one charge block for response spin 0: S13 (ranks 1 and 3, ``ci/basis.py:181-234``)
    or S135, which adds the whole rank-five shell (``cc/eom.py:1-40``, ``cc/s135_full.py:101-361``):
    every key of rank five with the block's charge and no net spin flip."""

    if space not in RESPONSE_SPACES:
        raise ValueError("response space must be 'S13' or 'S135'")
    if charge == 1:
        keys = [[(1, 0, 0, 0)], [(2, 1, 0, 0), (1, 0, 1, 1)]]
        five = [(3, 2, 0, 0), (2, 1, 1, 1), (1, 0, 2, 2)]
    elif charge == -1:
        keys = [[(0, 1, 0, 0)], [(1, 2, 0, 0), (0, 1, 1, 1)]]
        five = [(2, 3, 0, 0), (1, 2, 1, 1), (0, 1, 2, 2)]
    else:
        raise ValueError("response charge must be +1 or -1")
    return keys + [five] if space == "S135" else keys


@lru_cache(maxsize=128)
def response_coordinates(n: int, charge: int, space: str = "S13") -> CoordinateSet:
    """Strings of one charge block (built once per site count, charge and space, read-only):
    rank by rank, keys whose group counts exceed n hold no string and are left out."""

    keys = [[k for k in rank if max(k) <= n] for rank in response_keys(charge, space)]
    return _frozen_set(coordinate_set(n, [rank for rank in keys if rank]))


def reference_response_order(coords: CoordinateSet) -> np.ndarray:
    """Positions of ``coords`` (a response set built by ``response_coordinates``) in the
    reference's vector order: the S1 and S3 strings in the same order as here (both sort them by
    mask value), then the rank-five shell as the reference enumerates it
    (``component_tensor.py:384-409,439-470``): its keys in ascending tuple order and, within a
    key, the lexicographic index combinations of every slot group raveled with the last group
    fastest.  ``vector_here[order]`` is the vector in the reference's order."""

    from itertools import combinations
    from math import comb

    n = coords.n
    blocks = sorted((b for b in coords.blocks if sum(b.key) == 5), key=lambda b: tuple(b.key))
    head = np.sort(np.concatenate([b.positions for b in coords.blocks if sum(b.key) != 5] or
                                  [np.zeros(0, np.int64)]))
    parts = [head]
    for b in blocks:
        counts = [int(c) for c in b.key]
        dims = [comb(n, c) for c in counts]
        rank_of = [{cmb: r for r, cmb in enumerate(combinations(range(n), c))} for c in counts]
        flat = np.zeros(b.positions.size, dtype=np.int64)
        col = 0
        for g, c in enumerate(counts):  # block indices hold group 0's indices first, each group sorted
            ranks = np.array([rank_of[g][tuple(int(v) for v in row)] for row in b.indices[:, col:col + c]],
                             dtype=np.int64) if c else np.zeros(b.positions.size, dtype=np.int64)
            flat = flat * dims[g] + ranks
            col += c
        order = np.empty(b.positions.size, dtype=np.int64)
        order[flat] = b.positions
        parts.append(order)
    return np.concatenate(parts)


# ----------------------------------------------------------------------------- amplitude tilde map


@dataclass
class TildeMap:
    """Anti-linear amplitude orbit map ``t_partner = phase * conj(t)`` (``cc/symmetry.py:19-164``).

    ``unpack`` and ``pack`` are written as gathers so the same arrays drive the
    device kernels: ``t[i] = alpha[i] x[k0[i]] + beta[i] x[k1[i]]`` and
    ``x[c] = Re(w[c] t[src[c]])``.
    """

    partner: np.ndarray
    phase: np.ndarray
    alpha: np.ndarray
    beta: np.ndarray
    k0: np.ndarray
    k1: np.ndarray
    pack_src: np.ndarray
    pack_w: np.ndarray

    @property
    def dimension(self) -> int:
        return int(self.partner.size)

    @property
    def real_size(self) -> int:
        return int(self.pack_src.size)

    def unpack(self, x) -> np.ndarray:
        x = np.asarray(x, dtype=np.float64)
        if x.shape != (self.real_size,) or not np.all(np.isfinite(x)):
            raise ValueError("independent tilde variables have the wrong shape")
        return self.alpha * x[self.k0] + self.beta * x[self.k1]

    def pack(self, t, *, check: bool = True) -> np.ndarray:
        t = np.asarray(t, dtype=np.complex128)
        if t.shape != (self.dimension,) or not np.all(np.isfinite(t)):
            raise ValueError("amplitudes have the wrong shape")
        if check and self.defect(t) > 2.0e-10:
            raise ValueError("amplitudes do not satisfy the tilde constraint")
        return np.real(self.pack_w * t[self.pack_src])

    def defect(self, t) -> float:
        t = np.asarray(t, dtype=np.complex128)
        return float(np.max(np.abs(t[self.partner] - self.phase * t.conj()), initial=0.0))


def _self_scale(eta: complex) -> complex:
    scale = complex(np.sqrt(eta / abs(eta)))
    if abs(scale - eta * scale.conjugate()) > abs(-scale - eta * (-scale).conjugate()):
        scale = -scale
    return scale


def amplitude_tilde_map(coords: CoordinateSet, mode_tilde: ModeTilde) -> TildeMap:
    """Induced determinant tilde map with the sorting sign (``cc/symmetry.py:150-187``)."""

    dim = coords.dimension
    partner = np.empty(dim, dtype=np.int64)
    phase = np.empty(dim, dtype=np.complex128)
    start = 0
    for deg in np.unique(coords.degrees):
        sel = np.flatnonzero(coords.degrees == deg)
        modes = coords.modes[sel, :deg]
        mapped = mode_tilde.partners[modes]
        inversions = np.zeros(sel.size, dtype=np.int64)
        for a in range(deg):
            for b in range(a + 1, deg):
                inversions += mapped[:, a] > mapped[:, b]
        factor = np.prod(mode_tilde.factors[modes], axis=1)
        factor = np.where(inversions % 2 == 1, -factor, factor)
        keys = _order_key(modes)
        target = _order_key(np.sort(mapped, axis=1))
        loc = np.searchsorted(keys, target)
        if np.any(loc >= keys.size) or np.any(keys[np.minimum(loc, keys.size - 1)] != target):
            raise ValueError("cluster basis is not tilde closed")
        partner[sel] = sel[loc]
        phase[sel] = factor
        start += sel.size
    if np.any(partner[partner] != np.arange(dim)):
        raise ValueError("tilde partners must form an involution")
    if np.max(np.abs(phase[partner] * phase.conj() - 1.0), initial=0.0) > 2.0e-10:
        raise ValueError("tilde phases violate the double-tilde identity")
    independent = np.flatnonzero(np.arange(dim) <= partner)
    is_self = partner[independent] == independent
    counts = np.where(is_self, 1, 2)
    cursor = np.concatenate(([0], np.cumsum(counts)[:-1]))
    real_size = int(counts.sum())
    alpha = np.zeros(dim, dtype=np.complex128)
    beta = np.zeros(dim, dtype=np.complex128)
    k0 = np.zeros(dim, dtype=np.int64)
    k1 = np.zeros(dim, dtype=np.int64)
    pack_src = np.empty(real_size, dtype=np.int64)
    pack_w = np.empty(real_size, dtype=np.complex128)
    # self orbits: t_i = s_i x_c with the branch of _self_scale
    si, sc_ = independent[is_self], cursor[is_self]
    scale = _self_scales(phase[si])
    alpha[si], k0[si], k1[si] = scale, sc_, sc_
    pack_src[sc_], pack_w[sc_] = si, 1.0 / scale
    # pairs: t_i = x_c + i x_{c+1}, t_j = eta_i conj(t_i)
    pi_, pc = independent[~is_self], cursor[~is_self]
    pj = partner[pi_]
    alpha[pi_], beta[pi_], k0[pi_], k1[pi_] = 1.0, 1.0j, pc, pc + 1
    alpha[pj], beta[pj], k0[pj], k1[pj] = phase[pi_], -1.0j * phase[pi_], pc, pc + 1
    pack_src[pc], pack_w[pc] = pi_, 1.0
    pack_src[pc + 1], pack_w[pc + 1] = pi_, -1.0j
    return TildeMap(partner, phase, alpha, beta, k0, k1, pack_src, pack_w)


def _self_scales(eta: np.ndarray) -> np.ndarray:
    """Vectorised ``_self_scale``."""

    scale = np.sqrt(eta / np.abs(eta))
    flip = np.abs(scale - eta * scale.conj()) > np.abs(-scale - eta * (-scale).conj())
    return np.where(flip, -scale, scale)


# ----------------------------------------------------------------------------- problem


@dataclass
class FactorizedCCProblem:
    """CCSD problem data without a full sector (mirror of ``factorized.py:69-168``)."""

    level: str
    model: FermionicLindbladModel
    reference: HartreeFockReference
    coordinates: CoordinateSet
    tilde: TildeMap
    mode_tilde: ModeTilde
    rapidities: np.ndarray
    products: tuple[FactorizedModeProduct, ...]
    hf_iterations: int = 0

    @property
    def num_amplitudes(self) -> int:
        return self.coordinates.dimension

    @property
    def num_variables(self) -> int:
        return self.tilde.real_size

    @property
    def num_modes(self) -> int:
        return 4 * self.model.n_sites

    @property
    def complete_sector_dimension(self) -> int:
        return comb(2 * self.model.n_sites, self.model.n_sites) ** 2

    def rapidity_sums(self, coords: CoordinateSet | None = None) -> np.ndarray:
        """``Λ_μ = sum_{p in μ} λ_p`` for every string of a coordinate set."""

        c = self.coordinates if coords is None else coords
        lam = np.concatenate((self.rapidities, [0.0]))
        m = np.where(c.modes >= 0, c.modes, self.rapidities.size)
        return lam[m].sum(axis=1)

    def construction_record(self) -> dict:
        return {
            "level": self.level,
            "operator_convention": "physical_generator_L_equals_minus_i_paper_K",
            "num_modes": self.num_modes,
            "complex_amplitude_dimension": self.num_amplitudes,
            "independent_real_dimension": self.num_variables,
            "complete_charge_zero_sector_dimension": self.complete_sector_dimension,
            "complete_sector_state_vector_constructed": False,
            "dense_jacobian_constructed": False,
        }


def factorized_cc_problem(model: FermionicLindbladModel, level: str = "CCSD", *, reference_mode: str = "full_hf") -> FactorizedCCProblem:
    """Full-HF CCSD lowering (``factorized.py:435-465`` with its defaults)."""

    if not isinstance(model, FermionicLindbladModel):
        raise TypeError("factorized CC requires a FermionicLindbladModel")
    if level != "CCSD":
        raise ValueError("this build lowers CCSD only")
    if reference_mode != "full_hf":
        raise ValueError("this build lowers reference_mode='full_hf' only")
    if any(sorted((p % 2, q % 2)) != sorted((r % 2, s % 2)) for p, q, r, s, _ in model.interaction.terms):
        raise ValueError("the current component coordinates require separate up/down conservation")
    hf = solve_ghf(
        model.physical_E,
        model.interaction,
        model.physical_gamma_loss,
        model.physical_gamma_gain,
        tolerance=5.0e-13,
        max_iterations=400,
    )
    if not hf.converged or hf.self_consistency_residual > 1e-10 or hf.lyapunov_residual > 1e-10 or hf.physicality_defect > 1e-10:
        raise RuntimeError("the default general full-HF reference did not converge")
    reference = full_hartree_fock_reference(model, hf.fock, hf.gamma)
    n = model.n_sites
    mode_tilde = tilde_map_from_references(reference.up, reference.down)
    for spin, ref in enumerate(reference.mode_references):
        expected = np.where(np.arange(2 * n) < n, 1, -1)
        if not np.array_equal(ref.mode_charges, expected):
            raise ValueError("mode charges violate the frozen PH block order")
    coords = static_coordinates(n)
    tilde = amplitude_tilde_map(coords, mode_tilde)
    rapidities = np.concatenate([ref.rapidities for ref in reference.mode_references]).astype(np.complex128)
    products = tuple(interaction_products(reference) + counterterm_products(model, reference))
    return FactorizedCCProblem("CCSD", model, reference, coords, tilde, mode_tilde, rapidities, products, hf.iterations)
