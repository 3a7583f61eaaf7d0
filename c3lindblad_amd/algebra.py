"""Symbolic exterior-algebra compiler for the CCSD residual, its JVP and the EOM sigma.

The reference evaluates ``exp(-T) K exp(T)|0>`` product by product with the
exact linear similarity transform ``exp(-T) F exp(T) = F + [F, T]`` of every
linear particle-hole form ``F = C + A`` (creator part ``C``, annihilator part
``A``); ``[F, T] = ι_A T`` is a pure creation polynomial
(``src/c3l/cc/factorized.py:1-20,554-700``; exterior-form version
``src/c3l/cc/component_tensor.py:764-816``).  Applying the factors of one
product right-to-left to the vacuum therefore only ever needs

* left wedge with a constant creator vector ``u`` (``C``),
* left wedge with a *shift* ``ι_a T`` (rank 1 from T1, rank 3 from T2),
* the interior product ``ι_a``, an odd derivation that distributes over the
  wedge with sign ``(-1)^(rank of the factors to its left)``.

This module performs that expansion once, on the host and symbolically.
Every resulting term is ``coefficient * atom_1 ∧ ... ∧ atom_k`` where an
atom is a constant creator vector or an amplitude block with an ordered chain
of interior contractions.  Terms are put in a canonical atom order (with the
graded sign) and merged.  ``emit_plan`` then turns a term list into flat
arrays: the contraction jobs that build every atom from its parent, and, per
output block, the list of (coefficient*shuffle-sign, atom, slot→position)
entries that evaluate the wedge directly at canonical output coordinates.
The device kernels and the CPU port both execute these arrays.

Tensor convention.  An atom with remaining slot groups ``g_1 <= ... <= g_r``
is stored padded, ``n^r`` entries, antisymmetric within each group, such
that the element is ``(1/prod k_g!) sum x[q_1..q_r] bar P_q1 ... bar P_qr``.
Then the canonical coefficient of a sorted string is the padded value at the
sorted indices, ``ι_a`` on the first slot ``s`` of group ``g`` is
``(-1)^s sum_q a[q] x[.., q@s, ..]``, and a wedge evaluated at a canonical
coordinate is the sum over group-respecting shuffles of the permutation sign
times the product of atom values.
"""

from __future__ import annotations

from collections import defaultdict
from dataclasses import dataclass, field
from itertools import combinations

import numpy as np

from .lowering import STATIC_KEYS, BlockKey, FactorizedModeProduct, LinearModeOperator

N_GROUPS = 4
MAX_RANK = 5  # highest rank of an atom or output block (the rank-five shell of the S135 response)
# U: constant creator vector, T: amplitude block, C: creation tensor (a merged
# sum of wedges of U/T atoms), D: tangent block, X: EOM input block.
BASE_ORDER = {"U": 0, "T": 1, "C": 2, "D": 3, "X": 4}


def key_slots(key: BlockKey) -> tuple[int, ...]:
    return tuple(g for g, count in enumerate(key) for _ in range(count))


@dataclass(frozen=True, order=True)
class Atom:
    """Constant creator ``U`` or amplitude/tangent/input block with contractions."""

    base: str
    ident: int  # vector id for U, block id for T/D/X
    slots: tuple[int, ...]  # base slot groups
    contractions: tuple[tuple[int, int], ...] = ()  # (group, vector id), innermost first

    def sort_key(self):
        return (BASE_ORDER[self.base], self.ident, self.contractions)

    @property
    def remaining(self) -> tuple[int, ...]:
        left = list(self.slots)
        for g, _ in self.contractions:
            left.remove(g)
        return tuple(left)

    @property
    def rank(self) -> int:
        return len(self.slots) - len(self.contractions)

    def contract(self, group: int, vec: int) -> "Atom | None":
        if group not in self.remaining:
            return None
        return Atom(self.base, self.ident, self.slots, self.contractions + ((group, vec),))

    def parent(self) -> "Atom":
        return Atom(self.base, self.ident, self.slots, self.contractions[:-1])


Term = tuple[Atom, ...]


class VectorTable:
    """Deduplicated complex n-vectors referenced by ``U`` atoms and contractions."""

    def __init__(self, n: int):
        self.n = n
        self.vectors: list[np.ndarray] = []
        self._index: dict[bytes, int] = {}

    def add(self, vector: np.ndarray) -> int:
        v = np.ascontiguousarray(vector, dtype=np.complex128)
        key = v.tobytes()
        if key not in self._index:
            self._index[key] = len(self.vectors)
            self.vectors.append(v)
        return self._index[key]

    def array(self) -> np.ndarray:
        if not self.vectors:
            return np.zeros((0, self.n), dtype=np.complex128)
        return np.stack(self.vectors)


@dataclass(frozen=True)
class LinearData:
    creators: tuple[tuple[int, int], ...]  # (group, vector id)
    annihilators: tuple[tuple[int, int], ...]


def linear_data(op: LinearModeOperator, n: int, table: VectorTable) -> LinearData:
    """Split a linear PH form into per-group creator and annihilator vectors."""

    create = defaultdict(lambda: np.zeros(n, dtype=np.complex128))
    annihilate = defaultdict(lambda: np.zeros(n, dtype=np.complex128))
    for kind, mode, coefficient in op.components:
        g, local = divmod(mode, n)
        (create if kind == "create" else annihilate)[g][local] += coefficient
    creators = tuple((g, table.add(v)) for g, v in sorted(create.items()) if np.any(v != 0))
    annihilators = tuple((g, table.add(v)) for g, v in sorted(annihilate.items()) if np.any(v != 0))
    return LinearData(creators, annihilators)


def _canonical(atoms: Term) -> tuple[int, Term] | None:
    """Graded-commutative sort of atoms; ``None`` when the wedge vanishes."""

    items = list(atoms)
    sign = 1
    for i in range(len(items)):
        for j in range(len(items) - 1 - i):
            a, b = items[j], items[j + 1]
            if a.sort_key() > b.sort_key():
                items[j], items[j + 1] = b, a
                if (a.rank * b.rank) % 2:
                    sign = -sign
    for a, b in zip(items, items[1:]):
        if a == b and a.rank % 2 == 1:
            return None
    return sign, tuple(items)


def _add(store: dict, atoms: Term, coefficient: complex) -> None:
    if coefficient == 0.0:
        return
    canon = _canonical(atoms)
    if canon is None:
        return
    sign, ordered = canon
    store[ordered] += sign * coefficient


def _rank(atoms: Term) -> int:
    return sum(a.rank for a in atoms)


class TermCompiler:
    """Expand products of transformed linear forms into merged wedge terms."""

    def __init__(self, n: int, products: tuple[FactorizedModeProduct, ...], t_keys=STATIC_KEYS):
        self.n = n
        self.table = VectorTable(n)
        self.t_keys = tuple(t_keys)
        self._linear: dict[tuple, LinearData] = {}
        self.products = []
        for record in products:
            factors = tuple(self._lin(f) for f in record.factors)
            self.products.append((complex(record.coefficient), factors))

    def _lin(self, op: LinearModeOperator) -> LinearData:
        if op.components not in self._linear:
            self._linear[op.components] = linear_data(op, self.n, self.table)
        return self._linear[op.components]

    def _dot(self, a: int, u: int) -> complex:
        return complex(np.dot(self.table.vectors[a], self.table.vectors[u]))

    def apply_factor(self, state: dict, lin: LinearData, rank_limit: int) -> dict:
        out: dict = defaultdict(complex)
        for atoms, coefficient in state.items():
            r = _rank(atoms)
            if r + 1 <= rank_limit:
                for g, v in lin.creators:
                    _add(out, (Atom("U", v, (g,)),) + atoms, coefficient)
            for g, v in lin.annihilators:
                for b, key in enumerate(self.t_keys):
                    if key[g] == 0:
                        continue
                    shift = Atom("T", b, key_slots(key), ((g, v),))
                    if r + shift.rank <= rank_limit:
                        _add(out, (shift,) + atoms, coefficient)
            for g, v in lin.annihilators:
                prefix = 0
                for j, atom in enumerate(atoms):
                    sign = -1.0 if prefix % 2 else 1.0
                    if atom.base == "U":
                        if atom.slots[0] == g:
                            scalar = self._dot(v, atom.ident)
                            _add(out, atoms[:j] + atoms[j + 1 :], coefficient * sign * scalar)
                    else:
                        contracted = atom.contract(g, v)
                        if contracted is not None:
                            _add(out, atoms[:j] + (contracted,) + atoms[j + 1 :], coefficient * sign)
                    prefix += atom.rank
        return {k: v for k, v in out.items() if v != 0.0}

    def product_terms(self, coefficient: complex, factors, initial: Term, max_rank: int) -> dict:
        state = {initial: coefficient}
        count = len(factors)
        for step, lin in enumerate(reversed(factors)):
            remaining = count - step - 1
            state = self.apply_factor(state, lin, max_rank + remaining)
            if not state:
                break
        return state

    def static_terms(self, max_rank: int = 4) -> dict:
        """All terms of ``sum_products [exp(-T) prod F exp(T)]|0>`` up to ``max_rank``."""

        total: dict = defaultdict(complex)
        for coefficient, factors in self.products:
            for atoms, value in self.product_terms(coefficient, factors, (), max_rank).items():
                total[atoms] += value
        return {k: v for k, v in total.items() if v != 0.0}

    def response_terms(self, x_keys: tuple[BlockKey, ...], max_rank: int = 3) -> dict:
        """Connected terms of ``[K̄, X]|0>`` for every input block ``X_b`` (linear in X).

        Terms in which the input atom is never contracted reproduce
        ``X ∧ K̄|0>`` and cancel in the commutator
        (``component_tensor.py:2116-2223`` subtracts them explicitly).
        """

        total: dict = defaultdict(complex)
        for b, key in enumerate(x_keys):
            initial = (Atom("X", b, key_slots(key)),)
            for coefficient, factors in self.products:
                for atoms, value in self.product_terms(coefficient, factors, initial, max_rank).items():
                    if any(a.base == "X" and a.contractions for a in atoms):
                        total[atoms] += value
        return {k: v for k, v in total.items() if v != 0.0}


def tangent_terms(terms: dict) -> dict:
    """Differentiate every ``T`` atom in turn (``D`` marks the tangent block)."""

    out: dict = defaultdict(complex)
    for atoms, coefficient in terms.items():
        for j, atom in enumerate(atoms):
            if atom.base != "T":
                continue
            replaced = atoms[:j] + (Atom("D", atom.ident, atom.slots, atom.contractions),) + atoms[j + 1 :]
            _add(out, replaced, coefficient)
    return {k: v for k, v in out.items() if v != 0.0}


def factor_linear_terms(terms: dict, linear_base: str, first_index: int = 0) -> tuple[dict, list[tuple[BlockKey, dict]]]:
    """Rewrite terms linear in one ``linear_base`` atom as ``C_g ∧ L_g``.

    Every term contains exactly one atom ``L`` of ``linear_base``; moving it
    to the right end costs ``(-1)^(rank L * rank of the atoms after it)``.
    Terms are grouped by ``(L, block key of the rest)``; the rest of each
    group is summed into one creation tensor ``C_g`` (a new ``C`` atom,
    numbered from ``first_index`` so that several factorings share one
    creation program), which depends only on the fixed atoms and is
    materialised once per linearisation point.  Returns the two-factor terms
    and, per creation atom, its key and the terms that define it.
    """

    groups: dict = defaultdict(lambda: defaultdict(complex))
    for atoms, coefficient in terms.items():
        where = [i for i, a in enumerate(atoms) if a.base == linear_base]
        if len(where) != 1:
            raise ValueError("a linear term must contain exactly one input atom")
        j = where[0]
        lin = atoms[j]
        rest = atoms[:j] + atoms[j + 1 :]
        after = sum(a.rank for a in atoms[j + 1 :])
        sign = -1 if (lin.rank * after) % 2 else 1
        groups[(lin, term_key(rest))][rest] += sign * coefficient
    factored: dict = {}
    creations: list[tuple[BlockKey, dict]] = []
    for (lin, key), rest_terms in groups.items():
        rest_terms = {k: v for k, v in rest_terms.items() if v != 0.0}
        if not rest_terms:
            continue
        c = Atom("C", first_index + len(creations), key_slots(key))
        creations.append((key, rest_terms))
        factored[(c, lin)] = 1.0 + 0.0j
    return factored, creations


def term_key(atoms: Term) -> BlockKey:
    counts = [0] * N_GROUPS
    for a in atoms:
        for g in a.remaining:
            counts[g] += 1
    return tuple(counts)


# ----------------------------------------------------------------------------- plan emission


def _shuffles(atoms: Term, key: BlockKey) -> list[tuple[int, list[list[int]]]]:
    """Group-respecting assignments of output positions to atom slots, with signs."""

    slots = key_slots(key)
    group_positions = {g: [p for p, sg in enumerate(slots) if sg == g] for g in range(N_GROUPS)}
    # per group, the sequence of (atom index, slot index) needing positions
    per_group: dict[int, list[tuple[int, int]]] = {g: [] for g in range(N_GROUPS)}
    for j, a in enumerate(atoms):
        for s, g in enumerate(a.remaining):
            per_group[g].append((j, s))
    options_by_group = []
    for g in range(N_GROUPS):
        demand = per_group[g]
        positions = group_positions[g]
        if len(demand) != len(positions):
            return []
        # split positions among atoms in order of their slots, increasing within an atom
        atom_ids = [j for j, _ in demand]
        counts = defaultdict(int)
        for j in atom_ids:
            counts[j] += 1
        atoms_in_group = sorted(counts)
        assignments = [dict()]
        for j in atoms_in_group:
            new = []
            for assign in assignments:
                used = set(p for ps in assign.values() for p in ps)
                free = [p for p in positions if p not in used]
                for chosen in combinations(free, counts[j]):
                    d = dict(assign)
                    d[j] = list(chosen)
                    new.append(d)
            assignments = new
        options_by_group.append(assignments)
    results = []

    def rec(g, acc):
        if g == N_GROUPS:
            pos = [[] for _ in atoms]
            for assign in acc:
                for j, ps in assign.items():
                    pos[j].extend(ps)
            # slot order within an atom is group-sorted, and groups were visited in order
            order = [p for ps in pos for p in ps]
            inv = 0
            for a in range(len(order)):
                for b in range(a + 1, len(order)):
                    inv += order[a] > order[b]
            results.append((-1 if inv % 2 else 1, pos))
            return
        for assign in options_by_group[g]:
            rec(g + 1, acc + [assign])

    rec(0, [])
    return results


@dataclass
class BlockProgram:
    """Entries evaluating one output block at its canonical coordinates.

    ``target_atom >= 0`` marks a creation program: it evaluates the padded
    tensor of that ``C`` atom at every index tuple of ``key`` (the shuffle
    formula is valid for unsorted and repeated indices, giving the full
    antisymmetric tensor) and writes it into the atom workspace.
    ``scalar_index >= 0`` marks one degree-0 entry of an observable program,
    written to that position of its output vector.
    """

    key: BlockKey
    coefficients: np.ndarray  # complex [E]
    factor_start: np.ndarray  # int32 [E+1]
    factor_atom: np.ndarray  # int32 [F] atom id
    factor_positions: np.ndarray  # int32 [F, MAX_RANK] output positions per slot (-1 unused)
    target_atom: int = -1
    scalar_index: int = -1


@dataclass
class Plan:
    """Flat evaluation program shared by several actions.

    ``atoms[i]`` describes atom ``i`` stored at ``atom_offsets[i]`` of one
    complex workspace.  ``jobs[base]`` lists, per contraction depth, the jobs
    ``(atom, parent, slot, vector id, sign)`` that build the atoms of one base
    kind; a parent ``-1 - b`` is base block ``b`` itself.  ``programs[name]``
    holds the output block programs of one action.
    """

    n: int
    atoms: list[Atom]
    atom_offsets: np.ndarray
    workspace: int
    jobs: dict[str, list[np.ndarray]]
    programs: dict[str, list[BlockProgram]]
    vectors: np.ndarray
    n_terms: dict[str, int]

    def entries(self, name: str) -> int:
        return int(sum(b.coefficients.size for b in self.programs[name]))


def _block_program(term_list, key, index, target_atom=-1, scalar_index=-1) -> BlockProgram:
    coefficients, starts, f_atom, f_pos = [], [0], [], []
    for term_atoms, coefficient in term_list:
        for sign, positions in _shuffles(term_atoms, key):
            coefficients.append(sign * coefficient)
            for a, ps in zip(term_atoms, positions):
                f_atom.append(index[a])
                f_pos.append(ps + [-1] * (MAX_RANK - len(ps)))
            starts.append(len(f_atom))
    return BlockProgram(
        key,
        np.asarray(coefficients, dtype=np.complex128),
        np.asarray(starts, dtype=np.int32),
        np.asarray(f_atom, dtype=np.int32),
        np.asarray(f_pos, dtype=np.int32).reshape(-1, MAX_RANK),
        target_atom,
        scalar_index,
    )


def emit_plan(
    actions: dict[str, tuple[dict, tuple[BlockKey, ...]]],
    n: int,
    vectors: np.ndarray,
    creations: list[tuple[BlockKey, dict]] | None = None,
    scalars: list[dict] | None = None,
) -> Plan:
    """Lower merged term dictionaries (one per action) into flat arrays.

    ``creations[i]`` defines ``C`` atom ``i`` (its block key and the terms it
    sums); it becomes the ``creation`` program, run after the fixed atoms are
    loaded and before any program that reads ``C`` atoms.  ``scalars[i]``
    holds the degree-0 terms of observable ``i``; together they form the
    ``observable`` program, whose output ``i`` is that sum.
    """

    creations = creations or []
    scalars = scalars or []
    needed: dict[Atom, None] = {}

    def need(a: Atom) -> None:
        chain = a
        needed[chain] = None
        while chain.contractions:
            chain = chain.parent()
            if chain.contractions:
                needed[chain] = None

    by_action: dict[str, dict[BlockKey, list[tuple[Term, complex]]]] = {}
    for name, (terms, out_keys) in actions.items():
        by_key: dict[BlockKey, list[tuple[Term, complex]]] = defaultdict(list)
        for atoms, coefficient in terms.items():
            k = term_key(atoms)
            if k in out_keys:
                by_key[k].append((atoms, coefficient))
                for a in atoms:
                    need(a)
        by_action[name] = by_key
    for key, rest_terms in creations:
        for atoms in rest_terms:
            for a in atoms:
                need(a)
    for terms in scalars:
        for atoms in terms:
            for a in atoms:
                need(a)
    atoms = sorted(needed, key=lambda a: (len(a.contractions), a.sort_key()))
    index = {a: i for i, a in enumerate(atoms)}
    offsets = np.zeros(len(atoms), dtype=np.int64)
    cursor = 0
    for i, a in enumerate(atoms):
        offsets[i] = cursor
        cursor += n ** a.rank
    jobs: dict[str, list[np.ndarray]] = {}
    for base in ("T", "D", "X"):
        chosen = [a for a in atoms if a.base == base and a.contractions]
        depth_max = max((len(a.contractions) for a in chosen), default=0)
        levels = []
        for depth in range(1, depth_max + 1):
            rows = []
            for a in chosen:
                if len(a.contractions) != depth:
                    continue
                g, v = a.contractions[-1]
                parent = a.parent()
                slot = parent.remaining.index(g)
                pid = index[parent] if parent.contractions else -1 - parent.ident
                rows.append((index[a], pid, slot, v, -1 if slot % 2 else 1))
            levels.append(np.asarray(rows, dtype=np.int64).reshape(-1, 5))
        jobs[base] = levels
    programs: dict[str, list[BlockProgram]] = {}
    n_terms: dict[str, int] = {}
    for name, (terms, out_keys) in actions.items():
        programs[name] = [_block_program(by_action[name][k], k, index) for k in out_keys]
        n_terms[name] = sum(len(by_action[name][k]) for k in out_keys)
    if creations:
        blocks = []
        for i, (key, rest_terms) in enumerate(creations):
            target = index[Atom("C", i, key_slots(key))]
            blocks.append(_block_program(list(rest_terms.items()), key, index, target))
        programs["creation"] = blocks
        n_terms["creation"] = sum(len(t) for _, t in creations)
    if scalars:
        zero = (0, 0, 0, 0)
        programs["observable"] = [
            _block_program([(k, v) for k, v in terms.items() if term_key(k) == zero], zero, index, scalar_index=i)
            for i, terms in enumerate(scalars)
        ]
        n_terms["observable"] = sum(len(t) for t in scalars)
    return Plan(n, atoms, offsets, int(cursor), jobs, programs, vectors, n_terms)
