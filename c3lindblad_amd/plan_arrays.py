"""Flat arrays of a compiled contraction plan, in the exact form the C ABI takes.

``lower_plan`` places every atom and every padded base block in one complex
workspace and converts the symbolic plan of ``algebra.emit_plan`` into:

* ``bases``  -- ``(kind, offset, deg, src, sign)``: padded block element ``e`` is
  ``sign[e] * v[src[e]]`` of the loaded complex vector (``src < 0`` -> 0).
  The padding enumerates every within-group permutation of a canonical
  coordinate with its parity (the padded-tensor convention of
  ``component_tensor.py:302-337``);
* ``levels`` -- contraction jobs per kind and depth;
* ``blocks`` -- per program and output block: coordinates, destinations and
  the wedge entries (``fcode = rank | pos0<<3 | pos1<<6 | ... | pos4<<15``);
  creation programs write every padded index tuple of a ``C`` atom into the
  workspace (destinations are workspace offsets).

Storage order: a padded rank-r tensor stores index tuple ``(i_0, .., i_{r-1})`` at
``sum_s i_s n^s`` (first index fastest), and the coordinates of an output block
are listed in increasing destination.  The canonical string order (PH rank,
then integer mask) runs over the first index fastest, so consecutive
coordinates of a block write consecutive outputs and read consecutive elements
of the rank-r atoms indexed by all of their positions.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from itertools import permutations

import numpy as np

from .algebra import Plan, key_slots
from .lowering import BlockKey, CoordinateSet

KINDS = {"T": 0, "D": 1, "X": 2}
PROGRAMS = {
    "residual": 0,
    "reference": 1,
    "jvp": 2,
    "sigma": 3,
    "creation": 4,
    "observable": 5,
    "source_addition": 6,
    "source_removal": 7,
    "residual_low": 8,  # degree-0 terms plus 1/6 of the degree-1 terms (Simpson residual)
}


class DenseDest:
    """Destinations ``start .. start+size-1`` of a dense block (every index tuple of a padded
    tensor, in flat order) without storing them; converts to the index array on demand."""

    __slots__ = ("start", "size")

    def __init__(self, start: int, size: int):
        self.start = int(start)
        self.size = int(size)

    def __add__(self, shift) -> "DenseDest":
        return DenseDest(self.start + int(shift), self.size)

    def __array__(self, dtype=None, copy=None):
        return np.arange(self.start, self.start + self.size, dtype=dtype or np.int64)

    def max(self) -> int:
        return self.start + self.size - 1


def encode_factors(ranks: np.ndarray, positions: np.ndarray) -> np.ndarray:
    """``fcode = rank | pos0<<3 | pos1<<6 | pos2<<9 | pos3<<12 | pos4<<15`` (rank <= 5, positions < 5)."""

    code = ranks.astype(np.int32).copy()
    for s in range(positions.shape[1] if positions.size else 0):
        pos = np.where(positions[:, s] >= 0, positions[:, s], 0).astype(np.int32)
        code |= (pos << (3 + 3 * s)) * (s < ranks)
    return code


@dataclass
class LoweredPlan:
    n: int
    workspace: int
    vectors: np.ndarray
    const_offset: np.ndarray
    const_vec: np.ndarray
    bases: list = field(default_factory=list)
    levels: list = field(default_factory=list)
    blocks: list = field(default_factory=list)
    n_terms: dict = field(default_factory=dict)

    def entries(self, program: str) -> int:
        pid = PROGRAMS[program]
        return int(sum(b[4].size for b in self.blocks if b[0] == pid))


_scatter_cache: dict = {}


def padded_scatter(coords: CoordinateSet, key: BlockKey) -> tuple[np.ndarray, np.ndarray]:
    """Index/sign maps from a vector to the padded antisymmetric tensor of one block (shared,
    read-only, per coordinate set and key)."""

    cache_key = (id(coords), key)
    hit = _scatter_cache.get(cache_key)
    if hit is not None and hit[0] is coords:
        return hit[1], hit[2]
    src, sign = _padded_scatter(coords, key)
    src.setflags(write=False)
    sign.setflags(write=False)
    _scatter_cache[cache_key] = (coords, src, sign)
    return src, sign


def _padded_scatter(coords: CoordinateSet, key: BlockKey) -> tuple[np.ndarray, np.ndarray]:
    n = coords.n
    block = coords.block(key)
    deg = sum(key)
    size = n**deg
    src = np.full(size, -1, dtype=np.int64)
    sign = np.zeros(size, dtype=np.int8)
    if block is None or block.positions.size == 0:
        return src, sign
    slots = key_slots(key)
    groups = sorted(set(slots))
    group_slots = [[s for s, g in enumerate(slots) if g == grp] for grp in groups]
    idx = block.indices
    strides = n ** np.arange(deg, dtype=np.int64)  # first index fastest
    per_group_perms = []
    for gs in group_slots:
        perms = []
        for perm in permutations(range(len(gs))):
            inv = sum(1 for a in range(len(perm)) for b in range(a + 1, len(perm)) if perm[a] > perm[b])
            perms.append(([gs[p] for p in perm], -1 if inv % 2 else 1))
        per_group_perms.append(perms)

    def rec(level, mapping, parity):
        if level == len(group_slots):
            cols = np.empty((idx.shape[0], deg), dtype=np.int64)
            for target, source in mapping.items():
                cols[:, target] = idx[:, source]
            flat = cols @ strides
            src[flat] = block.positions
            sign[flat] = parity
            return
        gs = group_slots[level]
        for perm, s in per_group_perms[level]:
            m = dict(mapping)
            for target, source in zip(gs, perm):
                m[target] = source
            rec(level + 1, m, parity * s)

    rec(0, {}, 1)
    return src, sign


def lower_plan(
    plan: Plan,
    base_sets: dict[str, tuple[CoordinateSet, tuple[BlockKey, ...]]],
    out_sets: dict[str, CoordinateSet],
) -> LoweredPlan:
    """Lay out a symbolic plan for execution.

    ``base_sets[kind]`` is the coordinate set and ordered block keys of the
    vector loaded for atoms of that kind; ``out_sets[program]`` the
    coordinate set of the program's output vector.
    """

    n = plan.n
    cursor = plan.workspace
    base_offset: dict[tuple[str, int], int] = {}
    bases = []
    for kind, (coords, keys) in base_sets.items():
        for b, key in enumerate(keys):
            src, sign = padded_scatter(coords, key)
            base_offset[(kind, b)] = cursor
            bases.append((KINDS[kind], cursor, sum(key), src, sign))
            cursor += src.size
    const_offset = []
    const_vec = []
    for i, atom in enumerate(plan.atoms):
        if atom.base == "U":
            const_offset.append(plan.atom_offsets[i])
            const_vec.append(atom.ident)
    levels = []
    for kind, per_depth in plan.jobs.items():
        for jobs in per_depth:
            if jobs.size == 0:
                continue
            out_off = plan.atom_offsets[jobs[:, 0]]
            par_off = np.empty(jobs.shape[0], dtype=np.int64)
            par_rank = np.empty(jobs.shape[0], dtype=np.int32)
            for j, (a, pid, _, _, _) in enumerate(jobs):
                if pid >= 0:
                    par_off[j] = plan.atom_offsets[pid]
                else:
                    par_off[j] = base_offset[(kind, -1 - int(pid))]
                par_rank[j] = plan.atoms[a].rank + 1
            levels.append(
                (
                    KINDS[kind],
                    out_off.astype(np.int64),
                    par_off,
                    par_rank,
                    jobs[:, 2].astype(np.int32),
                    jobs[:, 3].astype(np.int32),
                    jobs[:, 4].astype(np.int32),
                )
            )
    targets = {bp.target_atom for bp in plan.programs.get("creation", [])}
    for bp in plan.programs.get("creation", []):
        if targets.intersection(int(a) for a in bp.factor_atom):
            raise ValueError("a creation tensor reads another creation tensor: one program launch cannot order them")
    blocks = []
    dense_coords: dict[int, tuple[np.ndarray, np.ndarray]] = {}  # every index tuple of a padded rank-deg tensor

    def dense(deg: int):
        if deg not in dense_coords:
            flat = np.arange(n**deg, dtype=np.int64)
            coord = (np.stack([(flat // n**d) % n for d in range(deg)], axis=1).astype(np.int32) if deg
                     else np.zeros((1, 0), np.int32))
            flat.setflags(write=False)
            coord.setflags(write=False)
            dense_coords[deg] = (flat, coord)
        return dense_coords[deg]

    for name, programs in plan.programs.items():
        coords = out_sets.get(name)
        for bp in programs:
            deg = sum(bp.key)
            if bp.target_atom >= 0:
                # creation tensor: every index tuple of its padded storage, written in place
                flat, coord = dense(deg)
                dest = DenseDest(plan.atom_offsets[bp.target_atom], flat.size)
            elif deg == 0:
                coord = np.zeros((1, 0), dtype=np.int32)
                dest = np.full(1, max(bp.scalar_index, 0), dtype=np.int64)
            else:
                block = coords.block(bp.key)
                if block is None or block.positions.size == 0:
                    continue
                order = np.argsort(block.positions, kind="stable")  # walk the output vector in order
                coord = block.indices[order].astype(np.int32)
                dest = block.positions[order].astype(np.int64)
            fcode = encode_factors(np.array([plan.atoms[a].rank for a in bp.factor_atom], dtype=np.int32),
                                   bp.factor_positions)
            foffset = plan.atom_offsets[bp.factor_atom] if bp.factor_atom.size else np.zeros(0, dtype=np.int64)
            blocks.append(
                (
                    PROGRAMS[name],
                    deg,
                    coord if coord.flags.c_contiguous else np.ascontiguousarray(coord),
                    dest,
                    bp.coefficients,
                    bp.factor_start,
                    foffset.astype(np.int64),
                    fcode.astype(np.int32),
                )
            )
    return LoweredPlan(
        n=n,
        workspace=int(cursor),
        vectors=plan.vectors,
        const_offset=np.asarray(const_offset, dtype=np.int64),
        const_vec=np.asarray(const_vec, dtype=np.int32),
        bases=bases,
        levels=levels,
        blocks=blocks,
        n_terms=dict(plan.n_terms),
    )


WORKSPACE_PROGRAMS = (PROGRAMS["creation"],)


def union_plan(parts: list[LoweredPlan], in_sizes: dict[int, int], out_sizes: dict[int, int]) -> LoweredPlan:
    """One plan evaluating several independent plans of equal ``n`` at once.

    ``parts[i]`` reads its input vectors at ``[i * in_sizes[kind], (i+1) * in_sizes[kind])`` of the
    concatenated input of that kind and writes program outputs at
    ``[i * out_sizes[program], ...)`` of the concatenated output (workspace programs write their
    own shifted workspace).  Workspaces, vector tables and constant atoms are concatenated; the
    contraction levels of every kind are merged depth by depth, so each level launch builds that
    depth's atoms of every part; blocks keep their per-part coefficients.  No part reads another
    part's atoms, so the union's results are exactly the parts' results side by side.
    """

    if not parts:
        raise ValueError("a union needs at least one plan")
    n = parts[0].n
    if any(p.n != n for p in parts):
        raise ValueError("united plans must share the site count")
    ws_base = np.cumsum([0] + [p.workspace for p in parts])
    vec_base = np.cumsum([0] + [p.vectors.shape[0] for p in parts])
    bases, blocks = [], []
    const_offset, const_vec = [], []
    per_kind_levels: dict[int, list[list]] = {}
    for i, p in enumerate(parts):
        w0, v0 = int(ws_base[i]), int(vec_base[i])
        const_offset.append(np.asarray(p.const_offset, dtype=np.int64) + w0)
        const_vec.append(np.asarray(p.const_vec, dtype=np.int32) + v0)
        for kind, offset, deg, src, sign in p.bases:
            shift = i * in_sizes[kind]
            bases.append((kind, int(offset) + w0, deg, np.where(src >= 0, src + shift, -1).astype(np.int64), sign))
        depth: dict[int, int] = {}
        for kind, out_off, par_off, par_rank, slot, vec, sign in p.levels:
            d = depth.get(kind, 0)
            depth[kind] = d + 1
            levels = per_kind_levels.setdefault(kind, [])
            while len(levels) <= d:
                levels.append([])
            levels[d].append((out_off + w0, par_off + w0, par_rank, slot, vec + v0, sign))
        for program, deg, coord, dest, coeff, fstart, foff, fcode in p.blocks:
            shift = w0 if program in WORKSPACE_PROGRAMS else i * out_sizes[program]
            blocks.append((program, deg, coord, dest + shift, coeff, fstart, foff + w0, fcode))
    levels = []
    for kind in sorted(per_kind_levels):
        for jobs in per_kind_levels[kind]:
            levels.append((kind, *[np.concatenate([j[k] for j in jobs]) for k in range(6)]))
    return LoweredPlan(
        n=n,
        workspace=int(ws_base[-1]),
        vectors=np.concatenate([p.vectors for p in parts], axis=0),
        const_offset=np.concatenate(const_offset),
        const_vec=np.concatenate(const_vec),
        bases=bases,
        levels=levels,
        blocks=blocks,
        n_terms={},
    )
