"""Sparsity patterns and the column colouring of the Jacobian.

Forward-mode differentiation gives ``J v`` per sweep. Columns that share no row
are structurally orthogonal, so one sweep seeded with the indicator of a whole
group of them returns every column of the group; choosing few groups is vertex
colouring of the column intersection graph. Restates modax/_sparsity.py:
  - ``normalize_sparsity``: mask, scipy sparse matrix or (nnz, 2) pairs to sorted
    row tuples, range-checked (:150-179);
  - ``colour_sparsity``: networkx greedy colouring, five strategies tried in the
    reference's order, the first strictly fewer colours kept (:33-39, 182-210);
  - ``CompressedJacobian``: slot of entry (r, c) = r * n_colours + colour[c] on
    the grid, or the packed (CSR) slot a sparse layout assigns (:42-140);
  - the structural-orthogonality check (:213-230).
"""

from __future__ import annotations

import functools
from dataclasses import dataclass

import numpy as np

STRATEGIES = (
    "largest_first",
    "smallest_last",
    "independent_set",
    "saturation_largest_first",
    "DSATUR",
)


@dataclass(frozen=True)
class CompressedJacobian:
    """Where each Jacobian entry lives in one trajectory's matrix buffer, and
    which colour sweep produces it. With no pattern every column is its own
    colour and the layout is the dense row-major matrix."""

    n_vars: int
    n_colours: int
    colour: tuple
    packed: tuple | None = None
    packed_diagonal: tuple | None = None
    n_slots: int | None = None

    @property
    def size(self) -> int:
        if self.n_slots is not None:
            return self.n_slots
        if self.packed is not None:
            return 1 + max(max(self.packed), max(self.packed_diagonal))
        return self.n_vars * self.n_colours

    @property
    def diagonal(self) -> tuple:
        return tuple(self.slot(i, i) for i in range(self.n_vars))

    def slot(self, row: int, col: int) -> int:
        grid = row * self.n_colours + self.colour[col]
        if self.packed is None:
            return grid
        s = self.packed[grid]
        if s < 0 and row == col:
            return self.packed_diagonal[row]
        if s < 0:
            raise ValueError(
                f"entry ({row}, {col}) is not in the pattern, so a packed layout has "
                "no slot for it; declare it in the sparsity pattern"
            )
        return s

    def store_slots(self):
        if self.packed is None:
            return None
        return np.asarray(self.packed, dtype=np.int32)

    def seed_table(self, min_width: int = 0) -> np.ndarray:
        """Row g is the indicator of colour g; the extra last row is zero."""
        width = max(self.n_vars, min_width)
        seeds = np.zeros((self.n_colours + 1, width))
        for c, g in enumerate(self.colour):
            seeds[g, c] = 1.0
        return seeds

    def stores(self, pattern=None):
        """``(row, colour, slot)`` for every slot a sweep writes, in the
        reference's write order (colour-major, rows ascending).

        The grid writes every ``(row, colour)`` position; a packed layout only
        the positions some declared entry claims."""
        out = []
        for g in range(self.n_colours):
            for row in range(self.n_vars):
                if self.packed is None:
                    out.append((row, g, row * self.n_colours + g))
                else:
                    s = self.packed[row * self.n_colours + g]
                    if s >= 0:
                        out.append((row, g, s))
        return out

    @property
    def needs_clear(self) -> bool:
        """A buffer with slots no sweep writes (fill-in, undeclared diagonal)
        must be zeroed before the sweeps (modax/rodas5P.py:375-378)."""
        if self.packed is None:
            return False
        claimed = sum(1 for s in self.packed if s >= 0)
        return self.size > claimed


def dense_jacobian(n_vars: int) -> CompressedJacobian:
    return CompressedJacobian(n_vars=n_vars, n_colours=n_vars, colour=tuple(range(n_vars)))


def normalize_sparsity(sparsity, n_vars: int) -> tuple:
    """A mask, a scipy sparse matrix or ``(row, col)`` pairs, as a hashable tuple
    of sorted column tuples per row."""
    if hasattr(sparsity, "tocoo"):
        coo = sparsity.tocoo()
        pairs = zip(coo.row.tolist(), coo.col.tolist())
    else:
        arr = np.asarray(sparsity)
        if arr.ndim == 2 and arr.shape == (n_vars, n_vars):
            rows, cols = np.nonzero(arr)
            pairs = zip(rows.tolist(), cols.tolist())
        elif arr.ndim == 2 and arr.shape[1] == 2:
            pairs = ((int(r), int(c)) for r, c in arr.tolist())
        else:
            raise ValueError(
                "sparsity must be an (n_vars, n_vars) mask, a scipy sparse matrix, or "
                f"an (nnz, 2) array of (row, col); got shape {arr.shape} for n_vars={n_vars}"
            )
    by_row = [set() for _ in range(n_vars)]
    for r, c in pairs:
        if not (0 <= r < n_vars and 0 <= c < n_vars):
            raise ValueError(f"sparsity entry ({r}, {c}) lies outside an {n_vars}x{n_vars} Jacobian")
        by_row[r].add(c)
    return tuple(tuple(sorted(cs)) for cs in by_row)


@functools.lru_cache(maxsize=None)
def colour_sparsity(pattern: tuple) -> CompressedJacobian:
    """Greedy colouring of the column intersection graph, fewest colours wins."""
    import networkx as nx

    n_vars = len(pattern)
    graph = nx.Graph()
    graph.add_nodes_from(range(n_vars))
    for cols in pattern:
        for a in range(len(cols)):
            for b in range(a + 1, len(cols)):
                graph.add_edge(cols[a], cols[b])
    best = None
    for strategy in STRATEGIES:
        colouring = nx.coloring.greedy_color(graph, strategy=strategy)
        if best is None or max(colouring.values()) < max(best.values()):
            best = colouring
    colour = tuple(int(best.get(c, 0)) for c in range(n_vars))
    compressed = CompressedJacobian(n_vars=n_vars, n_colours=max(colour) + 1, colour=colour)
    check_orthogonal(pattern, compressed)
    return compressed


def check_orthogonal(pattern, compressed: CompressedJacobian) -> None:
    """No row may hold two columns of the same colour (they would share a slot)."""
    for row, cols in enumerate(pattern):
        seen = {}
        for c in cols:
            g = compressed.colour[c]
            if g in seen:
                raise ValueError(
                    f"columns {seen[g]} and {c} share colour {g} but both have an entry "
                    f"in row {row}; the colouring is not structurally orthogonal for this pattern"
                )
            seen[g] = c


def check_superset(pattern: tuple, reads: np.ndarray) -> None:
    """The declared pattern must cover every entry the right-hand side reads.

    The reference documents this requirement (a subset silently corrupts
    entries whose columns share a colour, modax/rodas5P.py:1136-1140); here the
    traced callback says exactly which components each output reads, so a
    pattern that misses one is rejected instead."""
    n = len(pattern)
    missing = [
        (r, c) for r in range(n) for c in np.flatnonzero(reads[r]) if int(c) not in set(pattern[r])
    ]
    if missing:
        raise ValueError(
            f"the sparsity pattern omits {len(missing)} Jacobian entries the right-hand "
            f"side depends on, e.g. {missing[:5]}; the pattern must be a superset of them"
        )
