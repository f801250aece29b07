"""Jacobian sparsity: the caller's pattern, its column colouring, and where each
entry of the iteration matrix is stored.

A forward-mode sweep of the right-hand side seeded with a direction ``v``
returns ``J v``. If every column in a group of columns has its nonzeros in rows
no other column of the group touches, one sweep seeded with the sum of the
group's unit vectors returns all of them at once, and the pattern says which
row value belongs to which column. Grouping the columns into few such sets is
vertex colouring of the graph whose edges join columns sharing a row.

Interface kept from the reference (modax/_sparsity.py):
  - ``normalize_sparsity`` (:150-179): a mask, a scipy sparse matrix or
    ``(nnz, 2)`` coordinate pairs, in one canonical hashable form;
  - ``colour_sparsity`` (:182-210): the five networkx greedy strategies in the
    reference's order, the first one with the fewest colours kept;
  - ``CompressedJacobian`` (:42-140) and ``dense_jacobian`` (:143-147): the
    colouring plus where each entry lives, with the same public methods;
  - ``check_orthogonal`` (:213-230): a colouring that would put two entries of
    one row into one slot is rejected.

The storage side is its own object here (``SlotStorage``): the dense grid
needs none, and a sparse factorisation supplies one from its symbolic analysis
(``menodax_amd._sparse_direct.compressed_jacobian``).
"""

from __future__ import annotations

import functools
import itertools
from dataclasses import dataclass

import numpy as np

#: networkx greedy-colouring strategies, tried in this order; the first with
#: the fewest colours wins (the reference's choice, kept for equal colourings)
STRATEGIES = (
    "largest_first",
    "smallest_last",
    "independent_set",
    "saturation_largest_first",
    "DSATUR",
)


@dataclass(frozen=True)
class SlotStorage:
    """Storage of the iteration matrix when it is not the colour grid.

    ``grid[row * n_colours + g]`` is the slot of the entry of row ``row`` whose
    column has colour ``g`` (-1 when the row has no such declared entry);
    ``diagonal[i]`` the slot of ``(i, i)``, declared or not (``I/(h gamma)``
    lands on every diagonal); ``size`` the number of slots, which counts slots
    no Jacobian entry owns (a factorisation's fill-in)."""

    grid: tuple
    diagonal: tuple
    size: int


@dataclass(frozen=True)
class CompressedJacobian:
    """A column colouring of the Jacobian and the matrix slots its sweeps fill.

    Without ``storage`` the matrix is the colour grid: entry ``(r, c)`` at
    ``r * n_colours + colour[c]``. With every column its own colour that is the
    row-major dense matrix, so the dense solver is the uninformative end of the
    same layout rather than a special case."""

    n_vars: int
    n_colours: int
    colour: tuple
    storage: SlotStorage | None = None

    @property
    def size(self) -> int:
        """Elements of one trajectory's matrix."""
        return self.n_vars * self.n_colours if self.storage is None else self.storage.size

    @property
    def diagonal(self) -> tuple:
        """Slot of each diagonal entry (where ``1/(h gamma)`` is added)."""
        if self.storage is not None:
            return tuple(self.storage.diagonal)
        return tuple(i * self.n_colours + self.colour[i] for i in range(self.n_vars))

    def slot(self, row: int, col: int) -> int:
        """Slot of entry ``(row, col)``; a packed layout has none for an
        undeclared off-diagonal entry."""
        position = row * self.n_colours + self.colour[col]
        if self.storage is None:
            return position
        if row == col:
            return self.storage.diagonal[row]
        found = self.storage.grid[position]
        if found < 0:
            raise ValueError(
                f"entry ({row}, {col}) is not in the pattern, so a packed layout has no "
                "slot for it; declare it in the sparsity pattern")
        return found

    def store_slots(self):
        """Per grid position the slot its sweep value goes to (-1: not
        written), or None on the grid, where the positions are the slots."""
        if self.storage is None:
            return None
        return np.asarray(self.storage.grid, dtype=np.int32)

    def seed_table(self, min_width: int = 0) -> np.ndarray:
        """Sweep directions: row ``g`` is the indicator of colour ``g``; one
        extra zero row; at least ``min_width`` columns."""
        seeds = np.zeros((self.n_colours + 1, max(self.n_vars, min_width)))
        seeds[np.asarray(self.colour, dtype=np.int64), np.arange(self.n_vars)] = 1.0
        return seeds

    def stores(self):
        """``(row, colour, slot)`` of every value a sweep writes, colour by
        colour with rows ascending (the order the generated sweeps write in)."""
        n, c = self.n_vars, self.n_colours
        if self.storage is None:
            return [(row, g, row * c + g) for g in range(c) for row in range(n)]
        grid = self.storage.grid
        return [(row, g, grid[row * c + g]) for g in range(c) for row in range(n)
                if grid[row * c + g] >= 0]

    @property
    def needs_clear(self) -> bool:
        """True when some slot is written by no sweep (fill-in, an undeclared
        diagonal): the matrix must then be zeroed before the sweeps
        (modax/rodas5P.py:375-378)."""
        if self.storage is None:
            return False
        return self.storage.size > sum(1 for s in self.storage.grid if s >= 0)


def dense_jacobian(n_vars: int) -> CompressedJacobian:
    """Every column its own colour: the dense row-major matrix."""
    return CompressedJacobian(n_vars=n_vars, n_colours=n_vars, colour=tuple(range(n_vars)))


def _coordinates(sparsity, n_vars: int):
    """(rows, cols) of the declared entries, whatever the spelling."""
    if hasattr(sparsity, "tocoo"):
        coo = sparsity.tocoo()
        return np.asarray(coo.row), np.asarray(coo.col)
    arr = np.asarray(sparsity)
    if arr.ndim == 2 and arr.shape == (n_vars, n_vars):
        return np.nonzero(arr)
    if arr.ndim == 2 and arr.shape[1] == 2:
        return arr[:, 0], arr[:, 1]
    raise ValueError(
        "sparsity must be an (n_vars, n_vars) mask, a scipy sparse matrix, or an (nnz, 2) "
        f"array of (row, col); got shape {arr.shape} for n_vars={n_vars}")


def normalize_sparsity(sparsity, n_vars: int) -> tuple:
    """The declared entries as one tuple per row of its sorted, distinct
    columns: hashable, so kernels and analyses cache on it."""
    rows, cols = (np.asarray(a).astype(np.int64).ravel() for a in _coordinates(sparsity, n_vars))
    outside = (rows < 0) | (rows >= n_vars) | (cols < 0) | (cols >= n_vars)
    if outside.any():
        k = int(np.argmax(outside))
        raise ValueError(
            f"sparsity entry ({rows[k]}, {cols[k]}) lies outside an {n_vars}x{n_vars} Jacobian")
    keys = np.unique(rows * n_vars + cols)
    key_rows, key_cols = np.divmod(keys, n_vars)
    bounds = np.searchsorted(key_rows, np.arange(n_vars + 1))
    return tuple(tuple(int(c) for c in key_cols[bounds[r]:bounds[r + 1]]) for r in range(n_vars))


def _conflict_graph(pattern):
    """Columns as nodes, an edge wherever two columns share a row. Nodes are
    added in index order and edges row by row in ascending column pairs:
    networkx's greedy strategies break ties by that insertion order."""
    import networkx as nx

    graph = nx.Graph()
    graph.add_nodes_from(range(len(pattern)))
    graph.add_edges_from(itertools.chain.from_iterable(
        itertools.combinations(cols, 2) for cols in pattern))
    return graph


@functools.lru_cache(maxsize=None)
def colour_sparsity(pattern: tuple) -> CompressedJacobian:
    """The fewest-colour greedy colouring of the pattern's conflict graph."""
    import networkx as nx

    graph = _conflict_graph(pattern)
    candidates = [nx.coloring.greedy_color(graph, strategy=s) for s in STRATEGIES]
    # min() keeps the first of equally good colourings, i.e. the earliest strategy
    best = min(candidates, key=lambda colouring: max(colouring.values(), default=0))
    colour = tuple(int(best.get(c, 0)) for c in range(len(pattern)))
    compressed = CompressedJacobian(n_vars=len(pattern), n_colours=max(colour, default=0) + 1,
                                    colour=colour)
    check_orthogonal(pattern, compressed)
    return compressed


def check_orthogonal(pattern, compressed: CompressedJacobian) -> None:
    """Reject a colouring with two same-coloured columns in one row: their
    values would land in one sweep component and one slot."""
    colour = np.asarray(compressed.colour, dtype=np.int64)
    for row, cols in enumerate(pattern):
        if len(cols) < 2:
            continue
        cols = np.asarray(cols, dtype=np.int64)
        by_colour = cols[np.argsort(colour[cols], kind="stable")]
        clash = np.flatnonzero(np.diff(colour[by_colour]) == 0)
        if clash.size:
            a, b = int(by_colour[clash[0]]), int(by_colour[clash[0] + 1])
            raise ValueError(
                f"columns {a} and {b} share colour {int(colour[a])} but both have an entry in "
                f"row {row}; the colouring is not structurally orthogonal for this pattern")


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
