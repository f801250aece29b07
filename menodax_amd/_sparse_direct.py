"""A sparse direct solver compiled for one sparsity pattern.

The reference's interface (modax/_sparse_direct.py) is kept -- ``ORDERINGS``,
``fill_reducing_order`` (:91-152), ``fill_pattern`` (:155-184), ``analyse``
and ``SparseLULayout`` (:187-349), ``compressed_jacobian`` (:352-390), the
solver object and its cached constructors (:393-452), the unrolling
thresholds (:468-469) -- over this module's own construction:

  1. order    -- a symmetric fill-reducing permutation (``_amd.amd_order``) or
                 the natural order; the diagonal stays on the diagonal, where
                 I/(h gamma) guarantees a structural pivot;
  2. symbolic -- one up-looking pass over the permuted rows: row i of L+U is
                 row i of the pattern plus, for each column k < i that it
                 holds or gains on the way (taken in increasing order from a
                 heap), the columns right of the diagonal of the finished row
                 k. The same pass records, for every multiplier slot, the slots
                 of row i its update writes, so no later search is needed;
  3. layout   -- the rows concatenated in elimination order as one CSR image,
                 row and column origins translating back to the caller's
                 indices so the device code never permutes a vector;
  4. compile  -- the factorisation (inverse diagonal stored) and the two
                 triangular sweeps as HIP, straight-line with literal slots
                 below the thresholds and loops over ``__constant__`` index
                 tables above them.

There is no pivoting: the same pattern serves every trajectory, and Rodas5P
being a W-method, an inexact factorisation costs step-size control rather than
order (the controller rejects a step whose factorisation went bad).
"""

from __future__ import annotations

import bisect
import functools
import heapq
import itertools
from dataclasses import replace

import numpy as np

from ._amd import amd_order
from ._sparsity import CompressedJacobian, SlotStorage, colour_sparsity, normalize_sparsity

ORDERINGS = ("amd", "natural")

#: above these many multiply-adds the routines are table driven
#: (modax/_sparse_direct.py:468-469)
MAX_UNROLLED_SUBSTITUTIONS = 2048
MAX_UNROLLED_UPDATES = 8192


def fill_reducing_order(pattern, ordering: str = "amd") -> tuple:
    """Elimination order as caller indices: position i is eliminated i-th."""
    if ordering not in ORDERINGS:
        raise ValueError(f"unknown ordering {ordering!r}; expected one of {ORDERINGS}")
    if ordering == "natural":
        return tuple(range(len(pattern)))
    return amd_order(pattern)


def _symbolic_rows(rows):
    """Up-looking symbolic LU without pivoting.

    ``rows[i]`` holds the columns of row i (elimination indices). Returns each
    row's sorted L+U columns, diagonal included. Row i's lower columns are
    visited in increasing order; visiting k adds the finished row k's columns
    right of k, and any of those left of i is visited in turn. This is the
    least pattern closed under elimination -- what an unpivoted factorisation
    writes, and what elimination on the boolean pattern produces."""
    finished = []
    right_of_diagonal = []
    for i, given in enumerate(rows):
        members = set(given)
        members.add(i)
        pending = [k for k in members if k < i]
        heapq.heapify(pending)
        while pending:
            k = heapq.heappop(pending)
            for j in right_of_diagonal[k]:
                if j not in members:
                    members.add(j)
                    if j < i:
                        heapq.heappush(pending, j)
        cols = sorted(members)
        finished.append(cols)
        right_of_diagonal.append(cols[bisect.bisect_right(cols, i):])
    return finished


def fill_pattern(pattern) -> np.ndarray:
    """The exact L+U pattern of an unpivoted LU of ``pattern`` in its own
    order (diagonal included), as an (n, n) boolean matrix."""
    n = len(pattern)
    filled = np.zeros((n, n), dtype=bool)
    for i, cols in enumerate(_symbolic_rows(pattern)):
        filled[i, cols] = True
    return filled


class SparseLULayout:
    """L+U of one pattern as a CSR image in elimination order.

    Row i of the image is caller variable ``order[i]``. In each row the slots
    left of ``diag_ptr[i]`` hold L (unit diagonal implied), ``diag_ptr[i]``
    the diagonal (its inverse, once factorised) and the rest U."""

    def __init__(self, order, rows):
        self.n_vars = len(rows)
        self.order = tuple(int(v) for v in order)
        lengths = np.fromiter((len(r) for r in rows), dtype=np.int64, count=self.n_vars)
        self.row_ptr = np.zeros(self.n_vars + 1, dtype=np.int32)
        self.row_ptr[1:] = np.cumsum(lengths)
        self.col_ind = np.fromiter(itertools.chain.from_iterable(rows), dtype=np.int32,
                                   count=int(lengths.sum()))
        below = np.fromiter((bisect.bisect_left(r, i) for i, r in enumerate(rows)),
                            dtype=np.int32, count=self.n_vars)
        self.diag_ptr = (self.row_ptr[:-1] + below).astype(np.int32)
        self._updates = self._update_program(rows)

    def _update_program(self, rows):
        """For each multiplier slot p = (i, k): the slots of row i that row
        k's upper entries update, in row k's order (so the sources are read
        contiguously from row k and only destinations need an index)."""
        dst_start = np.zeros(self.nnz, dtype=np.int32)
        destination = []
        for i, cols in enumerate(rows):
            base = int(self.row_ptr[i])
            where = {c: base + j for j, c in enumerate(cols)}
            for p in range(base, int(self.diag_ptr[i])):
                k = int(self.col_ind[p])
                dst_start[p] = len(destination)
                upper_k = rows[k][int(self.diag_ptr[k] - self.row_ptr[k]) + 1:]
                destination.extend(where[c] for c in upper_k)
        upper_end = self.row_ptr[1:][self.col_ind].astype(np.int32)
        return dst_start, np.asarray(destination, dtype=np.int32), upper_end

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    @property
    def row_origin(self) -> np.ndarray:
        """Caller index of each elimination row."""
        return np.asarray(self.order, dtype=np.int32)

    @property
    def col_origin(self) -> np.ndarray:
        """Caller index of each slot's column."""
        return self.row_origin[self.col_ind]

    def slot_table(self) -> np.ndarray:
        """(n, n) caller (row, col) -> slot, -1 where the structure has nothing."""
        table = np.full((self.n_vars, self.n_vars), -1, dtype=np.int64)
        caller_row = np.repeat(self.row_origin, np.diff(self.row_ptr))
        table[caller_row, self.col_origin] = np.arange(self.nnz)
        return table

    def factorization_ops(self):
        """``(dst_start, destination, upper_end)`` indexed by multiplier slot:
        where its destinations start in ``destination``, and the end of the
        pivot row's upper part (its sources run from the pivot's slot + 1)."""
        return self._updates

    def flop_counts(self):
        """(multipliers, updates, substitutions): one factorisation's
        multipliers and multiply-add updates, one solve's multiply-adds."""
        lower = int((self.diag_ptr - self.row_ptr[:-1]).sum())
        upper = int((self.row_ptr[1:] - self.diag_ptr - 1).sum())
        return lower, int(self._updates[1].size), lower + upper

    def __repr__(self):
        return f"<SparseLULayout n_vars={self.n_vars} nnz={self.nnz} order={self.order}>"


def analyse(pattern, ordering: str = "amd") -> SparseLULayout:
    """Order, factorise symbolically, lay out L+U in CSR."""
    order = fill_reducing_order(pattern, ordering)
    position = np.empty(len(pattern), dtype=np.int64)
    position[np.asarray(order, dtype=np.int64)] = np.arange(len(pattern))
    permuted = [[int(position[c]) for c in pattern[v]] for v in order]
    return SparseLULayout(order, _symbolic_rows(permuted))


def compressed_jacobian(pattern, layout: SparseLULayout) -> CompressedJacobian:
    """The caller pattern's colouring, with each (row, colour) pointed at the
    CSR slot of its entry; fill-in slots belong to no column and stay
    unwritten by the sweeps."""
    colouring = colour_sparsity(pattern)
    n, width = layout.n_vars, colouring.n_colours
    table = layout.slot_table()
    rows = np.repeat(np.arange(n), [len(cols) for cols in pattern])
    cols = np.fromiter(itertools.chain.from_iterable(pattern), dtype=np.int64, count=rows.size)
    slots = table[rows, cols]
    if (slots < 0).any():
        k = int(np.argmax(slots < 0))
        raise ValueError(f"entry ({rows[k]}, {cols[k]}) is in the pattern but not the structure")
    grid = np.full(n * width, -1, dtype=np.int64)
    grid[rows * width + np.asarray(colouring.colour, dtype=np.int64)[cols]] = slots
    storage = SlotStorage(grid=tuple(int(s) for s in grid),
                          diagonal=tuple(int(s) for s in np.diagonal(table)), size=layout.nnz)
    return replace(colouring, storage=storage)


class SparseDirectSolver:
    """Layout, coloured slot map and the device source compiled for them."""

    ipiv_size = 1

    def __init__(self, layout: SparseLULayout, compressed: CompressedJacobian):
        self.layout = layout
        self.compressed = compressed
        self.ops = layout.factorization_ops()

    @property
    def nnz(self) -> int:
        return self.layout.nnz

    def unrolled(self):
        m, u, s = self.layout.flop_counts()
        return u <= MAX_UNROLLED_UPDATES, s <= MAX_UNROLLED_SUBSTITUTIONS

    def device_source(self) -> str:
        unroll_factor, unroll_solve = self.unrolled()
        diag = ", ".join(str(d) for d in self.compressed.diagonal)
        parts = [f"__device__ constexpr int DIAG[{self.layout.n_vars}] = {{{diag}}};"]
        parts.append(_factor_source(self.layout, self.ops, unroll_factor))
        parts.append(_solve_source(self.layout, unroll_solve))
        parts.append(_solve_t_source(self.layout, unroll_solve))
        return "\n".join(parts)

    def __repr__(self):
        m, u, s = self.layout.flop_counts()
        return (f"<SparseDirectSolver n_vars={self.layout.n_vars} nnz={self.nnz} "
                f"factorise={m + u} solve={s} colours={self.compressed.n_colours}>")


@functools.lru_cache(maxsize=None)
def sparse_direct_solver_for(pattern, ordering: str = "amd") -> SparseDirectSolver:
    layout = analyse(pattern, ordering)
    return SparseDirectSolver(layout, compressed_jacobian(pattern, layout))


def sparse_direct_solver(sparsity, n_vars: int, *, ordering: str = "amd") -> SparseDirectSolver:
    return sparse_direct_solver_for(normalize_sparsity(sparsity, n_vars), ordering)


# --- device source ------------------------------------------------------------
def _factor_source(layout, ops, unrolled: bool) -> str:
    """Up-looking LU in place; the diagonal left holding 1/U[i,i]. The reciprocal
    is formed in double and rounded to the storage type (the reference's
    typing), every update is one fma in the storage type."""
    n = layout.n_vars
    dst_start, destination, upper_end = ops
    pivot_of = layout.diag_ptr[layout.col_ind]
    row_ptr, diag_ptr = layout.row_ptr, layout.diag_ptr
    if unrolled:
        lines = ["template <class T>",
                 "__device__ __forceinline__ void sparse_factor(T* __restrict__ lu) {"]
        for i in range(n):
            for p in range(int(row_ptr[i]), int(diag_ptr[i])):
                above = int(pivot_of[p])
                lines.append(f"  {{ const T f = lu[{p}] * lu[{above}]; lu[{p}] = f;")
                q = int(dst_start[p])
                for s in range(above + 1, int(upper_end[p])):
                    d = int(destination[q])
                    lines.append(f"    lu[{d}] = mdx::fmaT(-f, lu[{s}], lu[{d}]);")
                    q += 1
                lines.append("  }")
            piv = int(diag_ptr[i])
            lines.append(f"  lu[{piv}] = static_cast<T>(1.0 / static_cast<double>(lu[{piv}]));")
        lines.append("}")
        return "\n".join(lines)
    tabs, (ROW, DIAG_, PIV, START, END, DST) = _pack(
        row_ptr, diag_ptr, pivot_of, dst_start, upper_end, destination)
    return "\n".join([
        _const_table("FACTOR_TAB", tabs),
        "template <class T>",
        "__device__ __forceinline__ void sparse_factor(T* __restrict__ lu) {",
        f"  for (int i = 0; i < {n}; ++i) {{",
        f"    const int pivot = FACTOR_TAB[{DIAG_} + i];",
        f"    for (int p = FACTOR_TAB[{ROW} + i]; p < pivot; ++p) {{",
        f"      const int above = FACTOR_TAB[{PIV} + p];",
        "      const T f = lu[p] * lu[above];",
        "      lu[p] = f;",
        f"      int q = {DST} + FACTOR_TAB[{START} + p];",
        f"      for (int s = above + 1; s < FACTOR_TAB[{END} + p]; ++s, ++q)",
        "        lu[FACTOR_TAB[q]] = mdx::fmaT(-f, lu[s], lu[FACTOR_TAB[q]]);",
        "    }",
        "    lu[pivot] = static_cast<T>(1.0 / static_cast<double>(lu[pivot]));",
        "  }",
        "}",
    ])


def _solve_source(layout, unrolled: bool) -> str:
    """Forward sweep through unit-L, back sweep through U times the stored
    inverse diagonal, reading and writing the right-hand side at the caller's
    indices."""
    n = layout.n_vars
    row_ptr, diag_ptr = layout.row_ptr, layout.diag_ptr
    origin, col_origin = layout.row_origin, layout.col_origin
    if unrolled:
        # the factorisation as any indexable type (a register array here)
        lines = ["template <class T, class L>",
                 "__device__ __forceinline__ void sparse_solve(const L& lu, T* __restrict__ r) {"]
        for i in range(n):
            row = int(origin[i])
            a, b = int(row_ptr[i]), int(diag_ptr[i])
            if a == b:
                continue
            body = "".join(f" acc = mdx::fmaT(-lu[{p}], r[{int(col_origin[p])}], acc);" for p in range(a, b))
            lines.append(f"  {{ T acc = r[{row}];{body} r[{row}] = acc; }}")
        for i in range(n - 1, -1, -1):
            row, piv = int(origin[i]), int(diag_ptr[i])
            body = "".join(
                f" acc = mdx::fmaT(-lu[{p}], r[{int(col_origin[p])}], acc);"
                for p in range(piv + 1, int(row_ptr[i + 1])))
            lines.append(f"  {{ T acc = r[{row}];{body} r[{row}] = acc * lu[{piv}]; }}")
        lines.append("}")
        return "\n".join(lines)
    tabs, (ROW, DIAG_, ORIG, CORIG) = _pack(row_ptr, diag_ptr, origin, col_origin)
    return "\n".join([
        _const_table("SOLVE_TAB", tabs),
        "template <class T>",
        "__device__ __forceinline__ void sparse_solve(const T* __restrict__ lu, T* __restrict__ r) {",
        f"  for (int i = 0; i < {n}; ++i) {{",
        f"    const int row = SOLVE_TAB[{ORIG} + i];",
        "    T acc = r[row];",
        f"    for (int p = SOLVE_TAB[{ROW} + i]; p < SOLVE_TAB[{DIAG_} + i]; ++p)",
        f"      acc = mdx::fmaT(-lu[p], r[SOLVE_TAB[{CORIG} + p]], acc);",
        "    r[row] = acc;",
        "  }",
        f"  for (int i = {n - 1}; i >= 0; --i) {{",
        f"    const int row = SOLVE_TAB[{ORIG} + i];",
        f"    const int pivot = SOLVE_TAB[{DIAG_} + i];",
        "    T acc = r[row];",
        f"    for (int p = pivot + 1; p < SOLVE_TAB[{ROW} + i + 1]; ++p)",
        f"      acc = mdx::fmaT(-lu[p], r[SOLVE_TAB[{CORIG} + p]], acc);",
        "    r[row] = acc * lu[pivot];",
        "  }",
        "}",
    ])


def _solve_t_source(layout, unrolled: bool) -> str:
    """The transposed solve ``M^T x = b`` with the same factors (M = L U in
    elimination order): ``U^T z = b`` forward, each row's pivot scaled by the
    stored inverse diagonal and its upper entries scattered onward, then
    ``L^T x = z`` backward, scattering the lower entries. For the adjoint; the
    factors are read through any indexable type (a register array, or the
    adjoint's per-lane LDS rows)."""
    n = layout.n_vars
    row_ptr, diag_ptr = layout.row_ptr, layout.diag_ptr
    origin, col_origin = layout.row_origin, layout.col_origin
    if unrolled:
        lines = ["template <class T, class L>",
                 "__device__ __forceinline__ void sparse_solve_t(const L& lu, T* __restrict__ r) {"]
        for i in range(n):
            row, piv = int(origin[i]), int(diag_ptr[i])
            body = "".join(f" r[{int(col_origin[p])}] = mdx::fmaT(-lu[{p}], z, r[{int(col_origin[p])}]);"
                           for p in range(piv + 1, int(row_ptr[i + 1])))
            lines.append(f"  {{ const T z = r[{row}] * lu[{piv}]; r[{row}] = z;{body} }}")
        for i in range(n - 1, -1, -1):
            row = int(origin[i])
            a, b = int(row_ptr[i]), int(diag_ptr[i])
            if a == b:
                continue
            body = "".join(f" r[{int(col_origin[p])}] = mdx::fmaT(-lu[{p}], x, r[{int(col_origin[p])}]);"
                           for p in range(a, b))
            lines.append(f"  {{ const T x = r[{row}];{body} }}")
        lines.append("}")
        return "\n".join(lines)
    tabs, (ROW, DIAG_, ORIG, CORIG) = _pack(row_ptr, diag_ptr, origin, col_origin)
    return "\n".join([
        _const_table("SOLVE_T_TAB", tabs),
        "template <class T, class L>",
        "__device__ __forceinline__ void sparse_solve_t(const L& lu, T* __restrict__ r) {",
        f"  for (int i = 0; i < {n}; ++i) {{",
        f"    const int row = SOLVE_T_TAB[{ORIG} + i];",
        f"    const int pivot = SOLVE_T_TAB[{DIAG_} + i];",
        "    const T z = r[row] * lu[pivot];",
        "    r[row] = z;",
        f"    for (int p = pivot + 1; p < SOLVE_T_TAB[{ROW} + i + 1]; ++p)",
        f"      r[SOLVE_T_TAB[{CORIG} + p]] = mdx::fmaT(-lu[p], z, r[SOLVE_T_TAB[{CORIG} + p]]);",
        "  }",
        f"  for (int i = {n - 1}; i >= 0; --i) {{",
        f"    const int row = SOLVE_T_TAB[{ORIG} + i];",
        "    const T x = r[row];",
        f"    for (int p = SOLVE_T_TAB[{ROW} + i]; p < SOLVE_T_TAB[{DIAG_} + i]; ++p)",
        f"      r[SOLVE_T_TAB[{CORIG} + p]] = mdx::fmaT(-lu[p], x, r[SOLVE_T_TAB[{CORIG} + p]]);",
        "  }",
        "}",
    ])


def _pack(*arrays):
    """One int32 table of everything and each piece's offset; all entries must
    be non-negative offsets (modax/_sparse_direct.py:625-639)."""
    offsets = np.cumsum([0] + [len(a) for a in arrays[:-1]])
    packed = np.concatenate([np.asarray(a, dtype=np.int64) for a in arrays])
    if packed.size and packed.min() < 0:
        raise ValueError("index tables must hold non-negative offsets only")
    if packed.size and packed.max() > np.iinfo(np.int32).max:
        raise ValueError("index table entry exceeds int32")
    return packed.astype(np.int32), tuple(int(o) for o in offsets)


def _const_table(name, values) -> str:
    body = ", ".join(str(int(v)) for v in values) if len(values) else "0"
    return f"__constant__ int {name}[{max(len(values), 1)}] = {{{body}}};"
