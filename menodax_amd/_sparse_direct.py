"""A sparse direct solver compiled for one sparsity pattern.

Restates the host analysis of modax/_sparse_direct.py and emits HIP device
code in place of its numba device functions:

  1. order   -- a symmetric fill-reducing permutation (``_amd.amd_order``) or
                the natural order (:91-152); the diagonal stays on the
                diagonal, where I/(h gamma) guarantees a structural pivot;
  2. fill    -- the exact L+U pattern of an unpivoted LU of the permuted
                pattern, by Gaussian elimination on the pattern alone
                (:155-184);
  3. layout  -- one CSR image of L+U in elimination order; row/column origins
                translate to the caller's indices so no permutation pass is
                needed (:187-297, 307-349);
  4. compile -- the up-looking factorisation (inverse diagonal stored) and the
                two triangular sweeps (:472-622), emitted straight-line with
                every slot a literal when small enough and table-driven from
                constant memory otherwise, at the reference's thresholds.

The colouring of the caller's pattern and the slot map that points each
coloured Jacobian entry at its CSR slot are ``compressed_jacobian`` (:352-390).
There is no pivoting: the same pattern serves every trajectory, and Rodas5P
being a W-method, an inexact factorisation costs step-size control rather
than order (the controller rejects a step whose factorisation went bad).
"""

from __future__ import annotations

import functools
from dataclasses import dataclass, replace

import numpy as np

from ._amd import amd_order
from ._sparsity import CompressedJacobian, colour_sparsity, normalize_sparsity

ORDERINGS = ("amd", "natural")

#: above these many multiply-adds the routines are table driven
#: (modax/_sparse_direct.py:468-469)
MAX_UNROLLED_SUBSTITUTIONS = 2048
MAX_UNROLLED_UPDATES = 8192


def fill_reducing_order(pattern, ordering: str = "amd") -> tuple:
    if ordering not in ORDERINGS:
        raise ValueError(f"unknown ordering {ordering!r}; expected one of {ORDERINGS}")
    n = len(pattern)
    if ordering == "natural":
        return tuple(range(n))
    return amd_order(pattern)


def fill_pattern(pattern) -> np.ndarray:
    """The exact L+U pattern of an unpivoted LU (diagonal forced present):
    eliminating column k, every row below with an entry there inherits row
    k's entries right of the diagonal."""
    n = len(pattern)
    filled = np.zeros((n, n), dtype=bool)
    for row, cols in enumerate(pattern):
        filled[row, list(cols)] = True
    np.fill_diagonal(filled, True)
    for k in range(n):
        below = np.flatnonzero(filled[k + 1 :, k]) + k + 1
        if below.size:
            filled[below, k + 1 :] |= filled[k, k + 1 :]
    return filled


@dataclass(frozen=True)
class SparseLULayout:
    """CSR image of L+U in elimination order: row i is caller variable
    ``order[i]``; L strictly left of ``diag_ptr``, U from it rightwards."""

    n_vars: int
    order: tuple
    row_ptr: np.ndarray
    col_ind: np.ndarray
    diag_ptr: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    @property
    def row_origin(self) -> np.ndarray:
        return np.asarray(self.order, dtype=np.int32)

    @property
    def col_origin(self) -> np.ndarray:
        return np.asarray(self.order, dtype=np.int32)[self.col_ind]

    def slot_table(self) -> np.ndarray:
        """(n, n) caller (row, col) -> slot, -1 where the structure has nothing."""
        table = np.full((self.n_vars, self.n_vars), -1, dtype=np.int64)
        origin = np.asarray(self.order, dtype=np.int64)
        for i in range(self.n_vars):
            a, b = int(self.row_ptr[i]), int(self.row_ptr[i + 1])
            table[origin[i], origin[self.col_ind[a:b]]] = np.arange(a, b)
        return table

    def factorization_ops(self):
        """For each multiplier slot p = (i, k): where its updates start in the
        destination list, the destination slots (in row k's upper order), and
        the end of row k's upper part -- so the inner loop reads its sources
        contiguously and only the destination needs an index."""
        n = self.n_vars
        dst_start = np.zeros(self.nnz, dtype=np.int64)
        upper_end = np.asarray(self.row_ptr[1:], dtype=np.int64)[np.asarray(self.col_ind, np.int64)]
        elim = np.full((n, n), -1, dtype=np.int64)
        for i in range(n):
            a, b = int(self.row_ptr[i]), int(self.row_ptr[i + 1])
            elim[i, self.col_ind[a:b]] = np.arange(a, b)
        dst = []
        for i in range(n):
            for p in range(int(self.row_ptr[i]), int(self.diag_ptr[i])):
                k = int(self.col_ind[p])
                dst_start[p] = len(dst)
                for q in range(int(self.diag_ptr[k]) + 1, int(self.row_ptr[k + 1])):
                    dst.append(int(elim[i, int(self.col_ind[q])]))
        return dst_start.astype(np.int32), np.asarray(dst, dtype=np.int32), upper_end.astype(np.int32)

    def flop_counts(self):
        """(multipliers, updates, substitutions): one factorisation's
        multipliers and multiply-add updates, one solve's multiply-adds."""
        lower = self.diag_ptr - self.row_ptr[:-1]
        upper = self.row_ptr[1:] - self.diag_ptr - 1
        multipliers = int(lower.sum())
        slot = np.arange(self.nnz)
        row = np.repeat(np.arange(self.n_vars), np.diff(self.row_ptr))
        lower_mask = slot < self.diag_ptr[row]
        updates = int(upper[self.col_ind[lower_mask]].sum())
        return multipliers, updates, multipliers + int(upper.sum())


def analyse(pattern, ordering: str = "amd") -> SparseLULayout:
    """Order, factorise symbolically, lay out L+U in CSR."""
    n = len(pattern)
    order = fill_reducing_order(pattern, ordering)
    fwd = np.asarray(order, dtype=np.int64)
    inv = np.empty(n, dtype=np.int64)
    inv[fwd] = np.arange(n)
    filled = fill_pattern(
        tuple(tuple(sorted(int(inv[c]) for c in pattern[int(fwd[i])])) for i in range(n))
    )
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    row_ptr[1:] = np.cumsum(filled.sum(axis=1))
    col_ind = np.concatenate([np.flatnonzero(filled[i]) for i in range(n)]).astype(np.int32)
    diag_ptr = np.array(
        [int(row_ptr[i]) + int(np.searchsorted(col_ind[row_ptr[i] : row_ptr[i + 1]], i))
         for i in range(n)],
        dtype=np.int32,
    )
    return SparseLULayout(n, tuple(int(i) for i in order), row_ptr, col_ind, diag_ptr)


def compressed_jacobian(pattern, layout: SparseLULayout) -> CompressedJacobian:
    """The caller pattern's colouring, with each (row, colour) pointed at the
    CSR slot of its entry; fill-in slots belong to no column and stay unwritten."""
    compressed = colour_sparsity(pattern)
    table = layout.slot_table()
    packed = np.full(layout.n_vars * compressed.n_colours, -1, dtype=np.int64)
    for row, cols in enumerate(pattern):
        for col in cols:
            s = int(table[row, col])
            if s < 0:
                raise ValueError(f"entry ({row}, {col}) is in the pattern but not the structure")
            packed[row * compressed.n_colours + compressed.colour[col]] = s
    diagonal = tuple(int(table[i, i]) for i in range(layout.n_vars))
    return replace(
        compressed, packed=tuple(int(s) for s in packed), packed_diagonal=diagonal,
        n_slots=layout.nnz,
    )


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


def oracle_tables(pattern, ordering: str = "amd"):
    """The structure as the C oracle takes it: (nnz, row_ptr, col_ind,
    diag_ptr, order, slot_of (n*n, caller indices), pattern mask (n*n))."""
    layout = analyse(pattern, ordering)
    n = layout.n_vars
    mask = np.zeros((n, n), dtype=np.int32)
    for r, cols in enumerate(pattern):
        mask[r, list(cols)] = 1
    return (layout.nnz, layout.row_ptr, layout.col_ind, layout.diag_ptr,
            np.asarray(layout.order, np.int32), layout.slot_table().astype(np.int32).ravel(),
            mask.ravel())


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
    ``L^T x = z`` backward, scattering the lower entries. For the adjoint."""
    n = layout.n_vars
    row_ptr, diag_ptr = layout.row_ptr, layout.diag_ptr
    origin, col_origin = layout.row_origin, layout.col_origin
    if unrolled:
        lines = ["template <class T>",
                 "__device__ __forceinline__ void sparse_solve_t(const T* __restrict__ lu, "
                 "T* __restrict__ r) {"]
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
        "template <class T>",
        "__device__ __forceinline__ void sparse_solve_t(const T* __restrict__ lu, T* __restrict__ r) {",
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
