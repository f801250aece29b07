"""Host analysis of the sparse direct solver and its colouring.

Mirrors the reference's CPU-side tests (tests/test_sparse_direct.py:144-316,
tests/test_sparsity.py:64-111): the symbolic fill against an unpivoted
elimination on numbers, the AMD fill goldens, the CSR invariants, a NumPy
replay of the compiled tables against numpy's solve, the footprint, colour
counts and the accepted pattern spellings. The replay here also reads the
*emitted device source*: the unrolled factorisation and sweeps are parsed back
and executed, so the code the GPU runs is what is checked.
"""

import re

import numpy as np
import pytest

from menodax_amd import _sparse_direct as SD
from tests import systems as S
from menodax_amd._sparsity import (
    CompressedJacobian,
    check_orthogonal,
    check_superset,
    colour_sparsity,
    dense_jacobian,
    normalize_sparsity,
)


def _arrow(n):
    mask = np.eye(n, dtype=bool)
    mask[0, :] = mask[:, 0] = True
    return mask


def _bordered_block(core, tails, length):
    n = core + tails * length
    mask = np.zeros((n, n), dtype=bool)
    mask[:core, :core] = True
    for t in range(tails):
        base = core + t * length
        parent = t % core
        for i in range(length):
            mask[base + i, base + i] = True
            if i:
                mask[base + i, base + i - 1] = True
            if i + 1 < length:
                mask[base + i, base + i + 1] = True
        mask[base, parent] = mask[parent, base] = True
    return mask


def _ring(n):
    return S.ring_mask(n)


def _matrix(mask, seed=0, shift=60.0):
    rng = np.random.default_rng(seed)
    dense = np.where(mask, rng.normal(size=mask.shape), 0.0)
    dense[np.diag_indices(len(mask))] += shift
    return dense


def _replay_tables(layout, dense):
    """The table-driven algorithm in NumPy (modax/_sparse_direct.py:507-521, 585-600)."""
    n = layout.n_vars
    dst_start, destination, upper_end = layout.factorization_ops()
    pivot_of = layout.diag_ptr[layout.col_ind]
    lu = np.zeros(layout.nnz)
    table = layout.slot_table()
    rows, cols = np.nonzero(table >= 0)
    lu[table[rows, cols]] = dense[rows, cols]
    for i in range(n):
        pivot = layout.diag_ptr[i]
        for p in range(layout.row_ptr[i], pivot):
            above = pivot_of[p]
            f = lu[p] * lu[above]
            lu[p] = f
            q = dst_start[p]
            for s in range(above + 1, upper_end[p]):
                lu[destination[q]] -= f * lu[s]
                q += 1
        lu[pivot] = 1.0 / lu[pivot]

    def solve(b):
        x = np.array(b, dtype=np.float64)
        o, co = layout.row_origin, layout.col_origin
        for i in range(n):
            acc = x[o[i]]
            for p in range(layout.row_ptr[i], layout.diag_ptr[i]):
                acc -= lu[p] * x[co[p]]
            x[o[i]] = acc
        for i in range(n - 1, -1, -1):
            piv = layout.diag_ptr[i]
            acc = x[o[i]]
            for p in range(piv + 1, layout.row_ptr[i + 1]):
                acc -= lu[p] * x[co[p]]
            x[o[i]] = acc * lu[piv]
        return x

    return solve


def _run_emitted(source, name, lu, r=None):
    """Execute an emitted unrolled device function body in Python (fma as a*b+c)."""
    start = source.index(f"void {name}(")
    body = source[source.index("{", start) + 1:]
    depth, end = 1, 0
    for k, ch in enumerate(body):
        depth += ch == "{"
        depth -= ch == "}"
        if depth == 0:
            end = k
            break
    body = body[:end]
    py = body.replace("mdx::fmaT", "_fma").replace("static_cast<T>", "")
    py = py.replace("static_cast<double>", "")
    py = re.sub(r"const T ", "", py)
    py = re.sub(r"\bT acc\b", "acc", py)
    stmts = [s.strip() for s in re.split(r"[;{}]", py) if s.strip()]
    env = {"lu": lu, "r": r, "_fma": lambda a, b, c: a * b + c}
    for st in stmts:
        exec(st, env)  # noqa: S102


# --- the symbolic factorisation ----------------------------------------------
def test_fill_pattern_is_exactly_what_an_unpivoted_lu_touches():
    mask = _ring(9)
    predicted = SD.fill_pattern(normalize_sparsity(mask, 9))
    work = _matrix(mask)
    for k in range(9):
        work[k + 1:, k] /= work[k, k]
        work[k + 1:, k + 1:] -= np.outer(work[k + 1:, k], work[k, k + 1:])
    touched = work != 0.0
    assert (touched == predicted).all()


def test_amd_undoes_the_worst_case_for_the_natural_order():
    pattern = normalize_sparsity(_arrow(12), 12)
    assert SD.analyse(pattern, "natural").nnz == 144
    amd = SD.analyse(pattern, "amd")
    assert amd.nnz == int(_arrow(12).sum())
    assert amd.order[-1] == 0


def test_a_bordered_block_system_needs_no_fill_at_all():
    mask = _bordered_block(5, 3, 7)
    amd = SD.analyse(normalize_sparsity(mask, len(mask)), "amd")
    assert amd.nnz == int(mask.sum())
    assert set(amd.order[-5:]) == set(range(5))


@pytest.mark.parametrize(
    "mask, expected",
    [
        (_arrow(12), 34),
        (_arrow(40), 118),
        (_ring(16), 74),
        (_ring(64), 314),
        (_bordered_block(5, 3, 7), 88),
        (_bordered_block(10, 8, 16), 484),
    ],
    ids=["arrow12", "arrow40", "ring16", "ring64", "bordered", "disco-like"],
)
def test_amd_fills_what_the_reference_records(mask, expected):
    """tests/test_sparse_direct.py:193-215: the fill goldens of AMD."""
    assert SD.analyse(normalize_sparsity(mask, len(mask)), "amd").nnz == expected


def test_amd_mass_eliminates_a_clique_and_sets_dense_rows_aside():
    """The published algorithm's special cases: a fully coupled block is one
    supervariable, eliminated at once with its members numbered before the
    pivot (the highest index, first off the LIFO degree list); a row denser
    than max(16, 10 sqrt(n)) is set aside and ordered last, here the hub of a
    300-variable arrow, which then fills nothing."""
    from menodax_amd._amd import amd_order

    clique = tuple(tuple(range(5)) for _ in range(5))
    assert amd_order(clique) == (0, 1, 2, 3, 4)
    mask = _arrow(300)
    pattern = normalize_sparsity(mask, 300)
    order = amd_order(pattern)
    assert sorted(order) == list(range(300)) and order[-1] == 0
    assert SD.analyse(pattern, "amd").nnz == int(mask.sum())


def test_amd_order_is_a_permutation_and_beats_natural_on_random_patterns():
    rng = np.random.default_rng(5)
    for n in (8, 20, 50):
        for _ in range(5):
            mask = rng.random((n, n)) < 3.0 / n
            mask |= np.eye(n, dtype=bool)
            pattern = normalize_sparsity(mask, n)
            order = SD.fill_reducing_order(pattern, "amd")
            assert sorted(order) == list(range(n))
            assert SD.analyse(pattern, "amd").nnz <= SD.analyse(pattern, "natural").nnz * 1.5


def test_the_ordering_is_validated():
    with pytest.raises(ValueError, match="unknown ordering"):
        SD.analyse(normalize_sparsity(_ring(4), 4), "minimum-degree-ish")
    assert "amd" in SD.ORDERINGS


# --- the layout ---------------------------------------------------------------
MASKS = [_ring(9), _arrow(8), _bordered_block(4, 2, 5), np.ones((5, 5), dtype=bool)]
IDS = ["ring", "arrow", "bordered", "dense"]


@pytest.mark.parametrize("mask", MASKS, ids=IDS)
def test_layout_is_a_consistent_csr_image(mask):
    n = len(mask)
    layout = SD.analyse(normalize_sparsity(mask, n), "natural")
    assert layout.row_ptr[0] == 0 and layout.row_ptr[-1] == layout.nnz
    assert (np.diff(layout.row_ptr) >= 1).all()
    for i in range(n):
        span = layout.col_ind[layout.row_ptr[i]:layout.row_ptr[i + 1]]
        assert (np.diff(span) > 0).all()
        assert layout.col_ind[layout.diag_ptr[i]] == i
    table = layout.slot_table()
    slots = [int(table[r, c]) for r, c in zip(*np.nonzero(mask))]
    assert min(slots) >= 0 and len(set(slots)) == len(slots) and max(slots) < layout.nnz


@pytest.mark.parametrize("ordering", ["natural", "amd"])
@pytest.mark.parametrize("mask", MASKS, ids=IDS)
def test_the_compiled_tables_factorise_the_matrix(mask, ordering):
    n = len(mask)
    layout = SD.analyse(normalize_sparsity(mask, n), ordering)
    dense = _matrix(mask)
    solve = _replay_tables(layout, dense)
    rng = np.random.default_rng(1)
    for _ in range(3):
        b = rng.normal(size=n)
        np.testing.assert_allclose(solve(b), np.linalg.solve(dense, b), rtol=1e-10)


@pytest.mark.parametrize("ordering", ["natural", "amd"])
@pytest.mark.parametrize("mask", MASKS + [S.arrow_structure(50)[4], S.vdp_sparsity(8, wrap=False)],
                         ids=IDS + ["block-arrow50", "vdp-chain16"])
def test_the_emitted_device_code_solves_the_matrix(mask, ordering):
    """The unrolled HIP source the kernel compiles, executed statement by statement."""
    n = len(mask)
    solver = SD.sparse_direct_solver_for(normalize_sparsity(mask, n), ordering)
    src = solver.device_source()
    dense = _matrix(mask, seed=3)
    lu = np.zeros(solver.nnz)
    table = solver.layout.slot_table()
    rows, cols = np.nonzero(table >= 0)
    lu[table[rows, cols]] = dense[rows, cols]
    _run_emitted(src, "sparse_factor", lu)
    rng = np.random.default_rng(4)
    b = rng.normal(size=n)
    r = b.copy()
    _run_emitted(src, "sparse_solve", lu, r)
    np.testing.assert_allclose(r, np.linalg.solve(dense, b), rtol=1e-9)
    # the transposed solve the adjoint uses, with the same factors
    r = b.copy()
    _run_emitted(src, "sparse_solve_t", lu, r)
    np.testing.assert_allclose(r, np.linalg.solve(dense.T, b), rtol=1e-9)


def test_table_driven_source_is_emitted_above_the_thresholds(monkeypatch):
    monkeypatch.setattr(SD, "MAX_UNROLLED_SUBSTITUTIONS", 0)
    monkeypatch.setattr(SD, "MAX_UNROLLED_UPDATES", 0)
    solver = SD.SparseDirectSolver(*_layout_and_map(_ring(6)))
    src = solver.device_source()
    assert "__constant__ int FACTOR_TAB" in src and "__constant__ int SOLVE_TAB" in src
    assert "__constant__ int SOLVE_T_TAB" in src


def _layout_and_map(mask):
    pattern = normalize_sparsity(mask, len(mask))
    layout = SD.analyse(pattern, "amd")
    return layout, SD.compressed_jacobian(pattern, layout)


def test_the_footprint_is_the_factorisation_and_nothing_more():
    pattern = normalize_sparsity(_ring(9), 9)
    layout = SD.analyse(pattern, "natural")
    compressed = SD.compressed_jacobian(pattern, layout)
    assert compressed.size == layout.nnz
    claimed = int((compressed.store_slots() >= 0).sum())
    assert claimed == int(_ring(9).sum()) < compressed.size
    assert compressed.needs_clear
    assert len(set(compressed.diagonal)) == 9


def test_the_pattern_may_be_coloured_as_tightly_as_it_is():
    mask = _bordered_block(5, 3, 7)
    pattern = normalize_sparsity(mask, len(mask))
    compressed = SD.compressed_jacobian(pattern, SD.analyse(pattern, "natural"))
    for cols in pattern:
        assert len({compressed.colour[c] for c in cols}) == len(cols)


def test_the_solver_is_cached_on_the_pattern():
    a = SD.sparse_direct_solver(_ring(6), 6, ordering="natural")
    b = SD.sparse_direct_solver(np.array(sorted(zip(*np.nonzero(_ring(6))))), 6, ordering="natural")
    assert a is b
    assert a.ipiv_size == 1 and a.nnz == a.compressed.size


def test_flop_counts_match_a_direct_count():
    layout = SD.analyse(normalize_sparsity(_ring(9), 9), "natural")
    m, u, s = layout.flop_counts()
    _, destination, _ = layout.factorization_ops()
    assert u == destination.size
    lower = int(sum(layout.diag_ptr - layout.row_ptr[:-1]))
    assert m == lower
    assert s == layout.nnz - 9


# --- colouring ---------------------------------------------------------------
def test_tridiagonal_needs_three_colours():
    n = 6
    band = np.zeros((n, n), dtype=bool)
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                band[i, j] = True
    compressed = colour_sparsity(normalize_sparsity(band, n))
    assert compressed.n_colours == 3 and compressed.size == n * 3


def test_chain_needs_two_colours():
    assert colour_sparsity(normalize_sparsity(S.CHAIN4_MASK, 4)).n_colours == 2


def test_no_pattern_is_the_dense_matrix():
    dense = dense_jacobian(4)
    assert dense.n_colours == 4
    assert [dense.slot(r, c) for r in range(4) for c in range(4)] == list(range(16))
    assert dense.diagonal == (0, 5, 10, 15)
    assert not dense.needs_clear


def test_colouring_is_checked_for_structural_orthogonality():
    bad = CompressedJacobian(n_vars=2, n_colours=1, colour=(0, 0))
    with pytest.raises(ValueError, match="not structurally orthogonal"):
        check_orthogonal(((0, 1), ()), bad)


def test_seed_table_covers_the_parameter_direction():
    compressed = colour_sparsity(normalize_sparsity(S.CHAIN4_MASK, 4))
    seeds = compressed.seed_table(min_width=9)
    assert seeds.shape == (compressed.n_colours + 1, 9)
    assert not seeds[-1].any()
    for c, g in enumerate(compressed.colour):
        assert seeds[g, c] == 1.0


def test_sparsity_accepts_several_spellings():
    from scipy import sparse as sp

    expected = normalize_sparsity(S.CHAIN4_MASK, 4)
    pairs = np.array(sorted(zip(*np.nonzero(S.CHAIN4_MASK))), dtype=np.int64)
    assert normalize_sparsity(pairs, 4) == expected
    assert normalize_sparsity(sp.csr_array(S.CHAIN4_MASK.astype(np.int8)), 4) == expected
    with pytest.raises(ValueError, match="lies outside"):
        normalize_sparsity(np.array([[0, 99]]), 4)
    with pytest.raises(ValueError, match="sparsity must be"):
        normalize_sparsity(np.zeros((3, 3, 3)), 4)


def test_a_pattern_missing_a_true_entry_is_rejected():
    from menodax_amd import _codegen

    traced = _codegen.trace(S.chain4, 4, 3)
    reads = _codegen.structural_pattern(traced)
    assert (reads == S.CHAIN4_MASK).all()
    short = S.CHAIN4_MASK.copy()
    short[2, 1] = False
    with pytest.raises(ValueError, match="omits 1 Jacobian entries"):
        check_superset(normalize_sparsity(short, 4), reads)
    check_superset(normalize_sparsity(S.CHAIN4_MASK, 4), reads)


def test_colouring_matches_networkx_greedy_best_of_five():
    """The colour vector is the reference's rule applied with networkx directly."""
    import networkx as nx

    mask = S.vdp_sparsity(8, wrap=False)
    pattern = normalize_sparsity(mask, 16)
    got = colour_sparsity(pattern)
    g = nx.Graph()
    g.add_nodes_from(range(16))
    for cols in pattern:
        for a in cols:
            for b in cols:
                if a < b:
                    g.add_edge(a, b)
    best = min((nx.coloring.greedy_color(g, strategy=s) for s in
                ("largest_first", "smallest_last", "independent_set",
                 "saturation_largest_first", "DSATUR")),
               key=lambda c: max(c.values()))
    assert got.n_colours == max(best.values()) + 1


# --- the emitted device code compiled for the host ----------------------------
_HOST_PRELUDE = r"""
#include <cmath>
#define __device__
#define __constant__ static const
#define __forceinline__ inline
#define __restrict__ __restrict
namespace mdx {
inline double fmaT(double a, double b, double c) { return std::fma(a, b, c); }
inline float fmaT(float a, float b, float c) { return std::fmaf(a, b, c); }
}
namespace sys {
"""
_HOST_EPILOGUE = r"""
}  // namespace sys
#define ENTRY(T, S)                                                              \
  extern "C" void factor_##S(T* lu) { sys::sparse_factor(lu); }                 \
  extern "C" void solve_##S(const T* lu, T* r) { sys::sparse_solve(lu, r); }    \
  extern "C" void solve_t_##S(const T* lu, T* r) { sys::sparse_solve_t(lu, r); }
ENTRY(double, d)
ENTRY(float, f)
"""


def _host_build(source, tmp_path, name):
    """The device source as a host shared library (g++, no contraction, so
    every multiply-add is the explicit fma the device code names)."""
    import ctypes
    import subprocess

    src = tmp_path / f"{name}.cpp"
    lib = tmp_path / f"lib{name}.so"
    src.write_text(_HOST_PRELUDE + source + _HOST_EPILOGUE)
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-fPIC", "-shared",
                    "-o", str(lib), str(src)], check=True)
    return ctypes.CDLL(str(lib))


def _host_run(lib, dtype, lu, b):
    import ctypes

    sfx = "d" if dtype == np.float64 else "f"
    ct = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    lu = np.array(lu, dtype=dtype, copy=True)
    getattr(lib, f"factor_{sfx}")(lu.ctypes.data_as(ctypes.POINTER(ct)))
    x = np.array(b, dtype=dtype, copy=True)
    xt = x.copy()
    getattr(lib, f"solve_{sfx}")(lu.ctypes.data_as(ctypes.POINTER(ct)),
                                 x.ctypes.data_as(ctypes.POINTER(ct)))
    getattr(lib, f"solve_t_{sfx}")(lu.ctypes.data_as(ctypes.POINTER(ct)),
                                   xt.ctypes.data_as(ctypes.POINTER(ct)))
    return lu, x, xt


@pytest.mark.parametrize("ordering", ["natural", "amd"])
@pytest.mark.parametrize("mask", [_ring(9), S.arrow_structure(24)[4], S.vdp_sparsity(8, wrap=False),
                                  _bordered_block(4, 2, 5)],
                         ids=["ring9", "block-arrow24", "vdp-chain16", "bordered"])
def test_table_driven_code_computes_what_the_unrolled_code_does(mask, ordering, tmp_path,
                                                                monkeypatch):
    """Both emissions of one analysis, compiled for the host: the factors and
    both solves agree bit for bit in both precisions (the reference requires
    the same of its two forms, tests/test_sparse_direct.py:385-415), and they
    solve the matrix (numpy's solve and its transpose)."""
    n = len(mask)
    pattern = normalize_sparsity(mask, n)
    solver = SD.sparse_direct_solver_for(pattern, ordering)
    assert solver.unrolled() == (True, True)
    unrolled = _host_build(solver.device_source(), tmp_path, "unrolled")
    monkeypatch.setattr(SD, "MAX_UNROLLED_SUBSTITUTIONS", 0)
    monkeypatch.setattr(SD, "MAX_UNROLLED_UPDATES", 0)
    assert solver.unrolled() == (False, False)
    table_src = solver.device_source()
    assert "FACTOR_TAB" in table_src and "SOLVE_T_TAB" in table_src
    tables = _host_build(table_src, tmp_path, "tables")
    dense = _matrix(mask, seed=8)
    lu0 = np.zeros(solver.nnz)
    table = solver.layout.slot_table()
    rows, cols = np.nonzero(table >= 0)
    lu0[table[rows, cols]] = dense[rows, cols]
    b = np.random.default_rng(9).normal(size=n)
    for dtype in (np.float64, np.float32):
        got_u = _host_run(unrolled, dtype, lu0, b)
        got_t = _host_run(tables, dtype, lu0, b)
        for u, t in zip(got_u, got_t):
            assert u.tobytes() == t.tobytes()
    _, x, xt = got_u
    _, x64, xt64 = _host_run(tables, np.float64, lu0, b)
    np.testing.assert_allclose(x64, np.linalg.solve(dense, b), rtol=1e-10)
    np.testing.assert_allclose(xt64, np.linalg.solve(dense.T, b), rtol=1e-10)
    np.testing.assert_allclose(x, np.linalg.solve(dense, b), rtol=1e-4)


def test_a_large_bordered_block_is_table_driven_by_itself(tmp_path):
    """150 variables with a dense 45-variable core: 29,475 updates and 2,190
    substitutions, over both thresholds without lowering them; the emitted
    tables solve the matrix."""
    fn, mask, _ = S.bordered_system()
    n = len(mask)
    solver = SD.sparse_direct_solver_for(normalize_sparsity(mask, n), "amd")
    m, u, s = solver.layout.flop_counts()
    assert u > SD.MAX_UNROLLED_UPDATES and s > SD.MAX_UNROLLED_SUBSTITUTIONS
    assert solver.unrolled() == (False, False)
    lib = _host_build(solver.device_source(), tmp_path, "bordered")
    dense = _matrix(mask, seed=2, shift=30.0)
    lu0 = np.zeros(solver.nnz)
    table = solver.layout.slot_table()
    rows, cols = np.nonzero(table >= 0)
    lu0[table[rows, cols]] = dense[rows, cols]
    b = np.random.default_rng(3).normal(size=n)
    _, x, xt = _host_run(lib, np.float64, lu0, b)
    np.testing.assert_allclose(x, np.linalg.solve(dense, b), rtol=1e-9)
    np.testing.assert_allclose(xt, np.linalg.solve(dense.T, b), rtol=1e-9)


# --- the product's analysis against the oracle's restatement -----------------
def _random_masks():
    rng = np.random.default_rng(21)
    out = []
    for n in (5, 12, 30, 60):
        for density in (1.5, 3.0):
            mask = rng.random((n, n)) < density / n  # unsymmetric on purpose
            out.append(mask)
    return out


@pytest.mark.parametrize("ordering", ["natural", "amd"])
@pytest.mark.parametrize("mask", MASKS + [S.arrow_structure(50)[4], S.vdp_sparsity(32, wrap=False),
                                          S.bordered_structure()[3]] + _random_masks())
def test_the_analysis_equals_the_oracles_elimination_on_the_boolean_pattern(mask, ordering):
    """The product's up-looking symbolic factorisation against the oracle's
    right-looking elimination on the boolean pattern (two independent
    constructions of one mathematically unique structure): CSR, diagonal
    positions and the caller slot table all equal, for the same order."""
    from oracle import oracle as O

    n = len(mask)
    pattern = normalize_sparsity(mask, n)
    layout = SD.analyse(pattern, ordering)
    nnz, row_ptr, col_ind, diag_ptr, order, slot_of, _ = O.sparse_structure(pattern, layout.order)
    assert nnz == layout.nnz
    np.testing.assert_array_equal(row_ptr, layout.row_ptr)
    np.testing.assert_array_equal(col_ind, layout.col_ind)
    np.testing.assert_array_equal(diag_ptr, layout.diag_ptr)
    np.testing.assert_array_equal(order, layout.row_origin)
    np.testing.assert_array_equal(slot_of, layout.slot_table().ravel())


def test_the_oracles_fill_is_what_numeric_elimination_touches():
    """The oracle's boolean elimination against unpivoted elimination on
    random numbers over unsymmetric patterns (no cancellation at these
    densities): every structural entry is touched and nothing else."""
    from oracle import oracle as O

    for mask in _random_masks():
        n = len(mask)
        order = tuple(np.random.default_rng(n).permutation(n))
        nnz, row_ptr, col_ind, *_ = O.sparse_structure(normalize_sparsity(mask, n), order)
        work = _matrix(mask, seed=n)[np.ix_(order, order)]
        for k in range(n):
            work[k + 1:, k] /= work[k, k]
            work[k + 1:, k + 1:] -= np.outer(work[k + 1:, k], work[k, k + 1:])
        touched = work != 0.0
        expected = np.zeros((n, n), dtype=bool)
        for i in range(n):
            expected[i, col_ind[row_ptr[i]:row_ptr[i + 1]]] = True
        assert (touched == expected).all()
