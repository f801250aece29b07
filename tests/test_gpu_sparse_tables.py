"""The table-driven sparse LU and solves on the GPU.

Above 8,192 factorisation updates or 2,048 substitutions the sparse solver is
emitted as loops over ``__constant__`` index tables instead of straight-line
code (modax/_sparse_direct.py:468-469, 472-622). The reference lowers both
thresholds to zero to run that form on patterns small enough to unroll and
requires it to equal the unrolled form bit for bit
(tests/test_sparse_direct.py:385-415). Here the same is asked of the forward
step (plain and refill launches, FP32 and FP64 LU), of the adjoint's
transposed solve, and of a 150-variable bordered block that crosses both
thresholds by itself -- each also bit-exact with the oracle.
"""

import numpy as np
import pytest

from menodax_amd import _kernels, _sparse_direct as SD, rodas5P
from menodax_amd._sparsity import normalize_sparsity
from oracle import oracle as O
from tests import systems as S

pytestmark = pytest.mark.gpu


def _same(hist, stats, ref_hist, ref_stats):
    for k in ref_stats:
        np.testing.assert_array_equal(np.asarray(stats[k]), np.asarray(ref_stats[k]), err_msg=k)
    np.testing.assert_array_equal(hist, ref_hist)


def _structure(mask, ordering):
    pattern = normalize_sparsity(mask, len(mask))
    return O.sparse_structure(pattern, SD.fill_reducing_order(pattern, ordering))


def _clear_kernels():
    _kernels.rodas5p_kernel.cache_clear()
    _kernels.rodas5p_adjoint_kernel.cache_clear()


@pytest.fixture
def tables(monkeypatch):
    """Every sparse solver built while this is active is table driven."""
    monkeypatch.setattr(SD, "MAX_UNROLLED_SUBSTITUTIONS", 0)
    monkeypatch.setattr(SD, "MAX_UNROLLED_UPDATES", 0)
    _clear_kernels()
    yield
    _clear_kernels()


def _unrolled_then_tables(request, run):
    """``run()`` with the unrolled solver, then with the table-driven one."""
    _clear_kernels()
    first = run()
    request.getfixturevalue("tables")
    return first, run()


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
@pytest.mark.parametrize("ordering", ["amd", "natural"])
def test_ring_with_fill_in(lu, ordering, request):
    mask = S.ring_mask(6)
    y0 = np.zeros((64, 6))
    y0[:, 0] = 1.0
    p = np.tile(np.array([50.0, 1.0]), (64, 1))
    p[:, 0] *= np.linspace(0.5, 1.5, 64)
    t = np.array([0.0, 0.01, 0.2])
    kw = dict(rtol=1e-9, atol=1e-11, first_step=1e-6, lu_precision=lu)
    ref, rst = O.rodas5p(O.ring6(), y0, t, p, layout=_structure(mask, ordering), **kw)

    def run():
        return [rodas5P.solve(S.ring6, y0, t, p, sparsity=mask, ordering=ordering,
                              return_stats=True, refill=refill, **kw) for refill in (False, True)]

    unrolled, tabled = _unrolled_then_tables(request, run)
    for (h_u, st_u), (h_t, st_t) in zip(unrolled, tabled):
        _same(h_t, st_t, h_u, st_u)
        _same(h_t, st_t, ref, rst)


@pytest.mark.parametrize("n", [24, 50])
def test_block_arrow(n, request):
    fn, mask, (rows, cols, vals, pidx) = S.arrow_system(n)
    y0, lam = S.arrow_ensemble(128, n=n)
    t = np.linspace(0.0, 10.0, 7)
    system = O.linquad(n, rows, cols, vals, 0.01, off_pidx=pidx, n_p=S.arrow_n_params(n))
    layout = _structure(mask, "amd")
    refs = {lu: O.rodas5p(system, y0, t, lam, layout=layout, rtol=1e-6, atol=1e-8,
                          lu_precision=lu) for lu in ("fp32", "fp64")}

    def run():
        return {(lu, refill): rodas5P.solve(fn, y0, t, lam, sparsity=mask, rtol=1e-6, atol=1e-8,
                                            lu_precision=lu, return_stats=True, refill=refill)
                for lu in ("fp32", "fp64") for refill in (False, True)}

    unrolled, tabled = _unrolled_then_tables(request, run)
    for key, (h_t, st_t) in tabled.items():
        _same(h_t, st_t, *unrolled[key])
        _same(h_t, st_t, *refs[key[0]])


def test_vdp_chain(request):
    """The 64-variable stiff VdP chain (the sparse configuration's system)."""
    fn, _ = S.vdp_lattice(32, mu=1.0, d=0.1, wrap=False)
    mask = S.vdp_sparsity(32, wrap=False)
    y0, mu = S.vdp_chain_ensemble(32, 128)
    t = np.linspace(0.0, 20.0, 5)
    kw = dict(rtol=1e-6, atol=1e-8)
    layout = _structure(mask, "amd")
    system = O.vdp_lattice(32, mu=1.0, d=0.1, wrap=False)
    refs = {lu: O.rodas5p(system, y0, t, mu, layout=layout, lu_precision=lu, **kw)
            for lu in ("fp32", "fp64")}

    def run():
        return {(lu, refill): rodas5P.solve(fn, y0, t, mu, sparsity=mask, lu_precision=lu,
                                            return_stats=True, refill=refill, **kw)
                for lu in ("fp32", "fp64") for refill in (False, True)}

    unrolled, tabled = _unrolled_then_tables(request, run)
    for key, (h_t, st_t) in tabled.items():
        _same(h_t, st_t, *unrolled[key])
        _same(h_t, st_t, *refs[key[0]])


def test_adjoint_transposed_solve(request):
    """The adjoint of the 20-variable sparse VdP chain runs the transposed
    sparse solve M^T x = b every stage; its table-driven form gives the same
    gradients bit for bit."""
    fn, _ = S.vdp_lattice(10, mu=1.0, d=0.1, wrap=False)
    mask = S.vdp_sparsity(10, wrap=False)
    rng = np.random.default_rng(0)
    mu = 10.0 ** rng.uniform(0.5, 1.5, size=(64, 1))
    y0 = rng.uniform(-2.0, 2.0, size=(64, 20))
    t = np.linspace(0.0, 3.0, 20)
    kw = dict(rtol=1e-10, atol=1e-12, lu_precision="fp64", sparsity=mask)
    g = rng.normal(size=(64, t.size, 20))

    def run():
        return rodas5P.adjoint_gradients(fn, y0, t, mu, g, **kw)

    (gy_u, gp_u), (gy_t, gp_t) = _unrolled_then_tables(request, run)
    assert np.all(np.isfinite(gy_u)) and np.all(np.isfinite(gp_u))
    np.testing.assert_array_equal(gy_t, gy_u)
    np.testing.assert_array_equal(gp_t, gp_u)


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
def test_bordered_block_over_the_thresholds(lu):
    """150 variables, a dense 45-variable core: table driven at the default
    thresholds (29,475 updates, 2,190 substitutions per step), plain and
    refill, bit-exact with the oracle."""
    _clear_kernels()
    fn, mask, (rows, cols, vals) = S.bordered_system()
    n = len(mask)
    solver = SD.sparse_direct_solver_for(normalize_sparsity(mask, n), "amd")
    assert solver.unrolled() == (False, False)
    y0, lam = S.bordered_ensemble(96, n)
    t = np.array([0.0, 0.01, 0.1, 1.0])
    kw = dict(rtol=1e-6, atol=1e-8, lu_precision=lu)
    ref, rst = O.rodas5p(O.linquad(n, rows, cols, vals, 0.01), y0, t, lam,
                         layout=_structure(mask, "amd"), **kw)
    for refill in (False, True):
        hist, st = rodas5P.solve(fn, y0, t, lam, sparsity=mask, return_stats=True,
                                 refill=refill, **kw)
        _same(hist, st, ref, rst)
