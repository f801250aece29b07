"""Parity of the HIP kernels with the oracle, on an MI355X.

The kernels and the oracle place every fused multiply-add in the same spot, use
the same deterministic power in the controller, and the right-hand sides here
are polynomial, so the bar is bit-exact: equal step counters for every
trajectory and identical histories (``np.array_equal``), in both LU precisions
and both launch shapes. Published reference goldens are checked on the GPU too.
"""

import numpy as np
import pytest

from menodax_amd import rodas5P, tsit5
from tests import systems as S
from oracle import oracle as O


def _structure(mask, n, ordering):
    """The oracle's own sparse structure for ``mask`` under the product's
    elimination order (the AMD permutation is the one input it takes)."""
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    pattern = normalize_sparsity(mask, n)
    return O.sparse_structure(pattern, SD.fill_reducing_order(pattern, ordering))

pytestmark = pytest.mark.gpu


def _same(hist, stats, ref_hist, ref_stats):
    for k in ref_stats:
        np.testing.assert_array_equal(np.asarray(stats[k]), ref_stats[k], err_msg=k)
    np.testing.assert_array_equal(hist, ref_hist)


# --- Rodas5P, dense -----------------------------------------------------------
@pytest.mark.parametrize("lu", ["fp32", "fp64"])
@pytest.mark.parametrize("refill", [False, True])
def test_robertson_ensemble_bit_exact(lu, refill):
    y0, p = S.robertson_lognormal_ensemble(2000, seed=1)
    t = S.robertson_benchmark_times()
    hist, st = rodas5P.solve(S.robertson, y0, t, p, rtol=1e-6, atol=1e-8, lu_precision=lu,
                             return_stats=True, refill=refill)
    ref, rst = O.rodas5p(O.robertson(), y0, t, p, rtol=1e-6, atol=1e-8, lu_precision=lu)
    _same(hist, st, ref, rst)


def test_robertson_published_single_trajectory():
    """benchmarks/stiff_robertson_divergence csv:2: 157 attempted, 10 rejected
    (initial state built as robertson.py:8-10 builds it)."""
    hist, st = rodas5P.solve(S.robertson, S.ROBERTSON_Y0[None], S.ROBERTSON_TIMES,
                             S.ROBERTSON_PARAMS[None], first_step=1e-4, rtol=1e-6, atol=1e-8,
                             lu_precision="fp32", return_stats=True)
    assert {k: int(v[0]) for k, v in st.items()} == dict(
        accepted_steps=147, rejected_steps=10, loop_steps=157)
    np.testing.assert_allclose(hist[0].sum(axis=-1), 1.0, rtol=1e-6)  # mass conservation


def test_robertson_divergence_distribution_on_gpu():
    y0, p = S.robertson_scenario(30000, 42, divergence=1.0)
    _, st = rodas5P.solve(S.robertson, y0, S.ROBERTSON_TIMES, p, first_step=1e-4, rtol=1e-6,
                          atol=1e-8, lu_precision="fp32", return_stats=True)
    _, rst = O.rodas5p(O.robertson(), y0, S.ROBERTSON_TIMES, p, first_step=1e-4, rtol=1e-6,
                       atol=1e-8, lu_precision="fp32")
    for k in rst:
        np.testing.assert_array_equal(st[k], rst[k])
    att = st["accepted_steps"] + st["rejected_steps"]
    assert att.mean() == pytest.approx(106.2629, rel=2e-4)  # published csv:6
    assert (att.min(), att.max()) == (32, 184)


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
@pytest.mark.parametrize("n_osc", [8, 12])
def test_stiff_vdp_ring_dense_bit_exact(lu, n_osc):
    """Stiff VdP ring (mu = 100), dense LU with pivoting: 16 variables take the
    register-resident unrolled LU, 24 the looped one."""
    fn, _ = S.vdp_lattice(n_osc)
    y0, p = S.vdp_scenario(n_osc, 70, 42, divergence=0.25)
    hist, st = rodas5P.solve(fn, y0, S.VDP_TIMES, p, first_step=1e-4, rtol=1e-6, atol=1e-8,
                             lu_precision=lu, return_stats=True)
    ref, rst = O.rodas5p(O.vdp_lattice(n_osc), y0, S.VDP_TIMES, p, first_step=1e-4, rtol=1e-6,
                         atol=1e-8, lu_precision=lu)
    _same(hist, st, ref, rst)


def test_forced_decay_closed_form_on_gpu():
    """tests/test_solver_options.py:159-176."""
    t = np.linspace(0.0, 1.0, 11)
    out = rodas5P.solve(S.forced_decay, np.zeros((1, 1)), t, np.array([[10.0, 5.0]]), rtol=1e-10,
                        atol=1e-12, first_step=1e-4, lu_precision="fp64")
    np.testing.assert_allclose(out[0, :, 0], S.forced_decay_exact(t, 10.0, 5.0), rtol=1e-8,
                               atol=1e-10)


def test_default_first_step_matches_explicit():
    """tests/test_solvers.py:71-96: first_step=None is 1e-6 of the window."""
    fn, _ = S.vdp_lattice(1, mu=1.0)
    y0, p = S.vdp_scenario(1, 8, 42, divergence=1.0)
    window = float(S.VDP_TIMES[-1] - S.VDP_TIMES[0])
    for solve in (rodas5P.solve, tsit5.solve):
        a = solve(fn, y0, S.VDP_TIMES, p, rtol=1e-6, atol=1e-8, first_step=window * 1e-6)
        b = solve(fn, y0, S.VDP_TIMES, p, rtol=1e-6, atol=1e-8)
        np.testing.assert_array_equal(a, b)


def test_tf_index_ends_each_trajectory_at_its_own_time():
    """tests/test_solver_options.py:179-206."""
    t = np.linspace(0.0, 1.0, 11)
    params = np.array([[10.0, 5.0, 1.0], [10.0, 5.0, 0.5]])
    out = rodas5P.solve(S.forced_decay, np.zeros((2, 1)), t, params, tf_index=2, rtol=1e-10,
                        atol=1e-12, first_step=1e-4, lu_precision="fp64")
    exact = S.forced_decay_exact(t, 10.0, 5.0)
    np.testing.assert_allclose(out[0, :, 0], exact, rtol=1e-8, atol=1e-10)
    stop = int(np.searchsorted(t, 0.5, side="right"))
    np.testing.assert_allclose(out[1, :stop, 0], exact[:stop], rtol=1e-8, atol=1e-10)
    assert (out[1, stop:, 0] == out[1, stop - 1, 0]).all()


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
def test_a_resumed_pilot_gives_the_unordered_solves_bits(lu):
    """``work_order="pilot"`` continues every Rodas5P trajectory from the step
    state its pilot left (``mdx_solve_launch_resumed``). Against the unordered
    refill solve, bit for bit with the same counters: trajectories that end
    inside the pilot (their own end time from the parameters, tf_index), ones
    that end after it, and a step budget (max_steps) reached inside the
    pilot's length and after it."""
    rng = np.random.default_rng(11)
    t = np.linspace(0.0, 1.0, 11)
    n = 300
    params = np.column_stack([10.0 * rng.uniform(0.5, 2.0, n), 5.0 * rng.uniform(0.5, 2.0, n),
                              rng.uniform(0.02, 1.0, n)])
    kw = dict(tf_index=2, rtol=1e-10, atol=1e-12, lu_precision=lu, return_stats=True)
    base, bst = rodas5P.solve(S.forced_decay, np.zeros((n, 1)), t, params, refill=True, **kw)
    assert (bst["accepted_steps"] < rodas5P.PILOT_STEPS).any()
    assert (bst["accepted_steps"] > rodas5P.PILOT_STEPS).any()
    hist, st = rodas5P.solve(S.forced_decay, np.zeros((n, 1)), t, params, work_order="pilot",
                             **kw)
    _same(hist, st, base, bst)
    y0, p = S.robertson_lognormal_ensemble(300, seed=6)
    tr = S.robertson_benchmark_times()
    for max_steps in (rodas5P.PILOT_STEPS // 2, 3 * rodas5P.PILOT_STEPS):
        kw = dict(rtol=1e-6, atol=1e-8, lu_precision=lu, max_steps=max_steps, return_stats=True)
        base, bst = rodas5P.solve(S.robertson, y0, tr, p, refill=True, **kw)
        hist, st = rodas5P.solve(S.robertson, y0, tr, p, work_order="pilot", **kw)
        _same(hist, st, base, bst)
        assert (st["loop_steps"] <= max_steps).all()


@pytest.mark.parametrize("placement", [None, "rolled", "lds"])
def test_a_resumed_tsit5_pilot_gives_the_unordered_solves_bits(placement):
    """Tsit5's pilot leaves each trajectory's state, its first-same-as-last
    stage and the controller's memory; the solve resumed from them gives the
    unordered refill solve's bits and counters: trajectories that end inside
    the 80-step pilot and after it, and a step budget inside and beyond the
    pilot's length. Lorenz (registers), and a 24-variable Van der Pol ring in
    the rolled step with its vectors in scratch and in LDS."""
    import torch

    from menodax_amd import _kernels, rodas5P as R5

    rng = np.random.default_rng(9)
    if placement is None:
        fn, nv = S.lorenz, 3
        y0, p = S.lorenz_rho_ensemble(300)
        t = np.linspace(0.0, 1.5, 21)
    else:
        fn, nv = S.vdp_lattice(12, mu=1.0)[0], 24
        y0 = rng.uniform(-2, 2, size=(300, nv))
        p = rng.uniform(0.5, 3.0, size=(300, 1))
        t = np.linspace(0.0, 1.6, 9)
    built = _kernels.tsit5_kernel(fn, nv, 1, _kernels.Tsit5Options(placement=placement))
    dev = torch.device("cuda")
    args = (built, torch.tensor(y0, device=dev), torch.tensor(t, device=dev),
            torch.tensor(p, device=dev), len(y0))
    for max_steps in (100000, R5.PILOT_STEPS // 2, 3 * R5.PILOT_STEPS // 2):
        kw = dict(dt0=0.0, rtol=1e-7, atol=1e-9, max_steps=max_steps)
        base, bst = R5.finish(R5.launch(*args, refill=True, **kw), True, True)
        hist, st = R5.finish(R5.launch(*args, order="pilot", **kw), True, True)
        _same(hist, st, base, bst)
        if max_steps == 100000:
            steps = bst["loop_steps"]
            assert (steps < R5.PILOT_STEPS).any() and (steps > R5.PILOT_STEPS).any()
        else:
            assert (st["loop_steps"] <= max_steps).all()


def test_a_large_tsit5_ensemble_claims_in_the_resumed_pilot_order_by_default(monkeypatch):
    """tsit5.solve, given no work_order, ranks an ensemble too large to be
    resident at once by the resumed pilot, with the bits of index order."""
    from menodax_amd import _kernels, _runtime

    built = _kernels.tsit5_kernel(S.lorenz, 3, 1, _kernels.Tsit5Options())
    n = _runtime.persistent_grid(built.refill) * built.block + 4096
    y0, p = S.lorenz_rho_ensemble(n)
    t = np.linspace(0.0, 1.0, 11)
    kw = dict(rtol=1e-6, atol=1e-8, return_stats=True)
    pilots = []
    real = rodas5P._pilot_order

    def spy(*a, **k):
        pilots.append(a[4])
        return real(*a, **k)

    monkeypatch.setattr(rodas5P, "_pilot_order", spy)
    hist, st = tsit5.solve(S.lorenz, y0, t, p, **kw)
    assert pilots == [n]
    ref, rst = tsit5.solve(S.lorenz, y0, t, p, refill=True, **kw)
    assert pilots == [n]
    _same(hist, st, ref, {k: np.asarray(v) for k, v in rst.items()})


def test_an_empty_ensemble_gives_empty_results():
    """N = 0 through every entry: the solves (plain, in the pilot's order,
    with sensitivities) and the adjoint return arrays with no trajectories."""
    t = S.robertson_benchmark_times()
    y0, p = np.zeros((0, 3)), np.zeros((0, 3))
    for kw in (dict(), dict(work_order="pilot"), dict(refill=True)):
        hist, st = rodas5P.solve(S.robertson, y0, t, p, return_stats=True, **kw)
        assert hist.shape == (0, t.size, 3)
        assert all(np.asarray(v).shape == (0,) for v in st.values())
        hist = tsit5.solve(S.lorenz, y0, t[:5], np.zeros((0, 1)), **kw)
        assert hist.shape == (0, 5, 3)
    state, sens = rodas5P.solve_sensitivities(S.robertson, y0, t, p)
    assert state.shape == (0, t.size, 3) and sens.shape == (0, t.size, 3, 6)
    gy, gp = rodas5P.adjoint_gradients(S.robertson, y0, t, p, np.zeros((0, t.size, 3)))
    assert gy.shape == (0, 3) and gp.shape == (0, 3)


def test_a_single_save_time_is_the_initial_state():
    """t_span of one time: the window is empty, the history is y0 and no step
    is taken, in both methods and launch shapes, as the oracle has it."""
    y0, p = S.robertson_lognormal_ensemble(100, seed=9)
    t = np.array([0.0])
    ref, rst = O.rodas5p(O.robertson(), y0, t, p, rtol=1e-6, atol=1e-8)
    for refill in (False, True):
        hist, st = rodas5P.solve(S.robertson, y0, t, p, rtol=1e-6, atol=1e-8,
                                 return_stats=True, refill=refill)
        _same(hist, st, ref, rst)
        assert (st["loop_steps"] == 0).all()
        np.testing.assert_array_equal(hist[:, 0], y0)
        yl, pl = S.lorenz_rho_ensemble(100)
        hist, st = tsit5.solve(S.lorenz, yl, t, pl, return_stats=True, refill=refill)
        np.testing.assert_array_equal(hist[:, 0], yl)
        assert (st["loop_steps"] == 0).all()


def test_a_zero_step_budget_holds_the_initial_state():
    """max_steps=0: every save holds y0, as the oracle has it, also when a
    pilot order is asked for (the pilot then takes no step and ranks nothing)."""
    y0, p = S.robertson_lognormal_ensemble(100, seed=9)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8, max_steps=0, return_stats=True)
    ref, rst = O.rodas5p(O.robertson(), y0, t, p, rtol=1e-6, atol=1e-8, max_steps=0)
    for extra in (dict(), dict(refill=True), dict(work_order="pilot")):
        hist, st = rodas5P.solve(S.robertson, y0, t, p, **kw, **extra)
        _same(hist, st, ref, rst)


def test_max_steps_exhaustion_holds_the_last_state():
    y0, p = S.robertson_lognormal_ensemble(100, seed=4)
    t = S.robertson_benchmark_times()
    hist, st = rodas5P.solve(S.robertson, y0, t, p, rtol=1e-6, atol=1e-8, max_steps=7,
                             return_stats=True)
    ref, rst = O.rodas5p(O.robertson(), y0, t, p, rtol=1e-6, atol=1e-8, max_steps=7)
    _same(hist, st, ref, rst)
    assert (st["loop_steps"] == 7).all()


def test_error_weights_and_pid_gains_bit_exact():
    y0, p = S.robertson_lognormal_ensemble(300, seed=5)
    t = S.robertson_benchmark_times()
    w = np.random.default_rng(0).uniform(0.0, 2.0, size=(300, 3))
    kw = dict(rtol=1e-7, atol=1e-9, error_weights=w, pcoeff=0.2, icoeff=0.6, dcoeff=0.1)
    hist, st = rodas5P.solve(S.robertson, y0, t, p, return_stats=True, lu_precision="fp64", **kw)
    ref, rst = O.rodas5p(O.robertson(), y0, t, p, lu_precision="fp64", weights=w,
                         **{k: v for k, v in kw.items() if k != "error_weights"})
    _same(hist, st, ref, rst)


@pytest.mark.parametrize("n", [1, 63, 65, 129])
def test_partial_waves_and_broadcast_inputs(n):
    """1-D y0 broadcast over a 2-D params (modax/_jax_common.py:12-46); partial
    last wavefronts."""
    _, p = S.robertson_lognormal_ensemble(n, seed=6)
    y0 = np.array([1.0, 0.0, 0.0])
    t = S.robertson_benchmark_times()
    hist, st = rodas5P.solve(S.robertson, y0, t, p, rtol=1e-6, atol=1e-8, return_stats=True)
    ref, rst = O.rodas5p(O.robertson(), np.tile(y0, (n, 1)), t, p, rtol=1e-6, atol=1e-8)
    assert hist.shape == (n, t.size, 3)
    _same(hist, st, ref, rst)


def test_torch_inputs_give_torch_outputs_and_both_layouts_agree():
    import torch

    from menodax_amd import _common, _kernels, _runtime

    y0, p = S.robertson_lognormal_ensemble(500, seed=8)
    t = S.robertson_benchmark_times()
    dev = torch.device("cuda", 0)
    out = rodas5P.solve(S.robertson, torch.tensor(y0, device=dev), torch.tensor(t, device=dev),
                        torch.tensor(p, device=dev), rtol=1e-6, atol=1e-8)
    assert isinstance(out, torch.Tensor) and out.shape == (500, t.size, 3)
    # the trajectory-major layout through the C ABI directly
    built = _kernels.rodas5p_kernel(S.robertson, 3, 3, _kernels.KernelOptions())
    y0d = torch.tensor(y0, device=dev)
    pd = torch.tensor(p, device=dev)
    td = torch.tensor(t, device=dev)
    hist = torch.empty((500, t.size, 3), dtype=torch.float64, device=dev)
    acc, rej, lp = (torch.empty(500, dtype=torch.int32, device=dev) for _ in range(3))
    _runtime.solve_launch(built.plain, torch.cuda.current_stream().cuda_stream, 0, 500, 3, 3,
                          t.size, _common.LAYOUT_TRAJECTORY_MAJOR, y0d.data_ptr(), td.data_ptr(),
                          pd.data_ptr(), 0.0, 1e-6, 1e-8, 100000, 0, hist.data_ptr(),
                          acc.data_ptr(), rej.data_ptr(), lp.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hist.cpu().numpy(), out.cpu().numpy())


# --- Rodas5P, sparse ------------------------------------------------------------
@pytest.mark.parametrize("lu", ["fp32", "fp64"])
@pytest.mark.parametrize("ordering", ["amd", "natural"])
def test_ring_with_fill_in_sparse_bit_exact(lu, ordering):
    """tests/test_sparse_direct.py:367-382: a pattern whose LU fills in, so the
    buffer must be cleared before the sweeps."""
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    mask = S.ring_mask(6)
    y0 = np.zeros((64, 6))
    y0[:, 0] = 1.0
    p = np.tile(np.array([50.0, 1.0]), (64, 1))
    p[:, 0] *= np.linspace(0.5, 1.5, 64)
    t = np.array([0.0, 0.01, 0.2])
    kw = dict(rtol=1e-9, atol=1e-11, first_step=1e-6, lu_precision=lu)
    hist, st = rodas5P.solve(S.ring6, y0, t, p, sparsity=mask, ordering=ordering,
                             return_stats=True, **kw)
    layout = _structure(mask, 6, ordering)
    ref, rst = O.rodas5p(O.ring6(), y0, t, p, layout=layout, **kw)
    _same(hist, st, ref, rst)
    dense = rodas5P.solve(S.ring6, y0, t, p, **kw)
    np.testing.assert_allclose(hist, dense, rtol=1e-6 if lu == "fp32" else 1e-8, atol=1e-10)


def test_chain_sparse_matches_closed_form():
    """tests/test_sparsity.py:113-127."""
    y0 = np.tile(np.array([1.0, 0.0, 0.0, 0.0]), (64, 1))
    p = np.tile(np.array([0.7, 0.4, 0.2]), (64, 1))
    t = np.array([0.0, 0.5, 2.0])
    kw = dict(rtol=1e-11, atol=1e-13, first_step=1e-4, lu_precision="fp64")
    dense = rodas5P.solve(S.chain4, y0, t, p, **kw)
    sparse = rodas5P.solve(S.chain4, y0, t, p, sparsity=S.CHAIN4_MASK, **kw)
    np.testing.assert_allclose(sparse, dense, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(sparse[0, -1], S.chain4_exact(2.0), rtol=1e-9, atol=1e-11)


def test_vdp_chain_full_window_mu_stratified_bit_exact():
    """The configured C3 window, t in [0, 2000], on 8 trajectories spread over
    mu in [10, 1000] (the 16,384-trajectory ensemble sorted by mu): the small-mu
    ones run into max_steps = 100,000 (their window holds ~125 relaxation
    periods of 32 oscillators out of phase), the large-mu ones finish; both
    are bit-exact with the oracle, tail fill included."""
    fn, _ = S.vdp_lattice(32, mu=1.0, d=0.1, wrap=False)
    mask = S.vdp_sparsity(32, wrap=False)
    y0, mu = S.vdp_chain_ensemble(32, 16384)
    idx = np.argsort(mu[:, 0], kind="stable")[np.linspace(0, 16383, 8).astype(int)]
    t = np.linspace(0.0, 2000.0, 11)
    print(f"vdp chain full window: mu {np.round(mu[idx, 0], 1)}", flush=True)
    hist, st = rodas5P.solve(fn, y0[idx], t, mu[idx], sparsity=mask, rtol=1e-6, atol=1e-8,
                             lu_precision="fp32", return_stats=True, max_steps=100000)
    layout = _structure(mask, 64, "amd")
    ref, rst = O.rodas5p(O.vdp_lattice(32, mu=1.0, d=0.1, wrap=False), y0[idx], t, mu[idx],
                         layout=layout, rtol=1e-6, atol=1e-8, lu_precision="fp32",
                         max_steps=100000)
    _same(hist, st, ref, rst)
    assert np.asarray(st["loop_steps"]).max() == 100000  # the cap is reached
    assert np.asarray(st["loop_steps"]).min() < 100000  # and not everywhere


@pytest.mark.parametrize("refill", [False, True])
def test_vdp_chain_sparse_fp32_bit_exact(refill):
    """The stiff VdP chain (64 variables, block-tridiagonal), sparse FP32 LU with AMD."""
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    fn, _ = S.vdp_lattice(32, mu=1.0, d=0.1, wrap=False)
    mask = S.vdp_sparsity(32, wrap=False)
    y0, mu = S.vdp_chain_ensemble(32, 256)
    t = np.linspace(0.0, 200.0, 11)
    hist, st = rodas5P.solve(fn, y0, t, mu, sparsity=mask, rtol=1e-6, atol=1e-8,
                             lu_precision="fp32", return_stats=True, refill=refill)
    layout = _structure(mask, 64, "amd")
    ref, rst = O.rodas5p(O.vdp_lattice(32, mu=1.0, d=0.1, wrap=False), y0, t, mu, layout=layout,
                         rtol=1e-6, atol=1e-8, lu_precision="fp32")
    _same(hist, st, ref, rst)


def test_block_arrow_sparse_bit_exact():
    """The uneven stiff ensemble's system, 50 variables, sparse FP32 LU."""
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    fn, mask, (rows, cols, vals, pidx) = S.arrow_system(50)
    y0, lam = S.arrow_ensemble(300)
    t = np.linspace(0.0, 10.0, 20)
    system = O.linquad(50, rows, cols, vals, 0.01, off_pidx=pidx, n_p=S.arrow_n_params(50))
    for lu in ("fp32", "fp64"):
        hist, st = rodas5P.solve(fn, y0, t, lam, sparsity=mask, rtol=1e-6, atol=1e-8,
                                 lu_precision=lu, return_stats=True, refill=True)
        layout = _structure(mask, 50, "amd")
        ref, rst = O.rodas5p(system, y0, t, lam, layout=layout, rtol=1e-6, atol=1e-8,
                             lu_precision=lu)
        _same(hist, st, ref, rst)


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
@pytest.mark.parametrize("placement", ["registers", "rolled", "lds"])
def test_every_placement_is_bit_exact(placement, lu):
    """Unrolled in registers, rolled through scratch, rolled through LDS: the
    placement moves data, never arithmetic, so each matches the oracle bit for
    bit in both launch shapes."""
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    fn, mask, (rows, cols, vals, pidx) = S.arrow_system(24)
    n_p = S.arrow_n_params(24)
    y0, lam = S.arrow_ensemble(200, n=24)
    t = np.linspace(0.0, 10.0, 7)
    system = O.linquad(24, rows, cols, vals, 0.01, off_pidx=pidx, n_p=n_p)
    layout = _structure(mask, 24, "amd")
    ref, rst = O.rodas5p(system, y0, t, lam, layout=layout, rtol=1e-6, atol=1e-8, lu_precision=lu)
    for refill in (False, True):
        hist, st = rodas5P.solve(fn, y0, t, lam, sparsity=mask, rtol=1e-6, atol=1e-8,
                                 lu_precision=lu, return_stats=True, refill=refill,
                                 placement=placement)
        _same(hist, st, ref, rst)


@pytest.mark.parametrize("block", [31, 48, 100])
@pytest.mark.parametrize("placement", ["registers", "lds"])
def test_workgroups_of_any_size_are_bit_exact(placement, block):
    """trajectories_per_block takes any positive count (modax/rodas5P.py:120-134).
    A workgroup filling part of a wavefront (31, 48 lanes) or one and a part
    (100) changes which lanes share a wave and an LDS block -- the [element][lane]
    rows then have that many lanes -- and never a trajectory's arithmetic:
    registers and the rolled LDS step, both launch shapes and a claim order,
    against the oracle; Tsit5's LDS step too."""
    from menodax_amd import _kernels, rodas5P as R5

    fn, mask, (rows, cols, vals, pidx) = S.arrow_system(24)
    n_p = S.arrow_n_params(24)
    y0, lam = S.arrow_ensemble(400, n=24)
    t = np.linspace(0.0, 10.0, 7)
    system = O.linquad(24, rows, cols, vals, 0.01, off_pidx=pidx, n_p=n_p)
    layout = _structure(mask, 24, "amd")
    ref, rst = O.rodas5p(system, y0, t, lam, layout=layout, rtol=1e-6, atol=1e-8,
                         lu_precision="fp32")
    kw = dict(sparsity=mask, rtol=1e-6, atol=1e-8, lu_precision="fp32", return_stats=True,
              placement=placement, trajectories_per_block=block)
    for extra in (dict(refill=False), dict(refill=True), dict(work_order="pilot")):
        hist, st = rodas5P.solve(fn, y0, t, lam, **kw, **extra)
        _same(hist, st, ref, rst)
    if placement != "lds":
        return
    import torch

    fv, _ = S.vdp_lattice(12, mu=1.0)
    rng = np.random.default_rng(9)
    y0v = rng.uniform(-2, 2, size=(300, 24))
    pv = rng.uniform(0.5, 3.0, size=(300, 1))
    tv = np.linspace(0.0, 2.0, 9)
    ref, rst = O.tsit5(O.vdp_lattice(12, mu=1.0), y0v, tv, pv, rtol=1e-7, atol=1e-9)
    built = _kernels.tsit5_kernel(fv, 24, 1, _kernels.Tsit5Options(
        placement="lds", trajectories_per_block=block))
    assert built.block == block
    dev = torch.device("cuda")
    for refill in (False, True):
        out = R5.launch(built, torch.tensor(y0v, device=dev), torch.tensor(tv, device=dev),
                        torch.tensor(pv, device=dev), 300, dt0=0.0, rtol=1e-7, atol=1e-9,
                        max_steps=100000, refill=refill)
        hist, st = R5.finish(out, True, True)
        _same(hist, st, ref, rst)


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
def test_the_tableau_in_registers_changes_no_bit(lu):
    """The unrolled step with its coefficients copied into registers once per
    kernel, against the oracle, in both launch shapes and for a 2-variable
    system too; the size rule picks it for these ensembles."""
    from menodax_amd import _kernels

    y0, p = S.robertson_lognormal_ensemble(3000, seed=2)
    t = S.robertson_benchmark_times()
    ref, rst = O.rodas5p(O.robertson(), y0, t, p, rtol=1e-6, atol=1e-8, lu_precision=lu)
    assert _kernels.tableau_in_registers(3, len(y0), 256)
    for tab in (None, True, False):
        for refill in (False, True):
            hist, st = rodas5P.solve(S.robertson, y0, t, p, rtol=1e-6, atol=1e-8,
                                     lu_precision=lu, return_stats=True, refill=refill,
                                     tableau_in_registers=tab)
            _same(hist, st, ref, rst)
    fn, _ = S.vdp_lattice(1, mu=100.0, d=0.1, wrap=False)
    y0v = np.random.default_rng(5).uniform(-2.0, 2.0, (500, 2))
    tv = np.linspace(0.0, 200.0, 11)
    refv, rstv = O.rodas5p(O.vdp_lattice(1, mu=100.0, d=0.1, wrap=False), y0v, tv, np.ones((500, 1)), rtol=1e-6,
                           atol=1e-8, lu_precision=lu)
    for tab in (True, False):
        hist, st = rodas5P.solve(fn, y0v, tv, np.ones((500, 1)), rtol=1e-6, atol=1e-8,
                                 lu_precision=lu, return_stats=True, tableau_in_registers=tab)
        _same(hist, st, refv, rstv)


@pytest.mark.parametrize("tab", [False, True])
def test_subnormal_stage_quotients_match_the_divisions(tab):
    """States of 1e-41..1e-34 put the FP32 back substitution's quotients in
    the float subnormal range, and some below half the smallest subnormal:
    the solves that produce them are redone with divisions, so every bit
    still matches the oracle, which always divides."""
    fn, _, (rows, cols, vals, pidx) = S.arrow_system(3)
    n_p = S.arrow_n_params(3)
    y0, lam = S.arrow_ensemble(512, n=3)
    y0 = y0 * 10.0 ** np.random.default_rng(7).uniform(-41.0, -34.0, (512, 1))
    t = np.linspace(0.0, 2.0, 5)
    system = O.linquad(3, rows, cols, vals, 0.01, off_pidx=pidx, n_p=n_p)
    ref, rst = O.rodas5p(system, y0, t, lam, rtol=1e-6, atol=1e-8, lu_precision="fp32")
    hist, st = rodas5P.solve(fn, y0, t, lam, rtol=1e-6, atol=1e-8, lu_precision="fp32",
                             return_stats=True, tableau_in_registers=tab)
    _same(hist, st, ref, rst)


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
@pytest.mark.parametrize("n_osc", [12, 20])
def test_the_wave_placement_is_bit_exact(lu, n_osc):
    """One trajectory per 64-lane workgroup, the lanes sharing out the dense
    factorisation and forward substitution: every trajectory matches the
    oracle bit for bit; for these ensemble sizes it is also the size rule's
    default placement."""
    fn, _ = S.vdp_lattice(n_osc)
    y0, p = S.vdp_scenario(n_osc, 300, 11, divergence=0.25)
    kw = dict(first_step=1e-4, rtol=1e-6, atol=1e-8, lu_precision=lu, return_stats=True)
    ref, rst = O.rodas5p(O.vdp_lattice(n_osc), y0, S.VDP_TIMES, p, first_step=1e-4, rtol=1e-6,
                         atol=1e-8, lu_precision=lu)
    for placement in ("wave", None):
        hist, st = rodas5P.solve(fn, y0, S.VDP_TIMES, p, placement=placement, **kw)
        _same(hist, st, ref, rst)


def _pivoting_blocks(n=20, size=200, seed=4):
    """dy/dt = A(p) y + 0.01 y*y with 2x2 blocks [[p_j, 1], [-b'_k, p_j]] (b' =
    LogUniform[1e2, 1e6]) chained by 0.5 couplings: for a fast block the
    iteration matrix's subdiagonal b' exceeds its diagonal 1/(h gamma) + |p_j|,
    so the dense factorisation swaps rows, for some blocks and not others."""
    rows, cols, vals, pidx = [], [], [], []
    for k in range(n // 2):
        rows += [2 * k, 2 * k + 1, 2 * k]
        cols += [2 * k + 1, 2 * k, (2 * k + 2) % n]
        vals += [1.0, -1.0, 0.5]
        pidx += [-1, n + k, -1]
    terms = [[f"p[{j}] * y[{j}]"] for j in range(n)]
    for r, c, v, k in zip(rows, cols, vals, pidx):
        terms[r].append(("-" if v < 0 else "") + f"p[{k}] * y[{c}]" if k >= 0
                        else f"{float(v)!r} * y[{c}]")
    fn = S.make_tuple_callback(
        "pivoting", [" + ".join(t) + f" + 0.01 * (y[{j}] * y[{j}])" for j, t in enumerate(terms)])
    rng = np.random.default_rng(seed)
    p = np.concatenate([-rng.uniform(5.0, 20.0, (size, n)),
                        10.0 ** rng.uniform(2.0, 6.0, (size, n // 2))], axis=1)
    y0 = rng.uniform(-1.0, 1.0, (size, n))
    system = O.linquad(n, rows, cols, vals, 0.01, off_pidx=pidx, n_p=n + n // 2)
    return fn, system, y0, p


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
def test_the_wave_placement_pivots_as_the_sequential_scan(lu):
    """A dense system whose factorisation swaps rows (a different set per
    trajectory and step): the wave placement's butterfly pivot search, row
    swaps and shared updates match the oracle's sequential LU bit for bit."""
    fn, system, y0, p = _pivoting_blocks()
    t = np.linspace(0.0, 0.2, 5)
    ref, rst = O.rodas5p(system, y0, t, p, rtol=1e-6, atol=1e-8, lu_precision=lu)
    for placement in ("wave", "rolled"):
        hist, st = rodas5P.solve(fn, y0, t, p, rtol=1e-6, atol=1e-8, lu_precision=lu,
                                 placement=placement, return_stats=True)
        _same(hist, st, ref, rst)


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
def test_the_wave_placement_of_a_joint_sensitivity_solve_changes_no_bit(lu):
    """The joint [y, S] solve (sensitivities in the parameter) on one 64-lane
    workgroup per trajectory, the size rule's choice here, against the rolled
    per-lane kernel: states and sensitivities bit-identical."""
    fn, _ = S.vdp_lattice(12)
    y0, p = S.vdp_scenario(12, 200, 5, divergence=0.25)
    kw = dict(first_step=1e-4, rtol=1e-6, atol=1e-8, lu_precision=lu, wrt_y0=False)
    got = rodas5P.solve_sensitivities(fn, y0, S.VDP_TIMES, p, **kw)
    want = rodas5P.solve_sensitivities(fn, y0, S.VDP_TIMES, p, placement="rolled", **kw)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


def test_a_claim_order_changes_no_bit():
    """The persistent launch handing trajectories out in a given order -- the
    costliest first, shuffled, or ranked by a short pilot -- computes every
    trajectory as the oracle does."""
    from menodax_amd import _common, _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    fn, mask, (rows, cols, vals, pidx) = S.arrow_system(24)
    n_p = S.arrow_n_params(24)
    y0, lam = S.arrow_ensemble(3000, n=24)
    t = np.linspace(0.0, 10.0, 7)
    system = O.linquad(24, rows, cols, vals, 0.01, off_pidx=pidx, n_p=n_p)
    layout = _structure(mask, 24, "amd")
    ref, rst = O.rodas5p(system, y0, t, lam, layout=layout, rtol=1e-6, atol=1e-8, lu_precision="fp32")
    kw = dict(sparsity=mask, rtol=1e-6, atol=1e-8, lu_precision="fp32", return_stats=True)
    cost = _common.cost_order(rst["accepted_steps"] + rst["rejected_steps"])
    shuffled = np.random.default_rng(3).permutation(len(y0))
    for order in (cost, shuffled, "pilot"):
        hist, st = rodas5P.solve(fn, y0, t, lam, work_order=order, **kw)
        _same(hist, st, ref, rst)
    y0l, pl = S.lorenz_rho_ensemble(1024)
    tl = np.linspace(0.0, 1.0, 100)
    ref, rst = O.tsit5(O.lorenz(), y0l, tl, pl, rtol=1e-6, atol=1e-8)
    for order in (np.random.default_rng(4).permutation(1024), "pilot"):
        hist, st = tsit5.solve(S.lorenz, y0l, tl, pl, rtol=1e-6, atol=1e-8, return_stats=True,
                               work_order=order)
        _same(hist, st, ref, rst)


def test_a_large_ensemble_claims_in_the_resumed_pilot_order_by_default(monkeypatch):
    """Given no work_order, a solve whose ensemble size picks the persistent
    launch (20,000 block-arrow trajectories of 50 variables against 16,384
    resident lanes) ranks its trajectories with the resumed pilot, and gives
    the bits of the same solve claimed in index order; a small ensemble, or an
    explicit refill=True, runs no pilot."""
    fn, mask, _ = S.arrow_system(50)
    y0, lam = S.arrow_ensemble(20000)
    t = np.linspace(0.0, 10.0, 20)
    kw = dict(sparsity=mask, rtol=1e-6, atol=1e-8, lu_precision="fp32", return_stats=True)
    pilots = []
    real = rodas5P._pilot_order

    def spy(*a, **k):
        pilots.append(a[4])  # the ensemble size
        return real(*a, **k)

    monkeypatch.setattr(rodas5P, "_pilot_order", spy)
    hist, st = rodas5P.solve(fn, y0, t, lam, **kw)
    assert pilots == [20000]
    ref, rst = rodas5P.solve(fn, y0, t, lam, refill=True, **kw)
    assert pilots == [20000]
    _same(hist, st, ref, {k: np.asarray(v) for k, v in rst.items()})
    rodas5P.solve(fn, y0[:1000], t, lam[:1000], **kw)
    assert pilots == [20000]


@pytest.mark.parametrize("ordered", [False, True])
@pytest.mark.parametrize("grid", [1, 7, 300])
def test_refill_claims_every_trajectory_exactly_once(grid, ordered):
    """The persistent launch's invariant. A lane claims until one claim lands
    at or past N and then stops, and the counter's atomic adds hand out
    disjoint ranges, so a final counter of N + (lanes in the grid) means the
    claims covered [0, N) once each and every lane stopped after exactly one
    failed claim. With that, the results equal the one-lane-per-trajectory
    kernel's (every trajectory solved, none twice)."""
    import torch

    from menodax_amd import _common, _kernels

    n = 3000
    y0, p = S.robertson_lognormal_ensemble(n, seed=4)
    t = S.robertson_benchmark_times()
    built = _kernels.rodas5p_kernel(S.robertson, 3, 3, _kernels.KernelOptions())
    dev = torch.device("cuda")
    times = torch.tensor(t, device=dev)
    y0_t, p_t = torch.tensor(y0, device=dev), torch.tensor(p, device=dev)
    kw = dict(dt0=0.0, rtol=1e-6, atol=1e-8, max_steps=100000)
    plain = rodas5P.launch(built, y0_t, times, p_t, n, refill=False, **kw)
    ws = torch.full((256,), 0xAB, dtype=torch.uint8, device=dev)  # the launch must clear it
    order = (torch.tensor(np.random.default_rng(grid).permutation(n), device=dev)
             if ordered else None)
    out = rodas5P.launch(built, y0_t, times, p_t, n, refill=True, grid=grid, workspace=ws,
                         order=order, **kw)
    torch.cuda.synchronize()
    counter = int(ws[:8].view(torch.int64).item())
    assert counter == n + grid * built.block
    for a, b in ((out.hist, plain.hist), (out.accepted, plain.accepted),
                 (out.rejected, plain.rejected), (out.loop, plain.loop)):
        assert torch.equal(a, b)
    assert _common.solver_stats(out.accepted, out.rejected, out.loop)["accepted_steps"].min() > 0


def test_array_rhs_stage_evaluations_are_bit_identical():
    """The array-form callback computing the same expressions drives the eight
    stage evaluations; the Jacobian still comes from ode_fn."""
    y0, p = S.robertson_lognormal_ensemble(500, seed=6)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8, return_stats=True)
    ref, rst = rodas5P.solve(S.robertson, y0, t, p, **kw)
    hist, st = rodas5P.solve(S.robertson, y0, t, p, array_rhs=S.robertson_array, **kw)
    _same(hist, st, ref, {k: np.asarray(v) for k, v in rst.items()})


# --- Tsit5 --------------------------------------------------------------------
@pytest.mark.parametrize("tab", [None, False, True])
@pytest.mark.parametrize("refill", [False, True])
def test_lorenz_rho_ensemble_bit_exact(refill, tab):
    """The Tsit5 configuration: rho ~ U[0, 28], 1,024 trajectories, 100 saves on
    [0, 1]; with the coefficients as constants or held in registers (the size
    rule's choice here)."""
    y0, p = S.lorenz_rho_ensemble(1024)
    t = np.linspace(0.0, 1.0, 100)
    hist, st = tsit5.solve(S.lorenz, y0, t, p, rtol=1e-6, atol=1e-8, return_stats=True,
                           refill=refill, tableau_in_registers=tab)
    ref, rst = O.tsit5(O.lorenz(), y0, t, p, rtol=1e-6, atol=1e-8)
    _same(hist, st, ref, rst)


@pytest.mark.parametrize("divergence, golden", [
    (0.1, (699.81274, 664, 735, 102.78706)),
    (1.0, (414.56557, 35, 734, 54.40924)),
])
def test_lorenz_divergence_is_the_published_distribution_on_gpu(divergence, golden):
    """benchmarks/nonstiff_lorenz_divergence csv:3, :12 -- 100,000 trajectories."""
    y0, p = S.lorenz_scenario(100000, 42, divergence=divergence)
    _, st = tsit5.solve(S.lorenz, y0, np.array(S.LORENZ_TIMES), p, first_step=1e-4, rtol=1e-6,
                        atol=1e-8, return_stats=True)
    att = st["accepted_steps"].astype(np.int64) + st["rejected_steps"]
    assert (att.mean(), att.min(), att.max(), st["rejected_steps"].mean()) == pytest.approx(golden)


def test_tsit5_unreached_saves_are_nan_and_counters_match():
    y0, p = S.lorenz_rho_ensemble(64)
    t = np.linspace(0.0, 1.0, 50)
    hist, st = tsit5.solve(S.lorenz, y0, t, p, rtol=1e-6, atol=1e-8, max_steps=5,
                           return_stats=True)
    ref, rst = O.tsit5(O.lorenz(), y0, t, p, rtol=1e-6, atol=1e-8, max_steps=5)
    for k in rst:
        np.testing.assert_array_equal(st[k], rst[k])
    np.testing.assert_array_equal(np.isnan(hist), np.isnan(ref))
    np.testing.assert_array_equal(hist[~np.isnan(hist)], ref[~np.isnan(ref)])


def test_tsit5_vdp_ring_twelve_steps():
    """nonstiff_vdp_divergence csv:2: 64-D ring, mu = 1, 12 steps, none rejected."""
    fn, _ = S.vdp_lattice(32, mu=1.0)
    y0, p = S.vdp_scenario(32, 128, 42, divergence=0.0)
    _, st = tsit5.solve(fn, y0, S.VDP_TIMES, p, first_step=1e-4, rtol=1e-6, atol=1e-8,
                        return_stats=True)
    assert (st["loop_steps"] == 12).all() and (st["rejected_steps"] == 0).all()


@pytest.mark.parametrize("placement", ["registers", "rolled", "lds"])
def test_tsit5_placements_are_bit_exact(placement):
    """The reference's shared/local backends as placements: registers, rolled
    through scratch, rolled through LDS -- the same bits as the oracle."""
    from menodax_amd import _kernels, rodas5P as R5

    fn, _ = S.vdp_lattice(12, mu=1.0)
    rng = np.random.default_rng(9)
    y0 = rng.uniform(-2, 2, size=(300, 24))
    p = rng.uniform(0.5, 3.0, size=(300, 1))
    t = np.linspace(0.0, 2.0, 9)
    ref, rst = O.tsit5(O.vdp_lattice(12, mu=1.0), y0, t, p, rtol=1e-7, atol=1e-9)
    built = _kernels.tsit5_kernel(fn, 24, 1, _kernels.Tsit5Options(placement=placement))
    import torch

    dev = torch.device("cuda")
    tt = torch.tensor(t, device=dev)
    for refill in (False, True):
        out = R5.launch(built, torch.tensor(y0, device=dev), tt, torch.tensor(p, device=dev), 300,
                        dt0=0.0, rtol=1e-7, atol=1e-9, max_steps=100000, refill=refill)
        hist, st = R5.finish(out, True, True)
        _same(hist, st, ref, rst)


def test_tsit5_backend_names():
    fn, _ = S.vdp_lattice(16, mu=1.0)
    y0, p = S.vdp_scenario(16, 8, 42, divergence=1.0)
    with pytest.raises(ValueError, match="shared backend requires"):
        tsit5.solve(fn, y0, S.VDP_TIMES, p, backend="shared")
    fn, _ = S.vdp_lattice(4, mu=1.0)
    y0, p = S.vdp_scenario(4, 64, 42, divergence=1.0)
    a = tsit5.solve(fn, y0, S.VDP_TIMES, p, backend="shared", first_step=1e-4, rtol=1e-6, atol=1e-8)
    b = tsit5.solve(fn, y0, S.VDP_TIMES, p, backend="local", first_step=1e-4, rtol=1e-6, atol=1e-8)
    np.testing.assert_array_equal(a, b)
