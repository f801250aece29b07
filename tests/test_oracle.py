"""The CPU oracle pinned against the reference's own published numbers and
against closed forms, before anything is held against it.

Goldens come from the reference's benchmark CSVs (RTX 4070 SUPER runs of the
reference solver):
  - benchmarks/nonstiff_lorenz_divergence/results-*.csv:2-12 -- Lorenz Tsit5,
    100,000 trajectories per divergence, step distributions;
  - benchmarks/stiff_robertson_divergence/results-*.csv:2-12 -- Robertson
    Rodas5P FP32 LU, 30,000 trajectories per divergence;
  - benchmarks/nonstiff_vdp_divergence/results-*.csv:2 -- 64-D VdP ring, Tsit5.
The Lorenz rows and the single-trajectory Robertson row (157 attempted, 10
rejected, with the initial state built as robertson.py:8-10 builds it) are
reproduced exactly. The Robertson FP32-LU ensemble rows depend on rounding:
with the FP32 LU the error estimate carries float noise, and the reference's
GPU compiler fuses multiply-adds where it chooses, so a few of 30,000
trajectories move by one step; the means agree to 2e-5 and the extremes to one
step.
"""

import math

import numpy as np
import pytest

from tests import systems as S
from oracle import oracle as O


def _structure(mask, n, ordering):
    """The oracle's own sparse structure for ``mask`` under the product's
    elimination order (the AMD permutation is the one input it takes)."""
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    pattern = normalize_sparsity(mask, n)
    return O.sparse_structure(pattern, SD.fill_reducing_order(pattern, ordering))

LORENZ_GOLDEN = {  # divergence: (mean attempted, min, max, mean rejected)
    0.0: (722.0, 722, 722, 104.0),
    0.1: (699.81274, 664, 735, 102.78706),
    0.5: (576.19952, 427, 735, 80.38524),
    1.0: (414.56557, 35, 734, 54.40924),
}
ROBERTSON_GOLDEN = {
    0.25: (123.61306666666667, 72, 174, 6.2366),
    1.0: (106.2629, 32, 184, 5.2309),
    2.0: (84.24176666666666, 18, 189, 4.5353),
}


def _dist(stats):
    att = stats["accepted_steps"].astype(np.int64) + stats["rejected_steps"]
    return float(att.mean()), int(att.min()), int(att.max()), float(stats["rejected_steps"].mean())


@pytest.mark.parametrize("divergence", sorted(LORENZ_GOLDEN))
def test_lorenz_divergence_distribution_is_the_published_one(divergence):
    n = 100000 if divergence else 1
    y0, p = S.lorenz_scenario(n, 42, divergence=divergence)
    _, st = O.tsit5(O.lorenz(), y0, np.array(S.LORENZ_TIMES), p, first_step=1e-4, rtol=1e-6,
                    atol=1e-8)
    mean, lo, hi, rej = _dist(st)
    g = LORENZ_GOLDEN[divergence]
    assert (lo, hi) == (g[1], g[2])
    assert mean == pytest.approx(g[0], rel=0, abs=1e-9)
    assert rej == pytest.approx(g[3], rel=0, abs=1e-9)


def test_robertson_single_trajectory_fp32_is_the_published_count():
    """stiff_robertson_divergence csv:2: 157 attempted, 10 rejected. The initial
    state is the reference's (1-eps)*alpha, eps, (1-eps)*(1-alpha), whose last
    component is 0.009000000000000008, not the literal 0.009 (with the literal
    the FP32 run rejects one step fewer)."""
    _, st = O.rodas5p(O.robertson(), S.ROBERTSON_Y0[None], S.ROBERTSON_TIMES,
                      S.ROBERTSON_PARAMS[None], first_step=1e-4, rtol=1e-6, atol=1e-8,
                      lu_precision="fp32")
    assert {k: int(v[0]) for k, v in st.items()} == dict(
        accepted_steps=147, rejected_steps=10, loop_steps=157)


@pytest.mark.parametrize("divergence", sorted(ROBERTSON_GOLDEN))
def test_robertson_divergence_distribution_matches_published(divergence):
    y0, p = S.robertson_scenario(30000, 42, divergence=divergence)
    _, st = O.rodas5p(O.robertson(), y0, S.ROBERTSON_TIMES, p, first_step=1e-4, rtol=1e-6,
                      atol=1e-8, lu_precision="fp32")
    mean, lo, hi, rej = _dist(st)
    g = ROBERTSON_GOLDEN[divergence]
    assert mean == pytest.approx(g[0], rel=2e-4)
    assert abs(lo - g[1]) <= 1 and abs(hi - g[2]) <= 1
    assert rej == pytest.approx(g[3], rel=2e-3)


def test_robertson_extreme_divergence_mean_matches_published():
    """Divergence 3.0 (csv:12): rates spread over six decades. The mean and the
    fastest trajectory match; the single slowest one (231 published, 212 here)
    is decided by float rounding in the FP32 LU and is not compared."""
    y0, p = S.robertson_scenario(30000, 42, divergence=3.0)
    _, st = O.rodas5p(O.robertson(), y0, S.ROBERTSON_TIMES, p, first_step=1e-4, rtol=1e-6,
                      atol=1e-8, lu_precision="fp32")
    mean, lo, _, rej = _dist(st)
    assert mean == pytest.approx(72.0164, rel=1e-4)
    assert lo == 14
    assert rej == pytest.approx(4.7871, rel=2e-3)


def test_vdp_ring_tsit5_takes_the_published_twelve_steps():
    """benchmarks/nonstiff_vdp_divergence csv:2: 64-D ring, mu = 1, 12 steps."""
    y0, p = S.vdp_scenario(32, 4, 42, divergence=0.0)
    _, st = O.tsit5(O.vdp_lattice(32, mu=1.0), y0, S.VDP_TIMES, p, first_step=1e-4, rtol=1e-6,
                    atol=1e-8)
    assert (st["loop_steps"] == 12).all() and (st["rejected_steps"] == 0).all()


def test_vdp_ring_tsit5_divergence_distribution():
    """nonstiff_vdp_divergence csv:3: divergence 0.25, 1000 trajectories."""
    y0, p = S.vdp_scenario(32, 1000, 42, divergence=0.25)
    _, st = O.tsit5(O.vdp_lattice(32, mu=1.0), y0, S.VDP_TIMES, p, first_step=1e-4, rtol=1e-6,
                    atol=1e-8)
    mean, lo, hi, rej = _dist(st)
    assert (mean, lo, hi, rej) == (27.109, 24, 31, 1.161)


def test_stiff_vdp_ring_rodas5p_takes_the_published_21_steps():
    """stiff_vdp_divergence csv:2: 96-D ring (mu = 100), dense FP32 LU, 21 steps,
    none rejected."""
    lu = "fp32"
    y0, p = S.vdp_scenario(48, 2, 42, divergence=0.0)
    _, st = O.rodas5p(O.vdp_lattice(48), y0, S.VDP_TIMES, p, first_step=1e-4, rtol=1e-6,
                      atol=1e-8, lu_precision=lu)
    assert (st["loop_steps"] == 21).all() and (st["rejected_steps"] == 0).all()


# --- closed forms -------------------------------------------------------------
@pytest.mark.parametrize("lu", ["fp32", "fp64"])
def test_forced_decay_reaches_the_closed_form(lu):
    """tests/test_solver_options.py:159-176 and test_nonautonomous.py:251-266:
    the df/dt stage term is what keeps order here."""
    t = np.linspace(0.0, 1.0, 11)
    h, _ = O.rodas5p(O.forced_decay(), np.zeros((1, 1)), t, np.array([[10.0, 5.0]]), rtol=1e-10,
                     atol=1e-12, first_step=1e-4, lu_precision=lu)
    exact = S.forced_decay_exact(t, 10.0, 5.0)
    if lu == "fp64":
        np.testing.assert_allclose(h[0, :, 0], exact, rtol=1e-8, atol=1e-10)
    else:
        assert np.max(np.abs(h[0, :, 0] - exact)) < 1e-5


def test_chain_closed_form_dense_and_sparse():
    """tests/test_sparsity.py:113-127 with the oracle's own sparse LU replay."""
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    y0 = np.array([[1.0, 0.0, 0.0, 0.0]])
    p = np.array([[0.7, 0.4, 0.2]])
    t = np.array([0.0, 0.5, 2.0])
    kw = dict(rtol=1e-11, atol=1e-13, first_step=1e-4, lu_precision="fp64")
    dense, _ = O.rodas5p(O.chain4(), y0, t, p, **kw)
    layout = _structure(S.CHAIN4_MASK, 4, "natural")
    sparse, _ = O.rodas5p(O.chain4(), y0, t, p, layout=layout, **kw)
    np.testing.assert_allclose(sparse, dense, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(sparse[0, -1], S.chain4_exact(2.0), rtol=1e-9, atol=1e-11)


def test_kaps_exact_solution():
    """reference/systems/python/kaps.py: y1 = exp(-2t), y2 = exp(-t)."""
    t = np.array([0.0, 0.5, 1.0, 2.0])
    h, _ = O.rodas5p(O.kaps(3, 1e-3), np.ones((1, 6)), t, np.ones((1, 1)), rtol=1e-9, atol=1e-11,
                     first_step=1e-4, lu_precision="fp64")
    exact = np.stack([np.exp(-2 * t), np.exp(-t)] * 3, axis=1)
    np.testing.assert_allclose(h[0], exact, rtol=1e-6, atol=1e-9)


def _one_step_error(run, exact, hs):
    """Local error of one accepted step of size h (max_steps=1 with a loose
    tolerance, the step's end state held in the last save by the tail fill)."""
    return [abs(run(h) - exact(h)) for h in hs]


def test_rodas5p_local_error_is_order_six():
    """An order-5 method's one-step error is O(h^6): halving h divides it by ~64.
    Non-autonomous on purpose, so the dt*d_i*df/dt stage terms are exercised."""
    lam, forcing = 2.0, 1.0

    def run(h):
        out, st = O.rodas5p(O.forced_decay(), np.array([[0.3]]), np.array([0.0, 10.0]),
                            np.array([[lam, forcing]]), rtol=1e3, atol=1e3, first_step=h,
                            max_steps=1, lu_precision="fp64")
        assert int(st["accepted_steps"][0]) == 1
        return out[0, -1, 0]

    def exact(h):
        return 0.3 * np.exp(-lam * h) + S.forced_decay_exact(h, lam, forcing)

    errs = _one_step_error(run, exact, (0.4, 0.2, 0.1))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all((rates > 5.4) & (rates < 6.6)), rates


def test_tsit5_local_error_is_order_six():
    w0 = np.sqrt(4.0 - 0.01)

    def run(h):
        out, st = O.tsit5(O.damped_osc(), np.array([[1.0, 0.0]]), np.array([0.0, 10.0]),
                          np.array([[4.0, 0.2]]), rtol=1e3, atol=1e3, first_step=h,
                          max_steps=1, fill=np.nan)
        assert int(st["accepted_steps"][0]) == 1
        return out[0, 1, 0] if np.isfinite(out[0, 1, 0]) else None

    # Tsit5 leaves unreached saves untouched, so read the state through a save
    # placed exactly at h instead.
    def run_at(h):
        out, st = O.tsit5(O.damped_osc(), np.array([[1.0, 0.0]]), np.array([0.0, h]),
                          np.array([[4.0, 0.2]]), rtol=1e3, atol=1e3, first_step=h,
                          max_steps=1)
        assert int(st["accepted_steps"][0]) == 1
        return out[0, 1, 0]

    del run

    def exact(h):
        return np.exp(-0.1 * h) * (np.cos(w0 * h) + 0.1 / w0 * np.sin(w0 * h))

    errs = _one_step_error(run_at, exact, (0.4, 0.2, 0.1))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all((rates > 5.4) & (rates < 6.6)), rates


def test_det_pow_is_accurate():
    """The controller's deterministic power against libm over the range it sees."""
    import ctypes

    lib = O._lib()
    lib.det_pow.restype = ctypes.c_double
    lib.det_pow.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(0)
    for x in 10.0 ** rng.uniform(-18, 18, 2000):
        for e in (-1.0 / 6.0, -1.0 / 5.0, 0.3, -0.05):
            got = lib.det_pow(x, e)
            assert got == pytest.approx(x**e, rel=4e-15)
    assert lib.det_pow(123.0, 0.0) == 1.0 and lib.det_pow(123.0, -0.0) == 1.0


def test_dense_lu_against_numpy():
    """The oracle's dense LU (through a one-step solve of a linear system) agrees
    with numpy's solve: y' = A y from y0 with one step is a linear map built
    from (I - h gamma A)^-1; instead check the closed form of the linear system."""
    from scipy.linalg import expm

    rows, cols, vals, pidx, _ = S.arrow_structure(10, seed=3)
    n_p = S.arrow_n_params(10)
    sysd = O.linquad(10, rows, cols, vals, 0.0, off_pidx=pidx, n_p=n_p)
    rng = np.random.default_rng(2)
    p = np.concatenate([-10.0 ** rng.uniform(-1, 2, size=10), rng.uniform(0.5, 3, size=n_p - 10)])
    y0 = rng.uniform(-1, 1, size=(1, 10))
    out, _ = O.rodas5p(sysd, y0, np.array([0.0, 0.5]), p[None], rtol=1e-11, atol=1e-13,
                       lu_precision="fp64")
    A = np.diag(p[:10])
    for r, c, v, k in zip(rows, cols, vals, pidx):
        A[r, c] += v * p[k] if k >= 0 else v
    np.testing.assert_allclose(out[0, -1], expm(0.5 * A) @ y0[0], rtol=1e-8, atol=1e-10)


def _uneven_sample(count):
    """``count`` trajectories spread over the 100,000-trajectory uneven stiff
    ensemble (50-variable block arrow), with its oracle system and AMD layout."""
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    _, mask, (rows, cols, vals, pidx) = S.arrow_system(50)
    sp = normalize_sparsity(mask, 50)
    system = O.linquad(50, rows, cols, vals, 0.01, off_pidx=pidx, n_p=S.arrow_n_params(50))
    layout = O.sparse_structure(sp, SD.fill_reducing_order(sp, "amd"))
    y0, p = S.arrow_ensemble(100000, n=50)
    idx = np.linspace(0, 99999, count).astype(int)
    return system, layout, y0[idx], p[idx], np.linspace(0.0, 10.0, 20)


def test_refined_fp32_solves_follow_the_fp64_solution():
    """Iterative refinement converges to the solution of the double system: one
    refinement of every FP32 stage solve brings the solve to the FP64-LU
    solve's trajectory (the same step counts, endpoints within 1e-7), where the
    unrefined FP32 solve, whose stage vectors carry FP32 rounding, lands up to
    1e-5 away on the long oscillating trajectories. A residual with a wrong
    sign or without its J x term would not converge and stays at the
    unrefined distance or beyond."""
    system, layout, y0, p, times = _uneven_sample(64)
    kw = dict(layout=layout, rtol=1e-6, atol=1e-8)
    h64, s64 = O.rodas5p(system, y0, times, p, lu_precision="fp64", **kw)
    h32, _ = O.rodas5p(system, y0, times, p, lu_precision="fp32", **kw)
    hr, sr = O.rodas5p(system, y0, times, p, lu_precision="fp32", lu_refine=True, **kw)
    d_ref = np.abs(hr[:, -1] - h64[:, -1]).max()
    d_plain = np.abs(h32[:, -1] - h64[:, -1]).max()
    assert d_plain > 1e-6
    assert d_ref < 1e-7 and d_ref < d_plain / 50
    assert np.array_equal(sr["accepted_steps"], s64["accepted_steps"])
    assert np.array_equal(sr["rejected_steps"], s64["rejected_steps"])


def test_refined_fp32_robertson_dense():
    """The same on the dense path: Robertson (C2 inputs), every step count equal
    to the FP64-LU solve's and the history within 5e-7 of it."""
    y0, p = S.robertson_lognormal_ensemble(256, seed=1)
    times = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8)
    h64, s64 = O.rodas5p(O.robertson(), y0, times, p, lu_precision="fp64", **kw)
    hr, sr = O.rodas5p(O.robertson(), y0, times, p, lu_precision="fp32", lu_refine=True, **kw)
    assert np.array_equal(sr["accepted_steps"], s64["accepted_steps"])
    assert np.abs(hr - h64).max() < 5e-7
    with pytest.raises(ValueError, match="FP32"):
        O.rodas5p(O.robertson(), y0, times, p, lu_precision="fp64", lu_refine=True)


# --- the BBN network: a non-polynomial, non-autonomous right-hand side ---------
def test_bbn_oracle_conserves_baryon_number():
    """Y_n + Y_p + 2 Y_d + 4 Y_4He = 1 is a linear invariant of the network, which
    Rosenbrock methods keep to rounding with an exact Jacobian: the FP64 solve
    holds it to 1e-12 over the prior box (a wrong sign or index in any rate
    breaks it)."""
    y0 = S.bbn_initial_state()
    p = S.bbn_grid(16)
    hist, _ = O.rodas5p(O.bbn(), np.tile(y0, (16, 1)), S.BBN_X_SPAN, p, lu_precision="fp64",
                        rtol=1e-8, atol=1e-12, first_step=0.1)
    np.testing.assert_allclose(hist @ np.array([1.0, 1.0, 2.0, 4.0]), 1.0, rtol=0, atol=1e-12)


def test_bbn_oracle_reproduces_the_consumers_calibration():
    """At Planck's eta_10 = 6.1 and N_eff = 3.044 the four-species network gives
    Y_P ~ 0.106 and D/H ~ 4.2e-6: the values the reference's consumer takes as its
    observations, calibrated to this model (examples/bbn_estimation/main.py:
    268-279). A tight FP64 solve lands within 0.5% of both."""
    y0 = S.bbn_initial_state()[None]
    p = np.array([[math.log10(6.1), 3.044]])
    hist, _ = O.rodas5p(O.bbn(), y0, S.BBN_X_SPAN, p, lu_precision="fp64", rtol=1e-10,
                        atol=1e-14, first_step=0.1)
    np.testing.assert_allclose(S.bbn_abundances(hist[0, -1]), S.BBN_OBSERVED, rtol=5e-3)


@pytest.mark.parametrize("x", [0.15, 1.0, 2.5, 2.6, 8.0, 120.0])
def test_bbn_oracle_derivatives_are_the_generated_ones(x):
    """The oracle's complex-step Jacobian and df/dx (an independent method) equal
    the code generator's forward-mode derivatives of the traced callback to
    rounding, on both sides of the g_star step at x = Q/m_e ~ 2.53."""
    from menodax_amd import _codegen as C

    traced = C.trace(S.bbn, 4, 2)
    y = np.array([0.3, 0.6, 1e-4, 0.02])
    p = np.array([0.8, 3.3])
    J, dT = O.jacobian(O.bbn(), y, x, p)
    Jg, dTg = C.jacobian_from_graph(traced, y, x, p)
    np.testing.assert_allclose(J, Jg, rtol=1e-13, atol=0)
    np.testing.assert_allclose(dT, dTg, rtol=1e-12, atol=0)
    np.testing.assert_allclose(O.rhs(O.bbn(), y, x, p), C.rhs_from_graph(traced, y, x, p),
                               rtol=1e-14, atol=0)
