"""Forward sensitivities on the CPU: the spec, the generated derivative code,
and the oracle's joint [y, S] Rodas5P against closed forms and differences.

Mirrors the reference's tests/test_sensitivity.py (cascade and forced-decay
closed forms, the error-control and coupling-block properties); the GPU kernel
is held against this oracle bit for bit in test_gpu_sensitivity.py.
"""

import numpy as np
import pytest

from menodax_amd import _codegen as cg
from tests import systems as S
from menodax_amd._sensitivity import SensitivitySpec, augmented_y0, split_augmented
from oracle import oracle as O
from oracle.callbacks import HostCallbacks

A, B = 0.7, 1.3
Y0 = np.array([2.0, 0.5])
T_SPAN = np.linspace(0.0, 2.0, 5)


# --- the spec ------------------------------------------------------------------
def test_spec_sizes_only_count_requested_blocks():
    s = SensitivitySpec(3, 4, wrt_y0=True, wrt_params=False)
    assert (s.n_sens, s.n_aug, s.direction_columns) == (3, 12, (-1, -1, -1))
    s = SensitivitySpec(3, 4, wrt_y0=False, wrt_params=True, param_columns=(2, 0))
    assert (s.n_sens, s.n_aug, s.direction_columns) == (2, 9, (2, 0))
    s = SensitivitySpec(3, 4, wrt_y0=True, wrt_params=True, error_control=False)
    assert (s.n_sens, s.n_aug, s.n_error) == (7, 24, 3)
    with pytest.raises(ValueError, match="at least one"):
        SensitivitySpec(3, 4, False, False)
    with pytest.raises(ValueError, match="repeat"):
        SensitivitySpec(3, 4, False, True, param_columns=(1, 1))
    with pytest.raises(ValueError, match="outside"):
        SensitivitySpec(3, 4, False, True, param_columns=(4,))


def test_augmented_layout_round_trips():
    import torch

    spec = SensitivitySpec(2, 3, True, True)
    y0 = torch.tensor([[1.0, 2.0], [3.0, 4.0]], dtype=torch.float64)
    z = augmented_y0(y0, spec)
    assert z.shape == (2, spec.n_aug)
    hist = z[:, None, :].expand(2, 4, spec.n_aug)
    state, sens = split_augmented(hist, spec)
    assert torch.equal(state[:, 0], y0)
    # S(t0): identity for the initial-state block, zero for the parameter block
    assert torch.equal(sens[0, 0, :, :2], torch.eye(2, dtype=torch.float64))
    assert torch.equal(sens[0, 0, :, 2:], torch.zeros(2, 3, dtype=torch.float64))


# --- the generated derivatives ---------------------------------------------------
def _fd(fn, x, d, h=1e-6):
    return (np.asarray(fn(x + h * d)) - np.asarray(fn(x - h * d))) / (2 * h)


@pytest.mark.parametrize("fn, n, n_p", [(S.robertson, 3, 3), (S.cascade, 2, 2),
                                        (S.forced_decay, 1, 2), (S.vdp_lattice(3)[0], 6, 1),
                                        (S.kaps(2, 1e-2)[0], 4, 1)])
def test_sensitivity_sweeps_against_central_differences(fn, n, n_p):
    tr = cg.trace(fn, n, n_p)
    rng = np.random.default_rng(5)
    for _ in range(5):
        y, p, t = rng.uniform(0.2, 1.5, n), rng.uniform(0.5, 2.0, n_p), rng.uniform(0.1, 1.0)
        s, pd, v = rng.normal(size=n), rng.normal(size=n_p), rng.normal(size=n)
        rhs, dt, cpl = cg.sensitivity_from_graph(tr, y, t, p, s, pd, v)
        f = lambda yy, pp, tt: np.array(cg.rhs_from_graph(tr, yy, tt, pp))  # noqa: E731
        h = 1e-6
        fd_rhs = (f(y + h * s, p + h * pd, t) - f(y - h * s, p - h * pd, t)) / (2 * h)
        np.testing.assert_allclose(rhs, fd_rhs, rtol=1e-6, atol=1e-8)
        sens = lambda yy, tt: cg.sensitivity_from_graph(tr, yy, tt, p, s, pd)[0]  # noqa: E731
        np.testing.assert_allclose(cpl, _fd(lambda yy: sens(yy, t), y, v), rtol=1e-6, atol=1e-8)
        np.testing.assert_allclose(dt, (sens(y, t + h) - sens(y, t - h)) / (2 * h),
                                   rtol=1e-6, atol=1e-8)


def test_augmented_trace_is_state_then_direction_blocks():
    tr = cg.trace(S.robertson, 3, 3)
    aug = cg.augmented_trace(tr, (-1, 2))
    assert aug.n_vars == 9
    rng = np.random.default_rng(2)
    z, p = rng.uniform(0.1, 1, 9), rng.uniform(0.5, 2, 3)
    out = np.array(cg.rhs_from_graph(aug, z, 0.3, p))
    np.testing.assert_array_equal(out[:3], cg.rhs_from_graph(tr, z[:3], 0.3, p))
    e2 = np.array([0.0, 0.0, 1.0])
    s0 = cg.sensitivity_from_graph(tr, z[:3], 0.3, p, z[3:6], np.zeros(3))[0]
    s1 = cg.sensitivity_from_graph(tr, z[:3], 0.3, p, z[6:9], e2)[0]
    np.testing.assert_array_equal(out[3:6], s0)
    np.testing.assert_array_equal(out[6:9], s1)


# --- the oracle's joint solve ------------------------------------------------------
def _joint_y0(y0, n, cols):
    n_dirs = len(cols)
    z = np.zeros(n * (1 + n_dirs))
    z[:n] = y0
    for d, c in enumerate(cols):
        if c < 0:
            z[n + d * n + d] = 1.0  # initial-state direction d is e_d
    return z


@pytest.mark.parametrize("lu", ["fp64", "fp32"])
def test_cascade_sensitivities_match_closed_form(lu):
    cb = HostCallbacks(S.cascade, 2, 2)
    cols = (-1, -1, 0, 1)
    hist, _ = O.rodas5p_callbacks(cb, _joint_y0(Y0, 2, cols), T_SPAN, [A, B],
                                  direction_columns=cols, rtol=1e-10, atol=1e-12, lu_precision=lu)
    z = hist[0]
    sens = z[:, 2:].reshape(len(T_SPAN), 4, 2).transpose(0, 2, 1)
    tol = dict(rtol=1e-6, atol=1e-9) if lu == "fp64" else dict(rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(z[:, :2], S.cascade_exact(T_SPAN, Y0, [A, B]), **tol)
    np.testing.assert_allclose(sens[:, :, :2], S.cascade_dy0(T_SPAN, Y0, [A, B]), **tol)
    np.testing.assert_allclose(sens[:, :, 2:], S.cascade_dparams(T_SPAN, Y0, [A, B]), **tol)


def test_forced_decay_sensitivities_need_the_time_derivative_rows():
    """d2f/dt dp is a nonzero constant here: the sensitivity rows of df/dt."""
    cb = HostCallbacks(S.forced_decay, 1, 2)
    t = np.linspace(0.0, 1.0, 11)
    cols = (0, 1)
    hist, _ = O.rodas5p_callbacks(cb, _joint_y0([0.0], 1, cols), t, [10.0, 5.0],
                                  direction_columns=cols, rtol=1e-8, atol=1e-10,
                                  lu_precision="fp64")
    z = hist[0]
    np.testing.assert_allclose(z[:, 0], S.forced_decay_exact(t, 10.0, 5.0), rtol=1e-7, atol=1e-12)
    exact = np.stack([S.forced_decay_dlam(t, 10.0, 5.0), S.forced_decay_dforcing(t, 10.0, 5.0)], -1)
    err = np.abs(z[:, 1:] - exact).max() / np.abs(exact).max()
    assert err < 1e-6


def test_without_error_control_the_joint_solve_keeps_the_plain_steps():
    """Zero weights and the n_vars divisor reproduce the plain error norm bit for
    bit, hence the plain step sequence and the plain state."""
    cb = HostCallbacks(S.robertson, 3, 3)
    y0, p = S.robertson_lognormal_ensemble(50, seed=4)
    t = S.robertson_benchmark_times()
    plain, ps = O.rodas5p_callbacks(cb, y0, t, p, rtol=1e-7, atol=1e-9, lu_precision="fp32")
    cols = (-1, -1, -1, 0, 1, 2)
    z0 = np.stack([_joint_y0(r, 3, cols) for r in y0])
    joint, js = O.rodas5p_callbacks(cb, z0, t, p, direction_columns=cols, error_control=False,
                                    rtol=1e-7, atol=1e-9, lu_precision="fp32")
    np.testing.assert_array_equal(joint[:, :, :3], plain)
    for k in ps:
        np.testing.assert_array_equal(js[k], ps[k])


def test_coupling_block_keeps_the_joint_step_count_near_the_plain_one():
    """Regression guard on the block-triangular coupling (tests/test_sensitivity.py:454-481
    of the reference): without it the joint solve needs many times the steps."""
    cb = HostCallbacks(S.cascade, 2, 2)
    _, ps = O.rodas5p_callbacks(cb, Y0, T_SPAN, [A, B], rtol=1e-8, atol=1e-10, lu_precision="fp64")
    cols = (0, 1)
    _, js = O.rodas5p_callbacks(cb, _joint_y0(Y0, 2, cols), T_SPAN, [A, B], direction_columns=cols,
                                rtol=1e-8, atol=1e-10, lu_precision="fp64")
    assert js["accepted_steps"][0] <= 2 * ps["accepted_steps"][0]


def test_stiff_bilinear_sensitivities_against_differences_of_the_plain_solve():
    """Robertson, bilinear in state and rate: d2f/dy dp is what the coupling block carries."""
    cb = HostCallbacks(S.robertson, 3, 3)
    y0 = np.array([1.0, 0.0, 0.0])
    p = np.array([0.04, 1e4, 3e7])
    t = np.array([0.0, 1e-3, 1e-1, 10.0, 40.0])
    kw = dict(rtol=1e-11, atol=1e-14, lu_precision="fp64")
    cols = (0, 1, 2)
    hist, _ = O.rodas5p_callbacks(cb, _joint_y0(y0, 3, cols), t, p, direction_columns=cols, **kw)
    sens = hist[0, :, 3:].reshape(len(t), 3, 3).transpose(0, 2, 1)
    for k in range(3):
        h = 1e-6 * p[k]
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (O.rodas5p_callbacks(cb, y0, t, pp, **kw)[0][0]
              - O.rodas5p_callbacks(cb, y0, t, pm, **kw)[0][0]) / (2 * h)
        scale = np.abs(fd).max(axis=0) + 1e-300
        assert np.all(np.abs(sens[:, :, k] - fd).max(axis=0) <= 2e-4 * scale)


def _callback_values(cb, n, n_p, y, p, s, q, v):
    """f, J, dT, sens_f, sens_dt and the coupling of a callback set at one point."""
    import ctypes

    D = ctypes.POINTER(ctypes.c_double)

    def call(addr, n_arrays, *arrays):
        proto = ctypes.CFUNCTYPE(None, *([D, ctypes.c_double] + [D] * n_arrays))
        proto(addr)(y.ctypes.data_as(D), 0.3, *[a.ctypes.data_as(D) for a in arrays])

    f, J, dT, sf, sd, cpl = (np.zeros(k) for k in (n, n * n, n, n, n, n))
    call(cb.addresses[0], 2, p, f)
    call(cb.addresses[1], 3, p, J, dT)
    call(cb.addresses[2], 4, p, s, q, sf)
    call(cb.addresses[3], 4, p, s, q, sd)
    call(cb.addresses[4], 5, p, s, q, v, cpl)
    return f, J, dT, sf, sd, cpl


@pytest.mark.parametrize("name", ["robertson", "vdp"])
def test_generated_sensitivity_derivatives_against_hand_derived_closed_forms(name):
    """The joint solve's derivative code -- J_y s + J_p q, its df/dt and the
    stage coupling D_y(J_y s + J_p q)[v] -- as the tracer emits it (what the
    GPU joint kernel is bit-exact with, through the oracle's callback solver)
    against closed forms derived by hand in oracle.c: equal to rounding at
    random points, and through the oracle's whole FP64-LU joint integration
    (every state and parameter direction) the same step counts on every
    trajectory and histories within 1e-9 relative. A wrong sign or a dropped
    second-derivative term moves the sensitivities by O(1)."""
    if name == "robertson":
        gen = HostCallbacks(S.robertson, 3, 3)
        y0, p = S.robertson_lognormal_ensemble(300, seed=3)
        t = S.robertson_benchmark_times()
        cols, kw = (-1, -1, -1, 0, 1, 2), dict(rtol=1e-6, atol=1e-8)
    else:
        gen = HostCallbacks(S.vdp_lattice(1, mu=1.0, d=0.1, wrap=False)[0], 2, 1)
        rng = np.random.default_rng(1)
        p = 10.0 ** rng.uniform(0.5, 1.5, size=(200, 1))
        y0 = rng.uniform(-2.0, 2.0, size=(200, 2))
        t = np.linspace(0.0, 3.0, 20)
        cols, kw = (-1, -1, 0), dict(rtol=1e-8, atol=1e-10)
    cf = O.ClosedFormCallbacks(name)
    n, n_p = cf.n, cf.n_p
    rng = np.random.default_rng(7)
    for _ in range(100):
        args = (rng.uniform(-2, 2, n), rng.uniform(0.5, 3.0, n_p), rng.normal(size=n),
                rng.normal(size=n_p), rng.normal(size=n))
        for a, b in zip(_callback_values(gen, n, n_p, *args), _callback_values(cf, n, n_p, *args)):
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-14 * max(np.abs(b).max(), 1.0))
    z0 = np.stack([_joint_y0(r, n, cols) for r in y0])
    a, sa = O.rodas5p_callbacks(gen, z0, t, p, direction_columns=cols, lu_precision="fp64", **kw)
    b, sb = O.rodas5p_callbacks(cf, z0, t, p, direction_columns=cols, lu_precision="fp64", **kw)
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    scale = np.abs(b).max(axis=1, keepdims=True)
    assert (np.abs(a - b) / scale).max() < 1e-9

