"""Forward sensitivities on the GPU: the joint [y, S] kernel against the
oracle's joint solve bit for bit, and the torch autograd rules (reverse mode
and forward mode) that contract S, against closed forms and differences."""

import numpy as np
import pytest

from menodax_amd import rodas5P, tsit5
from tests import systems as S
from menodax_amd._sensitivity import SensitivitySpec
from oracle import oracle as O
from oracle.callbacks import HostCallbacks


def _structure(mask, n, ordering):
    """The oracle's own sparse structure for ``mask`` under the product's
    elimination order (the AMD permutation is the one input it takes)."""
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    pattern = normalize_sparsity(mask, n)
    return O.sparse_structure(pattern, SD.fill_reducing_order(pattern, ordering))

pytestmark = pytest.mark.gpu

A, B = 0.7, 1.3
Y0 = np.array([2.0, 0.5])
T_SPAN = np.linspace(0.0, 2.0, 5)


def _joint_y0(y0, n, cols):
    y0 = np.atleast_2d(y0)
    z = np.zeros((y0.shape[0], n * (1 + len(cols))))
    z[:, :n] = y0
    for d, c in enumerate(cols):
        if c < 0:
            z[:, n + d * n + d] = 1.0
    return z


def _oracle_joint(cb, y0, t, p, spec, **kw):
    cols = spec.direction_columns
    hist, st = O.rodas5p_callbacks(cb, _joint_y0(y0, spec.n_vars, cols), t, p,
                                   direction_columns=cols, error_control=spec.error_control, **kw)
    n = spec.n_vars
    state = hist[:, :, :n]
    sens = hist[:, :, n:].reshape(hist.shape[0], hist.shape[1], len(cols), n).transpose(0, 1, 3, 2)
    return state, sens, st


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
@pytest.mark.parametrize("placement", ["registers", "rolled", "lds"])
def test_joint_robertson_bit_exact_with_the_oracle(placement, lu):
    y0, p = S.robertson_lognormal_ensemble(300, seed=3)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8, lu_precision=lu)
    state, sens, st = rodas5P.solve_sensitivities(S.robertson, y0, t, p, return_stats=True,
                                                  placement=placement, **kw)
    spec = SensitivitySpec(3, 3, True, True)
    ref_state, ref_sens, rst = _oracle_joint(HostCallbacks(S.robertson, 3, 3), y0, t, p, spec, **kw)
    for k in rst:
        np.testing.assert_array_equal(st[k], rst[k], err_msg=k)
    np.testing.assert_array_equal(state, ref_state)
    np.testing.assert_array_equal(sens, ref_sens)


def test_joint_sparse_parameter_columns_bit_exact_with_the_oracle():
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    fn, mask, _ = S.arrow_system(12)
    n_p = S.arrow_n_params(12)
    y0, lam = S.arrow_ensemble(128, n=12, lam_range=(1e-1, 1e3))
    t = np.linspace(0.0, 2.0, 5)
    cols = (0, 13)
    kw = dict(rtol=1e-6, atol=1e-8, lu_precision="fp32")
    state, sens, st = rodas5P.solve_sensitivities(fn, y0, t, lam, wrt_y0=False,
                                                  sens_param_columns=cols, sparsity=mask,
                                                  return_stats=True, **kw)
    spec = SensitivitySpec(12, n_p, False, True, param_columns=cols)
    layout = _structure(mask, 12, "amd")
    ref_state, ref_sens, rst = _oracle_joint(HostCallbacks(fn, 12, n_p), y0, t, lam, spec,
                                             layout=layout, **kw)
    for k in rst:
        np.testing.assert_array_equal(st[k], rst[k], err_msg=k)
    np.testing.assert_array_equal(state, ref_state)
    np.testing.assert_array_equal(sens, ref_sens)


@pytest.mark.parametrize("method", ["rodas5p", "tsit5"])
def test_cascade_sensitivities_match_closed_form(method):
    kw = dict(rtol=1e-10, atol=1e-12)
    if method == "rodas5p":
        state, sens = rodas5P.solve_sensitivities(S.cascade, Y0, T_SPAN, [A, B],
                                                  lu_precision="fp64", **kw)
    else:
        state, sens = tsit5.solve_sensitivities(S.cascade, Y0, T_SPAN, [A, B], **kw)
    np.testing.assert_allclose(state[0], S.cascade_exact(T_SPAN, Y0, [A, B]), rtol=1e-7)
    np.testing.assert_allclose(sens[0, :, :, :2], S.cascade_dy0(T_SPAN, Y0, [A, B]),
                               rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(sens[0, :, :, 2:], S.cascade_dparams(T_SPAN, Y0, [A, B]),
                               rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("method", ["rodas5p", "tsit5"])
def test_without_error_control_the_joint_state_is_the_plain_solve(method):
    y0, p = S.robertson_lognormal_ensemble(200, seed=8)
    t = S.robertson_benchmark_times() if method == "rodas5p" else np.linspace(0, 0.5, 6)
    kw = dict(rtol=1e-7, atol=1e-9, return_stats=True)
    mod = rodas5P if method == "rodas5p" else tsit5
    plain, ps = mod.solve(S.robertson, y0, t, p, **kw)
    state, _, js = mod.solve_sensitivities(S.robertson, y0, t, p, sens_error_control=False, **kw)
    np.testing.assert_array_equal(state, plain)
    for k in ps:
        np.testing.assert_array_equal(js[k], ps[k])


@pytest.mark.parametrize("method", ["rodas5p", "tsit5"])
def test_grad_matches_central_differences(method):
    import torch

    mod = rodas5P if method == "rodas5p" else tsit5
    kw = dict(rtol=1e-10, atol=1e-12)
    if method == "rodas5p":
        kw["lu_precision"] = "fp64"
    dev = torch.device("cuda")
    y0 = torch.tensor(Y0, device=dev, requires_grad=True)
    p = torch.tensor([A, B], dtype=torch.float64, device=dev, requires_grad=True)
    loss = (mod.solve(S.cascade, y0, T_SPAN, p, **kw) ** 2).sum()
    loss.backward()

    def value(y0v, pv):
        return float((mod.solve(S.cascade, y0v, T_SPAN, pv, **kw) ** 2).sum())

    for x, g, other, first in ((Y0, y0.grad, [A, B], True), (np.array([A, B]), p.grad, Y0, False)):
        for k in range(2):
            h = 1e-6 * abs(x[k])
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            fd = ((value(xp, other) - value(xm, other)) if first
                  else (value(other, xp) - value(other, xm))) / (2 * h)
            assert abs(float(g[k]) - fd) <= 1e-5 * abs(fd)


@pytest.mark.parametrize("method", ["rodas5p", "tsit5"])
def test_a_joint_solve_resumed_from_its_pilot_gives_the_same_gradients(method):
    """A differentiated solve with work_order="pilot" runs the joint [y, S]
    kernel from the pilot's augmented step state: the history, the counters
    and the gradients are the bits of the same joint solve in index order.
    The trajectories run past the 80-step pilot (Robertson's joint solves take
    122-146 steps), so they continue from its augmented state."""
    import torch

    dev = torch.device("cuda")
    rng = np.random.default_rng(12)
    if method == "rodas5p":
        solver, fn = rodas5P, S.robertson
        y0, p = S.robertson_lognormal_ensemble(300, seed=7)
        t = S.robertson_benchmark_times()
        kw = dict(rtol=1e-6, atol=1e-8)
    else:
        solver, fn = tsit5, S.cascade
        y0 = rng.uniform(0.5, 2.0, size=(300, 2))
        p = 10.0 ** rng.uniform(-1.0, 2.3, size=(300, 2))
        t = np.linspace(0.0, 2.0, 9)
        kw = dict(rtol=1e-10, atol=1e-12)
    out = []
    for order in ({"refill": True}, {"work_order": "pilot"}):
        yt = torch.tensor(y0, device=dev, requires_grad=True)
        pt = torch.tensor(p, device=dev, requires_grad=True)
        hist, st = solver.solve(fn, yt, t, pt, return_stats=True, **kw, **order)
        (hist * hist).sum().backward()
        out.append((hist.detach(), st, yt.grad, pt.grad))
    (h0, s0, gy0, gp0), (h1, s1, gy1, gp1) = out
    assert torch.equal(h0, h1) and torch.equal(gy0, gy1) and torch.equal(gp0, gp1)
    for k in s0:
        assert torch.equal(s0[k], s1[k]), k
    assert bool((s0["loop_steps"] > rodas5P.PILOT_STEPS).any())


def test_ensemble_gradients_are_per_trajectory_and_shared_params_sum():
    import torch

    dev = torch.device("cuda")
    y0 = np.array([[2.0, 0.5], [1.0, 1.5], [0.3, 0.2]])
    kw = dict(rtol=1e-10, atol=1e-12, lu_precision="fp64")
    yt = torch.tensor(y0, device=dev, requires_grad=True)
    shared = torch.tensor([A, B], dtype=torch.float64, device=dev, requires_grad=True)
    (rodas5P.solve(S.cascade, yt, T_SPAN, shared, **kw) ** 2).sum().backward()
    total = torch.zeros(2, dtype=torch.float64, device=dev)
    for r in range(3):
        yr = torch.tensor(y0[r], device=dev, requires_grad=True)
        pr = torch.tensor([A, B], dtype=torch.float64, device=dev, requires_grad=True)
        (rodas5P.solve(S.cascade, yr, T_SPAN, pr, **kw) ** 2).sum().backward()
        torch.testing.assert_close(yt.grad[r], yr.grad, rtol=1e-12, atol=0)
        total += pr.grad
    torch.testing.assert_close(shared.grad, total, rtol=1e-12, atol=0)


def test_forward_mode_jvp_contracts_the_sensitivities():
    import torch
    import torch.autograd.forward_ad as fwAD

    dev = torch.device("cuda")
    kw = dict(rtol=1e-9, atol=1e-11, lu_precision="fp64")
    y0, p = S.robertson_lognormal_ensemble(16, seed=2)
    t = np.array([0.0, 1e-2, 1.0, 40.0])
    dp = np.random.default_rng(0).normal(size=p.shape)
    state, sens = rodas5P.solve_sensitivities(S.robertson, y0, t, p, wrt_y0=False, **kw)
    expect = np.einsum("nsik,nk->nsi", sens, dp)
    with fwAD.dual_level():
        pd = fwAD.make_dual(torch.tensor(p, device=dev), torch.tensor(dp, device=dev))
        out = rodas5P.solve(S.robertson, torch.tensor(y0, device=dev), t, pd, **kw)
        primal, tangent = fwAD.unpack_dual(out)
    np.testing.assert_array_equal(primal.cpu().numpy(), state)
    np.testing.assert_allclose(tangent.cpu().numpy(), expect, rtol=1e-13, atol=1e-300)


def test_stats_stay_available_under_grad():
    import torch

    p = torch.tensor([A, B], dtype=torch.float64, device="cuda", requires_grad=True)
    hist, stats = rodas5P.solve(S.cascade, Y0, T_SPAN, p, rtol=1e-8, atol=1e-10,
                                return_stats=True)
    assert hist.requires_grad and not stats["accepted_steps"].requires_grad
    (hist ** 2).sum().backward()
    assert p.grad is not None and torch.isfinite(p.grad).all()


def test_a_small_sparse_joint_solve_runs_narrow_workgroups_with_the_same_result():
    """1,000 64-variable stiff rings with one parameter direction: 128 joint
    rows, whose stage vectors fit in LDS at 16 lanes, not at 64; the size rule's
    workgroup gives the 64-lane kernel's outputs bit for bit."""
    import torch

    from menodax_amd import _kernels

    dev = torch.device("cuda", 0)
    fn, _ = S.vdp_lattice(32)
    mask = S.vdp_sparsity(32)
    y0n, pn = S.vdp_scenario(32, 1000, divergence=0.5)
    y0, p = torch.tensor(y0n, device=dev), torch.tensor(pn, device=dev)
    t = torch.tensor(S.VDP_TIMES, device=dev)
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    assert _kernels.rodas5p_block(None, 128, "fp32", 1000, n_cu, None, mask) == 16
    kw = dict(first_step=1e-4, rtol=1e-6, atol=1e-8, sparsity=mask, wrt_y0=False)
    hist, sens = rodas5P.solve_sensitivities(fn, y0, t, p, **kw)
    hist64, sens64 = rodas5P.solve_sensitivities(fn, y0, t, p, trajectories_per_block=64, **kw)
    assert torch.equal(hist, hist64) and torch.equal(sens, sens64)
    assert torch.isfinite(sens).all()

    # the same joint kernels under autograd: a gradient of a plain solve
    def grad(block):
        q = p.clone().requires_grad_(True)
        h = rodas5P.solve(fn, y0, t, q, first_step=1e-4, rtol=1e-6, atol=1e-8, sparsity=mask,
                          trajectories_per_block=block)
        h[:, -1].sum().backward()
        return q.grad

    assert torch.equal(grad(None), grad(64))
