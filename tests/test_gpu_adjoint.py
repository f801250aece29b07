"""The discrete adjoint of Rodas5P on the GPU.

Pinned against central differences of the loss through the plain GPU solve
(h = 1e-6 relative, 1e-5 relative agreement),
against the forward-sensitivity gradient (the reference's own derivative,
_sensitivity.py, whose oracle restatement is bit-exact with the GPU in
test_gpu_sensitivity.py) and against a closed form; and for its own
machinery, gradients independent of the checkpoint interval.
"""

import numpy as np
import pytest

from menodax_amd import _kernels, rodas5P
from menodax_amd._sparsity import normalize_sparsity
from tests import systems as S

pytestmark = pytest.mark.gpu


def _vdp_problem(n_osc, n_traj, seed=0):
    """Stiff VdP (n = 2 n_osc), 20 saves, a target solved at a perturbed mu plus noise."""
    fn, _ = S.vdp_lattice(n_osc, mu=1.0, d=0.1, wrap=False)
    rng = np.random.default_rng(seed)
    mu = 10.0 ** rng.uniform(0.5, 1.5, size=(n_traj, 1))
    y0 = rng.uniform(-2.0, 2.0, size=(n_traj, 2 * n_osc))
    t = np.linspace(0.0, 3.0, 20)
    kw = dict(rtol=1e-10, atol=1e-12, lu_precision="fp64")
    target = rodas5P.solve(fn, y0, t, mu * 1.1, **kw)
    target = target + 0.01 * rng.normal(size=target.shape)
    return fn, y0, mu, t, target, kw


def _loss_and_cotangent(hist, target):
    r = hist - target
    return float((r * r).sum()), 2.0 * r


@pytest.mark.parametrize("n_osc", [1, 10])
def test_adjoint_matches_forward_sensitivities(n_osc):
    fn, y0, mu, t, target, kw = _vdp_problem(n_osc, 64)
    sparsity = None if n_osc == 1 else S.vdp_sparsity(n_osc, wrap=False)
    hist, sens = rodas5P.solve_sensitivities(fn, y0, t, mu, sparsity=sparsity, **kw)
    _, g = _loss_and_cotangent(hist, target)
    gdir = np.einsum("nsi,nsik->nk", g, sens)
    n = 2 * n_osc
    gy, gp = rodas5P.adjoint_gradients(fn, y0, t, mu, g, sparsity=sparsity, **kw)
    np.testing.assert_allclose(gy, gdir[:, :n], rtol=1e-5, atol=1e-7 * np.abs(gdir[:, :n]).max())
    np.testing.assert_allclose(gp, gdir[:, n:], rtol=1e-5, atol=1e-7 * np.abs(gdir[:, n:]).max())


def test_adjoint_matches_central_differences():
    fn, y0, mu, t, target, kw = _vdp_problem(1, 8, seed=1)
    hist = rodas5P.solve(fn, y0, t, mu, **kw)
    _, g = _loss_and_cotangent(hist, target)
    gy, gp = rodas5P.adjoint_gradients(fn, y0, t, mu, g, **kw)

    def per_traj_loss(y0v, muv):
        h = rodas5P.solve(fn, y0v, t, muv, **kw)
        return ((h - target) ** 2).sum(axis=(1, 2))

    h = 1e-6
    for k in range(2):
        d = np.zeros_like(y0)
        d[:, k] = h * np.maximum(np.abs(y0[:, k]), 1e-3)
        fd = (per_traj_loss(y0 + d, mu) - per_traj_loss(y0 - d, mu)) / (2 * d[:, k])
        np.testing.assert_allclose(gy[:, k], fd, rtol=1e-5, atol=1e-9 * np.abs(fd).max())
    dm = h * mu
    fd = (per_traj_loss(y0, mu + dm) - per_traj_loss(y0, mu - dm)) / (2 * dm[:, 0])
    np.testing.assert_allclose(gp[:, 0], fd, rtol=1e-5, atol=1e-9 * np.abs(fd).max())


def test_adjoint_of_a_non_autonomous_problem_against_the_closed_form():
    """d/dlam and d/dF of sum_s y(t_s) for y' = -lam y + F t: the df/dt term's share."""
    t = np.linspace(0.0, 1.0, 11)
    p = np.array([[10.0, 5.0]])
    g = np.ones((1, t.size, 1))
    _, gp = rodas5P.adjoint_gradients(S.forced_decay, [0.0], t, p, g, rtol=1e-10, atol=1e-12,
                                      lu_precision="fp64")
    exact = [S.forced_decay_dlam(t, 10.0, 5.0).sum(), S.forced_decay_dforcing(t, 10.0, 5.0).sum()]
    np.testing.assert_allclose(gp[0], exact, rtol=1e-6)


def test_gradients_do_not_depend_on_the_checkpoint_interval():
    fn, y0, mu, t, target, kw = _vdp_problem(1, 32, seed=2)
    hist = rodas5P.solve(fn, y0, t, mu, **kw)
    _, g = _loss_and_cotangent(hist, target)
    ref = rodas5P.adjoint_gradients(fn, y0, t, mu, g, ckpt_every=64, **kw)
    for c in (1, 7):
        got = rodas5P.adjoint_gradients(fn, y0, t, mu, g, ckpt_every=c, **kw)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])


def test_a_checkpoint_store_too_small_for_a_trajectory_is_sized_again(monkeypatch):
    """adjoint_gradients sizes the checkpoint store from max_steps within a
    memory budget, so its adjoint kernel's forward pass is the only one. With
    the budget cut to one checkpoint the trajectories outrun it: a forward
    solve then counts their steps and the adjoint runs again, with the same
    bits as the store sized at once."""
    fn, y0, mu, t, target, kw = _vdp_problem(1, 32, seed=6)
    hist = rodas5P.solve(fn, y0, t, mu, **kw)
    _, g = _loss_and_cotangent(hist, target)
    store_resized_bits(monkeypatch, lambda: rodas5P.adjoint_gradients(fn, y0, t, mu, g,
                                                                      ckpt_every=16, **kw))


def store_resized_bits(monkeypatch, run):
    """``run()`` (an adjoint_gradients call with ckpt_every=16 and the default
    max_steps) as sized from max_steps, then with a one-checkpoint budget:
    one launch, then a failed launch and a resized one, the same bits."""
    from menodax_amd import _autodiff

    runs = []
    real = _autodiff._adjoint_run

    def spy(job, y0_t, p_t, gbar, max_ckpt, must_fit):
        out = real(job, y0_t, p_t, gbar, max_ckpt, must_fit)
        runs.append((max_ckpt, out is not None))
        return out

    monkeypatch.setattr(_autodiff, "_adjoint_run", spy)
    ref = run()
    assert runs == [(100000 // 16 + 1, True)]  # room for max_steps accepted steps
    monkeypatch.setattr(_autodiff, "CHECKPOINT_STORE_BUDGET", 1)
    runs.clear()
    got = run()
    assert [r[1] for r in runs] == [False, True] and runs[0][0] == 1 and runs[1][0] > 1
    assert np.all(np.isfinite(ref[0])) and np.all(np.isfinite(ref[1]))
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])


@pytest.mark.parametrize("lu", ["fp64", "fp32"])
def test_adjoint_placements_give_the_same_bits(lu):
    # the unrolled register kernel (chosen for 2 variables) against the rolled
    # scratch kernel that larger systems use
    fn, y0, mu, t, target, kw = _vdp_problem(1, 64, seed=4)
    kw = dict(kw, lu_precision=lu)
    hist = rodas5P.solve(fn, y0, t, mu, **kw)
    _, g = _loss_and_cotangent(hist, target)
    ref = rodas5P.adjoint_gradients(fn, y0, t, mu, g, ckpt_every=16, placement="rolled", **kw)
    got = rodas5P.adjoint_gradients(fn, y0, t, mu, g, ckpt_every=16, **kw)
    assert np.all(np.isfinite(ref[0])) and np.all(np.isfinite(ref[1]))
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])


@pytest.mark.parametrize("lu,block", [("fp64", 32), ("fp64", 16), ("fp32", 64)])
def test_the_chain_adjoint_in_lds_gives_the_scratch_kernels_bits(lu, block):
    # the 20-variable chain's rolled adjoint with its stage vectors, their
    # adjoints and the factors in LDS (with the stage cache at 32 and 16 lanes,
    # without it at 64 FP32 lanes, where it does not fit), against the same
    # step in scratch; 96 trajectories leave the last workgroup partly empty
    import torch

    fn, y0, mu, t, target, kw = _vdp_problem(10, 96, seed=8)
    kw = dict(kw, lu_precision=lu, sparsity=S.vdp_sparsity(10, wrap=False), rtol=1e-8,
              atol=1e-10)
    hist = rodas5P.solve(fn, y0, t, mu, **kw)
    _, g = _loss_and_cotangent(hist, target)
    ref = rodas5P.adjoint_gradients(fn, y0, t, mu, g, ckpt_every=16, placement="rolled",
                                    trajectories_per_block=64, **kw)
    got = rodas5P.adjoint_gradients(fn, y0, t, mu, g, ckpt_every=16, placement="lds",
                                    trajectories_per_block=block, **kw)
    assert np.all(np.isfinite(ref[0])) and np.all(np.isfinite(ref[1]))
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])
    # by default a small ensemble takes the LDS placement, through either entry
    opts = rodas5P._adjoint_options(
        _kernels.KernelOptions(lu_precision=lu, sparsity=normalize_sparsity(kw["sparsity"], 20)),
        20, 96, None)
    assert (opts.placement, opts.trajectories_per_block, opts.adjoint_cache) == ("lds", 32, True)
    dflt = rodas5P.adjoint_gradients(fn, y0, t, mu, g, ckpt_every=16, **kw)
    np.testing.assert_array_equal(dflt[0], ref[0])
    np.testing.assert_array_equal(dflt[1], ref[1])
    dev = torch.device("cuda")
    y0t = torch.tensor(y0, device=dev, requires_grad=True)
    mut = torch.tensor(mu, device=dev, requires_grad=True)
    h = rodas5P.solve(fn, y0t, t, mut, adjoint=True, ckpt_every=16, **kw)
    (h * torch.tensor(g, device=dev)).sum().backward()
    np.testing.assert_array_equal(y0t.grad.cpu().numpy(), ref[0])
    np.testing.assert_array_equal(mut.grad.cpu().numpy(), ref[1])


def test_stiff_robertson_fp32_lu_adjoint_against_sensitivities():
    y0, p = S.robertson_lognormal_ensemble(64, seed=5)
    t = np.array([0.0, 1e-3, 1e-1, 10.0, 100.0])
    kw = dict(rtol=1e-9, atol=1e-14, lu_precision="fp32")
    hist, sens = rodas5P.solve_sensitivities(S.robertson, y0, t, p, **kw)
    g = np.random.default_rng(0).normal(size=hist.shape)
    gdir = np.einsum("nsi,nsik->nk", g, sens)
    gy, gp = rodas5P.adjoint_gradients(S.robertson, y0, t, p, g, **kw)
    scale = np.abs(gdir).max(axis=0)
    assert np.all(np.abs(gy - gdir[:, :3]).max(axis=0) <= 1e-3 * scale[:3])
    assert np.all(np.abs(gp - gdir[:, 3:]).max(axis=0) <= 1e-3 * scale[3:])


def test_torch_backward_through_the_adjoint():
    import torch

    fn, y0, mu, t, target, kw = _vdp_problem(1, 16, seed=3)
    dev = torch.device("cuda")
    y0t = torch.tensor(y0, device=dev, requires_grad=True)
    mut = torch.tensor(mu, device=dev, requires_grad=True)
    hist = rodas5P.solve(fn, y0t, t, mut, adjoint=True, **kw)
    loss = ((hist - torch.tensor(target, device=dev)) ** 2).sum()
    loss.backward()
    g = 2.0 * (hist.detach().cpu().numpy() - target)
    gy, gp = rodas5P.adjoint_gradients(fn, y0, t, mu, g, **kw)
    np.testing.assert_array_equal(y0t.grad.cpu().numpy(), gy)
    np.testing.assert_array_equal(mut.grad.cpu().numpy(), gp)


def test_adjoint_matches_central_differences_on_the_20_variable_chain():
    """The sparse (AMD) rolled adjoint of the 10-oscillator chain, every component
    of y0 and mu, 32 trajectories: against the forward sensitivities component by
    component to 1e-6 relative, and against central differences of the plain
    GPU solve (each component perturbed in all trajectories at once: they are
    independent) to 1e-5 relative. The differences are Richardson-extrapolated
    from relative steps 1e-4 and 5e-5 on solves at rtol 1e-12: the loss moves
    with the solve's accept/reject decisions, so small steps turn that noise
    into errors on the small components (gradients here span seven decades),
    while at these steps a plain central difference still carries an h^2
    error on the stiffest, most nonlinear components (1e-2 at 1e-4); hence
    also an absolute floor of 1e-8 of the largest component."""
    fn, y0, mu, t, target, kw = _vdp_problem(10, 32, seed=6)
    sparsity = S.vdp_sparsity(10, wrap=False)
    kw = dict(kw, sparsity=sparsity, rtol=1e-12, atol=1e-14)
    hist = rodas5P.solve(fn, y0, t, mu, **kw)
    _, g = _loss_and_cotangent(hist, target)
    gy, gp = rodas5P.adjoint_gradients(fn, y0, t, mu, g, **kw)
    hs, sens = rodas5P.solve_sensitivities(fn, y0, t, mu, **kw)
    gdir = np.einsum("nsi,nsik->nk", 2.0 * (hs - target), sens)
    got = np.concatenate([gy, gp], axis=1)
    np.testing.assert_allclose(got, gdir, rtol=1e-6, atol=1e-10 * np.abs(gdir).max())

    def per_traj_loss(y0v, muv):
        h = rodas5P.solve(fn, y0v, t, muv, **kw)
        return ((h - target) ** 2).sum(axis=(1, 2))

    def central(k, rel):
        if k < 20:
            d = np.zeros_like(y0)
            d[:, k] = rel * np.maximum(np.abs(y0[:, k]), 1e-3)
            return (per_traj_loss(y0 + d, mu) - per_traj_loss(y0 - d, mu)) / (2 * d[:, k])
        dm = rel * mu
        return (per_traj_loss(y0, mu + dm) - per_traj_loss(y0, mu - dm)) / (2 * dm[:, 0])

    fd = np.stack([(4.0 * central(k, 5e-5) - central(k, 1e-4)) / 3.0 for k in range(21)], axis=1)
    worst = np.max(np.abs(got - fd) / np.maximum(np.abs(fd), 1e-8 * np.abs(fd).max()))
    print(f"20-variable chain adjoint vs central differences: worst {worst:.2e}", flush=True)
    np.testing.assert_allclose(got, fd, rtol=1e-5, atol=1e-8 * np.abs(fd).max())


def test_adjoint_and_sensitivity_gap_shrinks_as_the_tolerance_tightens():
    """The adjoint differentiates the discrete solution exactly; the forward
    sensitivities approximate the continuous derivative within their own error
    control. Both converge to the derivative of the exact solution, so the gap
    between them falls as rtol tightens: by at least a decade from 1e-6 to
    1e-10, and at every step of 1e-6, 1e-8, 1e-10."""
    fn, y0, mu, t, target, _ = _vdp_problem(1, 64, seed=7)
    gaps = []
    for rtol in (1e-6, 1e-8, 1e-10):
        kw = dict(rtol=rtol, atol=rtol * 1e-2, lu_precision="fp64")
        hist, sens = rodas5P.solve_sensitivities(fn, y0, t, mu, **kw)
        _, g = _loss_and_cotangent(hist, target)
        gdir = np.einsum("nsi,nsik->nk", g, sens)
        gy, gp = rodas5P.adjoint_gradients(fn, y0, t, mu, g, **kw)
        got = np.concatenate([gy, gp], axis=1)
        gaps.append(float(np.max(np.abs(got - gdir).max(axis=1) / np.abs(gdir).max(axis=1))))
    print(f"adjoint - sensitivity gap at rtol 1e-6, 1e-8, 1e-10: {gaps}", flush=True)
    assert gaps[0] > gaps[1] > gaps[2]
    assert gaps[2] < gaps[0] / 10.0
