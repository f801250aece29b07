"""The discrete adjoint of Tsit5 on the GPU.

Pinned like the Rodas5P adjoint: against central differences of the loss
through the plain GPU solve (h = 1e-6 relative, 1e-5 relative agreement),
against the forward-sensitivity gradient of the Tsit5 joint solve, and against
the closed-form sensitivities of the linear cascade; for its own machinery,
gradients independent of the checkpoint interval and of the placement, and the
torch backward of ``solve(..., adjoint=True)`` equal to ``adjoint_gradients``.
"""

import numpy as np
import pytest

from menodax_amd import tsit5
from tests import systems as S

pytestmark = pytest.mark.gpu


def _lorenz_problem(n_traj, seed=0):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(5.0, 28.0, size=(n_traj, 1))
    y0 = np.column_stack([1.0 + rng.uniform(-0.1, 0.1, n_traj),
                          rng.uniform(-0.1, 0.1, n_traj), rng.uniform(0.0, 0.1, n_traj)])
    t = np.linspace(0.0, 1.0, 20)
    kw = dict(rtol=1e-10, atol=1e-12)
    target = tsit5.solve(S.lorenz, y0, t, rho * 1.05, **kw)
    target = target + 0.01 * rng.normal(size=target.shape)
    return y0, rho, t, target, kw


def _loss_and_cotangent(hist, target):
    r = hist - target
    return float((r * r).sum()), 2.0 * r


def test_a_checkpoint_store_too_small_for_a_trajectory_is_sized_again(monkeypatch):
    """The store sized from max_steps, and re-sized by a forward solve's step
    counts when a one-checkpoint budget is outrun: the same bits."""
    from tests.test_gpu_adjoint import store_resized_bits

    y0, rho, t, target, kw = _lorenz_problem(32, seed=6)
    hist = tsit5.solve(S.lorenz, y0, t, rho, **kw)
    _, g = _loss_and_cotangent(hist, target)
    store_resized_bits(monkeypatch, lambda: tsit5.adjoint_gradients(S.lorenz, y0, t, rho, g,
                                                                    ckpt_every=16, **kw))


def test_adjoint_matches_forward_sensitivities():
    y0, rho, t, target, kw = _lorenz_problem(64)
    hist, sens = tsit5.solve_sensitivities(S.lorenz, y0, t, rho, **kw)
    _, g = _loss_and_cotangent(hist, target)
    gdir = np.einsum("nsi,nsik->nk", g, sens)
    gy, gp = tsit5.adjoint_gradients(S.lorenz, y0, t, rho, g, **kw)
    np.testing.assert_allclose(gy, gdir[:, :3], rtol=1e-5, atol=1e-7 * np.abs(gdir[:, :3]).max())
    np.testing.assert_allclose(gp, gdir[:, 3:], rtol=1e-5, atol=1e-7 * np.abs(gdir[:, 3:]).max())


def test_rolled_adjoint_of_a_larger_system_matches_forward_sensitivities():
    # 10 variables: above the unrolled adjoint's size, the rolled scratch kernel
    fn = S.vdp_lattice(5, mu=1.0)[0]
    rng = np.random.default_rng(1)
    y0 = rng.uniform(-2.0, 2.0, size=(32, 10))
    mu = rng.uniform(0.5, 1.5, size=(32, 1))
    t = np.linspace(0.0, 2.0, 10)
    kw = dict(rtol=1e-10, atol=1e-12)
    hist, sens = tsit5.solve_sensitivities(fn, y0, t, mu, **kw)
    g = rng.normal(size=hist.shape)
    gdir = np.einsum("nsi,nsik->nk", g, sens)
    gy, gp = tsit5.adjoint_gradients(fn, y0, t, mu, g, **kw)
    np.testing.assert_allclose(gy, gdir[:, :10], rtol=1e-5, atol=1e-7 * np.abs(gdir[:, :10]).max())
    np.testing.assert_allclose(gp, gdir[:, 10:], rtol=1e-5, atol=1e-7 * np.abs(gdir[:, 10:]).max())


def test_adjoint_matches_central_differences():
    y0, rho, t, target, kw = _lorenz_problem(8, seed=2)
    hist = tsit5.solve(S.lorenz, y0, t, rho, **kw)
    _, g = _loss_and_cotangent(hist, target)
    gy, gp = tsit5.adjoint_gradients(S.lorenz, y0, t, rho, g, **kw)

    def per_traj_loss(y0v, rhov):
        h = tsit5.solve(S.lorenz, y0v, t, rhov, **kw)
        return ((h - target) ** 2).sum(axis=(1, 2))

    h = 1e-6
    for j in range(3):
        d = np.zeros_like(y0)
        d[:, j] = h * np.maximum(np.abs(y0[:, j]), 1e-3)
        fd = (per_traj_loss(y0 + d, rho) - per_traj_loss(y0 - d, rho)) / (2 * d[:, j])
        np.testing.assert_allclose(gy[:, j], fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
    d = h * rho
    fd = (per_traj_loss(y0, rho + d) - per_traj_loss(y0, rho - d)) / (2 * d[:, 0])
    np.testing.assert_allclose(gp[:, 0], fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())


def test_adjoint_of_the_cascade_against_the_closed_form():
    # loss = sum of every saved value: its gradient is the sum of the closed-form
    # sensitivities over saves and components
    y0 = np.array([[1.0, 0.5], [2.0, 0.1], [0.3, 1.0], [1.5, 1.5]])
    p = np.array([[0.7, 1.9], [1.3, 0.4], [2.0, 3.1], [0.5, 0.6]])
    t = np.linspace(0.0, 2.0, 9)
    g = np.ones((4, t.size, 2))
    gy, gp = tsit5.adjoint_gradients(S.cascade, y0, t, p, g, rtol=1e-12, atol=1e-14)
    for i in range(4):
        want_p = S.cascade_dparams(t, y0[i], p[i]).sum(axis=(0, 1))
        want_y = S.cascade_dy0(t, y0[i], p[i]).sum(axis=(0, 1))
        np.testing.assert_allclose(gp[i], want_p, rtol=1e-8)
        np.testing.assert_allclose(gy[i], want_y, rtol=1e-8)


def test_gradients_do_not_depend_on_the_checkpoint_interval():
    y0, rho, t, target, kw = _lorenz_problem(32, seed=3)
    hist = tsit5.solve(S.lorenz, y0, t, rho, **kw)
    _, g = _loss_and_cotangent(hist, target)
    ref = tsit5.adjoint_gradients(S.lorenz, y0, t, rho, g, ckpt_every=64, **kw)
    for c in (1, 7):
        got = tsit5.adjoint_gradients(S.lorenz, y0, t, rho, g, ckpt_every=c, **kw)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])


def test_adjoint_placements_give_the_same_bits():
    y0, rho, t, target, kw = _lorenz_problem(64, seed=4)
    hist = tsit5.solve(S.lorenz, y0, t, rho, **kw)
    _, g = _loss_and_cotangent(hist, target)
    ref = tsit5.adjoint_gradients(S.lorenz, y0, t, rho, g, ckpt_every=16, placement="rolled", **kw)
    got = tsit5.adjoint_gradients(S.lorenz, y0, t, rho, g, ckpt_every=16, **kw)
    assert np.all(np.isfinite(ref[0])) and np.all(np.isfinite(ref[1]))
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])


def test_torch_backward_through_the_adjoint():
    import torch

    y0, rho, t, target, kw = _lorenz_problem(16, seed=5)
    dev = torch.device("cuda")
    y0t = torch.tensor(y0, device=dev, requires_grad=True)
    rhot = torch.tensor(rho, device=dev, requires_grad=True)
    hist = tsit5.solve(S.lorenz, y0t, t, rhot, adjoint=True, **kw)
    loss = ((hist - torch.tensor(target, device=dev)) ** 2).sum()
    loss.backward()
    g = 2.0 * (hist.detach().cpu().numpy() - target)
    gy, gp = tsit5.adjoint_gradients(S.lorenz, y0, t, rho, g, **kw)
    np.testing.assert_array_equal(y0t.grad.cpu().numpy(), gy)
    np.testing.assert_array_equal(rhot.grad.cpu().numpy(), gp)
