"""A non-polynomial right-hand side on the GPU: the four-species Big Bang
nucleosynthesis network of the reference's consumer (examples/bbn_estimation/
main.py), with exp, sqrt, a 1.5 power and a comparison used as a 0/1 factor,
written once as a dual-form factory (tests/systems.make_bbn).

Parity bar against the oracle. The kernel evaluates exp and pow with the
device math library and the oracle with the host's. Both are faithful to about
an ulp, not bit-identical, so a trajectory is bit-exact only where no such
difference reaches a rounding that matters. The stated bar, at the consumer's
own settings (rtol 1e-4, atol 1e-9, first step 0.1, 256 steps at most):
  - at least 99% of the trajectories take exactly the oracle's step counts,
    and those agree to a small fraction of what the solve controls,
    c (atol + rtol |y|): c = 1e-3 with the FP64 LU; c = 0.1 with the FP32 LU,
    since a last-bit difference in f can flip the float rounding of a stage's
    right-hand side, which moves that stage by one float ulp (first GPU run:
    100% equal counts, 99.3% bit-identical, the rest within 2.8e-8, at most
    3% of the tolerance);
  - every trajectory, including one whose accept/reject sequence a last-bit
    difference flipped, lies within ten times the solve's own tolerance
    (atol + rtol |y|) of the oracle.
The share of bit-identical trajectories is printed.

The consumer (the reference's likelihood, main.py:226-235, 287-300): the
predicted abundances of a grid of universes under torch.func.vmap (one ensemble
launch, the same bits as the direct ensemble solve), and a chi^2 loss under
torch.compile whose gradient comes from the discrete adjoint, checked against
central differences to 1e-5 relative.
"""

import math

import numpy as np
import pytest

from menodax_amd import rodas5P
from tests import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lu", ["fp32", "fp64"])
def test_bbn_grid_against_the_oracle(lu):
    theta = S.bbn_grid(1024)
    y0 = np.tile(S.bbn_initial_state(), (theta.shape[0], 1))
    kw = dict(S.BBN_SOLVER, lu_precision=lu)
    hist, st = rodas5P.solve(S.bbn, y0, S.BBN_X_SPAN, theta, return_stats=True, **kw)
    ref, rst = O.rodas5p(O.bbn(), y0, S.BBN_X_SPAN, theta, **kw)
    same_steps = ((np.asarray(st["accepted_steps"]) == rst["accepted_steps"])
                  & (np.asarray(st["rejected_steps"]) == rst["rejected_steps"]))
    bit = np.all(hist == ref, axis=(1, 2))
    print(f"bbn {lu}: {same_steps.mean() * 100:.1f}% same step counts, "
          f"{bit.mean() * 100:.1f}% bit-identical to the oracle, "
          f"max abs {np.abs(hist - ref).max():.2e}", flush=True)
    assert same_steps.mean() >= 0.99
    c = 0.1 if lu == "fp32" else 1e-3
    tol = c * (kw["atol"] + kw["rtol"] * np.abs(ref[same_steps]))
    assert np.all(np.abs(hist[same_steps] - ref[same_steps]) <= tol)
    bound = 10.0 * (kw["atol"] + kw["rtol"] * np.abs(ref))
    assert np.all(np.abs(hist - ref) <= bound)


def _device():
    import torch

    return torch.device("cuda", 0)


def test_bbn_likelihood_under_vmap_is_one_ensemble_solve():
    """predict_abundances of one universe, vmapped over a grid: one launch whose
    results are the direct ensemble solve's, bit for bit; the best grid point
    sits next to the calibration point (eta_10 = 6.1, N_eff = 3.044)."""
    import torch
    from torch.func import vmap

    dev = _device()
    theta_np = S.bbn_grid(400)
    theta = torch.tensor(theta_np, device=dev)
    y0 = torch.tensor(S.bbn_initial_state(), device=dev)
    xs = torch.tensor(S.BBN_X_SPAN, device=dev)
    obs = torch.tensor(S.BBN_OBSERVED, device=dev)
    sig = torch.tensor(S.BBN_SIGMA, device=dev)

    def log_likelihood(th):
        sol = rodas5P.solve(S.bbn, y0, xs, th, **S.BBN_SOLVER)
        final = sol[0, -1]
        pred = torch.stack([4.0 * final[3], final[2] / final[1]])
        return -0.5 * (((pred - obs) / sig) ** 2).sum(), pred

    ll, pred = vmap(log_likelihood)(theta)
    direct = rodas5P.solve(S.bbn, np.tile(S.bbn_initial_state(), (400, 1)), S.BBN_X_SPAN,
                           theta_np, **S.BBN_SOLVER)
    np.testing.assert_array_equal(pred.cpu().numpy(), S.bbn_abundances(direct[:, -1]))
    best = theta_np[int(torch.argmax(ll).item())]
    assert abs(best[0] - math.log10(6.1)) < 0.05 and abs(best[1] - 3.044) < 0.25


def test_bbn_chi2_gradient_by_the_adjoint_under_compile():
    """d chi^2 / d(log10 eta_10, N_eff) for 8 universes by the adjoint kernel
    through a torch.compile'd loss (the same as eager, bit for bit), held
    against the reference's own derivative, the forward-sensitivity solve, to
    1e-5 relative at rtol 1e-8, and against central differences to 1e-5 at
    rtol 1e-12. Central differences need the tighter solve: chi^2 is a noisy
    function of the parameters at the level of the solve's accept/reject
    decisions (D/H enters divided by sigma = 3e-7), so at rtol 1e-8 they are
    off by up to 1e-2 at a relative step of 1e-6, worse at smaller steps,
    while the adjoint and the sensitivities agree to ~2e-7."""
    import torch

    dev = _device()
    rng = np.random.default_rng(5)
    theta_np = np.column_stack([rng.uniform(0.6, 0.9, 8), rng.uniform(2.5, 3.5, 8)])
    y0n = np.tile(S.bbn_initial_state(), (8, 1))
    y0 = torch.tensor(y0n, device=dev)
    xs = torch.tensor(S.BBN_X_SPAN, device=dev)
    obs = torch.tensor(S.BBN_OBSERVED, device=dev)
    sig = torch.tensor(S.BBN_SIGMA, device=dev)
    kw = dict(lu_precision="fp64", rtol=1e-8, atol=1e-12, first_step=0.1, max_steps=100000)

    def chi2_with(kw):
        def chi2(th):
            h = rodas5P.solve(S.bbn, y0, xs, th, adjoint=True, **kw)
            final = h[:, -1]
            pred = torch.stack([4.0 * final[:, 3], final[:, 2] / final[:, 1]], dim=1)
            return (((pred - obs) / sig) ** 2).sum()
        return chi2

    grads = []
    for fn in (torch.compile(chi2_with(kw)), chi2_with(kw)):
        th = torch.tensor(theta_np, device=dev, requires_grad=True)
        fn(th).backward()
        grads.append(th.grad.cpu().numpy())
    np.testing.assert_array_equal(grads[0], grads[1])

    # the forward-sensitivity gradient of the same loss
    hs, sens = rodas5P.solve_sensitivities(S.bbn, y0n, S.BBN_X_SPAN, theta_np, wrt_y0=False,
                                           **kw)
    fin, sf = hs[:, -1], sens[:, -1]
    pred = S.bbn_abundances(fin)
    dpred = np.stack([4.0 * sf[:, 3, :],
                      sf[:, 2, :] / fin[:, 1:2] - fin[:, 2:3] / fin[:, 1:2] ** 2 * sf[:, 1, :]],
                     axis=1)
    g_sens = np.einsum("nk,nkp->np", 2.0 * (pred - S.BBN_OBSERVED) / S.BBN_SIGMA ** 2, dpred)
    rel_sens = np.abs(grads[1] - g_sens) / np.abs(g_sens)

    # central differences against the adjoint, both at rtol 1e-12
    tight = dict(lu_precision="fp64", rtol=1e-12, atol=1e-16, first_step=0.1, max_steps=1000000)
    th = torch.tensor(theta_np, device=dev, requires_grad=True)
    chi2_with(tight)(th).backward()
    g_tight = th.grad.cpu().numpy()

    def per_universe(th_np):
        h = rodas5P.solve(S.bbn, y0n, S.BBN_X_SPAN, th_np, **tight)
        pred = S.bbn_abundances(h[:, -1])
        return (((pred - S.BBN_OBSERVED) / S.BBN_SIGMA) ** 2).sum(axis=1)

    fd = np.zeros_like(theta_np)
    for k in range(2):
        h = 1e-4 * np.abs(theta_np[:, k])
        up, dn = theta_np.copy(), theta_np.copy()
        up[:, k] += h
        dn[:, k] -= h
        fd[:, k] = (per_universe(up) - per_universe(dn)) / (2.0 * h)
    rel_fd = np.abs(g_tight - fd) / np.abs(fd)
    print(f"bbn adjoint vs forward sensitivities: max rel {rel_sens.max():.2e}; "
          f"vs central differences (rtol 1e-12): max rel {rel_fd.max():.2e}", flush=True)
    assert rel_sens.max() < 1e-5
    assert rel_fd.max() < 1e-5
