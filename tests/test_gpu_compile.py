"""Solves called from torch.compile'd code. A plain solve on tensors is traced
as the custom operator menodax::ensemble_solve (also with fullgraph=True), its
gradients through menodax::ensemble_adjoint; what only the eager path does
makes the compiler stop its graph at the solve and run it eagerly. Either way
results and gradients equal the eager ones."""

import numpy as np
import pytest

from menodax_amd import rodas5P, tsit5
from tests import systems as S

pytestmark = pytest.mark.gpu


def test_solve_inside_a_compiled_function():
    import torch

    dev = torch.device("cuda", 0)
    y0n, pn = S.robertson_lognormal_ensemble(256, seed=1)
    y0, p = torch.tensor(y0n, device=dev), torch.tensor(pn, device=dev)
    t = torch.tensor(S.robertson_benchmark_times(), device=dev)

    def f(y0, p):
        h = rodas5P.solve(S.robertson, y0, t, p, rtol=1e-6, atol=1e-8)
        return (h[:, -1] ** 2).sum(), h

    loss, h = torch.compile(f)(y0, p)
    eager_loss, eager = f(y0, p)
    assert torch.equal(h, eager)
    assert torch.equal(loss, eager_loss)


def test_gradients_through_a_compiled_function():
    import torch

    dev = torch.device("cuda", 0)
    y0n, rho = S.lorenz_rho_ensemble(64)
    t = np.linspace(0.0, 1.0, 10)

    def f(y0, p):
        h = tsit5.solve(S.lorenz, y0, t, p, rtol=1e-8, atol=1e-10, adjoint=True)
        return (h ** 2).sum()

    grads = []
    for fn in (torch.compile(f), f):
        y0 = torch.tensor(y0n, device=dev, requires_grad=True)
        p = torch.tensor(rho, device=dev, requires_grad=True)
        fn(y0, p).backward()
        grads.append((y0.grad, p.grad))
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])


def test_a_full_graph_holds_the_solve_as_one_operator():
    import torch

    dev = torch.device("cuda", 0)
    y0n, pn = S.robertson_lognormal_ensemble(512, seed=4)
    y0, p = torch.tensor(y0n, device=dev), torch.tensor(pn, device=dev)
    t = torch.tensor(S.robertson_benchmark_times(), device=dev)

    def rob(y0, p):
        h, st = rodas5P.solve(S.robertson, y0, t, p, rtol=1e-6, atol=1e-8, return_stats=True)
        return (h[:, -1] ** 2).sum(), h, st["accepted_steps"], st["rejected_steps"]

    fn, mask, _ = S.arrow_system(24)
    ya, lam = S.arrow_ensemble(300, n=24)
    ya, lam = torch.tensor(ya, device=dev), torch.tensor(lam, device=dev)
    ta = torch.linspace(0.0, 2.0, 5, dtype=torch.float64, device=dev)

    def arrow(y0, p):
        h = rodas5P.solve(fn, y0, ta, p, rtol=1e-6, atol=1e-8, sparsity=mask)
        return h.sum(), h

    yl, rho = S.lorenz_rho_ensemble(1024)
    yl, rho = torch.tensor(yl, device=dev), torch.tensor(rho, device=dev)
    tl = torch.linspace(0.0, 1.0, 100, dtype=torch.float64, device=dev)

    def lorenz(y0, p):
        h, st = tsit5.solve(S.lorenz, y0, tl, p, rtol=1e-6, atol=1e-8, return_stats=True)
        return h.sum(), h, st["accepted_steps"]

    for f, args in ((rob, (y0, p)), (arrow, (ya, lam)), (lorenz, (yl, rho))):
        got = torch.compile(f, fullgraph=True)(*args)
        want = f(*args)
        # the reduction is compiled code, summed in its own order
        torch.testing.assert_close(got[0], want[0], rtol=1e-13, atol=0)
        for k, (a, b) in enumerate(zip(got[1:], want[1:])):
            assert torch.equal(a, b), (f.__name__, k)


def test_full_graph_gradients_run_the_adjoint_operator():
    import torch

    dev = torch.device("cuda", 0)
    y0n, rho = S.lorenz_rho_ensemble(64)
    t = torch.linspace(0.0, 1.0, 10, dtype=torch.float64, device=dev)
    vdp_fn, _ = S.vdp_lattice(1, mu=10.0, d=0.1, wrap=False)
    yv = np.random.default_rng(2).uniform(-2.0, 2.0, (128, 2))
    tv = torch.linspace(0.0, 5.0, 6, dtype=torch.float64, device=dev)

    def lorenz(y0, p):
        h = tsit5.solve(S.lorenz, y0, t, p, rtol=1e-8, atol=1e-10, adjoint=True)
        return (h ** 2).sum()

    def vdp(y0, p):
        h = rodas5P.solve(vdp_fn, y0, tv, p, rtol=1e-8, atol=1e-10, adjoint=True)
        return (h ** 2).sum()

    for f, y, q in ((lorenz, y0n, rho), (vdp, yv, np.ones((128, 1)))):
        grads = []
        for fn in (torch.compile(f, fullgraph=True), f):
            y0 = torch.tensor(y, device=dev, requires_grad=True)
            p = torch.tensor(q, device=dev, requires_grad=True)
            fn(y0, p).backward()
            grads.append((y0.grad, p.grad))
        torch.testing.assert_close(grads[0][0], grads[1][0], rtol=1e-12, atol=0)
        torch.testing.assert_close(grads[0][1], grads[1][1], rtol=1e-12, atol=0)
