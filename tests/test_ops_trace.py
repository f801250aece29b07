"""Tracing a solve with torch.compile on the host: the compiler sees one
menodax::ensemble_solve node (the kernel launch, with a shape-only fake
implementation) and no graph break; what only the eager path does breaks the
graph instead. The traced graphs are recorded, not run (no GPU here)."""

import numpy as np
import pytest

from menodax_amd import rodas5P, tsit5
from tests import systems as S


def _trace(f, *args):
    import torch

    graphs = []

    def backend(gm, example_inputs):
        graphs.append(gm)
        n_out = len(next(n for n in gm.graph.nodes if n.op == "output").args[0])
        return lambda *a: tuple(torch.zeros(()) for _ in range(n_out))

    torch._dynamo.reset()
    torch.compile(f, fullgraph=True, backend=backend)(*args)
    return graphs


def _ops(gm):
    return [n.target for n in gm.graph.nodes if n.op == "call_function"]


def test_a_rodas5p_solve_is_one_operator_node():
    import torch

    y0 = torch.tensor(np.tile([1.0, 0.0, 0.0], (8, 1)))
    p = torch.tensor(np.tile([0.04, 3e7, 1e4], (8, 1)))
    t = torch.linspace(0.0, 1.0, 5, dtype=torch.float64)
    fn, mask, _ = S.arrow_system(6)

    def f(y0, p, ya, pa):
        h = rodas5P.solve(S.robertson, y0, t, p, rtol=1e-6, atol=1e-8)
        ha, st = rodas5P.solve(fn, ya, t, pa, sparsity=mask, return_stats=True)
        return (h[:, -1] ** 2).sum() + ha.sum(), st["accepted_steps"]

    ya, lam = S.arrow_ensemble(4, n=6)
    (gm,) = _trace(f, y0, p, torch.tensor(ya), torch.tensor(lam))
    solves = [n for n in gm.graph.nodes if n.target is torch.ops.menodax.ensemble_solve]
    assert len(solves) == 2
    # the solver is a constant of the graph; the two solvers differ
    keys = [n.args[3] for n in solves]
    assert all(isinstance(k, int) for k in keys) and keys[0] != keys[1]
    assert solves[0].args[4:] == (0.0, 1e-06, 1e-08, 100000, -1)


def test_a_tsit5_solve_and_a_broadcast_y0_trace_without_a_break():
    import torch

    y0 = torch.tensor([1.0, 0.0, 0.0], dtype=torch.float64)
    rho = torch.linspace(5.0, 28.0, 16, dtype=torch.float64)[:, None]
    t = torch.linspace(0.0, 1.0, 10, dtype=torch.float64)

    def f(y0, p):
        return tsit5.solve(S.lorenz, y0, t, p, rtol=1e-6, atol=1e-8, refill=True).sum()

    (gm,) = _trace(f, y0, rho)
    assert torch.ops.menodax.ensemble_solve in _ops(gm)
    node = next(n for n in gm.graph.nodes if n.target is torch.ops.menodax.ensemble_solve)
    assert node.args[-1] == 1
    assert tuple(node.args[0].meta["example_value"].shape) == (16, 3)


def test_what_only_the_eager_path_does_breaks_the_graph():
    import torch

    y0 = torch.tensor(np.tile([1.0, 0.0, 0.0], (8, 1)))
    p = torch.tensor(np.tile([0.04, 3e7, 1e4], (8, 1)))
    t = torch.linspace(0.0, 1.0, 5, dtype=torch.float64)

    def weighted(y0, p):
        return rodas5P.solve(S.robertson, y0, t, p, error_weights=[1.0, 2.0, 1.0]).sum()

    def sensitivities(y0, p):
        # gradients without adjoint=True come from the eager forward-sensitivity solve
        return rodas5P.solve(S.robertson, y0, t, p).sum()

    with pytest.raises(Exception, match="(?i)graph break|skip|disable"):
        _trace(weighted, y0, p)
    with pytest.raises(Exception, match="(?i)graph break|skip|disable"):
        _trace(sensitivities, y0, p.clone().requires_grad_(True))
