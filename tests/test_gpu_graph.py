"""Solves captured into a HIP graph replay bit-identically.

Each replay is checked against the eager solve, with other GPU work and
synchronisations between the replays; new inputs copied into the captured
buffers give the eager result for those inputs."""

import numpy as np
import pytest

from menodax_amd import rodas5P
from tests import systems as S
from menodax_amd.graph import CapturedSolve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("refill", [False, True])
def test_replays_match_the_eager_solve(refill):
    import torch

    y0, p = S.robertson_lognormal_ensemble(4096, seed=1)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8, lu_precision="fp32", refill=refill)
    eager, est = rodas5P.solve(S.robertson, y0, t, p, return_stats=True, **kw)
    cs = CapturedSolve(S.robertson, y0, t, p, **kw)
    # the persistent launch's counter: the eager launch's, which the graph reuses
    assert (cs._ws is not None) == refill
    for k in range(3):
        hist, st = cs.run(return_stats=True)
        np.testing.assert_array_equal(hist, eager)
        for name in est:
            np.testing.assert_array_equal(st[name], est[name])
        # unrelated eager work between replays
        torch.randn(1 << 16, device="cuda", dtype=torch.float64).cumsum(0).sum().item()
    assert cs.graph is not None


def test_new_inputs_through_the_captured_buffers():
    y0, p = S.robertson_lognormal_ensemble(2048, seed=2)
    y0b, pb = S.robertson_lognormal_ensemble(2048, seed=3)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8)
    cs = CapturedSolve(S.robertson, y0, t, p, **kw)
    first = cs.run().copy()
    np.testing.assert_array_equal(first, rodas5P.solve(S.robertson, y0, t, p, **kw))
    second = cs.run(y0=y0b, params=pb)
    np.testing.assert_array_equal(second, rodas5P.solve(S.robertson, y0b, t, pb, **kw))
    again = cs.run(y0=y0, params=p)
    np.testing.assert_array_equal(again, first)


def test_a_claim_order_set_between_replays():
    from menodax_amd import _common

    y0, p = S.robertson_lognormal_ensemble(4096, seed=4)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8)
    eager, est = rodas5P.solve(S.robertson, y0, t, p, return_stats=True, **kw)
    cs = CapturedSolve(S.robertson, y0, t, p, refill=True, work_order=np.arange(4096), **kw)
    hist, st = cs.run(return_stats=True)
    np.testing.assert_array_equal(hist, eager)
    cs.set_work_order(_common.cost_order(st["accepted_steps"] + st["rejected_steps"]))
    for _ in range(2):
        np.testing.assert_array_equal(cs.run(), eager)
    with pytest.raises(ValueError, match="permutation"):
        cs.set_work_order(np.zeros(4096, dtype=np.int64))


def test_a_wave_placement_solve_replays_bit_identically():
    """A small ensemble of a 24-variable dense system takes the wave
    placement (one trajectory per 64-lane workgroup, with barriers inside the
    kernel); its captured replays match the eager solve with new inputs too."""
    import torch

    fn, _ = S.vdp_lattice(12)
    y0, p = S.vdp_scenario(12, 200, 8, divergence=0.25)
    y0b, pb = S.vdp_scenario(12, 200, 9, divergence=0.25)
    kw = dict(first_step=1e-4, rtol=1e-6, atol=1e-8)
    cs = CapturedSolve(fn, y0, S.VDP_TIMES, p, **kw)
    assert cs.built.extra["placement"]["placement"] == "wave"
    eager = rodas5P.solve(fn, y0, S.VDP_TIMES, p, placement="rolled", **kw)
    for _ in range(2):
        np.testing.assert_array_equal(cs.run(), eager)
        torch.randn(1 << 16, device="cuda", dtype=torch.float64).cumsum(0).sum().item()
    np.testing.assert_array_equal(cs.run(y0=y0b, params=pb),
                                  rodas5P.solve(fn, y0b, S.VDP_TIMES, pb, placement="rolled", **kw))
