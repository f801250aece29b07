"""The autotuner times bit-identical candidates and returns the fastest."""

import numpy as np
import pytest

from menodax_amd import autotune, rodas5P
from tests import systems as S

pytestmark = pytest.mark.gpu


def test_tuned_options_reproduce_the_default_solve():
    fn, mask, _ = S.arrow_system(24)
    y0, lam = S.arrow_ensemble(2048, n=24, lam_range=(1e-1, 1e3))
    t = np.linspace(0.0, 2.0, 5)
    kw = dict(rtol=1e-6, atol=1e-8, lu_precision="fp32", sparsity=mask)
    best, table = autotune.tune(fn, y0, t, lam, blocks=(64, 128), **kw)
    assert best in table and best.seconds == min(c.seconds for c in table)
    assert {c.placement for c in table} == {"registers", "rolled", "lds"}
    opts = autotune.tuned_options(fn, y0, t, lam, blocks=(64, 128), **kw)
    tuned = rodas5P.solve(fn, y0, t, lam, **opts, **kw)
    np.testing.assert_array_equal(tuned, rodas5P.solve(fn, y0, t, lam, **kw))


def test_a_sample_is_timed_in_the_launch_shape_of_the_whole_ensemble():
    from menodax_amd import _kernels
    from menodax_amd._sparsity import normalize_sparsity

    fn, mask, _ = S.arrow_system(24)
    y0, lam = S.arrow_ensemble(4096, n=24, lam_range=(1e-1, 1e3))
    t = np.linspace(0.0, 2.0, 5)
    kw = dict(rtol=1e-6, atol=1e-8, lu_precision="fp32", sparsity=mask)
    best, table = autotune.tune(fn, y0, t, lam, sample=1024, blocks=(64, 128), **kw)
    pairs = [(c.placement, c.trajectories_per_block) for c in table]
    assert len(pairs) == len(set(pairs)), "one launch shape per candidate when sampling"
    for c in table:
        built = _kernels.rodas5p_kernel(fn, 24, lam.shape[1], _kernels.KernelOptions(
            lu_precision="fp32", sparsity=normalize_sparsity(mask, 24), placement=c.placement,
            trajectories_per_block=c.trajectories_per_block))
        assert c.refill == rodas5P._use_refill(built, 4096)


def test_a_small_system_is_timed_with_and_without_the_tableau_in_registers():
    y0, p = S.robertson_lognormal_ensemble(4096, seed=3)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8, lu_precision="fp32")
    best, table = autotune.tune(S.robertson, y0, t, p, blocks=(64,), **kw)
    regs = {c.tableau_in_registers for c in table if c.placement == "registers"}
    assert regs == {False, True}
    assert not any(c.tableau_in_registers for c in table if c.placement != "registers")
    opts = autotune.tuned_options(S.robertson, y0, t, p, blocks=(64,), **kw)
    np.testing.assert_array_equal(rodas5P.solve(S.robertson, y0, t, p, **opts, **kw),
                                  rodas5P.solve(S.robertson, y0, t, p, **kw))


def test_a_dense_system_above_the_unrolled_size_is_timed_in_the_wave_placement():
    fn, _ = S.vdp_lattice(12)
    y0, p = S.vdp_scenario(12, 256, 3, divergence=0.25)
    kw = dict(first_step=1e-4, rtol=1e-6, atol=1e-8, lu_precision="fp32")
    best, table = autotune.tune(fn, y0, S.VDP_TIMES, p, blocks=(16, 64), **kw)
    waves = [c for c in table if c.placement == "wave"]
    assert len(waves) == 1 and not waves[0].refill
    opts = autotune.tuned_options(fn, y0, S.VDP_TIMES, p, blocks=(16, 64), **kw)
    np.testing.assert_array_equal(rodas5P.solve(fn, y0, S.VDP_TIMES, p, **opts, **kw),
                                  rodas5P.solve(fn, y0, S.VDP_TIMES, p, **kw))
