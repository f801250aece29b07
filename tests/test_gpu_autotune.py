"""The autotuner times bit-identical candidates and returns the fastest."""

import numpy as np
import pytest

from menodax_amd import autotune, rodas5P, systems as S

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
