"""torch.func.vmap over single-trajectory solves lowers to one ensemble launch
(the reference's custom_vmap rule, modax/_jax_common.py:49-111)."""

import numpy as np
import pytest

from menodax_amd import rodas5P, tsit5
from tests import systems as S

pytestmark = pytest.mark.gpu


def test_vmap_over_y0_and_params_is_the_ensemble_solve():
    import torch
    from torch.func import vmap

    y0, p = S.robertson_lognormal_ensemble(64, seed=1)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8)
    ens = rodas5P.solve(S.robertson, y0, t, p, **kw)
    dev = torch.device("cuda")
    out = vmap(lambda yy, pp: rodas5P.solve(S.robertson, yy, t, pp, **kw))(
        torch.tensor(y0, device=dev), torch.tensor(p, device=dev))
    assert tuple(out.shape) == (64, 1, t.size, 3)
    np.testing.assert_array_equal(out[:, 0].cpu().numpy(), ens)


def test_vmap_over_params_only_broadcasts_y0_and_keeps_stats():
    import torch
    from torch.func import vmap

    _, p = S.robertson_lognormal_ensemble(32, seed=2)
    t = S.robertson_benchmark_times()
    y0 = torch.tensor(S.ROBERTSON_Y0, device="cuda")
    kw = dict(rtol=1e-6, atol=1e-8, return_stats=True)
    hist, stats = vmap(lambda pp: rodas5P.solve(S.robertson, y0, t, pp, **kw))(
        torch.tensor(p, device="cuda"))
    ref, rstats = rodas5P.solve(S.robertson, S.ROBERTSON_Y0, t, p, **kw)
    np.testing.assert_array_equal(hist[:, 0].cpu().numpy(), ref)
    np.testing.assert_array_equal(stats["accepted_steps"][:, 0].cpu().numpy(),
                                  rstats["accepted_steps"])


def test_tsit5_vmap_and_nested_ensembles_raise():
    import torch
    from torch.func import vmap

    y0, p = S.lorenz_rho_ensemble(16)
    t = np.linspace(0.0, 1.0, 5)
    ens = tsit5.solve(S.lorenz, y0, t, p, rtol=1e-6, atol=1e-8)
    out = vmap(lambda yy, pp: tsit5.solve(S.lorenz, yy, t, pp, rtol=1e-6, atol=1e-8))(
        torch.tensor(y0, device="cuda"), torch.tensor(p, device="cuda"))
    np.testing.assert_array_equal(out[:, 0].cpu().numpy(), ens)
    nested = torch.tensor(np.stack([y0, y0]), device="cuda")
    with pytest.raises(NotImplementedError, match="already-ensembled"):
        vmap(lambda yy: tsit5.solve(S.lorenz, yy, t, p))(nested)
