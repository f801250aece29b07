"""The Rodas5P save hook, run inside the launch at every save.

Mirrors the reference's tests/test_save_hook.py: the hook sees the same
dense-output states the history holds, so an accumulation it performs in-kernel
matches the same accumulation done on the saved history afterwards (a trapezoid
integral, and the energy stored per save); save_history=False still returns the
final state. Every placement and both launch shapes agree bit for bit.
"""

import numpy as np
import pytest

from menodax_amd import rodas5P
from tests import systems as S

pytestmark = pytest.mark.gpu

N_SAVE = 33


def _problem():
    t_span = np.linspace(0.0, 4.0, N_SAVE)
    y0 = np.array([[1.0, 0.0], [0.5, 1.0], [-1.0, 0.3]])
    params = np.array([[4.0, 0.2], [9.0, 0.5], [1.0, 0.0]])
    return y0, t_span, params


SETTINGS = dict(rtol=1e-8, atol=1e-10, lu_precision="fp64")


@pytest.mark.parametrize("placement", [None, "rolled", "lds"])
def test_hook_accumulates_what_the_history_integrates(placement):
    y0, t_span, params = _problem()
    hist = rodas5P.solve(S.damped_oscillator, y0, t_span, params, **SETTINGS)
    final, acc = rodas5P.solve(S.damped_oscillator, y0, t_span, params,
                               save_hook=S.trapezoid_hook(t_span), hook_size=3,
                               save_history=False, placement=placement, **SETTINGS)
    assert final.shape == (3, 1, 2)
    np.testing.assert_array_equal(final[:, 0], hist[:, -1])
    np.testing.assert_allclose(acc[:, 0], np.trapezoid(hist[..., 0], t_span, axis=1), rtol=1e-12)
    np.testing.assert_allclose(acc[:, 1], np.trapezoid(hist[..., 1], t_span, axis=1), rtol=1e-12)
    assert np.all(acc[:, 2] == N_SAVE)


@pytest.mark.parametrize("placement", [None, "rolled", "lds"])
def test_hook_can_store_derived_saves_beside_the_history(placement):
    y0, t_span, params = _problem()
    plain = rodas5P.solve(S.damped_oscillator, y0, t_span, params, **SETTINGS)
    for refill in (False, True):
        hist, energy = rodas5P.solve(S.damped_oscillator, y0, t_span, params,
                                     save_hook=S.energy_hook, hook_size=N_SAVE,
                                     placement=placement, refill=refill, **SETTINGS)
        assert hist.shape == (3, N_SAVE, 2)
        np.testing.assert_array_equal(hist, plain)
        expected = 0.5 * (hist[..., 1] ** 2 + params[:, :1] * hist[..., 0] ** 2)
        np.testing.assert_allclose(energy, expected, rtol=1e-13, atol=1e-15)
        # the undamped trajectory conserves energy; the hooked values show it
        np.testing.assert_allclose(energy[2], energy[2, 0], rtol=1e-6)
