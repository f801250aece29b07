"""Tsit5 on MI355X: ``solve(ode_fn, y0, t_span, params, ...)``.

The API is the reference's (modax/tsit5.py:623-709). ``backend`` keeps the
reference's values: there, "shared" put the nine per-trajectory vectors in
shared memory and "local" in thread-local arrays (tsit5.py:92-139). Here the
vectors live in VGPRs in both cases -- 9 n doubles per lane fits the register
file for every system the shared backend accepted -- so the two names select
the same kernel and are bit-identical by construction; the reference's size
check on "shared" is kept so the same calls fail the same way.
"""

from __future__ import annotations

from . import _common, _kernels
from ._kernels import Tsit5Options
from .rodas5P import finish, launch

_SHARED_MAX_NVARS = 16


def _check_backend(n_system: int, backend: str) -> None:
    """modax/tsit5.py:123-139."""
    if backend == "local" or backend == "auto":
        return
    if backend == "shared":
        if n_system > _SHARED_MAX_NVARS:
            raise ValueError(
                f"shared backend requires a system size <= {_SHARED_MAX_NVARS}; got {n_system}"
            )
        return
    raise ValueError(f"backend must be 'auto', 'shared', or 'local'; got {backend!r}")


def solve(
    ode_fn,
    y0,
    t_span,
    params,
    *,
    rtol=1e-8,
    atol=1e-10,
    first_step=None,
    max_steps=100000,
    return_stats=False,
    error_weights=None,
    pcoeff=0.0,
    icoeff=1.0,
    dcoeff=0.0,
    backend="auto",
    sens_error_control=True,
    sens_param_columns=None,
    trajectories_per_block=None,
    refill=None,
    stream=None,
):
    """Integrate an ensemble with Tsit5; see the module docstring."""
    del sens_error_control, sens_param_columns
    dev = _common.device()
    y0_t, p_t, n, n_vars = _common.normalize_y0_params(y0, params, dev)
    _check_backend(n_vars, backend)
    options = Tsit5Options(
        pcoeff=float(pcoeff),
        icoeff=float(icoeff),
        dcoeff=float(dcoeff),
        trajectories_per_block=_kernels.trajectories_per_block_or_default(trajectories_per_block),
        unit_weights=error_weights is None,
    )
    built = _kernels.tsit5_kernel(ode_fn, n_vars, p_t.shape[1], options)
    times = _common.as_f64_device(t_span, dev).contiguous()
    if times.ndim != 1 or times.numel() < 1:
        raise ValueError("t_span must be a non-empty 1-D array of save times")
    out = launch(
        built, y0_t, times, p_t, n,
        dt0=_common.initial_step(first_step), rtol=rtol, atol=atol, max_steps=max_steps,
        weights=_common.error_weights_soa(error_weights, n, n_vars, dev), refill=refill,
        stream=stream,
    )
    return finish(out, return_stats, _common.wants_numpy(y0, params, t_span))
