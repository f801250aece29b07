"""Tsit5 on MI355X: ``solve(ode_fn, y0, t_span, params, ...)``.

The API is the reference's (modax/tsit5.py:623-709). ``backend`` keeps the
reference's values and size checks: there, "shared" put the nine
per-trajectory vectors in shared memory and "local" in thread-local arrays
(tsit5.py:92-139). Here where the vectors live is a placement chosen by size
(``_kernels.tsit5_placement``): in VGPRs up to 16 variables -- which covers
every system the shared backend accepted, so both names select the same kernel
there -- and above that a rolled stage loop with the stage vectors, stage
argument and state in LDS as far as they fit. Placements are bit-identical.
Differentiation and ``torch.func.vmap`` work as for ``rodas5P.solve``:
gradients come from the forward-sensitivity joint solve by default, or, with
``adjoint=True``, from the discrete-adjoint kernel (checkpointed every
``ckpt_every`` accepted steps), which ``adjoint_gradients`` also runs directly.
``tableau_in_registers`` holds the method's coefficients in registers for the
unrolled step (``None`` = by system and ensemble size, as for Rodas5P); the
bits are the same either way.
"""

from __future__ import annotations

from dataclasses import replace

from . import _autodiff, _common, _kernels
from ._kernels import Tsit5Options
from .rodas5P import _tableau, _use_refill, finish, launch

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
    work_order=None,
    adjoint=False,
    ckpt_every=64,
    tableau_in_registers=None,
):
    """Integrate an ensemble with Tsit5; see the module docstring."""
    if _autodiff.under_vmap(y0, t_span, params):
        # torch.func.vmap over single-trajectory solves: one ensemble launch
        kwargs = dict(rtol=rtol, atol=atol, first_step=first_step, max_steps=max_steps,
                      return_stats=True, error_weights=error_weights, pcoeff=pcoeff,
                      icoeff=icoeff, dcoeff=dcoeff, backend=backend,
                      trajectories_per_block=trajectories_per_block, refill=refill, stream=stream,
                      tableau_in_registers=tableau_in_registers)

        def run(y, t, p):
            hist, stats = solve(ode_fn, y, t, p, **kwargs)
            return (hist, stats["accepted_steps"], stats["rejected_steps"], stats["loop_steps"])

        dev = _common.device()
        hist, acc, rej, loop = _autodiff.vmap_solve(
            run, *(_common.as_f64_device(x, dev) for x in (y0, t_span, params)))
        if return_stats:
            return hist, _common.solver_stats(acc, rej, loop)
        return hist
    dev = _common.device()
    y0_t, p_t, n, n_vars = _common.normalize_y0_params(y0, params, dev)
    n_params = p_t.shape[1]
    options = Tsit5Options(
        pcoeff=float(pcoeff),
        icoeff=float(icoeff),
        dcoeff=float(dcoeff),
        trajectories_per_block=_block(trajectories_per_block, n_vars, n),
        unit_weights=error_weights is None,
        tableau_in_registers=_tableau(tableau_in_registers, n_vars, n, None, None),
    )
    times = _common.as_f64_device(t_span, dev).contiguous()
    if times.ndim != 1 or times.numel() < 1:
        raise ValueError("t_span must be a non-empty 1-D array of save times")
    launch_kw = dict(dt0=_common.initial_step(first_step), rtol=rtol, atol=atol,
                     max_steps=max_steps, refill=refill, stream=stream,
                     order=_common.work_order_arg(work_order, n, dev))
    wrt_y0, wrt_params = _autodiff.wanted(y0, params)
    if (wrt_y0 or wrt_params) and adjoint:
        _check_backend(n_vars, backend)
        job = _autodiff.AdjointJob(
            built=_kernels.tsit5_kernel(ode_fn, n_vars, n_params, options),
            adjoint=_kernels.tsit5_adjoint_kernel(ode_fn, n_vars, n_params, options),
            times=times, n=n, n_vars=n_vars, n_params=n_params,
            weights=_common.error_weights_soa(error_weights, n, n_vars, dev),
            launch_kw=launch_kw, ckpt_every=int(ckpt_every))
        hist, counters = _autodiff.adjoint_solve(job, y0_t, p_t)
        if return_stats:
            return hist, _common.solver_stats(*counters)
        return hist
    if wrt_y0 or wrt_params:
        job = _joint_job(ode_fn, options, backend, times, n, n_vars, n_params, error_weights,
                         dev, launch_kw)
        state, counters = _autodiff.differentiable_solve(
            job, y0_t, p_t, wrt_y0, wrt_params, sens_error_control, sens_param_columns)
        if return_stats:
            return state, _common.solver_stats(*counters)
        return state
    _check_backend(n_vars, backend)
    built = _kernels.tsit5_kernel(ode_fn, n_vars, n_params, options)
    if launch_kw["order"] is None and refill is None and _use_refill(built, n):
        # the resumed pilot's claim order, as for rodas5P.solve
        launch_kw = dict(launch_kw, order="pilot")
    out = launch(built, y0_t.detach(), times, p_t.detach(), n,
                 weights=_common.error_weights_soa(error_weights, n, n_vars, dev), **launch_kw)
    return finish(out, return_stats, _common.wants_numpy(y0, params, t_span))


def _joint_job(ode_fn, options, backend, times, n, n_vars, n_params, error_weights, dev,
               launch_kw):
    w = None
    if error_weights is not None:
        w = _common.as_f64_device(error_weights, dev)
        w = w.expand(n, n_vars) if w.ndim == 1 else w

    def build(spec, unit_weights):
        # the joint system is what the backend has to hold (modax/tsit5.py:123-139)
        _check_backend(spec.n_aug, backend)
        return _kernels.tsit5_kernel(ode_fn, n_vars, n_params,
                                     replace(options, sens=spec, unit_weights=unit_weights))

    return _autodiff.JointJob(build=build, times=times, n=n, n_vars=n_vars, n_params=n_params,
                              weights=w, launch_kw=launch_kw)


def adjoint_gradients(ode_fn, y0, t_span, params, hist_cotangent, *, ckpt_every=64,
                      rtol=1e-8, atol=1e-10, first_step=None, max_steps=100000,
                      error_weights=None, pcoeff=0.0, icoeff=1.0, dcoeff=0.0, backend="auto",
                      trajectories_per_block=None, stream=None, placement=None):
    """Gradients of a loss of the Tsit5 solve's history by the discrete
    adjoint; arguments and return as ``rodas5P.adjoint_gradients``:
    ``hist_cotangent`` is dL/dhist ``(N, n_save, n_vars)`` for the solve with
    the same options, the result ``(dL/dy0 (N, n_vars), dL/dparams (N,
    n_params))``. Accepted steps are differentiated with their sizes held
    fixed; saves the solve never reached (NaN in its history) carry no
    derivative. ``placement`` chooses where the kernels keep their per-lane
    vectors (by size when None); every placement gives the same bits."""
    dev = _common.device()
    y0_t, p_t, n, n_vars = _common.normalize_y0_params(y0, params, dev)
    n_params = p_t.shape[1]
    _check_backend(n_vars, backend)
    options = Tsit5Options(
        pcoeff=float(pcoeff), icoeff=float(icoeff), dcoeff=float(dcoeff),
        trajectories_per_block=_kernels.trajectories_per_block_or_default(trajectories_per_block),
        unit_weights=error_weights is None, placement=placement,
    )
    times = _common.as_f64_device(t_span, dev).contiguous()
    launch_kw = dict(dt0=_common.initial_step(first_step), rtol=rtol, atol=atol,
                     max_steps=max_steps, refill=False, stream=stream)
    weights = _common.error_weights_soa(error_weights, n, n_vars, dev)
    job = _autodiff.AdjointJob(
        built=_kernels.tsit5_kernel(ode_fn, n_vars, n_params, options),
        adjoint=_kernels.tsit5_adjoint_kernel(ode_fn, n_vars, n_params, options),
        times=times, n=n, n_vars=n_vars, n_params=n_params, weights=weights,
        launch_kw=launch_kw, ckpt_every=int(ckpt_every))
    y0_d, p_d = y0_t.detach(), p_t.detach()
    g = _common.as_f64_device(hist_cotangent, dev)
    if tuple(g.shape) != (n, times.numel(), n_vars):
        raise ValueError(f"hist_cotangent must be (N, n_save, n_vars) = "
                         f"{(n, times.numel(), n_vars)}; got {tuple(g.shape)}")
    gy, gp = _autodiff.adjoint_gradients(job, y0_d, p_d, g)
    if _common.wants_numpy(y0, params, t_span, hist_cotangent):
        return gy.cpu().numpy(), gp.cpu().numpy()
    return gy, gp


def solve_sensitivities(ode_fn, y0, t_span, params, *, wrt_y0=True, wrt_params=True,
                        sens_error_control=True, sens_param_columns=None, return_stats=False,
                        rtol=1e-8, atol=1e-10, first_step=None, max_steps=100000,
                        error_weights=None, pcoeff=0.0, icoeff=1.0, dcoeff=0.0, backend="auto",
                        trajectories_per_block=None, refill=None, stream=None):
    """One joint ``[y, S]`` Tsit5 solve: the state ``(N, n_save, n_vars)`` and
    ``S`` ``(N, n_save, n_vars, n_sens)``; see ``rodas5P.solve_sensitivities``."""
    dev = _common.device()
    y0_t, p_t, n, n_vars = _common.normalize_y0_params(y0, params, dev)
    n_params = p_t.shape[1]
    options = Tsit5Options(
        pcoeff=float(pcoeff), icoeff=float(icoeff), dcoeff=float(dcoeff),
        trajectories_per_block=_kernels.trajectories_per_block_or_default(trajectories_per_block),
        unit_weights=error_weights is None,
    )
    times = _common.as_f64_device(t_span, dev).contiguous()
    launch_kw = dict(dt0=_common.initial_step(first_step), rtol=rtol, atol=atol,
                     max_steps=max_steps, refill=refill, stream=stream)
    job = _joint_job(ode_fn, options, backend, times, n, n_vars, n_params, error_weights, dev,
                     launch_kw)
    job.spec = _autodiff.SensitivitySpec(
        n_vars, n_params, bool(wrt_y0), bool(wrt_params), bool(sens_error_control),
        None if sens_param_columns is None else tuple(int(c) for c in sens_param_columns))
    out, state, sens = _autodiff.joint_solve(job, y0_t.detach(), p_t.detach())
    stats = _common.solver_stats(out.accepted, out.rejected, out.loop)
    if _common.wants_numpy(y0, params, t_span):
        state, sens = state.cpu().numpy(), sens.cpu().numpy()
        stats = {k: v.cpu().numpy() for k, v in stats.items()}
    return (state, sens, stats) if return_stats else (state, sens)


def _block(requested, n_vars, n_traj):
    """``_kernels.tsit5_block`` for this device (its compute units are only
    asked for when the rule needs them: a rolled step and no caller's size)."""
    if requested is not None or n_vars <= _kernels.UNROLLED_MAX_VARS:
        return _kernels.tsit5_block(requested, n_vars, n_traj, 0)
    return _kernels.tsit5_block(requested, n_vars, n_traj, _common.n_cu())


def _traced(ode_fn, y0, t_span, params, *, rtol=1e-8, atol=1e-10, first_step=None,
            max_steps=100000, return_stats=False, error_weights=None, pcoeff=0.0, icoeff=1.0,
            dcoeff=0.0, backend="auto", sens_error_control=True, sens_param_columns=None,
            trajectories_per_block=None, refill=None, stream=None, work_order=None,
            adjoint=False, ckpt_every=64, tableau_in_registers=None):
    """``solve`` while ``torch.compile`` traces it (see rodas5P._traced)."""
    from . import _ops

    if error_weights is not None or stream is not None or work_order is not None:
        return NotImplemented
    if hasattr(y0, "shape") and len(y0.shape) in (1, 2):
        _check_backend(y0.shape[-1], backend)
    n_vars = y0.shape[-1] if hasattr(y0, "shape") and len(y0.shape) in (1, 2) else 0
    n_traj = 1
    for x in (y0, params):
        if hasattr(x, "shape") and len(x.shape) == 2:
            n_traj = max(n_traj, int(x.shape[0]))
    options = (("pcoeff", float(pcoeff)), ("icoeff", float(icoeff)), ("dcoeff", float(dcoeff)),
               ("trajectories_per_block", _block(trajectories_per_block, n_vars, n_traj)),
               ("unit_weights", True))
    return _ops.traced_solve("tsit5", ode_fn, y0, t_span, params, options=options,
                             sparsity=None, tableau=tableau_in_registers, rtol=rtol, atol=atol,
                             first_step=first_step, max_steps=max_steps, refill=refill,
                             adjoint=adjoint, ckpt_every=ckpt_every, return_stats=return_stats)


# Traced by torch.compile, a plain solve is one custom operator (menodax::
# ensemble_solve); everything else runs eagerly (see _common.eager_only).
solve = _common.compilable(solve, _traced)
adjoint_gradients = _common.eager_only(adjoint_gradients)
solve_sensitivities = _common.eager_only(solve_sensitivities)
