"""Differentiating a solve: forward sensitivities behind ``torch.autograd``.

The counterpart of modax/_sensitivity.py:331-413 (``make_sensitivity_solver``).
There a ``jax.custom_jvp`` rule integrates the joint ``[y, S]`` system whenever
JAX differentiates the solve, and contracts ``S`` with the input tangents; JAX
transposes that contraction for reverse mode. Here a ``torch.autograd.Function``
does the same for both of torch's modes:

* its ``forward`` integrates the joint system for the inputs that need a
  derivative (``requires_grad``, or a forward-mode dual tangent) and keeps ``S``;
* ``backward`` contracts ``S`` with the output cotangent (the transpose);
* ``jvp`` contracts it with the input tangents.

Only the blocks asked for are integrated, so a solve differentiated with
respect to ``params`` alone carries no initial-state columns. A 1-D ``y0`` or
``params`` broadcast over the ensemble is expanded before the Function, so its
gradient is the sum over trajectories, as for a shared argument in the
reference (``_match_shape``).
"""

from __future__ import annotations

from dataclasses import dataclass

from . import _common
from ._sensitivity import (
    SensitivitySpec,
    augmented_error_weights,
    augmented_y0,
    split_augmented,
)


def _has_tangent(x) -> bool:
    torch = _common._torch()
    if not isinstance(x, torch.Tensor):
        return False
    if x.requires_grad:
        return True
    try:
        from torch.autograd import forward_ad

        return forward_ad.unpack_dual(x).tangent is not None
    except RuntimeError:  # no forward-AD level is active
        return False


def wanted(y0, params):
    """Which of ``y0`` and ``params`` a derivative is asked of."""
    return _has_tangent(y0), _has_tangent(params)


@dataclass
class JointJob:
    """Everything a joint solve needs besides the two differentiable inputs."""

    build: object          # spec -> (Built, needs weights) for the joint system
    times: object
    n: int
    n_vars: int
    n_params: int
    weights: object        # user error weights (N, n_vars) tensor or None
    launch_kw: dict
    spec: SensitivitySpec = None
    stats: tuple = None


def joint_solve(job: JointJob, y0_t, p_t):
    """Integrate the joint system for ``job.spec``; returns the launch outputs
    and ``(state (N, n_save, n), S (N, n_save, n, n_sens))``."""
    import torch

    from .rodas5P import launch

    spec = job.spec
    y0_aug = augmented_y0(y0_t, spec)
    need_w = job.weights is not None or not spec.error_control
    built = job.build(spec, not need_w)
    w_aug = None
    if need_w:
        base = job.weights if job.weights is not None else torch.ones(
            (job.n, job.n_vars), dtype=torch.float64, device=y0_t.device)
        w_aug = _common.soa(augmented_error_weights(base, spec))
    out = launch(built, y0_aug, job.times, p_t, job.n, weights=w_aug, **job.launch_kw)
    state, sens = split_augmented(out.ensemble_view(), spec)
    return out, state, sens


def _function():
    import torch

    class SensitivitySolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, y0_t, p_t, job):
            out, state, sens = joint_solve(job, y0_t.detach(), p_t.detach())
            state = state.contiguous()
            sens = sens.contiguous()
            ctx.sens = sens
            ctx.spec = job.spec
            ctx.n_params = job.n_params
            acc, rej, loop = out.accepted, out.rejected, out.loop
            ctx.mark_non_differentiable(acc, rej, loop)
            return state, acc, rej, loop

        @staticmethod
        def backward(ctx, g_state, *_):
            spec, sens = ctx.spec, ctx.sens
            gdir = torch.einsum("nsi,nsik->nk", g_state.to(sens.dtype), sens)
            gy0 = gdir[:, : spec.n_y0_dirs] if spec.wrt_y0 else None
            gp = None
            if spec.wrt_params:
                gp = gdir.new_zeros((gdir.shape[0], ctx.n_params))
                gp[:, list(spec.param_seed_columns)] = gdir[:, spec.n_y0_dirs:]
            return gy0, gp, None

        @staticmethod
        def jvp(ctx, dy0, dp, _):
            spec, sens = ctx.spec, ctx.sens
            tangent = torch.zeros(sens.shape[:3], dtype=sens.dtype, device=sens.device)
            if spec.wrt_y0 and dy0 is not None:
                tangent += torch.einsum("nsij,nj->nsi", sens[..., : spec.n_y0_dirs],
                                        dy0.to(sens.dtype))
            if spec.wrt_params and dp is not None:
                cols = list(spec.param_seed_columns)
                tangent += torch.einsum("nsik,nk->nsi", sens[..., spec.n_y0_dirs:],
                                        dp.to(sens.dtype)[:, cols])
            return tangent, None, None, None

    return SensitivitySolve


_FN = None


def differentiable_solve(job: JointJob, y0_t, p_t, wrt_y0, wrt_params, error_control,
                         param_columns):
    """The solve as a differentiable function of ``y0_t`` and ``p_t`` (both
    ``(N, ...)``, possibly expanded views of the caller's tensors)."""
    global _FN
    if _FN is None:
        _FN = _function()
    job.spec = SensitivitySpec(job.n_vars, job.n_params, bool(wrt_y0), bool(wrt_params),
                               bool(error_control),
                               None if param_columns is None else tuple(param_columns))
    state, acc, rej, loop = _FN.apply(y0_t, p_t, job)
    return state, (acc, rej, loop)
