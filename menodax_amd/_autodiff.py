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


# --- the discrete adjoint -------------------------------------------------------
#: accepted steps between checkpoints of the adjoint's forward pass
DEFAULT_CKPT_EVERY = 64


@dataclass
class AdjointJob:
    """A plain solve and the adjoint kernel that differentiates it."""

    built: object          # forward kernels (Built)
    adjoint: object        # adjoint kernel (Built)
    times: object
    n: int
    n_vars: int
    n_params: int
    weights: object        # (n_vars, N) error weights or None
    launch_kw: dict
    ckpt_every: int = DEFAULT_CKPT_EVERY


#: the most device memory a checkpoint store sized without the solve's step
#: counts may take: 8 GiB, and at most a quarter of what is free
CHECKPOINT_STORE_BUDGET = 8 << 30


def _checkpoint_store_guess(n, n_vars, ckpt_every, max_steps) -> int:
    """Checkpoints to provide when the solve's accepted-step counts are not
    known: enough for ``max_steps`` accepted steps (then no trajectory can
    outrun the store), or what ``CHECKPOINT_STORE_BUDGET`` holds."""
    import torch

    from . import _runtime

    bound = max(int(max_steps), 0) // ckpt_every + 1
    if n == 0:
        return 1
    one = _runtime.adjoint_workspace_bytes(n, n_vars, ckpt_every, 1)
    per = _runtime.adjoint_workspace_bytes(n, n_vars, ckpt_every, 2) - one
    free, _ = torch.cuda.mem_get_info()
    budget = min(CHECKPOINT_STORE_BUDGET, free // 4)
    return max(1, min(bound, 1 + (budget - one) // per))


def adjoint_gradients(job: AdjointJob, y0_t, p_t, gbar, accepted=None):
    """dL/dy0 ``(N, n_vars)`` and dL/dparams ``(N, n_params)`` for the cotangent
    ``gbar`` ``(N, n_save, n_vars)`` of the solve's history. ``accepted`` are
    that solve's per-trajectory accepted-step counts, which size the
    checkpoint store; without them the store is sized by
    ``_checkpoint_store_guess`` and the adjoint kernel's own forward pass is
    the only one, unless a trajectory outruns that store: then a forward
    solve counts the steps and the adjoint runs again (the gradients do not
    depend on the store's size)."""
    c = int(job.ckpt_every)
    if c < 1:
        raise ValueError("ckpt_every must be positive")
    if accepted is not None:
        return _adjoint_run(job, y0_t, p_t, gbar, int(accepted.max().item()) // c + 1 if job.n
                            else 1, must_fit=True)
    guess = _checkpoint_store_guess(job.n, job.n_vars, c, job.launch_kw["max_steps"])
    out = _adjoint_run(job, y0_t, p_t, gbar, guess, must_fit=False)
    if out is not None:
        return out
    from .rodas5P import launch

    fwd = launch(job.built, y0_t, job.times, p_t, job.n, weights=job.weights, **job.launch_kw)
    return _adjoint_run(job, y0_t, p_t, gbar, int(fwd.accepted.max().item()) // c + 1,
                        must_fit=True)


def _adjoint_run(job: AdjointJob, y0_t, p_t, gbar, max_ckpt, must_fit):
    """One adjoint launch with room for ``max_ckpt`` checkpoints per
    trajectory; None when a trajectory outran them and ``must_fit`` is false."""
    import torch

    from . import _runtime

    dev = y0_t.device
    n, nv, np_ = job.n, job.n_vars, job.n_params
    c = int(job.ckpt_every)
    ws_bytes = _runtime.adjoint_workspace_bytes(n, nv, c, max_ckpt)
    ws = torch.empty(max(ws_bytes, 8), dtype=torch.uint8, device=dev)
    y0_s = _common.soa(y0_t)
    p_s = _common.soa(p_t) if np_ else torch.zeros((1, max(n, 1)), dtype=torch.float64, device=dev)
    g_s = gbar.to(device=dev, dtype=torch.float64).permute(1, 2, 0).contiguous()
    gy = torch.empty((nv, n), dtype=torch.float64, device=dev)
    gp = torch.empty((max(np_, 1), n), dtype=torch.float64, device=dev)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    kw = job.launch_kw
    stream = kw.get("stream")
    s = torch.cuda.current_stream(dev) if stream is None else stream
    _runtime.adjoint_launch(
        job.adjoint.plain, s.cuda_stream, n, nv, np_, job.times.numel(), _common.LAYOUT_SOA,
        y0_s.data_ptr(), job.times.data_ptr(), p_s.data_ptr(), kw["dt0"], kw["rtol"], kw["atol"],
        kw["max_steps"], 0 if job.weights is None else job.weights.data_ptr(), g_s.data_ptr(),
        gy.data_ptr(), gp.data_ptr(), status.data_ptr(), c, max_ckpt, ws.data_ptr(), ws_bytes)
    bad = int((status != 0).sum().item())
    if bad:
        if not must_fit:
            return None
        raise RuntimeError(f"adjoint: {bad} trajectories outran their checkpoint store; "
                           "their forward pass took more accepted steps than the primal solve")
    return gy.t(), gp[:np_].t()


def _adjoint_function():
    import torch

    from .rodas5P import launch

    class AdjointSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, y0_t, p_t, job):
            y0_d, p_d = y0_t.detach(), p_t.detach()
            out = launch(job.built, y0_d, job.times, p_d, job.n, weights=job.weights,
                         **job.launch_kw)
            ctx.job = job
            ctx.save_for_backward(y0_d, p_d, out.accepted)
            hist = out.ensemble_view()
            acc, rej, loop = out.accepted, out.rejected, out.loop
            ctx.mark_non_differentiable(acc, rej, loop)
            return hist, acc, rej, loop

        @staticmethod
        def backward(ctx, g_hist, *_):
            y0_d, p_d, accepted = ctx.saved_tensors
            gy, gp = adjoint_gradients(ctx.job, y0_d, p_d, g_hist, accepted)
            need_y0, need_p = ctx.needs_input_grad[:2]
            return (gy if need_y0 else None), (gp if need_p else None), None

    return AdjointSolve


_ADJ_FN = None


def adjoint_solve(job: AdjointJob, y0_t, p_t):
    """The solve as a differentiable function of ``y0_t``/``p_t`` whose
    reverse mode runs the discrete-adjoint kernel."""
    global _ADJ_FN
    if _ADJ_FN is None:
        _ADJ_FN = _adjoint_function()
    hist, acc, rej, loop = _ADJ_FN.apply(y0_t, p_t, job)
    return hist, (acc, rej, loop)


# --- vmap -------------------------------------------------------------------------
def under_vmap(*xs) -> bool:
    """Whether any argument is a tensor batched by ``torch.func.vmap``."""
    import torch

    is_batched = getattr(torch._C._functorch, "is_batchedtensor", None)
    return is_batched is not None and any(
        isinstance(x, torch.Tensor) and is_batched(x) for x in xs)


def _vmap_function():
    import torch

    class VmapSolve(torch.autograd.Function):
        """A solve whose ``torch.func.vmap`` over single-trajectory calls is one
        ensemble launch (modax/_jax_common.py:49-111): the batched ``y0``
        and/or ``params`` become the ensemble's trajectories, an unbatched one
        is broadcast, a batched ``t_span`` is taken as one shared grid (its
        first row). Each call's result keeps its own leading axis of 1."""

        generate_vmap_rule = False

        @staticmethod
        def forward(y0, t_span, params, run):
            return run(y0, t_span, params)

        @staticmethod
        def setup_context(ctx, inputs, output):
            pass

        @staticmethod
        def vmap(info, in_dims, y0, t_span, params, run):
            b = info.batch_size
            y_dim, t_dim, p_dim = in_dims[:3]

            def ensemble(x, dim, name):
                if dim is None:
                    if x.ndim != 1:
                        raise NotImplementedError(
                            f"vmap with an unbatched ensemble-shaped {name} is not supported; "
                            "call the solver with batched y0/params directly instead")
                    return x.expand(b, x.shape[0])
                x = x.movedim(dim, 0)
                if x.ndim != 2:
                    raise NotImplementedError(
                        f"vmap over an already-ensembled {name} is not supported; "
                        "call the solver with batched y0/params directly instead")
                return x

            if t_dim is not None:
                t_span = t_span.movedim(t_dim, 0)[0]
            out = run(ensemble(y0, y_dim, "y0"), t_span, ensemble(params, p_dim, "params"))
            # one call's result is (1, n_save, n_vars); stats (1,)
            return tuple(o[:, None] for o in out), tuple(0 for _ in out)

    return VmapSolve


_VMAP_FN = None


def vmap_solve(run, y0, t_span, params):
    """Apply ``run(y0 (N, n), t_span, params (N, n_p)) -> tuple of outputs``
    through the vmap-aware Function."""
    global _VMAP_FN
    if _VMAP_FN is None:
        _VMAP_FN = _vmap_function()
    return _VMAP_FN.apply(y0, t_span, params, run)
