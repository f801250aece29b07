"""The ensemble solve as torch custom operators, so that a solve inside
``torch.compile(..., fullgraph=True)`` is one opaque node of the traced graph.

This is the role the reference's JAX primitive plays under ``jax.jit``
(modax/_jax_numba_custom_call.py:66-371: a primitive with a shape-only
abstract evaluation whose lowering is the kernel launch, and the derivative
rule registered on it, modax/rodas5P.py:1204-1237):

- ``menodax::ensemble_solve`` launches a registered solver on
  ``(y0, t_span, params)`` and returns the structure-of-arrays history
  ``(n_save, n_vars, N)`` and the accepted / rejected / loop counters; its fake
  implementation only allocates those shapes;
- ``menodax::ensemble_adjoint`` runs the discrete-adjoint kernel for a
  cotangent of that history and returns the gradients in ``y0`` and
  ``params``; it is the registered backward of ``ensemble_solve``, so a
  compiled graph differentiates a solve through the adjoint.

A solver -- method, callback and the options that select a kernel -- is
registered once, at trace time, by ``solver_key``, which the compiler
evaluates as a constant; the operators build or reuse the kernels for the
shapes they are given (the size rules that depend on the ensemble size are
applied there, as in an eager solve).
"""

from __future__ import annotations

from dataclasses import dataclass, replace

import torch

from . import _autodiff, _common, _kernels

_SOLVERS: list = []
_INDEX: dict = {}


@dataclass(frozen=True)
class _Solver:
    method: str            # "rodas5p" or "tsit5"
    ode_fn: object
    n_vars: int
    n_params: int
    options: object        # KernelOptions or Tsit5Options, tableau choice unresolved
    tableau: object        # the caller's tableau_in_registers (None: size rule)
    ckpt_every: int
    block: object = None   # the caller's trajectories_per_block (None: size rules)

    def forward(self, n, wave_ok=False):
        """The forward kernels for an ensemble of ``n``; ``wave_ok`` lets the
        size rule pick the wave placement (a plain launch, no refill)."""
        from .rodas5P import _tableau

        opts = replace(self.options, tableau_in_registers=_tableau(
            self.tableau, self.n_vars, n, getattr(self.options, "placement", None),
            getattr(self.options, "max_registers", None)))
        if self.method == "rodas5p":
            if not wave_ok:
                return _kernels.rodas5p_kernel(self.ode_fn, self.n_vars, self.n_params, opts)
            from .rodas5P import _wave

            opts = replace(opts, placement=_wave(opts, self.n_vars, n))
            return _kernels.rodas5p_kernel(self.ode_fn, self.n_vars, self.n_params, opts)
        return _kernels.tsit5_kernel(self.ode_fn, self.n_vars, self.n_params, opts)

    def adjoint(self, n):
        """The adjoint kernel for an ensemble of ``n``."""
        if self.method == "rodas5p":
            from .rodas5P import _adjoint_options

            return _kernels.rodas5p_adjoint_kernel(
                self.ode_fn, self.n_vars, self.n_params,
                _adjoint_options(self.options, self.n_vars, n, self.block))
        return _kernels.tsit5_adjoint_kernel(self.ode_fn, self.n_vars, self.n_params,
                                             self.options)


@torch.compiler.assume_constant_result
def solver_key(method, ode_fn, n_vars, n_params, sparsity, options, tableau, ckpt_every):
    """The registry index of a solver; ``options`` is a tuple of
    ``(name, value)`` pairs of the kernel options (without ``sparsity``, which
    is normalised here). Evaluated once per trace and kept as a constant."""
    from ._sparsity import normalize_sparsity

    kw = dict(options)
    if sparsity is not None:
        kw["sparsity"] = normalize_sparsity(sparsity, int(n_vars))
    block = kw.get("trajectories_per_block")
    kw["trajectories_per_block"] = _kernels.trajectories_per_block_or_default(block)
    opts = (_kernels.KernelOptions if method == "rodas5p" else _kernels.Tsit5Options)(**kw)
    solver = _Solver(method, ode_fn, int(n_vars), int(n_params), opts, tableau,
                     int(ckpt_every), block)
    key = _INDEX.get(solver)
    if key is None:
        key = len(_SOLVERS)
        _SOLVERS.append(solver)
        _INDEX[solver] = key
    return key


def _refill(code):
    return None if code < 0 else bool(code)


@torch.library.custom_op("menodax::ensemble_solve", mutates_args=(), device_types="cuda")
def ensemble_solve(y0: torch.Tensor, t_span: torch.Tensor, params: torch.Tensor, solver: int,
                   dt0: float, rtol: float, atol: float, max_steps: int,
                   refill: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """One ensemble launch of the registered ``solver``: ``y0`` ``(N, n_vars)``,
    ``params`` ``(N, n_params)``, float64 on the GPU; ``refill`` -1 decides by
    size, 0 / 1 forces the launch shape."""
    from .rodas5P import launch

    s = _SOLVERS[solver]
    n = y0.shape[0]
    out = launch(s.forward(n, wave_ok=not refill > 0), y0, t_span.contiguous(), params, n,
                 dt0=dt0, rtol=rtol, atol=atol, max_steps=max_steps, refill=_refill(refill))
    return out.hist, out.accepted, out.rejected, out.loop


@ensemble_solve.register_fake
def _(y0, t_span, params, solver, dt0, rtol, atol, max_steps, refill):
    n, n_vars = y0.shape
    counter = y0.new_empty((n,), dtype=torch.int32)
    return (y0.new_empty((t_span.shape[0], n_vars, n)), counter, torch.empty_like(counter),
            torch.empty_like(counter))


@torch.library.custom_op("menodax::ensemble_adjoint", mutates_args=(), device_types="cuda")
def ensemble_adjoint(g_hist: torch.Tensor, y0: torch.Tensor, t_span: torch.Tensor,
                     params: torch.Tensor, accepted: torch.Tensor, solver: int, dt0: float,
                     rtol: float, atol: float, max_steps: int
                     ) -> tuple[torch.Tensor, torch.Tensor]:
    """Gradients in ``y0`` and ``params`` of <g_hist, history> for the solve
    ``ensemble_solve`` ran; ``g_hist`` is ``(n_save, n_vars, N)`` and
    ``accepted`` that solve's accepted-step counts."""
    s = _SOLVERS[solver]
    n = y0.shape[0]
    job = _autodiff.AdjointJob(
        built=s.forward(n), adjoint=s.adjoint(n), times=t_span.contiguous(), n=n,
        n_vars=s.n_vars, n_params=s.n_params, weights=None,
        launch_kw=dict(dt0=dt0, rtol=rtol, atol=atol, max_steps=max_steps, refill=None,
                       stream=None, order=None),
        ckpt_every=s.ckpt_every)
    gy, gp = _autodiff.adjoint_gradients(job, y0, params, g_hist.permute(2, 0, 1), accepted)
    return gy.contiguous(), gp.contiguous()


@ensemble_adjoint.register_fake
def _(g_hist, y0, t_span, params, accepted, solver, dt0, rtol, atol, max_steps):
    return torch.empty_like(y0), torch.empty_like(params)


def _setup(ctx, inputs, output):
    y0, t_span, params, solver, dt0, rtol, atol, max_steps, _ = inputs
    ctx.save_for_backward(y0, t_span, params, output[1])
    ctx.cfg = (solver, dt0, rtol, atol, max_steps)


def _backward(ctx, g_hist, *_):
    y0, t_span, params, accepted = ctx.saved_tensors
    if g_hist is None:
        return (None,) * 9
    gy, gp = torch.ops.menodax.ensemble_adjoint(g_hist.contiguous(), y0, t_span, params,
                                                accepted, *ctx.cfg)
    need_y0, _, need_p = ctx.needs_input_grad[:3]
    return ((gy if need_y0 else None), None, (gp if need_p else None)) + (None,) * 6


ensemble_solve.register_autograd(_backward, setup_context=_setup)


def _broadcast(y0, params):
    """``normalize_y0_params`` in torch operations the compiler traces."""
    y0 = y0.to(torch.float64)
    params = params.to(torch.float64)
    if y0.ndim not in (1, 2) or params.ndim not in (1, 2):
        raise ValueError("y0 must have shape (n_vars,) or (N, n_vars) and params shape "
                         "(n_params,) or (N, n_params)")
    if y0.ndim == 2 and params.ndim == 2 and params.shape[0] != y0.shape[0]:
        raise ValueError("params must have shape (n_params,) or (N, n_params) when y0 has "
                         "shape (N, n_vars)")
    n = y0.shape[0] if y0.ndim == 2 else (params.shape[0] if params.ndim == 2 else 1)
    if y0.ndim == 1:
        y0 = y0.expand(n, y0.shape[0])
    if params.ndim == 1:
        params = params.expand(n, params.shape[0])
    return y0, params


def traced_solve(method, ode_fn, y0, t_span, params, *, options, sparsity, tableau, rtol, atol,
                 first_step, max_steps, refill, adjoint, ckpt_every, return_stats):
    """The solve as the compiler traces it (called while compiling only):
    the history ``(N, n_save, n_vars)`` and, with ``return_stats``, the
    counters; gradients in ``y0``/``params`` come from the adjoint kernel.
    Returns NotImplemented when the call needs the eager path (the caller
    then runs it eagerly, a graph break)."""
    if not all(isinstance(x, torch.Tensor) for x in (y0, t_span, params)):
        return NotImplemented
    if (y0.requires_grad or params.requires_grad) and not adjoint:
        # the eager default differentiates through forward sensitivities
        return NotImplemented
    y0, params = _broadcast(y0, params)
    key = solver_key(method, ode_fn, y0.shape[1], params.shape[1], sparsity, options, tableau,
                     ckpt_every)
    hist, acc, rej, loop = torch.ops.menodax.ensemble_solve(
        y0, t_span.to(torch.float64), params, key, _common.initial_step(first_step),
        float(rtol), float(atol), int(max_steps), -1 if refill is None else int(bool(refill)))
    hist = hist.permute(2, 0, 1)
    if return_stats:
        return hist, _common.solver_stats(acc, rej, loop)
    return hist
