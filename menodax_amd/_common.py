"""Host-side helpers shared by the two solvers: argument normalisation, the
structure-of-arrays device layout, outputs and stats.

Restates modax/_jax_common.py:12-46 (``normalize_y0_params``),
modax/_numba_common.py:55-81 (error weights, first-step sentinel) and
modax/_numba_common.py:133-138 (the stats dict) for torch tensors on the GPU.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

LAYOUT_TRAJECTORY_MAJOR = 0
LAYOUT_SOA = 1


def _torch():
    import torch

    return torch


def eager_only(fn):
    """``fn`` run eagerly when a ``torch.compile``d caller reaches it: the
    compiler stops its graph at the call (a graph break) instead of tracing
    into the launch, whose stream handles and device pointers exist only at
    run time. Results are unchanged; ``fullgraph=True`` callers get the
    compiler's error for a graph break."""
    try:
        import torch

        return torch.compiler.disable(fn, recursive=True)
    except (ImportError, AttributeError):
        return fn


def compilable(eager, traced):
    """``eager`` as the public entry: it runs eagerly, except while
    ``torch.compile`` traces a call that ``traced`` expresses with the custom
    operators of ``_ops`` (``traced`` returns NotImplemented otherwise, and the
    compiler stops its graph at the eager call: a graph break)."""
    import functools

    import torch

    disabled = eager_only(eager)

    @functools.wraps(eager)
    def entry(*args, **kwargs):
        if torch.compiler.is_compiling():
            out = traced(*args, **kwargs)
            if out is not NotImplemented:
                return out
        return disabled(*args, **kwargs)

    return entry


def device():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError(
            "menodax_amd needs a ROCm GPU: torch.cuda.is_available() is False "
            "(there is no CPU fallback)"
        )
    return torch.device("cuda", torch.cuda.current_device())


def n_cu():
    """Compute units of the current device."""
    torch = _torch()
    return torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count


def as_f64_device(x, dev):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


def normalize_y0_params(y0, params, dev):
    """Broadcast ``y0``/``params`` to a common ``(N, ...)`` ensemble layout
    (modax/_jax_common.py:12-46): either may be 1-D, a 2-D one fixes N, two 2-D
    ones must agree."""
    y0_arr = as_f64_device(y0, dev)
    p_arr = as_f64_device(params, dev)
    if y0_arr.ndim not in (1, 2) or p_arr.ndim not in (1, 2):
        raise ValueError(
            "y0 must have shape (n_vars,) or (N, n_vars) and params shape "
            f"(n_params,) or (N, n_params); got y0.shape={tuple(y0_arr.shape)} and "
            f"params.shape={tuple(p_arr.shape)}"
        )
    if y0_arr.ndim == 2:
        n = y0_arr.shape[0]
        if p_arr.ndim == 2 and p_arr.shape[0] != n:
            raise ValueError(
                "params must have shape (n_params,) or (N, n_params) when y0 has "
                f"shape (N, n_vars); got y0.shape={tuple(y0_arr.shape)} and "
                f"params.shape={tuple(p_arr.shape)}"
            )
    elif p_arr.ndim == 2:
        n = p_arr.shape[0]
    else:
        n = 1
    if y0_arr.ndim == 1:
        y0_arr = y0_arr.expand(n, y0_arr.shape[0])
    if p_arr.ndim == 1:
        p_arr = p_arr.expand(n, p_arr.shape[0])
    return y0_arr, p_arr, n, y0_arr.shape[1]


def soa(x):
    """``(N, k)`` -> a contiguous ``(k, N)`` tensor (no copy when ``x`` is
    already the transpose of one)."""
    return x.t().contiguous()


def initial_step(first_step):
    """The ``dt0`` scalar: non-positive means 1e-6 of the window, decided in the
    kernel (modax/_numba_common.py:72-81)."""
    return 0.0 if first_step is None else float(first_step)


def error_weights_soa(error_weights, n, n_vars, dev):
    """``None`` means unit weights (the kernel is then built not to read any);
    otherwise a ``(n_vars,)`` or ``(N, n_vars)`` array, returned as
    ``(n_vars, N)`` (modax/_numba_common.py:55-69)."""
    if error_weights is None:
        return None
    w = as_f64_device(error_weights, dev)
    if w.ndim == 1:
        w = w.expand(n, n_vars)
    if tuple(w.shape) != (n, n_vars):
        raise ValueError(f"error_weights must be (n_vars,) or (N, n_vars); got {tuple(w.shape)}")
    return soa(w)


def solver_stats(accepted, rejected, loop_steps):
    """modax/_numba_common.py:133-138."""
    return {"accepted_steps": accepted, "rejected_steps": rejected, "loop_steps": loop_steps}


def wants_numpy(*xs):
    torch = _torch()
    return not any(isinstance(x, torch.Tensor) for x in xs)


@dataclass
class Outputs:
    hist: object  # (n_save, n_vars, N) storage
    accepted: object
    rejected: object
    loop: object
    hook_out: object = None  # (hook_size, N) storage when a save hook runs
    #: the launch's workspace (a persistent kernel's refill counter), for a
    #: caller that launches again into the same outputs
    workspace: object = None

    def ensemble_view(self):
        """``(N, n_save, n_vars)`` view of the structure-of-arrays history."""
        return self.hist.permute(2, 0, 1)


def allocate_outputs(n, n_save, n_vars, dev):
    torch = _torch()
    return Outputs(
        hist=torch.empty((n_save, n_vars, n), dtype=torch.float64, device=dev),
        accepted=torch.empty(n, dtype=torch.int32, device=dev),
        rejected=torch.empty(n, dtype=torch.int32, device=dev),
        loop=torch.empty(n, dtype=torch.int32, device=dev),
    )


def work_order(order, n, dev):
    """A claim order for the persistent launch as an int64 device tensor,
    checked to be a permutation of the trajectory indices 0..n-1 (the kernel
    reads it without bounds beyond skipping out-of-range entries)."""
    import torch

    o = torch.as_tensor(order, dtype=torch.int64, device=dev).reshape(-1).contiguous()
    if o.numel() != n:
        raise ValueError(f"work_order must hold {n} trajectory indices; got {o.numel()}")
    if n and not torch.equal(torch.sort(o).values, torch.arange(n, dtype=torch.int64, device=dev)):
        raise ValueError("work_order must be a permutation of the trajectory indices 0..N-1")
    return o


def work_order_arg(value, n, dev):
    """``solve``'s ``work_order`` as ``launch`` takes it: None, ``"pilot"``, or
    a checked permutation tensor."""
    if value is None or (isinstance(value, str) and value == "pilot"):
        return value
    if isinstance(value, str):
        raise ValueError(f"work_order must be a permutation or 'pilot', got {value!r}")
    return work_order(value, n, dev)


def cost_order(attempted):
    """The trajectories by decreasing cost (attempted steps of an earlier
    solve, e.g. ``stats["accepted_steps"] + stats["rejected_steps"]``): a
    ``work_order`` that claims the longest trajectories first."""
    import numpy as np
    import torch

    if isinstance(attempted, torch.Tensor):
        return torch.argsort(attempted.to(torch.int64), descending=True, stable=True)
    return np.argsort(-np.asarray(attempted, dtype=np.int64), kind="stable")
