"""Repeated solves as one captured HIP graph.

A caller that solves the same system many times -- the same shapes and
options, new initial states or parameters each time (a sampler's likelihood,
a parameter sweep) -- pays the host-side setup of every ``solve`` call once
here: the kernel is built and its device buffers allocated at construction,
the launch (the refill counter's reset kernel and the solve kernel) is captured
into a graph on the first ``run``, and every ``run`` after that copies the new
inputs into the captured buffers and replays the graph.

Nothing in the captured launch allocates or calls ``hipMemsetAsync``: the
refill counter is cleared by a kernel, so replays see a zeroed counter.
Replays give the same bits as the eager solve (tests/test_gpu_graph.py).
"""

from __future__ import annotations

from dataclasses import replace

from . import _common, _kernels
from .rodas5P import _tableau, _wave, launch
from ._kernels import KernelOptions, trajectories_per_block_or_default
from ._sparsity import normalize_sparsity


class CapturedSolve:
    """``CapturedSolve(ode_fn, y0, t_span, params, **options)`` then
    ``run(y0=None, params=None)`` -> the ``(N, n_save, n_vars)`` history (a view
    of the captured output buffer: copy it before the next ``run`` if it must
    be kept). Options are ``rodas5P.solve``'s launch options."""

    def __init__(self, ode_fn, y0, t_span, params, *, rtol=1e-8, atol=1e-10, first_step=None,
                 max_steps=100000, pcoeff=0.0, icoeff=1.0, dcoeff=0.0, lu_precision="fp32",
                 trajectories_per_block=None, sparsity=None, ordering="amd", tf_index=None,
                 max_registers=None, refill=None, placement=None, work_order=None,
                 tableau_in_registers=None):
        import torch

        self._torch = torch
        dev = _common.device()
        y0_t, p_t, n, n_vars = _common.normalize_y0_params(y0, params, dev)
        if sparsity is not None:
            sparsity = normalize_sparsity(sparsity, n_vars)
        options = KernelOptions(
            pcoeff=float(pcoeff), icoeff=float(icoeff), dcoeff=float(dcoeff),
            lu_precision=lu_precision,
            trajectories_per_block=trajectories_per_block_or_default(trajectories_per_block),
            sparsity=sparsity, ordering=ordering,
            tf_index=-1 if tf_index is None else int(tf_index), max_registers=max_registers,
            placement=placement,
            tableau_in_registers=_tableau(tableau_in_registers, n_vars, n, placement,
                                          max_registers))
        if work_order is None and not refill:
            options = replace(options, placement=_wave(options, n_vars, n))
        self.n, self.n_vars, self.n_params = n, n_vars, p_t.shape[1]
        self.built = _kernels.rodas5p_kernel(ode_fn, n_vars, self.n_params, options)
        # static, structure-of-arrays inputs the graph reads
        self._y0 = _common.soa(y0_t.detach()).clone()
        self._p = _common.soa(p_t.detach()).clone()
        self._times = _common.as_f64_device(t_span, dev).contiguous().clone()
        self._kw = dict(dt0=_common.initial_step(first_step), rtol=rtol, atol=atol,
                        max_steps=max_steps, refill=refill)
        if isinstance(work_order, str):
            raise ValueError("a captured solve takes an explicit work_order (a permutation); "
                             "a pilot's device sort is not captured")
        if work_order is not None:
            # the claim order lives in a buffer the graph reads: set_work_order
            # replaces its contents between replays
            self._kw["order"] = _common.work_order(work_order, n, dev).clone()
        # eager launch: loads the module, allocates outputs and workspace
        self._out = launch(self.built, self._y0.t(), self._times, self._p.t(), n, **self._kw)
        self._ws = self._out.workspace
        self._graph = None
        self._as_numpy = _common.wants_numpy(y0, params, t_span)

    def _capture(self):
        torch = self._torch
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(stream):
            with torch.cuda.graph(graph, stream=stream):
                launch(self.built, self._y0.t(), self._times, self._p.t(), self.n,
                       outputs=self._out, workspace=self._ws, stream=stream, **self._kw)
        torch.cuda.current_stream().wait_stream(stream)
        self._graph = graph

    def run(self, y0=None, params=None, return_stats=False):
        torch = self._torch
        for new, buf in ((y0, self._y0), (params, self._p)):
            if new is None:
                continue
            x = _common.as_f64_device(new, buf.device)
            if x.ndim == 1:
                x = x.expand(self.n, x.shape[0])
            if tuple(x.shape) != (self.n, buf.shape[0]):
                raise ValueError(f"expected shape {(self.n, buf.shape[0])}, got {tuple(x.shape)}")
            buf.copy_(x.t())
        if self._graph is None:
            self._capture()
        self._graph.replay()
        hist = self._out.ensemble_view()
        stats = _common.solver_stats(self._out.accepted, self._out.rejected, self._out.loop)
        if self._as_numpy:
            torch.cuda.synchronize()
            hist = hist.cpu().numpy()
            stats = {k: v.cpu().numpy() for k, v in stats.items()}
        return (hist, stats) if return_stats else hist

    def set_work_order(self, order):
        """Replace the claim order of a solve built with ``work_order`` (a
        permutation of the trajectories, e.g. ``_common.cost_order`` of the
        last run's step counts); takes effect from the next ``run``."""
        if "order" not in self._kw:
            raise ValueError("this CapturedSolve was built without a work_order")
        buf = self._kw["order"]
        buf.copy_(_common.work_order(order, self.n, buf.device))

    @property
    def graph(self):
        return self._graph

