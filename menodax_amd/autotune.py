"""Autotuning the Rodas5P launch: placement, workgroup size, launch shape.

The placements (registers / rolled scratch / rolled LDS, and for a dense
matrix above the unrolled size one trajectory per 64-lane workgroup), for the
unrolled step of a small system whether the tableau is held in registers, the
trajectories per workgroup and the two launch shapes (one lane per trajectory, persistent
refill) move data and work, never arithmetic, so every candidate computes the
same bits. Which is fastest depends on the system size, the register and LDS
footprint and the spread of step counts, and is measured rather than guessed:
``tune`` times each candidate on a sample of the ensemble, checks that they all
agree bit for bit, and returns the winner's options, cached per (callback,
sizes, options) for the process.

A sample is a different regime from the whole ensemble for the launch shape: a
persistent grid sized for the GPU would give every sampled trajectory a lane of
its own. So when timing a sample of k out of N trajectories, the launch shape
is the one ``solve`` picks for N (persistent refill once N exceeds the lanes
the persistent grid keeps resident), and a persistent launch runs on its
resident grid scaled by k/N, each lane working through as many trajectories as
it would in the whole solve. That keeps the work per lane, not the number of
workgroups sharing a compute unit: a sample times a placement with several
resident workgroups per CU without their contention, so it ranks such
placements optimistically. With no sampling both shapes are timed, or with
``shapes="solve"`` only the one ``solve`` picks -- exact, at the cost of a
whole-ensemble launch per candidate.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, replace

from . import _common, _kernels
from ._kernels import PLACEMENTS, KernelOptions, rodas5p_placement
from ._sparsity import normalize_sparsity
from . import _runtime
from .rodas5P import _use_refill, launch

_CACHE: dict = {}


@dataclass(frozen=True)
class Choice:
    placement: str
    trajectories_per_block: int
    refill: bool
    seconds: float
    tableau_in_registers: bool = False


def candidates(n_vars: int, lu_precision: str, blocks=(64, 128, 256), dense=False,
               lu_refine=False):
    """(placement, block, tableau in registers) worth timing: the unrolled
    register kernel only where it can plausibly fit, with and without the
    tableau in registers where that fits too, the LDS placement only at block
    sizes where something fits in LDS, and for a dense matrix above the
    unrolled size the wave placement (one trajectory per 64-lane workgroup)."""
    out = []
    k_bytes = 8 if lu_refine else None
    if (dense and not lu_refine
            and _kernels.DENSE_UNROLL_MAX < n_vars <= _kernels.WAVE_MAX_VARS):
        out.append((_kernels.WAVE, 64, False))
    for placement in PLACEMENTS:
        if placement == "registers" and n_vars > 32:
            continue
        tabs = (False, True) if (placement == "registers"
                                 and n_vars <= _kernels.TABLEAU_REGS_MAX_VARS) else (False,)
        for b in blocks:
            if placement == "lds" and rodas5p_placement(n_vars, lu_precision, b, "lds",
                                                        k_bytes=k_bytes)["lds_bytes"] == 0:
                continue
            out.extend((placement, b, tab) for tab in tabs)
    return out


def tune(ode_fn, y0, t_span, params, *, sample=None, reps=2, rtol=1e-8, atol=1e-10,
         first_step=None, max_steps=100000, lu_precision="fp32", sparsity=None, ordering="amd",
         pcoeff=0.0, icoeff=1.0, dcoeff=0.0, tf_index=None, blocks=(64, 128, 256), shapes=None,
         lu_refine=False):
    """Time every candidate on ``sample`` trajectories (all of them by default;
    see the module docstring for the launch shape of a sample) and return
    ``(best Choice, [Choice for every candidate])``. Raises if two candidates
    disagree in any bit."""
    import numpy as np
    import torch

    dev = _common.device()
    y0_t, p_t, n, n_vars = _common.normalize_y0_params(y0, params, dev)
    n_full = n
    if sample is not None and sample < n:
        y0_t, p_t, n = y0_t[:sample], p_t[:sample], int(sample)
    n_params = p_t.shape[1]
    if sparsity is not None:
        sparsity = normalize_sparsity(sparsity, n_vars)
    base = KernelOptions(pcoeff=float(pcoeff), icoeff=float(icoeff), dcoeff=float(dcoeff),
                         lu_precision=lu_precision, sparsity=sparsity, ordering=ordering,
                         tf_index=-1 if tf_index is None else int(tf_index),
                         lu_refine=bool(lu_refine))
    if shapes not in (None, "solve"):
        raise ValueError(f"shapes must be None or 'solve', got {shapes!r}")
    key = (ode_fn, n_vars, n_params, base, n, n_full, tuple(blocks), shapes)
    if key in _CACHE:
        return _CACHE[key]
    times = _common.as_f64_device(t_span, dev).contiguous()
    kw = dict(dt0=_common.initial_step(first_step), rtol=rtol, atol=atol, max_steps=max_steps)
    results, reference = [], None
    for placement, block, tab in candidates(n_vars, lu_precision, blocks,
                                            dense=sparsity is None, lu_refine=lu_refine):
        built = _kernels.rodas5p_kernel(
            ode_fn, n_vars, n_params,
            replace(base, placement=placement, trajectories_per_block=block,
                    tableau_in_registers=tab))
        launches = [(False, 0), (True, 0)]
        if not built.refill:
            launches = [(False, 0)]  # the wave placement has no persistent launch
        elif n < n_full or shapes == "solve":
            if not _use_refill(built, n_full):
                launches = [(False, 0)]
            elif n < n_full:
                launches = [(True, max(1, round(_runtime.persistent_grid(built.refill)
                                                * n / n_full)))]
            else:
                launches = [(True, 0)]
        for refill, grid in launches:
            out = launch(built, y0_t, times, p_t, n, refill=refill, grid=grid, **kw)
            torch.cuda.synchronize()
            hist = out.hist.cpu().numpy()
            if reference is None:
                reference = hist
            elif not np.array_equal(hist, reference, equal_nan=True):
                raise RuntimeError(f"placement {placement}, block {block}, tableau in registers "
                                   f"{tab}, refill {refill} changed the result; placements "
                                   "must be bit-identical")
            ws = out.workspace
            t0 = time.perf_counter()
            for _ in range(reps):
                launch(built, y0_t, times, p_t, n, refill=refill, grid=grid, outputs=out,
                       workspace=ws, **kw)
            torch.cuda.synchronize()
            results.append(Choice(placement, block, refill, (time.perf_counter() - t0) / reps,
                                  tab))
    best = min(results, key=lambda c: c.seconds)
    _CACHE[key] = (best, results)
    return best, results


def tuned_options(ode_fn, y0, t_span, params, **kw) -> dict:
    """``solve`` keyword arguments for the fastest candidate (see ``tune``)."""
    best, _ = tune(ode_fn, y0, t_span, params, **kw)
    return {"placement": best.placement, "trajectories_per_block": best.trajectories_per_block,
            "refill": best.refill, "tableau_in_registers": best.tableau_in_registers}
