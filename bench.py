#!/usr/bin/env python
"""Ensemble Rodas5P throughput on one MI355X.

Workload (the metric's configuration): the stiff Robertson ensemble, 65,536
trajectories, rates k1 = 0.04 e^u1, k2 = 3e7 e^u2, k3 = 1e4 e^u3 with
u ~ N(0, 0.3^2) per trajectory (seed 1), y0 = (1, 0, 0), t in [0, 1e5] with 50
saves ([0] and 49 log-spaced from 1e-5), rtol 1e-6, atol 1e-8, dense LU,
double-precision state. A "step" of this benchmark is one whole ensemble solve
with inputs resident in HBM; it is timed for the FP32-LU kernel (the headline
``value``) and for the FP64-LU kernel beside it.

The JSON line also carries
  - ``roofline``: the solve kernel's algorithmic FP64/FP32 flops (counted from
    the generated code and the kernel's own step counters) over its average
    duration measured with HIP events on the launch stream, as FP64-equivalent
    TFLOP/s against the FP64 vector peak -- the kernels are VALU bound, with no
    shared contraction for the matrix cores and HBM touched once per input
    and save;
  - ``cpu_baseline``: the C oracle (OpenMP over trajectories, the same
    algorithm and inputs) on the host's cores;
  - ``endpoint_error``: the error of both LU precisions against a tight
    (rtol 1e-12, FP64) oracle solve on a sample of trajectories;
  - ``reference_config``: the reference's own published benchmark inputs
    (100,000 identical Robertson trajectories, benchmarks/stiff_robertson_scaling)
    for comparison with its RTX 4070 SUPER numbers;
  - ``uneven_stiff``: the 100,000-trajectory uneven stiff ensemble (50-variable
    block-arrow system, step counts spread over two decades): the FP32-LU
    persistent-refill solve against the plain FP64 one-lane solve, with the
    endpoint error against FP64 oracle solves;
  - ``adjoint``: gradients of a squared-error loss over 20 saves with respect to
    mu and y0 for 8,192 stiff Van der Pol trajectories (2 variables, and the
    20-variable chain with its sparse LU), FP64, by the discrete-adjoint kernel
    (checkpoint every 64 accepted steps) and by the forward-sensitivity joint
    solve, with their agreement and a central-difference check on a sample.

``python bench.py --gpus N`` with N > 1 exits: the benchmark runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_TRAJ = 65536
RTOL, ATOL = 1e-6, 1e-8
REF_RTX4070_FP32_TRAJ_PER_S = 100000 / 35.97e-3  # stiff_robertson_scaling csv:12
REF_RTX4070_FP64_TRAJ_PER_S = 100000 / 45.59e-3  # csv:23


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "rodas5p_robertson65536_hbm_counters.txt")


def profiled_traffic(n):
    """HBM bytes per launch of the FP32-LU solve kernel, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated); only for the
    ensemble size that profile was taken at."""
    if n != N_TRAJ or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


COUNTER_PROFILE = os.path.join(ROOT, "profiles", "rodas5p_valu_occupancy_lds_counters.txt")


def profiled_counters(n):
    """The FP32-LU solve kernel's hardware-counter figures (executed VALU FLOP/s,
    occupancy, VALU busy, LDS bank conflicts) from the committed rocprofv3
    summary; only for the ensemble size that profile was taken at."""
    if n != N_TRAJ or not os.path.exists(COUNTER_PROFILE):
        return None
    out = {}
    with open(COUNTER_PROFILE) as f:
        for line in f:
            if line.startswith("c2_") and ":" in line:
                key, val = line.split(":", 1)
                out[key[3:]] = float(val)
    if not out:
        return None
    out["source"] = "profiles/rodas5p_valu_occupancy_lds_counters.txt"
    return out


def time_solves(run, steps, warmup, torch):
    """Warm up, then time exactly ``steps`` solves bracketed by synchronisation;
    per-launch HIP events on the launch stream give the kernel's own time."""
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        out = run()
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    event_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return out, wall, float(np.mean(event_ms))


#: the bound the uneven stiff solves are held to: endpoint error against the
#: FP64 oracle
ERROR_TARGET = 1e-6


def error_stats(got, ref):
    """Endpoint error of ``got`` against ``ref`` (both (k, n_vars), the final
    save): per trajectory the largest component error, absolute and relative
    (relative over components with |ref| >= 1e-3, where a relative figure means
    something), summarised over the sample, with the share of trajectories
    whose absolute error is within ERROR_TARGET."""
    d = np.abs(got - ref)
    per_abs = d.max(axis=1)
    big = np.abs(ref) >= 1e-3
    rel = np.where(big, d / np.where(big, np.abs(ref), 1.0), 0.0).max(axis=1)
    return {"sample": int(got.shape[0]),
            "max_abs": float(per_abs.max()), "p99_abs": float(np.quantile(per_abs, 0.99)),
            "median_abs": float(np.median(per_abs)),
            "max_rel": float(rel.max()), "p99_rel": float(np.quantile(rel, 0.99)),
            "share_within_target": float(np.mean(per_abs <= ERROR_TARGET))}


def spread(n, k):
    """``k`` trajectory indices spread evenly over ``0..n-1``."""
    return np.unique(np.linspace(0, n - 1, min(n, k)).astype(np.int64))


def c2_endpoint_errors(hists, y0, params, times, O):
    """C2's endpoint error for each timed kernel: against the FP64-LU oracle at
    the same tolerance over the whole ensemble (the reference's FP64 result;
    the FP64 kernel matches it bit for bit), and against a tight FP64 oracle
    (rtol 1e-12) on 4,096 trajectories spread over it."""
    same, _ = O.rodas5p(O.robertson(), y0, times, params, rtol=RTOL, atol=ATOL,
                        lu_precision="fp64")
    idx = spread(y0.shape[0], 4096)
    tight, _ = O.rodas5p(O.robertson(), y0[idx], times, params[idx], rtol=1e-12, atol=1e-14,
                         lu_precision="fp64")
    out = {"target": f"max abs endpoint error vs the FP64 oracle <= {ERROR_TARGET:g}"}
    for key, hist in hists.items():
        vs_same = error_stats(hist[:, -1], same[:, -1])
        out[key] = {"vs_fp64_oracle_same_tol": vs_same,
                    "vs_tight_oracle": error_stats(hist[idx, -1], tight[:, -1]),
                    "target_met": vs_same["max_abs"] <= ERROR_TARGET}
    return out


UNEVEN_VARIANTS = ("fp32_refill", "fp32_plain", "fp64_refill", "fp64_plain",
                   "fp32_refined_refill")


def uneven_stiff(torch, dev, O, log_fn, n=100000, err_sample=4096, tight_sample=1024,
                 variants=UNEVEN_VARIANTS):
    """The uneven stiff ensemble: dy/dt = A y + 0.01 y*y, 50 variables, A block
    arrow (2x2 rotation blocks plus a dense hub row), eigenvalue real parts
    -LogUniform down to a per-trajectory stiffness ceiling up to 1e5, t in
    [0, 10] with 20 saves, rtol 1e-6 atol 1e-8. Times the FP32-LU refill
    kernel and the plain FP64 kernel (and the other two combinations), and the
    FP32 LU with refined stage solves (FP64 stage vectors) on the refill
    kernel; ``variants`` narrows that to some of them (profiling runs), which
    leaves out the comparison and the endpoint error unless the headline
    variants are among them. The endpoint error of every variant is taken
    against the FP64-LU oracle at the same tolerance on ``err_sample``
    trajectories spread over the ensemble, and against a tight oracle (rtol
    1e-10) on ``tight_sample``."""
    from menodax_amd import _kernels, _sparse_direct as SD, rodas5P, roofline
    from tests import systems as S
    from menodax_amd._sparsity import normalize_sparsity

    n_vars = 50
    fn, mask, (rows, cols, vals, pidx) = S.arrow_system(n_vars)
    n_p = S.arrow_n_params(n_vars)
    sp = normalize_sparsity(mask, n_vars)
    y0, p = S.arrow_ensemble(n, n=n_vars)
    times = np.linspace(0.0, 10.0, 20)
    y0_t = torch.tensor(np.ascontiguousarray(y0.T), device=dev).t()
    p_t = torch.tensor(np.ascontiguousarray(p.T), device=dev).t()
    t_t = torch.tensor(times, device=dev)
    res = {}
    hists = {}

    def timed(key, built, refill):
        outs = rodas5P.launch(built, y0_t, t_t, p_t, n, dt0=0.0, rtol=RTOL, atol=ATOL,
                              max_steps=100000, refill=refill)
        ws = outs.workspace

        def run():
            return rodas5P.launch(built, y0_t, t_t, p_t, n, dt0=0.0, rtol=RTOL, atol=ATOL,
                                  max_steps=100000, refill=refill, outputs=outs, workspace=ws)

        o, wall, ev_ms = time_solves(run, 2, 0, torch)
        ms = wall / 2 * 1e3
        att = (o.accepted + o.rejected).cpu().numpy().astype(np.int64)
        work = roofline.solve_work(built, int(att.sum()), n * (len(times) - 1))
        res[key] = {"ms": ms, "trajectories_per_s": n / (ms / 1e3),
                    "accepted_steps_per_s": float(o.accepted.sum().item()) / (ms / 1e3),
                    "roofline_frac": work.roof_seconds() / (ev_ms / 1e3),
                    "placement": built.extra.get("placement", {}).get("placement"),
                    "block": built.block}
        hists[key] = o.ensemble_view().cpu().numpy()
        log_fn(f"uneven stiff {key}: {ms:.1f} ms")
        return o, outs, ws, att

    att = None
    for lu in ("fp32", "fp64"):
        built = _kernels.rodas5p_kernel(fn, n_vars, n_p, _kernels.KernelOptions(lu_precision=lu,
                                                                                sparsity=sp))
        for refill in (True, False):
            key = f"{lu}_{'refill' if refill else 'plain'}"
            if key not in variants:
                continue
            o, outs, ws, a = timed(key, built, refill)
            if key == "fp32_refill":
                att = a
                # a repeated solve of the same ensemble: the queue hands out the
                # trajectories by decreasing step count of this solve (not the
                # headline figure, which knows nothing in advance)
                from menodax_amd import _common

                order = _common.work_order(_common.cost_order(o.accepted + o.rejected), n, dev)

                def run_ordered(built=built, outs=outs, ws=ws, order=order):
                    return rodas5P.launch(built, y0_t, t_t, p_t, n, dt0=0.0, rtol=RTOL,
                                          atol=ATOL, max_steps=100000, refill=True,
                                          outputs=outs, workspace=ws, order=order)

                o2, wall2, _ = time_solves(run_ordered, 2, 0, torch)
                ms2 = wall2 / 2 * 1e3
                res["fp32_refill_cost_ordered"] = {
                    "ms": ms2, "trajectories_per_s": n / (ms2 / 1e3),
                    "order": "decreasing attempted steps of an earlier solve of the same ensemble",
                    "bit_identical": bool(np.array_equal(o2.ensemble_view().cpu().numpy(),
                                                         hists[key]))}
                log_fn(f"uneven stiff fp32_refill cost-ordered: {ms2:.1f} ms")

                # a first solve that ranks its trajectories itself: a pilot of
                # every trajectory, the device sort, then the ordered solve,
                # all inside the timed region
                def run_piloted(built=built, outs=outs, ws=ws):
                    return rodas5P.launch(built, y0_t, t_t, p_t, n, dt0=0.0, rtol=RTOL,
                                          atol=ATOL, max_steps=100000, refill=True,
                                          outputs=outs, workspace=ws, order="pilot")

                o3, wall3, _ = time_solves(run_piloted, 2, 0, torch)
                ms3 = wall3 / 2 * 1e3
                res["fp32_refill_piloted"] = {
                    "ms": ms3, "trajectories_per_s": n / (ms3 / 1e3),
                    "order": f"a {rodas5P.PILOT_STEPS}-step pilot of every trajectory ranks them "
                             "by steps left; pilot and sort inside the timed region",
                    "bit_identical": bool(np.array_equal(o3.ensemble_view().cpu().numpy(),
                                                         hists[key]))}
                log_fn(f"uneven stiff fp32_refill piloted: {ms3:.1f} ms")
    from menodax_amd import autotune

    refined_tune = None
    if "fp32_refined_refill" in variants:
        # the refined FP32 LU: its stage vectors are FP64, so which placement
        # and workgroup keep them on chip is measured, on the whole ensemble in
        # the launch shape solve picks (every candidate checked bit-identical)
        best_r, cands_r = autotune.tune(fn, y0_t, t_t, p_t, reps=1, rtol=RTOL, atol=ATOL,
                                        shapes="solve", max_steps=100000, lu_precision="fp32",
                                        sparsity=mask, lu_refine=True, blocks=(32, 64, 128))
        refined_tune = {"chosen": {"placement": best_r.placement,
                                   "trajectories_per_block": best_r.trajectories_per_block,
                                   "refill": best_r.refill, "ms": best_r.seconds * 1e3},
                        "candidates": [{"placement": c.placement,
                                        "block": c.trajectories_per_block, "refill": c.refill,
                                        "ms": c.seconds * 1e3} for c in cands_r]}
        built_r = _kernels.rodas5p_kernel(fn, n_vars, n_p, _kernels.KernelOptions(
            lu_precision="fp32", sparsity=sp, lu_refine=True, placement=best_r.placement,
            trajectories_per_block=best_r.trajectories_per_block))
        _, outs_r, ws_r, _ = timed("fp32_refined_refill", built_r, best_r.refill)
        res["fp32_refined_refill"]["tuned"] = refined_tune
        if best_r.refill:
            # the same first solve ranked by the pilot: about half of the refined
            # solve is its slowest trajectory's own latency, which a late claim
            # adds to the rest of the ensemble's
            def run_refined_piloted():
                return rodas5P.launch(built_r, y0_t, t_t, p_t, n, dt0=0.0, rtol=RTOL, atol=ATOL,
                                      max_steps=100000, refill=True, outputs=outs_r,
                                      workspace=ws_r, order="pilot")

            o4, wall4, _ = time_solves(run_refined_piloted, 2, 0, torch)
            ms4 = wall4 / 2 * 1e3
            res["fp32_refined_refill_piloted"] = {
                "ms": ms4, "trajectories_per_s": n / (ms4 / 1e3),
                "order": f"a {rodas5P.PILOT_STEPS}-step pilot of every trajectory ranks them by "
                         "steps left; pilot and sort inside the timed region",
                "bit_identical": bool(np.array_equal(o4.ensemble_view().cpu().numpy(),
                                                     hists["fp32_refined_refill"]))}
            log_fn(f"uneven stiff fp32_refined_refill piloted: {ms4:.1f} ms")
    if not {"fp32_refill", "fp64_plain"} <= set(res):
        return {"variants": res}
    # the autotuner over placement x workgroup size on the whole ensemble, in
    # the launch shape solve picks for it, every candidate checked
    # bit-identical, the fastest kept
    best, cands = autotune.tune(fn, y0_t, t_t, p_t, reps=1, rtol=RTOL, atol=ATOL, shapes="solve",
                                max_steps=100000, lu_precision="fp32", sparsity=mask)
    default_block = _kernels.trajectories_per_block_or_default(None)
    default = next(c for c in cands if c.placement == res["fp32_refill"]["placement"]
                   and c.trajectories_per_block == default_block)
    tuned = {"sample": n, "chosen": {"placement": best.placement,
                                     "trajectories_per_block": best.trajectories_per_block,
                                     "refill": best.refill, "sample_ms": best.seconds * 1e3},
             "default_sample_ms": default.seconds * 1e3,
             "candidates": [{"placement": c.placement, "block": c.trajectories_per_block,
                             "refill": c.refill, "ms": c.seconds * 1e3} for c in cands]}
    log_fn(f"autotune: {best.placement} block {best.trajectories_per_block} refill {best.refill} "
           f"{best.seconds*1e3:.1f} ms on {n} (default {default.seconds*1e3:.1f} ms)")
    steps = att
    edges = np.logspace(np.log10(max(1, steps.min())), np.log10(steps.max() + 1), 9)
    counts, _ = np.histogram(steps, bins=edges)
    pairs = [lu for lu in ("fp32", "fp64") if f"{lu}_refill" in hists and f"{lu}_plain" in hists]
    same = (all(np.array_equal(hists[f"{lu}_refill"], hists[f"{lu}_plain"]) for lu in pairs)
            if pairs else None)
    # endpoint error against the oracle: FP64 LU at the same tolerance on a
    # spread sample, and a tight FP64 solve on a smaller one
    system = O.linquad(n_vars, rows, cols, vals, 0.01, off_pidx=pidx, n_p=n_p)
    layout = O.sparse_structure(sp, SD.fill_reducing_order(sp, "amd"))
    idx = spread(n, err_sample)
    t0 = time.perf_counter()
    same_tol, _ = O.rodas5p(system, y0[idx], times, p[idx], layout=layout, rtol=RTOL, atol=ATOL,
                            lu_precision="fp64")
    tidx = spread(n, tight_sample)
    tight, _ = O.rodas5p(system, y0[tidx], times, p[tidx], layout=layout, rtol=1e-10, atol=1e-12,
                         lu_precision="fp64", max_steps=1000000)
    log_fn(f"uneven stiff oracle errors: {time.perf_counter() - t0:.1f} s")
    err = {"target": f"max abs endpoint error vs the FP64 oracle <= {ERROR_TARGET:g}",
           "sample": f"{idx.size} trajectories spread evenly over the ensemble "
                     f"({tidx.size} for the tight oracle, rtol 1e-10)"}
    for key in ("fp32_refill", "fp32_refined_refill", "fp64_plain"):
        if key not in hists:
            continue
        e = error_stats(hists[key][idx, -1], same_tol[:, -1])
        err[key] = {"vs_fp64_oracle_same_tol": e,
                    "vs_tight_oracle": error_stats(hists[key][tidx, -1], tight[:, -1]),
                    "target_met": e["max_abs"] <= ERROR_TARGET}
    fp64_ms = res["fp64_plain"]["ms"]
    target = {"criterion": "FP32-LU persistent refill >= 1.5x the plain FP64 one-lane kernel in "
                           "trajectories/s, with endpoint error vs the FP64 oracle <= 1e-6"}
    for key in ("fp32_refill", "fp32_refill_piloted", "fp32_refined_refill",
                "fp32_refined_refill_piloted"):
        if key in res:
            # a piloted solve is bit-identical to its unordered one (checked),
            # so it has that one's error
            base = key.replace("_piloted", "")
            if base != key and not res[key]["bit_identical"]:
                continue
            sp_ = fp64_ms / res[key]["ms"]
            target[key] = {"speedup_over_fp64_plain": sp_,
                           "max_abs_error_vs_fp64_oracle":
                               err[base]["vs_fp64_oracle_same_tol"]["max_abs"],
                           "met": bool(sp_ >= 1.5 and err[base]["target_met"])}
    return {
        "workload": "100,000 trajectories of dy/dt = A y + 0.01 y*y, 50 vars, A block arrow "
                    "(196 nnz) with per-trajectory stiffness up to 1e5, t in [0,10], 20 saves, "
                    "rtol 1e-6 atol 1e-8, fp64 state, sparse LU (AMD)",
        **res,
        "fp32_refill_over_fp64_plain": fp64_ms / res["fp32_refill"]["ms"],
        "fp32_refill_piloted_over_fp64_plain": fp64_ms / res["fp32_refill_piloted"]["ms"],
        "speed_and_error_target": target,
        "refill_bit_identical_to_plain": same,
        "attempted_steps": {"p1": float(np.percentile(steps, 1)),
                            "median": float(np.median(steps)),
                            "p99": float(np.percentile(steps, 99)), "max": int(steps.max()),
                            "p99_over_p1": float(np.percentile(steps, 99) / np.percentile(steps, 1)),
                            "histogram_edges": [float(e) for e in edges],
                            "histogram_counts": [int(c) for c in counts]},
        "endpoint_error": err,
        "autotune": tuned,
    }


def lorenz_block(torch, dev, n=1024):
    """C1: Tsit5 on the Lorenz ensemble -- rho ~ U[0, 28] (seed 0), y0 = (1, 0, 0),
    t in [0, 1] with 100 saves, rtol 1e-6 atol 1e-8."""
    from menodax_amd import _kernels, rodas5P
    from tests import systems as S

    y0, p = S.lorenz_rho_ensemble(n)
    t = torch.tensor(np.linspace(0.0, 1.0, 100), device=dev)
    y0_t = torch.tensor(np.ascontiguousarray(y0.T), device=dev).t()
    p_t = torch.tensor(np.ascontiguousarray(p.T), device=dev).t()
    built = _kernels.tsit5_kernel(S.lorenz, 3, 1, _kernels.Tsit5Options(
        tableau_in_registers=rodas5P._tableau(None, 3, n, None, None)))
    outs = rodas5P.launch(built, y0_t, t, p_t, n, dt0=0.0, rtol=RTOL, atol=ATOL,
                          max_steps=100000, refill=False)

    def run():
        return rodas5P.launch(built, y0_t, t, p_t, n, dt0=0.0, rtol=RTOL, atol=ATOL,
                              max_steps=100000, refill=False, outputs=outs)

    o, wall, ev_ms = time_solves(run, 20, 3, torch)
    att = (o.accepted + o.rejected).float()
    # endpoint error over the whole ensemble: the FP64 oracle at the same
    # tolerance (the same arithmetic) and a tight one
    from oracle import oracle as O

    tn = np.linspace(0.0, 1.0, 100)
    same, _ = O.tsit5(O.lorenz(), y0, tn, p, rtol=RTOL, atol=ATOL)
    tight, _ = O.tsit5(O.lorenz(), y0, tn, p, rtol=1e-12, atol=1e-14)
    yg = o.ensemble_view()[:, -1].cpu().numpy()
    return {"workload": "Tsit5, Lorenz, 1,024 trajectories, rho ~ U[0,28], y0=(1,0,0), t in [0,1], "
                        "100 saves, rtol 1e-6 atol 1e-8, fp64",
            "ms": wall / 20 * 1e3, "kernel_ms": ev_ms,
            "trajectories_per_s": n / (wall / 20),
            "attempted_mean": float(att.mean().item()), "attempted_max": int(att.max().item()),
            "endpoint_error": {"vs_fp64_oracle_same_tol": error_stats(yg, same[:, -1]),
                               "vs_tight_oracle": error_stats(yg, tight[:, -1])}}


def vdp_chain_block(torch, dev, log_fn, n=16384):
    """C3: Rodas5P, sparse FP32 LU (AMD) on 32 diffusively coupled Van der Pol
    oscillators (64 vars, open chain), mu ~ LogUniform[10, 1000], y0 ~ U[-2, 2],
    t in [0, 2000] with 11 saves, rtol 1e-6 atol 1e-8, max_steps 100,000."""
    from menodax_amd import _kernels, rodas5P, roofline
    from tests import systems as S
    from menodax_amd._sparsity import normalize_sparsity

    fn, _ = S.vdp_lattice(32, mu=1.0, d=0.1, wrap=False)
    mask = normalize_sparsity(S.vdp_sparsity(32, wrap=False), 64)
    y0, mu = S.vdp_chain_ensemble(32, n)
    t = torch.tensor(np.linspace(0.0, 2000.0, 11), device=dev)
    y0_t = torch.tensor(np.ascontiguousarray(y0.T), device=dev).t()
    p_t = torch.tensor(np.ascontiguousarray(mu.T), device=dev).t()
    res = {"workload": "Rodas5P sparse (AMD) FP32 LU, 32-oscillator Van der Pol chain (64 vars), "
                       f"{n} trajectories, mu ~ LogUniform[10,1000], t in [0,2000], 11 saves, "
                       "rtol 1e-6 atol 1e-8, max_steps 100000"}
    built = _kernels.rodas5p_kernel(fn, 64, 1, _kernels.KernelOptions(lu_precision="fp32",
                                                                     sparsity=mask))
    for refill in (True, False):
        t0 = time.perf_counter()
        o = rodas5P.launch(built, y0_t, t, p_t, n, dt0=0.0, rtol=RTOL, atol=ATOL,
                           max_steps=100000, refill=refill)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        att = (o.accepted + o.rejected).cpu().numpy().astype(np.int64)
        work = roofline.solve_work(built, int(att.sum()), n * 10)
        key = "refill" if refill else "plain"
        res[key] = {"ms": sec * 1e3, "trajectories_per_s": n / sec,
                    "accepted_steps_per_s": float(o.accepted.sum().item()) / sec,
                    "roofline_frac": work.roof_seconds() / sec,
                    "attempted_mean": float(att.mean()), "attempted_max": int(att.max()),
                    "hit_max_steps": int((o.loop.cpu().numpy() >= 100000).sum())}
        log_fn(f"vdp chain {key}: {sec*1e3:.1f} ms")
    # parity and endpoint error on 16 trajectories spread over mu (sorted by
    # mu, evenly spaced): the oracle with the same sparse FP32 LU and tolerance
    # (the same arithmetic, so the same bits), and the FP64-LU oracle. Against
    # FP64 the endpoint says little for the small-mu trajectories: over
    # [0, 2000] they run through up to ~125 relaxation periods of 32
    # oscillators out of phase, the FP32 and FP64 step sequences part, and a
    # phase difference moves the endpoint by O(1) (both stop at the same
    # max_steps); the large-mu trajectories agree to the tolerance.
    from menodax_amd import _sparse_direct as SD
    from oracle import oracle as O

    att = (o.accepted + o.rejected).cpu().numpy()
    idx = np.argsort(mu[:, 0], kind="stable")[spread(n, 16)]
    system = O.vdp_lattice(32, mu=1.0, d=0.1, wrap=False)
    layout = O.sparse_structure(mask, SD.fill_reducing_order(mask, "amd"))
    tn = np.linspace(0.0, 2000.0, 11)
    t0 = time.perf_counter()
    same, st = O.rodas5p(system, y0[idx], tn, mu[idx], layout=layout, rtol=RTOL, atol=ATOL,
                         lu_precision="fp32")
    same64, _ = O.rodas5p(system, y0[idx], tn, mu[idx], layout=layout, rtol=RTOL, atol=ATOL,
                          lu_precision="fp64")
    log_fn(f"vdp chain oracle: {time.perf_counter() - t0:.1f} s")
    hist = o.ensemble_view()[idx].cpu().numpy()
    e64 = error_stats(hist[:, -1], same64[:, -1])
    e64["per_trajectory_max_abs_by_mu"] = [
        [float(mu[i, 0]), float(np.max(np.abs(hist[k, -1] - same64[k, -1])))]
        for k, i in enumerate(idx)]
    res["endpoint_error"] = {
        "sample": "16 trajectories spread over mu",
        "mu": [float(v) for v in mu[idx, 0]],
        "vs_oracle_same_tol_fp32_lu": error_stats(hist[:, -1], same[:, -1]),
        "history_bit_identical_to_oracle": bool(np.array_equal(hist, same)),
        "counters_equal_to_oracle": bool(np.array_equal(att[idx], st["accepted_steps"]
                                                         + st["rejected_steps"])),
        "vs_fp64_oracle_same_tol": e64}
    return res


def adjoint_block(torch, log_fn, n=8192, ckpt_every=64):
    """C4: d(loss)/d(mu, y0) for stiff Van der Pol ensembles."""
    from menodax_amd import rodas5P
    from tests import systems as S

    out = {"workload": f"{n} stiff Van der Pol trajectories, mu ~ LogUniform[10, 100], "
                       "y0 ~ U[-2, 2], t in [0, 3], 20 saves, loss = sum of squared errors vs a "
                       "target solved at 1.1 mu plus noise (seed 0), fp64, rtol 1e-8 atol 1e-10",
           "ckpt_every": ckpt_every}
    for n_osc in (1, 10):
        fn, _ = S.vdp_lattice(n_osc, mu=1.0, d=0.1, wrap=False)
        nv = 2 * n_osc
        sparsity = None if n_osc == 1 else S.vdp_sparsity(n_osc, wrap=False)
        rng = np.random.default_rng(0)
        mu = 10.0 ** rng.uniform(1.0, 2.0, size=(n, 1))
        y0 = rng.uniform(-2.0, 2.0, size=(n, nv))
        t = np.linspace(0.0, 3.0, 20)
        kw = dict(rtol=1e-8, atol=1e-10, lu_precision="fp64", sparsity=sparsity)
        target = rodas5P.solve(fn, y0, t, mu * 1.1, **kw) + 0.01 * rng.normal(size=(n, 20, nv))
        dev = torch.device("cuda", 0)
        y0d, mud = torch.tensor(y0, device=dev), torch.tensor(mu, device=dev)
        tgt = torch.tensor(target, device=dev)

        def run_fwd():
            return rodas5P.solve(fn, y0d, t, mud, **kw)

        hist = run_fwd()
        g = 2.0 * (hist - tgt)

        def run_adj():
            return rodas5P.adjoint_gradients(fn, y0d, t, mud, g, ckpt_every=ckpt_every, **kw)

        def run_sens():
            return rodas5P.solve_sensitivities(fn, y0d, t, mud, **kw)

        def clock(f, reps=3):
            f()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                r = f()
            torch.cuda.synchronize()
            return r, (time.perf_counter() - t0) / reps

        _, t_fwd = clock(run_fwd)
        (gy, gp), t_adj = clock(run_adj)
        (hs, sens), t_sens = clock(run_sens)
        gdir = torch.einsum("nsi,nsik->nk", 2.0 * (hs - tgt), sens)
        ref = torch.cat([gdir[:, :nv], gdir[:, nv:]], 1)
        got = torch.cat([gy, gp], 1)
        rel_sens = float(((got - ref).abs().max(1).values / ref.abs().max(1).values).max())
        # central differences of the loss in mu and every y0 component on 8
        # trajectories spread over the ensemble, against the adjoint of the
        # same tight solve (rtol 1e-12: the adjoint differentiates the discrete
        # solve, so both sides take the same one). Each difference is
        # Richardson-extrapolated from relative steps h and h/2 at h = 1e-4,
        # 1e-5 and 1e-6, and for each component the level that agrees best
        # with the next smaller one is kept -- a choice made from the
        # differences alone: large steps leave a truncation error where the
        # loss bends sharply (the relaxation jumps), small ones turn the solve's
        # accept/reject decisions into noise. Every perturbed solve is one
        # trajectory of a single batched launch.
        idx = np.linspace(0, n - 1, 8).astype(int)
        kwt = dict(kw, rtol=1e-12, atol=1e-14)
        yk, mk, tk = y0[idx], mu[idx], target[idx]
        k = idx.size
        hk = rodas5P.solve(fn, yk, t, mk, **kwt)
        gyk, gpk = rodas5P.adjoint_gradients(fn, yk, t, mk, 2.0 * (hk - tk), **kwt)
        steps = (1e-4, 5e-5, 1e-5, 5e-6, 1e-6, 5e-7)
        ys, ms, width = [], [], []
        for rel in steps:
            for j in range(nv + 1):
                d = np.zeros_like(yk)
                dm = np.zeros_like(mk)
                if j < nv:
                    d[:, j] = rel * np.maximum(np.abs(yk[:, j]), 1e-3)
                    width.append(d[:, j])
                else:
                    dm[:, 0] = rel * mk[:, 0]
                    width.append(dm[:, 0])
                ys += [yk + d, yk - d]
                ms += [mk + dm, mk - dm]
        hp = rodas5P.solve(fn, np.concatenate(ys), t, np.concatenate(ms), **kwt)
        lp = ((hp - np.tile(tk, (len(ys), 1, 1))) ** 2).sum((1, 2)).reshape(len(steps), nv + 1,
                                                                           2, k)
        dif = (lp[:, :, 0] - lp[:, :, 1]) / (2 * np.asarray(width).reshape(len(steps), nv + 1, k))
        rich = (4.0 * dif[1::2] - dif[0::2]) / 3.0            # (levels, nv + 1, k)
        pick = np.argmin(np.abs(rich[:-1] - rich[1:]), axis=0)  # best-agreeing level
        fd = np.take_along_axis(rich, pick[None], axis=0)[0].T  # (k, nv + 1)
        gk = np.concatenate([gyk, gpk], axis=1)
        scale = np.abs(fd).max(axis=1, keepdims=True)
        rel_fd = float((np.abs(gk - fd).max(axis=1) / scale[:, 0]).max())
        big = np.abs(fd) >= 1e-3 * scale
        rel_fd_comp = float((np.abs(gk - fd)[big] / np.abs(fd)[big]).max())
        key = f"vdp{nv}"
        out[key] = {
            "n_vars": nv, "lu": "dense" if sparsity is None else "sparse (AMD)",
            "forward_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3,
            "forward_sensitivity_ms": t_sens * 1e3,
            "gradients_per_s": n / (t_fwd + t_adj),
            "adjoint_vs_sensitivity_max_rel": rel_sens,
            "adjoint_vs_central_difference_max_rel": rel_fd,
            "adjoint_vs_central_difference_componentwise_max_rel": rel_fd_comp,
            "central_difference_check": "8 trajectories, mu and every y0 component, rtol 1e-12, "
                                        "Richardson-extrapolated differences at relative steps "
                                        "1e-4, 1e-5 or 1e-6 (the level agreeing best with the "
                                        "next); max_rel normwise per trajectory (max |g - fd| / "
                                        "max |fd|), componentwise over components >= 1e-3 of "
                                        "the trajectory's largest",
        }
        log_fn(f"adjoint {key}: fwd {t_fwd*1e3:.1f} ms adj {t_adj*1e3:.1f} ms sens {t_sens*1e3:.1f} ms "
               f"rel {rel_sens:.2e} fd {rel_fd:.2e} (componentwise {rel_fd_comp:.2e})")
    out["tsit5_lorenz3"] = tsit5_adjoint_line(torch, log_fn, n, ckpt_every)
    return out


def tsit5_adjoint_line(torch, log_fn, n, ckpt_every):
    """The Tsit5 discrete adjoint on a Lorenz ensemble of the same size:
    rho ~ U[5, 28], y0 near (1, 0, 0), t in [0, 1], 20 saves, the loss against a
    target solved at 1.05 rho plus noise, rtol 1e-8 atol 1e-10."""
    from menodax_amd import tsit5
    from tests import systems as S

    rng = np.random.default_rng(0)
    rho = rng.uniform(5.0, 28.0, size=(n, 1))
    y0 = np.column_stack([1.0 + rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n),
                          rng.uniform(0.0, 0.1, n)])
    t = np.linspace(0.0, 1.0, 20)
    kw = dict(rtol=1e-8, atol=1e-10)
    target = tsit5.solve(S.lorenz, y0, t, rho * 1.05, **kw) + 0.01 * rng.normal(size=(n, 20, 3))
    dev = torch.device("cuda", 0)
    y0d, rhod, tgt = (torch.tensor(x, device=dev) for x in (y0, rho, target))

    def clock(f, reps=3):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = f()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) / reps

    hist, t_fwd = clock(lambda: tsit5.solve(S.lorenz, y0d, t, rhod, **kw))
    g = 2.0 * (hist - tgt)
    (gy, gp), t_adj = clock(lambda: tsit5.adjoint_gradients(S.lorenz, y0d, t, rhod, g,
                                                             ckpt_every=ckpt_every, **kw))
    (hs, sens), t_sens = clock(lambda: tsit5.solve_sensitivities(S.lorenz, y0d, t, rhod, **kw))
    gdir = torch.einsum("nsi,nsik->nk", 2.0 * (hs - tgt), sens)
    got = torch.cat([gy, gp], 1)
    rel_sens = float(((got - gdir).abs().max(1).values / gdir.abs().max(1).values).max())
    k = 4
    h = 1e-6 * rho[:k]
    lp = ((tsit5.solve(S.lorenz, y0[:k], t, rho[:k] + h, **kw) - target[:k]) ** 2).sum((1, 2))
    lm = ((tsit5.solve(S.lorenz, y0[:k], t, rho[:k] - h, **kw) - target[:k]) ** 2).sum((1, 2))
    fd = (lp - lm) / (2 * h[:, 0])
    rel_fd = float(np.max(np.abs(gp[:k, 0].cpu().numpy() - fd) / np.abs(fd)))
    log_fn(f"adjoint tsit5 lorenz: fwd {t_fwd*1e3:.1f} ms adj {t_adj*1e3:.1f} ms "
           f"sens {t_sens*1e3:.1f} ms rel {rel_sens:.2e} fd {rel_fd:.2e}")
    return {"method": "Tsit5", "n_vars": 3, "workload": f"{n} Lorenz trajectories, rho ~ U[5, 28], "
            "y0 ~ (1, 0, 0) + U[-0.1, 0.1], t in [0, 1], 20 saves, loss vs a target at 1.05 rho "
            "plus noise, rtol 1e-8 atol 1e-10",
            "forward_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3,
            "forward_sensitivity_ms": t_sens * 1e3, "gradients_per_s": n / (t_fwd + t_adj),
            "adjoint_vs_sensitivity_max_rel": rel_sens,
            "adjoint_vs_central_difference_max_rel": rel_fd}


# The reference's published RTX 4070 SUPER times (ms) for its own benchmark
# scripts, and where they are recorded (paths under the reference's benchmarks/).
REF_PUBLISHED_MS = {
    "lorenz_tsit5_n100000": (93.62, "nonstiff_lorenz_scaling/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:11"),
    "lorenz_tsit5_n1000": (3.26, "nonstiff_lorenz_scaling/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:7"),
    "vdp_ring_nonstiff_n64_tsit5": (2.095, "nonstiff_vdp_dimensionality/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:10"),
    "vdp_ring_nonstiff_n128_tsit5": (5.284, "nonstiff_vdp_dimensionality/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:11"),
    "vdp_ring_n32_sparse_fp32": (1.60, "stiff_vdp_dimensionality/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:9"),
    "vdp_ring_n32_dense_fp32": (43.44, "stiff_vdp_dimensionality/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:20"),
    "vdp_ring_n32_dense_fp64": (52.98, "stiff_vdp_dimensionality/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:31"),
    "vdp_ring_n64_sparse_fp32": (5.94, "stiff_vdp_dimensionality/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:10"),
    "vdp_ring_n64_dense_fp32": (380.06, "stiff_vdp_dimensionality/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:21"),
    "vdp_ring_n64_dense_fp64": (334.02, "stiff_vdp_dimensionality/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:32"),
    "vdp_ring_n64_value_and_grad_sparse": (14.69, "stiff_vdp_gradient/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:10"),
    "vdp_ring_n64_value_and_grad_dense": (515.08, "stiff_vdp_gradient/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:21"),
}


def reference_benchmarks(torch, log_fn, reps=5):
    """The reference's own benchmark points with published times, run through
    the public API on device-resident inputs (one warm-up, then the mean of
    ``reps`` synchronised calls, as its time_blocked does): Lorenz Tsit5 on
    identical trajectories (y0 = (1000, -500, 500), rho = 0.5, t in [0, 5],
    first_step 1e-4); Tsit5 on the non-stiff Van der Pol ring (mu = 1, 64 and
    128 variables, 1,000 identical trajectories); the stiff Van der Pol ring (mu = 100, d = 10, 1,000
    identical trajectories, saves at 0, 0.25, ..., 1, first_step 1e-4) with
    the sparse FP32, dense FP32 and dense FP64 LU; and the value and gradient
    of the summed final state with respect to every trajectory's parameter by
    forward sensitivities (the reference's jax.value_and_grad), sparse and
    dense. rtol 1e-6, atol 1e-8 throughout. The dense rings are also run with
    the autotuner's choice (reported next to the default's time). Every row
    carries the endpoint error of its first 8 trajectories against the FP64
    oracle at the same tolerance and against a tight FP64 oracle (rtol 1e-10);
    the value-and-gradient rows also the gradient against central differences
    of the plain solve."""
    from menodax_amd import _sparse_direct as SD, autotune, rodas5P, tsit5
    from menodax_amd._sparsity import normalize_sparsity
    from oracle import oracle as O
    from tests import systems as S

    dev = torch.device("cuda", 0)
    K = 8

    def errors(got, method, system, y0k, tk, pk, first_step, layout=None):
        """``got`` (K, n_vars) at t_end against the oracle's FP64 solves."""
        if method == "tsit5":
            same, _ = O.tsit5(system, y0k, tk, pk, first_step=first_step, rtol=1e-6, atol=1e-8)
            tight, _ = O.tsit5(system, y0k, tk, pk, first_step=first_step, rtol=1e-10,
                               atol=1e-12, max_steps=10000000)
        else:
            same, _ = O.rodas5p(system, y0k, tk, pk, first_step=first_step, rtol=1e-6, atol=1e-8,
                                lu_precision="fp64", layout=layout)
            tight, _ = O.rodas5p(system, y0k, tk, pk, first_step=first_step, rtol=1e-10,
                                 atol=1e-12, lu_precision="fp64", layout=layout,
                                 max_steps=1000000)
        return {"vs_fp64_oracle_same_tol": error_stats(got, same[:, -1]),
                "vs_tight_oracle": error_stats(got, tight[:, -1])}

    def clock(f):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    out = {}

    def record(key, ms, **extra):
        pub, src = REF_PUBLISHED_MS[key]
        out[key] = {"ms": ms, "published_rtx4070_super_ms": pub, "speedup": pub / ms,
                    "source": src, **extra}
        log_fn(f"reference benchmark {key}: {ms:.3f} ms (published {pub} ms)")

    kw = dict(first_step=1e-4, rtol=1e-6, atol=1e-8)
    t_l = torch.tensor(S.LORENZ_TIMES, device=dev)
    for n in (100000, 1000):
        y0 = torch.tensor(np.tile(S.LORENZ_Y0, (n, 1)), device=dev)
        p = torch.tensor(np.tile(S.LORENZ_PARAMS, (n, 1)), device=dev)
        ms = clock(lambda: tsit5.solve(S.lorenz, y0, t_l, p, **kw))
        got = tsit5.solve(S.lorenz, y0, t_l, p, **kw)[:K, -1].cpu().numpy()
        record(f"lorenz_tsit5_n{n}", ms,
               endpoint_error=errors(got, "tsit5", O.lorenz(), y0[:K].cpu().numpy(),
                                     np.asarray(S.LORENZ_TIMES), p[:K].cpu().numpy(), 1e-4))
    t_v = torch.tensor(S.VDP_TIMES, device=dev)
    for n_osc in (32, 64):
        # the non-stiff ring (mu = 1) under Tsit5
        fn, _ = S.vdp_lattice(n_osc, mu=1.0)
        y0n, pn = S.vdp_scenario(n_osc, 1000, divergence=0.0)
        y0, p = torch.tensor(y0n, device=dev), torch.tensor(pn, device=dev)
        ms = clock(lambda: tsit5.solve(fn, y0, t_v, p, **kw))
        got = tsit5.solve(fn, y0, t_v, p, **kw)[:K, -1].cpu().numpy()
        record(f"vdp_ring_nonstiff_n{2 * n_osc}_tsit5", ms,
               endpoint_error=errors(got, "tsit5", O.vdp_lattice(n_osc, mu=1.0), y0n[:K],
                                     S.VDP_TIMES, pn[:K], 1e-4))
    for n_osc in (16, 32):
        fn, _ = S.vdp_lattice(n_osc)
        mask = S.vdp_sparsity(n_osc)
        y0n, pn = S.vdp_scenario(n_osc, 1000, divergence=0.0)
        y0, p = torch.tensor(y0n, device=dev), torch.tensor(pn, device=dev)
        nv = 2 * n_osc
        cases = (("sparse_fp32", dict(lu_precision="fp32", sparsity=mask)),
                 ("dense_fp32", dict(lu_precision="fp32")),
                 ("dense_fp64", dict(lu_precision="fp64")))
        ring = normalize_sparsity(mask, nv)
        ring_layout = O.sparse_structure(ring, SD.fill_reducing_order(ring, "amd"))
        for name, extra in cases:
            ms = clock(lambda: rodas5P.solve(fn, y0, t_v, p, **kw, **extra))
            got = rodas5P.solve(fn, y0, t_v, p, **kw, **extra)[:K, -1].cpu().numpy()
            err = errors(got, "rodas5p", O.vdp_lattice(n_osc), y0n[:K], S.VDP_TIMES, pn[:K],
                         1e-4, layout=ring_layout if "sparsity" in extra else None)
            tuned = {}
            if name.startswith("dense"):
                # the autotuner's placement, workgroup size and launch shape
                # (tuning, like compilation, outside the timed calls)
                opts = autotune.tuned_options(fn, y0, t_v, p, reps=1, blocks=(4, 8, 16, 32, 64),
                                              **kw, **extra)
                tuned = {"default_ms": ms, "tuned": opts}
                ms = clock(lambda: rodas5P.solve(fn, y0, t_v, p, **kw, **extra, **opts))
            record(f"vdp_ring_n{nv}_{name}", ms, endpoint_error=err, **tuned)
        if n_osc == 32:
            for name, extra in (("sparse", dict(sparsity=mask)), ("dense", {})):
                def value_and_grad(extra=extra):
                    hist, sens = rodas5P.solve_sensitivities(fn, y0, t_v, p, wrt_y0=False,
                                                             **kw, **extra)
                    return hist[:, -1].sum(), sens[:, -1, :, 0].sum(dim=1)
                ms = clock(value_and_grad)
                hist, sens = rodas5P.solve_sensitivities(fn, y0[:K], t_v, p[:K], wrt_y0=False,
                                                         **kw, **extra)
                err = errors(hist[:, -1].cpu().numpy(), "rodas5p", O.vdp_lattice(n_osc),
                             y0n[:K], S.VDP_TIMES, pn[:K], 1e-4,
                             layout=ring_layout if extra else None)
                # the gradient of the summed final state in mu against central
                # differences of tight FP64 solves (rtol 1e-12, relative step
                # 1e-5): at the benchmark's own rtol 1e-6 and FP32 LU the
                # difference quotient is dominated by the solves' noise
                grad = sens[:, -1, :, 0].sum(dim=1).cpu().numpy()
                h = 1e-5 * pn[:K]
                tight = dict(first_step=1e-4, rtol=1e-12, atol=1e-14, lu_precision="fp64",
                             max_steps=1000000, sparsity=extra.get("sparsity"))
                up = rodas5P.solve(fn, y0n[:K], S.VDP_TIMES, pn[:K] + h, **tight)
                dn = rodas5P.solve(fn, y0n[:K], S.VDP_TIMES, pn[:K] - h, **tight)
                fd = (up[:, -1].sum(axis=1) - dn[:, -1].sum(axis=1)) / (2 * h[:, 0])
                err["gradient_vs_tight_central_difference_max_rel"] = float(
                    np.max(np.abs(grad - fd) / np.abs(fd)))
                record(f"vdp_ring_n{nv}_value_and_grad_{name}", ms,
                       method="forward sensitivities (joint solve)", endpoint_error=err)
    return out


# The reference's Robertson divergence rows (30,000 trajectories, FP32 LU):
# unsorted and pre-sorted by attempted steps, most first (its sort_by_steps).
REF_DIVERGENCE_MS = {
    1.0: (15.00275800935924, 9.008343098685145,
          "stiff_robertson_divergence/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:6, :17"),
    3.0: (11.780368979088962, 6.499434006400406,
          "stiff_robertson_divergence/results-NVIDIA_GeForce_RTX_4070_SUPER.csv:12, :23"),
}


def robertson_divergence(torch, log_fn, reps=5):
    """The reference's stiff Robertson divergence benchmark (robertson
    make_scenario at divergence 1.0 and 3.0, seed 42, 30,000 trajectories,
    first_step 1e-4, rtol 1e-6, atol 1e-8, FP32 LU) through the engine's launch
    shapes: the one-lane-per-trajectory kernel on the ensemble as given, on the
    ensemble pre-sorted by attempted steps (the reference's sorted row: the
    inputs reordered, most steps first), the persistent refill kernel, the
    refill kernel claiming trajectories by decreasing step count of an earlier
    solve, and the refill kernel after a pilot ranks them (pilot and sort in the
    timed call). One warm-up, then the mean of ``reps`` synchronised solves."""
    from menodax_amd import _common, rodas5P
    from tests import systems as S

    dev = torch.device("cuda", 0)
    t = torch.tensor(S.ROBERTSON_TIMES, device=dev)
    kw = dict(first_step=1e-4, rtol=1e-6, atol=1e-8, lu_precision="fp32")

    def clock(f):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    out = {}
    for div, (pub_unsorted, pub_sorted, src) in REF_DIVERGENCE_MS.items():
        y0n, pn = S.robertson_scenario(30000, 42, divergence=div)
        y0, p = torch.tensor(y0n, device=dev), torch.tensor(pn, device=dev)
        base, st = rodas5P.solve(S.robertson, y0, t, p, return_stats=True, refill=False, **kw)
        steps = (st["accepted_steps"] + st["rejected_steps"]).cpu().numpy()
        order = np.argsort(-steps, kind="stable")
        y0s, ps = y0[order], p[order]
        claim = _common.cost_order(st["accepted_steps"] + st["rejected_steps"])
        row = {"published_rtx4070_super_ms": {"unsorted": pub_unsorted, "sorted": pub_sorted},
               "source": src, "attempted_mean": float(steps.mean()),
               "attempted_min": int(steps.min()), "attempted_max": int(steps.max())}
        row["plain_ms"] = clock(lambda: rodas5P.solve(S.robertson, y0, t, p, refill=False, **kw))
        row["plain_presorted_ms"] = clock(
            lambda: rodas5P.solve(S.robertson, y0s, t, ps, refill=False, **kw))
        row["refill_ms"] = clock(lambda: rodas5P.solve(S.robertson, y0, t, p, refill=True, **kw))
        row["refill_cost_ordered_ms"] = clock(
            lambda: rodas5P.solve(S.robertson, y0, t, p, refill=True, work_order=claim, **kw))
        row["refill_piloted_ms"] = clock(
            lambda: rodas5P.solve(S.robertson, y0, t, p, refill=True, work_order="pilot", **kw))
        same = rodas5P.solve(S.robertson, y0, t, p, refill=True, work_order=claim, **kw)
        row["bit_identical_across_shapes"] = bool(torch.equal(same, base))
        row["speedup_unsorted"] = pub_unsorted / row["plain_ms"]
        row["speedup_sorted"] = pub_sorted / row["plain_presorted_ms"]
        out[f"divergence_{div:g}"] = row
        log_fn(f"robertson divergence {div:g}: plain {row['plain_ms']:.3f} presorted "
               f"{row['plain_presorted_ms']:.3f} refill {row['refill_ms']:.3f} ordered "
               f"{row['refill_cost_ordered_ms']:.3f} piloted {row['refill_piloted_ms']:.3f} ms "
               f"(published {pub_unsorted:.2f} / {pub_sorted:.2f} ms)")
    return out


def wave_ring(torch, log_fn, reps=3):
    """The reference's 64-variable stiff Van der Pol ring, 1,000 identical
    trajectories, dense LU: the per-lane kernel (placement "lds") against the
    wave placement, FP32 and FP64 LU, one warm-up and ``reps`` timed solves."""
    from menodax_amd import rodas5P
    from tests import systems as S

    dev = torch.device("cuda", 0)
    fn, _ = S.vdp_lattice(32)
    y0n, pn = S.vdp_scenario(32, 1000, divergence=0.0)
    y0, p = torch.tensor(y0n, device=dev), torch.tensor(pn, device=dev)
    t = torch.tensor(S.VDP_TIMES, device=dev)
    out = {}
    for lu in ("fp32", "fp64"):
        for placement in ("lds", "wave"):
            kw = dict(first_step=1e-4, rtol=1e-6, atol=1e-8, lu_precision=lu,
                      placement=placement)
            rodas5P.solve(fn, y0, t, p, **kw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                rodas5P.solve(fn, y0, t, p, **kw)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / reps * 1e3
            out[f"{lu}_{placement}_ms"] = ms
            log_fn(f"wave ring: {lu} {placement} {ms:.2f} ms")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_TRAJ)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget for the CPU oracle baseline sample")
    ap.add_argument("--refill", choices=("auto", "on", "off"), default="auto")
    ap.add_argument("--skip-reference-config", action="store_true")
    ap.add_argument("--skip-uneven-stiff", action="store_true",
                    help="leave out the 100k uneven stiff ensemble comparison")
    ap.add_argument("--skip-adjoint", action="store_true",
                    help="leave out the adjoint-gradient block")
    ap.add_argument("--skip-vdp-chain", action="store_true",
                    help="leave out the 64-variable Van der Pol chain ensemble")
    ap.add_argument("--only-wave-ring", action="store_true",
                    help="run and print only the 64-variable dense Van der Pol ring, per-lane "
                         "kernel against the wave placement (for profiling)")
    ap.add_argument("--only-uneven-stiff", action="store_true",
                    help="run and print only the uneven stiff ensemble comparison (for profiling)")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="run and print only the adjoint-gradient block (for profiling)")
    ap.add_argument("--uneven-variants", default=",".join(UNEVEN_VARIANTS),
                    help="with --only-uneven-stiff: comma-separated subset of "
                         + ", ".join(UNEVEN_VARIANTS) + " (profiling runs)")
    ap.add_argument("--lu", choices=("both", "fp32", "fp64"), default="both",
                    help="profile runs: time one LU precision only (the JSON line needs both)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, copy MIB MiB device-to-device a few times "
                         "(a plain copy of known size to scale HBM byte counters by)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(
            "bench.py measures one GPU: trajectories are independent, so a multi-GPU run "
            "is N/G trajectories per GPU with no exchange; run with --gpus 1",
            file=sys.stderr,
        )
        sys.exit(2)

    scipy_pool = None
    if not (args.no_cpu_baseline or args.only_uneven_stiff or args.only_wave_ring
            or args.only_adjoint or args.calibrate_copy
            or args.lu != "both"):
        # worker processes for the SciPy CPU baseline, spawned while this
        # process has not yet initialised the GPU
        from oracle import oracle as O
        from oracle import scipy_baseline as SB

        scipy_pool = SB.Pool(O.max_threads())

    import torch

    from menodax_amd import _kernels, _runtime, roofline, rodas5P
    from tests import systems as S

    if args.only_wave_ring:
        print(json.dumps({"wave_ring": wave_ring(torch, log)}), flush=True)
        return

    if args.only_adjoint:
        print(json.dumps({"adjoint": adjoint_block(torch, log)}), flush=True)
        return

    if args.only_uneven_stiff:
        from oracle import oracle as O

        variants = tuple(v for v in args.uneven_variants.split(",") if v)
        unknown = set(variants) - set(UNEVEN_VARIANTS)
        if unknown:
            ap.error(f"unknown --uneven-variants {sorted(unknown)}")
        print(json.dumps({"uneven_stiff": uneven_stiff(torch, torch.device("cuda", 0), O, log,
                                                       variants=variants)}), flush=True)
        return

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: copied {args.calibrate_copy} MiB three times")
        return
    arch, n_cu = _runtime.device_info()
    log(f"bench: device {arch}, {n_cu} CUs, N={args.n}")

    y0_np, p_np = S.robertson_lognormal_ensemble(args.n, seed=1)
    times_np = S.robertson_benchmark_times()
    # inputs resident in HBM in the structure-of-arrays layout; the (N, k)
    # views handed to the launcher are transposes of contiguous storage, so the
    # launch makes no copies
    y0_s = torch.tensor(np.ascontiguousarray(y0_np.T), device=dev)
    p_s = torch.tensor(np.ascontiguousarray(p_np.T), device=dev)
    times = torch.tensor(times_np, device=dev)
    y0_v, p_v = y0_s.t(), p_s.t()
    refill = {"auto": None, "on": True, "off": False}[args.refill]

    results = {}
    # the headline FP32-LU kernel, the FP64-LU kernel, and the FP32 LU with
    # refined stage solves (FP64 stage vectors)
    variants = (("fp32", "fp32", False), ("fp64", "fp64", False), ("fp32_refined", "fp32", True))
    if args.lu != "both":
        variants = tuple(v for v in variants if v[0] == args.lu)
    for lu, lu_prec, refine in variants:
        built = _kernels.rodas5p_kernel(S.robertson, 3, 3, _kernels.KernelOptions(
            lu_precision=lu_prec, lu_refine=refine,
            tableau_in_registers=rodas5P._tableau(None, 3, args.n, None, None)))
        info = _runtime.kernel_info(built.plain)
        use_refill = rodas5P._use_refill(built, args.n) if refill is None else refill
        outputs = rodas5P.launch(built, y0_v, times, p_v, args.n, dt0=0.0, rtol=RTOL, atol=ATOL,
                                 max_steps=100000, refill=use_refill)
        ws = outputs.workspace

        def run(built=built, outputs=outputs, ws=ws, use_refill=use_refill):
            return rodas5P.launch(built, y0_v, times, p_v, args.n, dt0=0.0, rtol=RTOL, atol=ATOL,
                                  max_steps=100000, refill=use_refill, outputs=outputs,
                                  workspace=ws)

        log(f"bench: timing lu={lu} refill={use_refill} vgprs={info['vgprs']} "
            f"scratch={info['scratch_bytes']}B blocks/CU={info['blocks_per_cu']}")
        out, wall, event_ms = time_solves(run, args.steps, args.warmup, torch)
        acc = out.accepted.cpu().numpy().astype(np.int64)
        rej = out.rejected.cpu().numpy().astype(np.int64)
        hist = out.ensemble_view().cpu().numpy()
        attempted = int((acc + rej).sum())
        saves = args.n * (times_np.size - 1)
        work = roofline.solve_work(built, attempted, saves)
        t_kernel = event_ms / 1e3
        results[lu] = dict(
            built=built, info=info, refill=use_refill, wall=wall, event_ms=event_ms,
            ms_per_step=wall / args.steps * 1e3, accepted=int(acc.sum()), rejected=int(rej.sum()),
            attempted=attempted, work=work, t_kernel=t_kernel, hist=hist,
            max_attempted=int((acc + rej).max()), min_attempted=int((acc + rej).min()),
        )
        log(f"bench: lu={lu} {results[lu]['ms_per_step']:.3f} ms/solve (events {event_ms:.3f} ms), "
            f"{args.n / (results[lu]['ms_per_step'] / 1e3):.4g} traj/s, "
            f"mean attempted {attempted / args.n:.2f}")

    if args.lu != "both":
        log("bench: single-precision profile run, no JSON line")
        return
    r32, r64 = results["fp32"], results["fp64"]
    traj_s = args.n / (r32["ms_per_step"] / 1e3)
    w = r32["work"]
    roof_t = w.roof_seconds()
    fp64_equiv = w.fp64 + w.fp32 * roofline.P64_TFLOPS / roofline.P32_TFLOPS
    achieved = fp64_equiv / r32["t_kernel"] / 1e12

    line = {
        "metric": "trajectories/s + accepted steps/s, Rodas5P FP32-LU vs FP64 at rtol 1e-6, 1 GPU",
        "value": traj_s,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state, f32 LU",
        "data": "synthetic (seeded lognormal Robertson rates)",
        "config": {
            "workload": "Robertson Rodas5P dense LU, lognormal rates sigma 0.3, 50 log-spaced "
                        "saves to t=1e5, rtol 1e-6 atol 1e-8",
            "n_traj": args.n,
            "n_vars": 3,
            "n_save": int(times_np.size),
            "lu_precision": "fp32",
            "kernel": "refill" if r32["refill"] else "one lane per trajectory",
            "tableau_in_registers": bool(r32["built"].source.count("#define MDX_TAB_REGS 1")),
        },
        "accepted_steps_per_s": r32["accepted"] / (r32["ms_per_step"] / 1e3),
        "fp64_lu": {
            "trajectories_per_s": args.n / (r64["ms_per_step"] / 1e3),
            "accepted_steps_per_s": r64["accepted"] / (r64["ms_per_step"] / 1e3),
            "ms_per_step": r64["ms_per_step"],
        },
        "fp32_over_fp64": r64["ms_per_step"] / r32["ms_per_step"],
        "fp32_refined_lu": {
            "trajectories_per_s": args.n / (results["fp32_refined"]["ms_per_step"] / 1e3),
            "accepted_steps_per_s": results["fp32_refined"]["accepted"]
            / (results["fp32_refined"]["ms_per_step"] / 1e3),
            "ms_per_step": results["fp32_refined"]["ms_per_step"],
            "over_fp64": r64["ms_per_step"] / results["fp32_refined"]["ms_per_step"],
            "note": "FP32 factorisation and solves, each stage solve refined once against the "
                    "FP64 iteration matrix, stage vectors in FP64 (lu_refine=True)",
        },
        "steps_per_trajectory": {
            "fp32": {"attempted_mean": r32["attempted"] / args.n, "min": r32["min_attempted"],
                     "max": r32["max_attempted"], "rejected_mean": r32["rejected"] / args.n},
            "fp64": {"attempted_mean": r64["attempted"] / args.n, "min": r64["min_attempted"],
                     "max": r64["max_attempted"], "rejected_mean": r64["rejected"] / args.n},
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": roofline.P64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / roofline.P64_TFLOPS,
            "traffic": profiled_traffic(args.n),
            "traffic_source": "profiles/rodas5p_robertson65536_hbm_counters.txt",
            "counters": profiled_counters(args.n),
            "kernel": "mdx_rodas5p_refill" if r32["refill"] else "mdx_rodas5p",
            "kernel_ms": r32["t_kernel"] * 1e3,
            "flops_fp64": w.fp64,
            "flops_fp32": w.fp32,
            "special_fn_calls": w.special,
            "roof_ms": roof_t * 1e3,
            "note": "achieved = (fp64 flops + fp32 flops * 78.6/157.3) / kernel time: "
                    "FP64-equivalent TFLOP/s against the FP64 vector peak; frac = roof time "
                    "/ kernel time",
            "hbm_algorithmic_bytes": roofline.hbm_bytes(args.n, 3, 3, int(times_np.size), False),
            "vgprs": r32["info"]["vgprs"],
            "scratch_bytes": r32["info"]["scratch_bytes"],
        },
    }

    # endpoint error against the FP64 oracle (whole ensemble) and a tight one
    from oracle import oracle as O

    line["endpoint_error"] = c2_endpoint_errors(
        {k: results[k]["hist"] for k in ("fp32", "fp64", "fp32_refined")}, y0_np, p_np,
        times_np, O)

    if not args.no_cpu_baseline:
        threads = O.max_threads()
        # a bounded sample of the same workload: time growing slices until the
        # budget is spent, report the rate of the last one
        n_cpu = 1024
        while True:
            t0 = time.perf_counter()
            O.rodas5p(O.robertson(), y0_np[:n_cpu], times_np, p_np[:n_cpu], rtol=RTOL, atol=ATOL,
                      lu_precision="fp32", threads=threads)
            dt = time.perf_counter() - t0
            if dt > args.cpu_seconds / 3 or n_cpu >= args.n:
                break
            n_cpu = min(args.n, n_cpu * 4)
        # the whole ensemble within the budget: repeat it for a steadier rate
        reps = 1
        if n_cpu >= args.n and dt < args.cpu_seconds / 3:
            reps = max(1, min(100, int(args.cpu_seconds / 3 / dt)))
            t0 = time.perf_counter()
            for _ in range(reps):
                O.rodas5p(O.robertson(), y0_np[:n_cpu], times_np, p_np[:n_cpu], rtol=RTOL,
                          atol=ATOL, lu_precision="fp32", threads=threads)
            dt = (time.perf_counter() - t0) / reps
        line["cpu_baseline"] = {
            "value": n_cpu / dt,
            "unit": "trajectories/s",
            "cores": threads,
            "kind": "port",
            "sample": f"first {n_cpu} trajectories of the same ensemble"
                      + (f" (all of it, solved {reps} times)" if reps > 1 else "")
                      + f", FP32-LU, C oracle with OpenMP, {dt:.3f} s per solve",
        }
        # the reference's own kind of CPU baseline: SciPy solve_ivp per
        # trajectory over a pool of single-threaded worker processes (started
        # before this process touched the GPU)
        pool = scipy_pool
        try:
            n_sp, dt_sp, _ = SB.timed_sample(pool, "robertson", y0_np, times_np, p_np,
                                             args.cpu_seconds / 3, rtol=RTOL, atol=ATOL)
            line["cpu_baseline_scipy"] = {
                "value": n_sp / dt_sp, "unit": "trajectories/s", "cores": threads,
                "kind": "port", "method": "LSODA",
                "sample": f"first {n_sp} trajectories of the same ensemble, scipy.integrate."
                          f"solve_ivp (LSODA) per trajectory over {threads} single-threaded "
                          f"worker processes, NumPy right-hand side, {dt_sp:.2f} s",
            }
            y0l, rhol = S.lorenz_rho_ensemble(1024)
            n_l, dt_l, _ = SB.timed_sample(pool, "lorenz", y0l, np.linspace(0.0, 1.0, 100), rhol,
                                           args.cpu_seconds / 3, method="RK45", rtol=RTOL,
                                           atol=ATOL)
            lorenz_scipy = {"value": n_l / dt_l, "unit": "trajectories/s", "cores": threads,
                            "kind": "port", "method": "RK45",
                            "sample": f"{n_l} trajectories, {dt_l:.2f} s"}
        finally:
            pool.close()

    if not args.skip_reference_config:
        # the reference's own published inputs (stiff_robertson_scaling, N=100000)
        n_ref = 100000
        y0r = torch.tensor(np.ascontiguousarray(np.tile(S.ROBERTSON_Y0, (n_ref, 1)).T), device=dev)
        pr = torch.tensor(np.ascontiguousarray(np.tile(S.ROBERTSON_PARAMS, (n_ref, 1)).T), device=dev)
        tr = torch.tensor(S.ROBERTSON_TIMES, device=dev)
        ref = {}
        for lu in ("fp32", "fp64"):
            built = _kernels.rodas5p_kernel(S.robertson, 3, 3, _kernels.KernelOptions(
                lu_precision=lu, tableau_in_registers=rodas5P._tableau(None, 3, n_ref, None, None)))
            outs = rodas5P.launch(built, y0r.t(), tr, pr.t(), n_ref, dt0=1e-4, rtol=RTOL, atol=ATOL,
                                  max_steps=100000, refill=False)

            def run_ref(built=built, outs=outs):
                return rodas5P.launch(built, y0r.t(), tr, pr.t(), n_ref, dt0=1e-4, rtol=RTOL,
                                      atol=ATOL, max_steps=100000, refill=False, outputs=outs)

            o, wall, _ = time_solves(run_ref, 5, 1, torch)
            acc = int(o.accepted.sum().item())
            ms = wall / 5 * 1e3
            ref[lu] = {"ms": ms, "trajectories_per_s": n_ref / (ms / 1e3),
                       "accepted_steps_per_s": acc / (ms / 1e3),
                       "attempted_per_traj": float((o.accepted + o.rejected).float().mean().item())}
        ref["fp32"]["vs_rtx4070_super"] = ref["fp32"]["trajectories_per_s"] / REF_RTX4070_FP32_TRAJ_PER_S
        ref["fp64"]["vs_rtx4070_super"] = ref["fp64"]["trajectories_per_s"] / REF_RTX4070_FP64_TRAJ_PER_S
        line["reference_config"] = {
            "workload": "100,000 identical Robertson trajectories, y0=(0.891,0.1,0.009), "
                        "p=(0.04,1e4,3e7), 5 saves to 1e5, first_step 1e-4, rtol 1e-6, atol 1e-8",
            "published_rtx4070_super_ms": {"fp32": 35.97, "fp64": 45.59},
            **ref,
        }

    if not args.skip_reference_config:
        line["reference_benchmarks"] = reference_benchmarks(torch, log)
        line["reference_benchmarks"]["robertson_divergence"] = robertson_divergence(torch, log)
    if not args.skip_uneven_stiff:
        line["uneven_stiff"] = uneven_stiff(torch, dev, O, log)
    if not args.skip_adjoint:
        line["adjoint"] = adjoint_block(torch, log)
    line["lorenz_tsit5"] = lorenz_block(torch, dev)
    if not args.no_cpu_baseline:
        line["lorenz_tsit5"]["cpu_baseline_scipy"] = lorenz_scipy
    if not args.skip_vdp_chain:
        line["vdp_chain"] = vdp_chain_block(torch, dev, log)

    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
