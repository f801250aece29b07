"""CPU baseline with SciPy: each trajectory of the ensemble integrated by
``scipy.integrate.solve_ivp``, the trajectories spread over a pool of worker
processes.

Measurement infrastructure, not the product: only bench.py's CPU-baseline leg
runs it. It is this build's counterpart of the reference's SciPy baseline
(reference/solvers/python/scipy_solve_ivp.py:149-255): a ``spawn`` pool of
single-threaded workers (one BLAS/OpenMP thread each, :29-35), ``solve_ivp``
per trajectory with ``t_eval`` at the save times, LSODA by default (:171),
rows an integration did not reach left NaN (:75-98). The reference evaluates a
JAX-compiled right-hand side; JAX is absent here, so the right-hand sides are
NumPy restatements of the benchmark systems below, and the pool runs nothing
but NumPy and SciPy (no GPU library is imported in a worker).
"""

from __future__ import annotations

import multiprocessing as mp
import os
import time

import numpy as np

_SINGLE_THREAD = ("OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "OMP_NUM_THREADS")


def robertson(t, y, p):
    """Robertson with rates p = (k1, k3, k2) as the ensemble stores them."""
    return np.array([-p[0] * y[0] + p[1] * y[1] * y[2],
                     p[0] * y[0] - p[1] * y[1] * y[2] - p[2] * y[1] ** 2,
                     p[2] * y[1] ** 2])


def lorenz(t, y, p):
    """Lorenz with sigma = 10, beta = 8/3 and rho = p[0]."""
    return np.array([10.0 * (y[1] - y[0]), y[0] * (p[0] - y[2]) - y[1],
                     y[0] * y[1] - (8.0 / 3.0) * y[2]])


SYSTEMS = {"robertson": robertson, "lorenz": lorenz}


def _one(task):
    from scipy.integrate import solve_ivp

    name, method, rtol, atol, first_step, times, y0, p = task
    fun = SYSTEMS[name]
    sol = solve_ivp(lambda t, y: fun(t, y, p), (float(times[0]), float(times[-1])), y0,
                    method=method, t_eval=times, rtol=rtol, atol=atol, first_step=first_step)
    out = np.full((times.size, y0.size), np.nan)
    ys = np.asarray(sol.y, dtype=np.float64).T
    out[:ys.shape[0]] = ys
    return out


def _init_worker():
    for key in _SINGLE_THREAD:
        os.environ[key] = "1"


class Pool:
    """A spawn pool of ``workers`` single-threaded processes; started once and
    reused, so the timed solves exclude process start-up and imports."""

    def __init__(self, workers: int):
        self.workers = int(workers)
        saved = {k: os.environ.get(k) for k in _SINGLE_THREAD}
        try:
            for key in _SINGLE_THREAD:
                os.environ[key] = "1"
            self._pool = mp.get_context("spawn").Pool(self.workers, initializer=_init_worker)
        finally:
            for key, val in saved.items():
                if val is None:
                    os.environ.pop(key, None)
                else:
                    os.environ[key] = val

    def solve(self, system, y0, times, params, *, method="LSODA", rtol=1e-6, atol=1e-8,
              first_step=None):
        """``(N, n_save, n_vars)`` for the ensemble ``y0`` ``(N, n_vars)``,
        ``params`` ``(N, n_params)`` at the save times."""
        times = np.asarray(times, dtype=np.float64)
        tasks = [(system, method, rtol, atol, first_step, times,
                  np.asarray(y0[i], np.float64), np.asarray(params[i], np.float64))
                 for i in range(len(y0))]
        chunk = max(1, len(tasks) // (self.workers * 4))
        return np.stack(self._pool.map(_one, tasks, chunksize=chunk))

    def close(self):
        self._pool.terminate()
        self._pool.join()


def timed_sample(pool, system, y0, times, params, budget_s, *, start=256, **kw):
    """Solve growing leading slices of the ensemble until one takes more than
    ``budget_s`` (or the whole ensemble is done), after one untimed warm-up on
    a small slice; returns ``(trajectories solved, seconds, history)``."""
    pool.solve(system, y0[:min(len(y0), 16)], times, params[:min(len(y0), 16)], **kw)
    n = min(start, len(y0))
    while True:
        t0 = time.perf_counter()
        hist = pool.solve(system, y0[:n], times, params[:n], **kw)
        dt = time.perf_counter() - t0
        if dt > budget_s or n >= len(y0):
            return n, dt, hist
        n = min(len(y0), n * 4)
