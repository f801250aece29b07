"""ctypes front end of the CPU oracle (``oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg, always as the checker or the reported CPU
baseline, never as the product. ``menodax_amd`` does not import this module.

Each ``System`` here names one of the oracle's hand-written right-hand sides
(closed-form Jacobian included). The product integrates the matching Python
``ode_fn`` from ``tests/systems.py``; the two descriptions are written
independently, which is what makes a parity test between them mean something.

Parity status of each oracle function (how it is pinned, see the tests named):
  - ``rodas5p`` dense, LU fp32: reference goldens -- Robertson 157 attempted /
    10 rejected exactly (benchmarks/stiff_robertson_divergence csv:2), the
    seeded divergence distributions (csv:3-12) to 2e-5 in the mean, the 96-D
    stiff VdP ring's 21 steps (stiff_vdp_divergence csv:2); closed forms
    (forced decay, chain, Kaps). tests/test_oracle.py
  - ``rodas5p`` dense, LU fp64: closed forms at tight tolerance; convergence order.
  - ``rodas5p`` sparse: against its own dense path on the same system (the
    reference's dense-vs-sparse tests, tests/test_sparse_direct.py:347-382)
    and the closed-form chain (tests/test_sparsity.py:114-128).
  - ``tsit5``: reference goldens -- Lorenz 722 attempted / 104 rejected
    (benchmarks/nonstiff_lorenz_divergence csv:2) and the seeded divergence
    distributions; VdP 64-D ring 12 steps (nonstiff_vdp_divergence csv:2).
  - ``jacobian``: against central differences of ``rhs``.
  - ``rodas5p_callbacks`` (a callback system, optionally carrying forward
    sensitivities): the plain path against ``rodas5p`` on the same system bit
    for bit; the joint path by closed-form sensitivities (decay, cascade,
    forced decay: tests/test_sensitivity.py:42-110 of the reference), by
    central differences of the plain solve, and by the joint solve's state
    being bit-identical to the plain solve when the sensitivities are kept
    out of step-size control (modax/_sensitivity.py:116-127, 275-293). The
    callbacks' derivative code is the code generator's (oracle/callbacks.py),
    pinned on its own in tests/test_codegen.py.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_LIB = None

KIND = dict(
    lorenz=0,
    robertson=1,
    forced_decay=2,
    decay=3,
    chain4=4,
    ring6=5,
    vdp_lattice=6,
    linquad=7,
    damped_osc=8,
    kaps=9,
    bbn=11,
)


def build(force: bool = False) -> str:
    """Compile ``liboracle.so`` with the recipe in ``oracle/Makefile``."""
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(
        _LIB_PATH
    ) < os.path.getmtime(os.path.join(_HERE, "oracle.c")):
        subprocess.run(["make", "-B" if force else "-s", "liboracle.so"], cwd=_HERE, check=True)
    return _LIB_PATH


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _LIB = ctypes.CDLL(_LIB_PATH)
        _LIB.oracle_max_threads.restype = ctypes.c_int
    return _LIB


@dataclass
class System:
    """One of the oracle's hand-written systems and its configuration."""

    kind: str
    n: int
    n_p: int
    n_osc: int = 0
    wrap: int = 0
    coupling_range: int = 1
    mu: float = 0.0
    d: float = 0.0
    omega: float = 0.0
    quad: float = 0.0
    off_rows: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    off_cols: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    off_vals: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    off_pidx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    inv_eps: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float64))

    def _pack(self):
        ints = np.array(
            [KIND[self.kind], self.n, self.n_p, self.n_osc, self.wrap, self.coupling_range,
             len(self.off_rows)],
            dtype=np.int32,
        )
        dbls = np.array([self.mu, self.d, self.omega, self.quad], dtype=np.float64)
        pidx = self.off_pidx if len(self.off_pidx) == len(self.off_rows) else np.full(
            len(self.off_rows), -1)
        off_rc = np.ascontiguousarray(
            np.concatenate([self.off_rows, self.off_cols, pidx]).astype(np.int32)
        )
        if off_rc.size == 0:
            off_rc = np.zeros(1, np.int32)
        off_vals = np.ascontiguousarray(self.off_vals, dtype=np.float64)
        if off_vals.size == 0:
            off_vals = np.zeros(1, np.float64)
        inv_eps = np.ascontiguousarray(self.inv_eps, dtype=np.float64)
        return ints, dbls, off_rc, off_vals, inv_eps


def lorenz():
    return System("lorenz", 3, 1)


def robertson():
    return System("robertson", 3, 3)


def forced_decay():
    return System("forced_decay", 1, 2)


def decay():
    return System("decay", 1, 1)


def chain4():
    return System("chain4", 4, 3)


def ring6():
    return System("ring6", 6, 2)


def damped_osc():
    return System("damped_osc", 2, 2)


def vdp_lattice(n_osc, *, mu=100.0, d=10.0, omega=1.0, coupling_range=1, wrap=True):
    """reference/systems/python/vdp.py:253-290; ``wrap=False`` is the open chain."""
    return System(
        "vdp_lattice", 2 * n_osc, 1, n_osc=n_osc, wrap=int(wrap),
        coupling_range=coupling_range, mu=mu, d=d, omega=omega,
    )


def linquad(n, off_rows, off_cols, off_vals, quad, off_pidx=None, n_p=None):
    """f_j = p_j y_j + sum of off-diagonal terms + quad * y_j*y_j, a term being
    a * y_c, or (a * p_k) * y_c where ``off_pidx`` names a parameter k."""
    rows = np.asarray(off_rows, np.int32)
    return System(
        "linquad", n, n if n_p is None else n_p, quad=quad,
        off_rows=rows, off_cols=np.asarray(off_cols, np.int32),
        off_vals=np.asarray(off_vals, np.float64),
        off_pidx=np.full(len(rows), -1, np.int32) if off_pidx is None
        else np.asarray(off_pidx, np.int32),
    )


def bbn():
    """The four-species BBN network (examples/bbn_estimation/main.py:74-161),
    x = Q/T the independent variable, parameters (log10 eta_10, N_eff);
    derivatives by the complex step."""
    return System("bbn", 4, 2)


def kaps(n_pairs, epsilon_min):
    """reference/systems/python/kaps.py:26-48 (one scale parameter p[0])."""
    eps = np.array(
        [10.0 ** (np.log10(epsilon_min) * i / (n_pairs - 1)) for i in range(n_pairs)]
    )
    return System("kaps", 2 * n_pairs, 1, inv_eps=1.0 / eps)


def sparse_structure(pattern, order):
    """The sparse structure ``rodas5p(..., layout=)`` takes, built here from the
    pattern and an elimination order alone, independently of the product's
    symbolic analysis. Restates the reference's:
      - fill by Gaussian elimination on the boolean pattern, right-looking,
        diagonal forced present (modax/_sparse_direct.py:155-184), on the
        pattern permuted *before* the fill (:321-329);
      - CSR of L+U in elimination order with ascending columns and the
        diagonal's position per row (:330-349);
      - the caller (row, col) -> slot table (:226-233).
    ``pattern`` is a mask or rows of column indices; ``order[i]`` is the caller
    variable eliminated i-th. Returns ``(nnz, row_ptr, col_ind, diag_ptr,
    order, slot_of (n*n, caller indices), mask (n*n))``.

    The order is an input: the AMD permutation itself is parity-unpinned (the
    reference's SuiteSparse AMD is unavailable), so the tests hand the oracle
    the product's permutation and pin the fill, layout and arithmetic here."""
    if isinstance(pattern, np.ndarray) and pattern.ndim == 2 and pattern.dtype == bool:
        mask = pattern.copy()
    else:
        n0 = len(pattern)
        mask = np.zeros((n0, n0), dtype=bool)
        for r, cols in enumerate(pattern):
            mask[r, list(cols)] = True
    n = mask.shape[0]
    perm = np.asarray(order, dtype=np.int64)
    if sorted(perm.tolist()) != list(range(n)):
        raise ValueError("order is not a permutation")
    work = mask[np.ix_(perm, perm)].copy()
    work[np.arange(n), np.arange(n)] = True
    for k in range(n):
        rows = np.flatnonzero(work[k + 1:, k]) + k + 1
        work[rows, k + 1:] |= work[k, k + 1:]
    counts = work.sum(axis=1)
    row_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    col_ind = np.nonzero(work)[1].astype(np.int32)  # row-major: ascending columns per row
    diag_ptr = (row_ptr[:-1] + np.array([work[i, :i].sum() for i in range(n)])).astype(np.int32)
    slot_of = np.full((n, n), -1, dtype=np.int32)
    er, ec = np.nonzero(work)
    slot_of[perm[er], perm[ec]] = np.arange(er.size, dtype=np.int32)
    return (int(row_ptr[-1]), row_ptr, col_ind, diag_ptr, perm.astype(np.int32),
            slot_of.ravel(), mask.astype(np.int32).ravel())


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _ensemble(system, y0, params, weights):
    y0 = np.ascontiguousarray(np.atleast_2d(np.asarray(y0, np.float64)))
    params = np.ascontiguousarray(np.atleast_2d(np.asarray(params, np.float64)))
    n = max(y0.shape[0], params.shape[0])
    if y0.shape[0] != n:
        y0 = np.ascontiguousarray(np.broadcast_to(y0, (n, y0.shape[1])))
    if params.shape[0] != n:
        params = np.ascontiguousarray(np.broadcast_to(params, (n, params.shape[1])))
    if weights is None:
        weights = np.ones((n, system.n))
    weights = np.ascontiguousarray(np.broadcast_to(np.asarray(weights, np.float64), (n, system.n)))
    return n, y0, params, weights


def rodas5p(system, y0, times, params, *, weights=None, first_step=None, rtol=1e-8,
            atol=1e-10, max_steps=100000, lu_precision="fp32", pcoeff=0.0, icoeff=1.0,
            dcoeff=0.0, tf_index=None, layout=None, threads=0, lu_refine=False):
    """Ensemble Rodas5P on the CPU. ``layout`` is a sparse ``(nnz, row_ptr, col_ind,
    diag_ptr, order, slot_of, pattern)`` tuple, or ``None`` for the dense LU.
    ``lu_refine`` (FP32 LU only): one step of iterative refinement of every stage
    solve against the double iteration matrix, stage vectors kept in double."""
    if lu_refine and lu_precision != "fp32":
        raise ValueError("lu_refine refines FP32 solves")
    lib = _lib()
    n_traj, y0, params, weights = _ensemble(system, y0, params, weights)
    times = np.ascontiguousarray(times, dtype=np.float64)
    n = system.n
    hist = np.zeros((n_traj, times.size, n))
    acc = np.zeros(n_traj, np.int32)
    rej = np.zeros(n_traj, np.int32)
    loop = np.zeros(n_traj, np.int32)
    ints, dbls, off_rc, off_vals, inv_eps = system._pack()
    if layout is None:
        nnz = 0
        dummy = np.zeros(1, np.int32)
        row_ptr = col_ind = diag_ptr = order = slot_of = pattern = dummy
    else:
        nnz, row_ptr, col_ind, diag_ptr, order, slot_of, pattern = layout
        row_ptr, col_ind, diag_ptr, order, slot_of, pattern = (
            np.ascontiguousarray(a, dtype=np.int32)
            for a in (row_ptr, col_ind, diag_ptr, order, slot_of, pattern)
        )
    I, D = ctypes.c_int, ctypes.c_double
    pi = lambda a: _ptr(a, ctypes.c_int)  # noqa: E731
    pd = lambda a: _ptr(a, ctypes.c_double)  # noqa: E731
    rc = lib.oracle_rodas5p(
        pi(ints), pd(dbls), pi(off_rc), pd(off_vals), pd(inv_eps), I(n_traj), pd(y0),
        pd(times), I(times.size), pd(params), pd(weights),
        D(0.0 if first_step is None else float(first_step)), D(rtol), D(atol), I(max_steps),
        I((2 if lu_refine else 1) if lu_precision == "fp32" else 0), D(pcoeff), D(icoeff),
        D(dcoeff), I(-1 if tf_index is None else int(tf_index)), I(int(nnz)), pi(row_ptr),
        pi(col_ind),
        pi(diag_ptr), pi(order), pi(slot_of), pi(pattern), pd(hist), pi(acc), pi(rej),
        pi(loop), I(int(threads)),
    )
    if rc != 0:
        raise ValueError(f"oracle rejected the sparse structure (code {rc})")
    return hist, dict(accepted_steps=acc, rejected_steps=rej, loop_steps=loop)


class ClosedFormCallbacks:
    """The hand-derived closed-form callbacks in ``oracle.c`` (f, J, the
    sensitivity rows J_y s + J_p q, their df/dt and the coupling
    D_y(J_y s + J_p q)[v]) for ``"robertson"`` (3 variables, 3 rates) or
    ``"vdp"`` (one Van der Pol oscillator, mu = 1.0 * p[0]), in the form
    ``rodas5p_callbacks`` takes: an independent stand-in for the derivative
    graphs of ``callbacks.HostCallbacks``."""

    SIZES = {"robertson": (3, 3), "vdp": (2, 1)}

    def __init__(self, name):
        lib = _lib()
        self.n, self.n_p = self.SIZES[name]
        dt = "zero_sens_dt" if name == "robertson" else "vdp_sens_dt"
        names = (f"{name}_f", f"{name}_jac", f"{name}_sens_f", dt, f"{name}_sens_cpl")
        self.addresses = tuple(ctypes.cast(getattr(lib, f"oracle_cf_{x}"), ctypes.c_void_p).value
                               for x in names)


def rodas5p_callbacks(cb, y0, times, params, *, direction_columns=(), error_control=True,
                      weights=None, first_step=None, rtol=1e-8, atol=1e-10, max_steps=100000,
                      lu_precision="fp32", pcoeff=0.0, icoeff=1.0, dcoeff=0.0, tf_index=None,
                      layout=None, threads=0):
    """Ensemble Rodas5P of a ``callbacks.HostCallbacks`` system. With
    ``direction_columns`` (per direction the parameter it seeds, -1 for an
    initial-state direction) it integrates the joint ``[y, S]`` system:
    ``y0`` (and ``weights``) must then be ``n * (1 + n_dirs)`` wide -- the
    caller builds ``S(t0)`` -- and the history comes back that wide."""
    lib = _lib()
    n, n_p = cb.n, cb.n_p
    cols = np.ascontiguousarray(list(direction_columns) or [0], dtype=np.int32)
    n_sens = len(direction_columns)
    na = n * (1 + n_sens)
    y0 = np.ascontiguousarray(np.atleast_2d(np.asarray(y0, np.float64)))
    params = np.ascontiguousarray(np.atleast_2d(np.asarray(params, np.float64)))
    n_traj = max(y0.shape[0], params.shape[0])
    y0 = np.ascontiguousarray(np.broadcast_to(y0, (n_traj, na)))
    params = np.ascontiguousarray(np.broadcast_to(params, (n_traj, n_p)))
    if weights is None:
        weights = np.ones(na)
        if not error_control:
            weights[n:] = 0.0
    weights = np.ascontiguousarray(np.broadcast_to(np.asarray(weights, np.float64), (n_traj, na)))
    times = np.ascontiguousarray(times, dtype=np.float64)
    hist = np.zeros((n_traj, times.size, na))
    acc = np.zeros(n_traj, np.int32)
    rej = np.zeros(n_traj, np.int32)
    loop = np.zeros(n_traj, np.int32)
    if layout is None:
        nnz = 0
        dummy = np.zeros(1, np.int32)
        row_ptr = col_ind = diag_ptr = order = slot_of = pattern = dummy
    else:
        nnz, row_ptr, col_ind, diag_ptr, order, slot_of, pattern = layout
        row_ptr, col_ind, diag_ptr, order, slot_of, pattern = (
            np.ascontiguousarray(a, dtype=np.int32)
            for a in (row_ptr, col_ind, diag_ptr, order, slot_of, pattern)
        )
    I, D, V = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    pi = lambda a: _ptr(a, ctypes.c_int)  # noqa: E731
    pd = lambda a: _ptr(a, ctypes.c_double)  # noqa: E731
    rc = lib.oracle_rodas5p_cb(
        *(V(a) for a in cb.addresses), I(n), I(n_p), I(n_sens), pi(cols),
        I(na if error_control else n), I(n_traj), pd(y0), pd(times), I(times.size), pd(params),
        pd(weights), D(0.0 if first_step is None else float(first_step)), D(rtol), D(atol),
        I(max_steps), I(1 if lu_precision == "fp32" else 0), D(pcoeff), D(icoeff), D(dcoeff),
        I(-1 if tf_index is None else int(tf_index)), I(int(nnz)), pi(row_ptr), pi(col_ind),
        pi(diag_ptr), pi(order), pi(slot_of), pi(pattern), pd(hist), pi(acc), pi(rej),
        pi(loop), I(int(threads)),
    )
    if rc != 0:
        raise ValueError(f"oracle rejected the call (code {rc})")
    return hist, dict(accepted_steps=acc, rejected_steps=rej, loop_steps=loop)


def tsit5(system, y0, times, params, *, weights=None, first_step=None, rtol=1e-8, atol=1e-10,
          max_steps=100000, pcoeff=0.0, icoeff=1.0, dcoeff=0.0, threads=0, fill=np.nan):
    """Ensemble Tsit5 on the CPU; saves never reached hold ``fill``."""
    lib = _lib()
    n_traj, y0, params, weights = _ensemble(system, y0, params, weights)
    times = np.ascontiguousarray(times, dtype=np.float64)
    hist = np.full((n_traj, times.size, system.n), fill)
    acc = np.zeros(n_traj, np.int32)
    rej = np.zeros(n_traj, np.int32)
    loop = np.zeros(n_traj, np.int32)
    ints, dbls, off_rc, off_vals, inv_eps = system._pack()
    I, D = ctypes.c_int, ctypes.c_double
    pi = lambda a: _ptr(a, ctypes.c_int)  # noqa: E731
    pd = lambda a: _ptr(a, ctypes.c_double)  # noqa: E731
    lib.oracle_tsit5(
        pi(ints), pd(dbls), pi(off_rc), pd(off_vals), pd(inv_eps), I(n_traj), pd(y0), pd(times),
        I(times.size), pd(params), pd(weights), D(0.0 if first_step is None else float(first_step)),
        D(rtol), D(atol), I(max_steps), D(pcoeff), D(icoeff), D(dcoeff), pd(hist), pi(acc),
        pi(rej), pi(loop), I(int(threads)),
    )
    return hist, dict(accepted_steps=acc, rejected_steps=rej, loop_steps=loop)


def rhs(system, y, t, p):
    lib = _lib()
    ints, dbls, off_rc, off_vals, inv_eps = system._pack()
    y = np.ascontiguousarray(y, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    f = np.zeros(system.n)
    lib.oracle_rhs(
        _ptr(ints, ctypes.c_int), _ptr(dbls, ctypes.c_double), _ptr(off_rc, ctypes.c_int),
        _ptr(off_vals, ctypes.c_double), _ptr(inv_eps, ctypes.c_double),
        _ptr(y, ctypes.c_double), ctypes.c_double(t), _ptr(p, ctypes.c_double),
        _ptr(f, ctypes.c_double),
    )
    return f


def jacobian(system, y, t, p):
    """Closed-form ``(J, df/dt)`` of the oracle's system."""
    lib = _lib()
    ints, dbls, off_rc, off_vals, inv_eps = system._pack()
    y = np.ascontiguousarray(y, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    J = np.zeros((system.n, system.n))
    dT = np.zeros(system.n)
    lib.oracle_jacobian(
        _ptr(ints, ctypes.c_int), _ptr(dbls, ctypes.c_double), _ptr(off_rc, ctypes.c_int),
        _ptr(off_vals, ctypes.c_double), _ptr(inv_eps, ctypes.c_double),
        _ptr(y, ctypes.c_double), ctypes.c_double(t), _ptr(p, ctypes.c_double),
        _ptr(J, ctypes.c_double), _ptr(dT, ctypes.c_double),
    )
    return J, dT


def max_threads() -> int:
    return int(_lib().oracle_max_threads())
