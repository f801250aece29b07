"""Reference test systems and their seeded ensembles, restated.

The right-hand sides here are ordinary Python callbacks in the form ``solve``
accepts (a tuple of scalars from constant indices), written with the same
arithmetic order as the reference's own systems so that generated HIP code
evaluates them the same way. The ensemble generators reproduce the reference's
``make_scenario`` draws call for call with ``numpy.random.default_rng(seed)``,
so the same seed gives the same trajectories.

Sources (reference repository):
  - Lorenz:        reference/systems/python/lorenz.py:136-209
  - Robertson:     reference/systems/python/robertson.py:6-109
  - VdP lattice:   reference/systems/python/vdp.py:224-359
  - Kaps:          reference/systems/python/kaps.py:23-60
  - test callbacks: tests/test_solver_options.py:142-147,
    tests/test_sparsity.py:36-61, tests/test_sparse_direct.py:333-342,
    tests/test_save_hook.py:23-24

The benchmark ensembles that have no reference generator (the uneven stiff
ensemble, the VdP chain with log-uniform mu) are defined here with fixed seeds.
"""

from __future__ import annotations

import math

import numpy as np

# --------------------------------------------------------------------------
# Lorenz (lorenz.py)
# --------------------------------------------------------------------------
LORENZ_TIMES = (0.0, 5.0)
LORENZ_Y0 = np.array([1000.0, -500.0, 500.0])
LORENZ_PARAMS = np.array([0.5])


def lorenz(y, t, p):
    sigma = 10.0
    beta = 8.0 / 3.0
    return (
        sigma * (y[1] - y[0]),
        y[0] * (p[0] - y[2]) - y[1],
        y[0] * y[1] - beta * y[2],
    )


def lorenz_params(size, seed=42):
    rng = np.random.default_rng(seed)
    return 0.5 * (1.0 + 0.05 * (2.0 * rng.random((size, 1)) - 1.0))


def lorenz_scenario(size, seed=42, *, divergence=1.0):
    """lorenz.py:170-209: radial ICs ``Y0 * (1 - t)``, t ~ U(0, divergence)."""
    divergence = _check_divergence(divergence)
    params = LORENZ_PARAMS + divergence * (lorenz_params(size, seed) - LORENZ_PARAMS)
    rng = np.random.default_rng(seed)
    t = rng.uniform(0.0, divergence, size=(size, 1))
    return LORENZ_Y0 * (1.0 - t), params


def lorenz_rho_ensemble(size, seed=0):
    """Tsit5 benchmark ensemble: rho ~ U[0, 28], y0 = (1, 0, 0), t in [0, 1]."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.0, 28.0, size=(size, 1))
    y0 = np.tile(np.array([1.0, 0.0, 0.0]), (size, 1))
    return y0, rho


# --------------------------------------------------------------------------
# Robertson (robertson.py)
# --------------------------------------------------------------------------
ROBERTSON_TIMES = np.array((0.0, 1e-6, 1e-2, 1e2, 1e5))
_ALPHA = 0.99
_EPS = 0.1
ROBERTSON_Y0 = np.array([(1 - _EPS) * _ALPHA, _EPS, (1 - _EPS) * (1 - _ALPHA)])
#: (k1, k3, k2): the reference orders the rates this way (robertson.py:13, 17-22).
ROBERTSON_PARAMS = np.array([0.04, 1e4, 3e7])


def robertson(y, t, p):
    return (
        -p[0] * y[0] + p[1] * y[1] * y[2],
        p[0] * y[0] - p[1] * y[1] * y[2] - p[2] * y[1] ** 2,
        p[2] * y[1] ** 2,
    )


def robertson_array(y, t, p, out):
    """Robertson as an array-form right-hand side (``array_rhs``): the same
    expressions in the same order as ``robertson``, written into ``out``."""
    out[0] = -p[0] * y[0] + p[1] * y[1] * y[2]
    out[1] = p[0] * y[0] - p[1] * y[1] * y[2] - p[2] * y[1] ** 2
    out[2] = p[2] * y[1] ** 2


def robertson_initial_conditions(size, seed=42, *, divergence=1.0):
    divergence = _check_divergence(divergence)
    if divergence == 0.0:
        return np.broadcast_to(ROBERTSON_Y0, (size, 3)).copy()
    rng = np.random.default_rng(seed)
    alpha = _ALPHA * rng.uniform(0.0, 1.0, size)
    eps_blend = rng.uniform(0.0, min(divergence / 4.0, 1.0), size)
    eps = 10.0 ** ((1.0 - eps_blend) * np.log10(_EPS) + eps_blend * np.log10(1e-8))
    return np.column_stack([(1 - eps) * alpha, eps, (1 - eps) * (1 - alpha)])


def robertson_scenario(size, seed=42, *, divergence=1.0):
    """robertson.py:86-109: log-uniform rate scales and mass-conserving ICs."""
    divergence = _check_divergence(divergence)
    if divergence == 0.0:
        params = np.broadcast_to(ROBERTSON_PARAMS, (size, 3)).copy()
    else:
        rng = np.random.default_rng(seed)
        scales = 10.0 ** rng.uniform(-divergence, divergence, size=(size, 3))
        params = ROBERTSON_PARAMS * scales
    return robertson_initial_conditions(size, seed, divergence=divergence), params


def robertson_lognormal_ensemble(size, seed=1, sigma=0.3):
    """Stiff Robertson benchmark ensemble: k1 = 0.04 e^u1, k2 = 3e7 e^u2,
    k3 = 1e4 e^u3 with u ~ N(0, sigma^2); y0 = (1, 0, 0); returned in the
    reference's (k1, k3, k2) parameter order."""
    rng = np.random.default_rng(seed)
    u = rng.normal(0.0, sigma, size=(size, 3))
    k1 = 0.04 * np.exp(u[:, 0])
    k2 = 3e7 * np.exp(u[:, 1])
    k3 = 1e4 * np.exp(u[:, 2])
    y0 = np.tile(np.array([1.0, 0.0, 0.0]), (size, 1))
    return y0, np.column_stack([k1, k3, k2])


def robertson_benchmark_times():
    """[0] and 49 log-spaced save times from 1e-5 to 1e5 (50 saves)."""
    return np.concatenate([[0.0], np.logspace(-5.0, 5.0, 49)])


# --------------------------------------------------------------------------
# Van der Pol lattice (vdp.py), ring or open chain
# --------------------------------------------------------------------------
VDP_TIMES = np.array((0.0, 0.25, 0.5, 0.75, 1.0))


def vdp_neighbours(n_osc, osc, coupling_range=1, wrap=True):
    """vdp.py:235-250; an open chain drops the neighbours past either end."""
    if coupling_range < 1:
        raise ValueError("coupling_range must be at least 1")
    found = []
    for offset in range(1, coupling_range + 1):
        for other in (osc - offset, osc + offset):
            if wrap:
                other %= n_osc
            elif not 0 <= other < n_osc:
                continue
            if other != osc:
                found.append(other)
    return found


def make_tuple_callback(name, components):
    """``def name(y, t, p): return ((c0) + 0.0, ...)`` from expression strings,
    the form the reference's reference systems are written in
    (reference/systems/python/_tuple_codegen.py:17-30)."""
    body = ",\n        ".join(f"({expr}) + 0.0" for expr in components)
    source = f"def {name}(y, t, p):\n    return (\n        {body},\n    )\n"
    namespace: dict = {}
    exec(compile(source, f"<generated {name}>", "exec"), namespace)  # noqa: S102
    fn = namespace[name]
    fn._generated_source = source
    return fn


def vdp_lattice(n_osc, *, mu=100.0, d=10.0, omega=1.0, coupling_range=1, wrap=True):
    """vdp.py:253-290; ``wrap=False`` gives the open chain."""
    y0 = np.array([2.0, 0.0] * n_osc)
    strength = d / coupling_range
    components = []
    for osc in range(n_osc):
        base = 2 * osc
        x, v = f"y[{base}]", f"y[{base + 1}]"
        others = vdp_neighbours(n_osc, osc, coupling_range, wrap)
        laplacian = " + ".join(
            [f"y[{2 * other}]" for other in others] + [f"-{float(len(others))!r} * {x}"]
        )
        components += [
            v,
            f"p[0] * {mu!r} * (1.0 - {x} * {x}) * {v}"
            f" + -{omega * omega!r} * {x}"
            f" + {strength!r} * ({laplacian})",
        ]
    return make_tuple_callback("ode_fn", components), y0


def vdp_sparsity(n_osc, coupling_range=1, wrap=True):
    """vdp.py:293-307."""
    n_vars = 2 * n_osc
    mask = np.zeros((n_vars, n_vars), dtype=bool)
    for osc in range(n_osc):
        x, v = 2 * osc, 2 * osc + 1
        mask[x, v] = True
        mask[v, x] = mask[v, v] = True
        for other in vdp_neighbours(n_osc, osc, coupling_range, wrap):
            mask[v, 2 * other] = True
    return mask


def vdp_params(size, seed=42):
    rng = np.random.default_rng(seed)
    return np.array(1.0 + 0.2 * (2.0 * rng.random((size, 1)) - 1.0))


def vdp_scenario(n_osc, size, seed=42, *, divergence=1.0):
    """vdp.py:325-359."""
    n_vars = 2 * n_osc
    divergence = _check_divergence(divergence)
    rng = np.random.default_rng(seed)
    amplitudes = rng.uniform(0.25, 3.0, size=(size, n_osc))
    signs = rng.choice(np.array([-1.0, 1.0]), size=(size, n_osc))
    target_x = amplitudes * signs
    position_divergence = min(divergence, 1.0)
    x = 2.0 + position_divergence * (target_x - 2.0)
    v = rng.normal(0.0, 2.0 * divergence, size=(size, n_osc))
    y0 = np.empty((size, n_vars))
    y0[:, 0::2] = x
    y0[:, 1::2] = v
    base_params = vdp_params(size, seed)
    param_center = max(divergence, 1.0)
    params = np.maximum(param_center + divergence * (base_params - 1.0), 1e-6)
    return y0, params


def vdp_chain_ensemble(n_osc, size, seed=3, mu_range=(10.0, 1000.0)):
    """Stiff VdP chain benchmark ensemble: mu ~ LogUniform[mu_range] carried in
    p[0] (the lattice is built with mu=1.0, coupling 0.1); y0 ~ U[-2, 2]."""
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.uniform(np.log(mu_range[0]), np.log(mu_range[1]), size=(size, 1)))
    y0 = rng.uniform(-2.0, 2.0, size=(size, 2 * n_osc))
    return y0, mu


# --------------------------------------------------------------------------
# Small test callbacks from the reference's tests
# --------------------------------------------------------------------------
def forced_decay(y, t, p):
    """tests/test_solver_options.py:142-143."""
    return (-p[0] * y[0] + p[1] * t,)


def forced_decay_exact(t, lam, forcing):
    return forcing * (lam * t - 1.0 + np.exp(-lam * t)) / lam**2


def cascade(y, t, p):
    """Linear cascade y0' = -a y0, y1' = a y0 - b y1 (tests/test_sensitivity.py:42-43):
    closed-form solution and sensitivities in both rates and both initial states."""
    return (-p[0] * y[0], p[0] * y[0] - p[1] * y[1])


def cascade_exact(t, y0, p):
    """``y(t)`` as ``(n_save, 2)``."""
    a, b = p
    c0, c1 = y0
    k = a / (b - a)
    return np.stack([c0 * np.exp(-a * t), (c1 - k * c0) * np.exp(-b * t) + k * c0 * np.exp(-a * t)],
                    axis=-1)


def cascade_dparams(t, y0, p):
    """``dy/d(a, b)`` as ``(n_save, 2, 2)``: differentiate ``cascade_exact``."""
    a, b = p
    c0, c1 = y0
    k = a / (b - a)
    dk_da, dk_db = b / (b - a) ** 2, -a / (b - a) ** 2
    ea, eb = np.exp(-a * t), np.exp(-b * t)
    d_da = np.stack([-t * c0 * ea, c0 * (dk_da * (ea - eb) - k * t * ea)], axis=-1)
    d_db = np.stack([np.zeros_like(t), c0 * dk_db * (ea - eb) - t * (c1 - k * c0) * eb], axis=-1)
    return np.stack([d_da, d_db], axis=-1)


def cascade_dy0(t, y0, p):
    """``dy/d(c0, c1)`` as ``(n_save, 2, 2)``."""
    a, b = p
    k = a / (b - a)
    ea, eb = np.exp(-a * t), np.exp(-b * t)
    return np.stack([np.stack([ea, k * (ea - eb)], axis=-1),
                     np.stack([np.zeros_like(t), eb], axis=-1)], axis=-1)


def forced_decay_dlam(t, lam, forcing):
    """d/dlam of ``forced_decay_exact``."""
    return forcing * ((t - t * np.exp(-lam * t)) * lam
                      - 2.0 * (lam * t - 1.0 + np.exp(-lam * t))) / lam**3


def forced_decay_dforcing(t, lam, _forcing):
    """d/dF of ``forced_decay_exact``."""
    return (lam * t - 1.0 + np.exp(-lam * t)) / lam**2


def decay(y, t, p):
    return (-p[0] * y[0],)


def chain4(y, t, p):
    """tests/test_sparsity.py:40-46."""
    return (
        -p[0] * y[0],
        p[0] * y[0] - p[1] * y[1],
        p[1] * y[1] - p[2] * y[2],
        p[2] * y[2],
    )


CHAIN4_MASK = np.zeros((4, 4), dtype=bool)
for _r, _c in [(0, 0), (1, 0), (1, 1), (2, 1), (2, 2), (3, 2)]:
    CHAIN4_MASK[_r, _c] = True


def chain4_exact(t, a=0.7, b=0.4, c=0.2):
    """tests/test_sparsity.py:49-61."""
    y0 = np.exp(-a * t)
    y1 = a / (b - a) * (np.exp(-a * t) - np.exp(-b * t))
    y2 = a * b * (
        np.exp(-a * t) / ((b - a) * (c - a))
        + np.exp(-b * t) / ((a - b) * (c - b))
        + np.exp(-c * t) / ((a - c) * (b - c))
    )
    return np.array([y0, y1, y2, 1.0 - y0 - y1 - y2])


def ring6(y, t, p):
    """tests/test_sparse_direct.py:333-342: six-point periodic diffusion with decay."""
    return (
        p[0] * (y[1] - 2.0 * y[0] + y[5]) - p[1] * y[0],
        p[0] * (y[2] - 2.0 * y[1] + y[0]) - p[1] * y[1],
        p[0] * (y[3] - 2.0 * y[2] + y[1]) - p[1] * y[2],
        p[0] * (y[4] - 2.0 * y[3] + y[2]) - p[1] * y[3],
        p[0] * (y[5] - 2.0 * y[4] + y[3]) - p[1] * y[4],
        p[0] * (y[0] - 2.0 * y[5] + y[4]) - p[1] * y[5],
    )


def ring_mask(n):
    mask = np.zeros((n, n), dtype=bool)
    for i in range(n):
        mask[i, i] = mask[i, (i + 1) % n] = mask[i, (i - 1) % n] = True
    return mask


def damped_oscillator(y, t, p):
    """tests/test_save_hook.py:23-24."""
    return (y[1], -p[0] * y[0] - p[1] * y[1])


def trapezoid_hook(t_span):
    """A save hook integrating the state over the save grid by the trapezoid
    rule, and counting the saves (tests/test_save_hook.py:48-52)."""
    t = np.asarray(t_span, dtype=np.float64)
    w = np.zeros_like(t)
    w[1:] += 0.5 * np.diff(t)
    w[:-1] += 0.5 * np.diff(t)

    def integrate_state(save_idx, y, t_, p_row, acc):
        acc[0] += w[save_idx] * y[0]
        acc[1] += w[save_idx] * y[1]
        acc[2] += 1.0

    return integrate_state


def energy_hook(save_idx, y, t, p_row, acc):
    """A save hook storing the oscillator's energy per save (tests/test_save_hook.py:80-81)."""
    acc[save_idx] = 0.5 * (y[1] * y[1] + p_row[0] * y[0] * y[0])


def kaps(n_pairs, epsilon_min):
    """kaps.py:26-48; exact solution y1 = exp(-2t), y2 = exp(-t) for p[0] = 1."""
    epsilon = [10.0 ** (np.log10(epsilon_min) * i / (n_pairs - 1)) for i in range(n_pairs)]
    components = []
    for pair, eps in enumerate(epsilon):
        base = 2 * pair
        inv_eps = 1.0 / float(eps)
        y1, y2 = f"y[{base}]", f"y[{base + 1}]"
        components += [
            f"p[0] * (-{inv_eps + 2.0!r} * {y1} + {inv_eps!r} * ({y2} * {y2}))",
            f"p[0] * (({y1} - {y2}) - {y2} * {y2})",
        ]
    fn = make_tuple_callback("ode_fn", components)
    return fn, np.array([1.0, 1.0] * n_pairs)


# --------------------------------------------------------------------------
# The uneven stiff ensemble: dy/dt = A y + q (y*y), A block-arrow
# --------------------------------------------------------------------------
def arrow_structure(n=50, seed=7):
    """The block-arrow matrix of the uneven stiff ensemble.

    Variables 0..n-3 pair into 2x2 blocks ``[[a, w], [-w, b]]``, variable n-2
    stands alone and n-1 is the hub, which every other variable feeds (the hub
    row is dense; the hub feeds nothing back). The diagonal entries and the
    block rotation rates ``w`` are per-trajectory parameters: p[0..n-1] the
    diagonal, p[n + k] the rate of block k. A is block lower triangular, so its
    eigenvalues are those of the blocks, (a + b)/2 +- sqrt(((a - b)/2)^2 - w^2),
    and the hub's diagonal. The declared sparsity mask is the full block arrow
    (hub row and hub column), a superset of the true pattern.

    Returns ``(rows, cols, vals, pidx, mask)`` with one term per off-diagonal
    entry: ``vals * y[col]``, or ``(vals * p[pidx]) * y[col]`` when pidx >= 0.
    """
    rng = np.random.default_rng(seed)
    rows, cols, vals, pidx = [], [], [], []
    n_blocks = (n - 2) // 2
    for k in range(n_blocks):
        b = 2 * k
        rows += [b, b + 1]
        cols += [b + 1, b]
        vals += [1.0, -1.0]
        pidx += [n + k, n + k]
    for j in range(n - 1):
        rows.append(n - 1)
        cols.append(j)
        vals.append(float(rng.uniform(-0.2, 0.2)))
        pidx.append(-1)
    mask = np.eye(n, dtype=bool)
    for k in range(n_blocks):
        b = 2 * k
        mask[b, b + 1] = mask[b + 1, b] = True
    mask[n - 1, :] = True
    mask[:, n - 1] = True
    return (np.array(rows), np.array(cols), np.array(vals), np.array(pidx), mask)


def arrow_system(n=50, quad=0.01, seed=7):
    """``ode_fn`` for dy/dt = A(p) y + quad * y*y over the block arrow; the
    parameter row has n diagonal entries then (n - 2) // 2 block rates."""
    rows, cols, vals, pidx, mask = arrow_structure(n, seed)
    terms = [[f"p[{j}] * y[{j}]"] for j in range(n)]
    for r, c, v, k in zip(rows, cols, vals, pidx):
        if k >= 0:
            terms[r].append(("" if v > 0 else "-") + f"p[{k}] * y[{c}]")
        else:
            terms[r].append(f"{float(v)!r} * y[{c}]")
    components = [
        " + ".join(t) + f" + {float(quad)!r} * (y[{j}] * y[{j}])" for j, t in enumerate(terms)
    ]
    fn = make_tuple_callback("arrow", components)
    return fn, mask, (rows, cols, vals, pidx)


def arrow_n_params(n=50):
    return n + (n - 2) // 2


def arrow_ensemble(size, n=50, seed=11, lam_range=(1e-1, 1e5), rate_range=(0.05, 200.0)):
    """Per-trajectory parameters of the uneven stiff ensemble.

    Real parts: each trajectory draws a stiffness ceiling uniformly between
    1e1 and 1e5 in log scale and its n diagonal entries as
    -LogUniform[lam_range[0], ceiling]. Rotation: each trajectory draws one rate
    scale LogUniform over ``rate_range`` and each block's rate as that scale
    times U[0.5, 1.5]. The step count of a trajectory grows with the number of
    oscillations its slowly decaying blocks go through in the window, which is
    what spreads the ensemble's step counts over two decades. y0 ~ U[-1, 1].
    """
    rng = np.random.default_rng(seed)
    n_blocks = (n - 2) // 2
    lo = math.log10(lam_range[0])
    hi = math.log10(lam_range[1])
    ceiling = rng.uniform(lo + 2.0, hi, size=(size, 1))
    u = rng.uniform(0.0, 1.0, size=(size, n))
    lam = -(10.0 ** (lo + u * (ceiling - lo)))
    scale = np.exp(rng.uniform(np.log(rate_range[0]), np.log(rate_range[1]), size=(size, 1)))
    rates = scale * rng.uniform(0.5, 1.5, size=(size, n_blocks))
    y0 = rng.uniform(-1.0, 1.0, size=(size, n))
    return y0, np.concatenate([lam, rates], axis=1)


def _check_divergence(divergence):
    if not np.isfinite(divergence) or divergence < 0.0:
        raise ValueError("divergence must be finite and non-negative")
    return float(divergence)


# --------------------------------------------------------------------------
# A bordered-block system whose sparse LU exceeds the unrolling thresholds
# --------------------------------------------------------------------------
def bordered_structure(core=45, tails=7, length=15, seed=5):
    """A dense ``core`` x ``core`` block with ``tails`` tridiagonal chains of
    ``length`` variables hanging off it (chain t's first variable coupled both
    ways to core variable t % core): the shape of the reference's
    Einstein-Boltzmann-like bordered patterns (tests/test_sparse_direct.py:
    175-190), with a core large enough that one factorisation takes more than
    8,192 multiply-add updates and one solve more than 2,048 substitutions, so
    the solver is emitted table-driven without lowering any threshold.

    The diagonal is per trajectory (p[0..n-1], negative); off-diagonal entries
    are fixed, small enough that A stays diagonally dominant. Returns
    ``(rows, cols, vals, mask)`` of the off-diagonal terms ``vals * y[col]``."""
    rng = np.random.default_rng(seed)
    n = core + tails * length
    mask = np.eye(n, dtype=bool)
    mask[:core, :core] = True
    for t in range(tails):
        base = core + t * length
        for i in range(length):
            if i:
                mask[base + i, base + i - 1] = mask[base + i - 1, base + i] = True
        mask[base, t % core] = mask[t % core, base] = True
    rows, cols = np.nonzero(mask & ~np.eye(n, dtype=bool))
    vals = rng.uniform(-1.0, 1.0, size=rows.size) / np.sqrt(core)
    return rows, cols, vals, mask


def bordered_system(core=45, tails=7, length=15, quad=0.01, seed=5):
    """``ode_fn`` of dy/dt = A(p) y + quad * y*y over the bordered pattern."""
    rows, cols, vals, mask = bordered_structure(core, tails, length, seed)
    n = len(mask)
    terms = [[f"p[{j}] * y[{j}]"] for j in range(n)]
    for r, c, v in zip(rows, cols, vals):
        terms[r].append(f"{float(v)!r} * y[{c}]")
    components = [" + ".join(t) + f" + {float(quad)!r} * (y[{j}] * y[{j}])"
                  for j, t in enumerate(terms)]
    return make_tuple_callback("bordered", components), mask, (rows, cols, vals)


def bordered_ensemble(size, n, seed=13):
    """Diagonal entries -LogUniform[1, 1e4] (a stiff spread), y0 ~ U[-1, 1]."""
    rng = np.random.default_rng(seed)
    lam = -(10.0 ** rng.uniform(0.0, 4.0, size=(size, n)))
    y0 = rng.uniform(-1.0, 1.0, size=(size, n))
    return y0, lam


# --------------------------------------------------------------------------
# A four-species Big Bang nucleosynthesis network (n, p, D, 4He): a stiff,
# non-autonomous right-hand side with exp, sqrt, a 1.5 power and a
# comparison used as a 0/1 factor. The physics and constants are those of
# the reference's consumer, examples/bbn_estimation/main.py:44-60 (constants)
# and :74-161 (rates and network); the independent variable is x = Q/T.
# --------------------------------------------------------------------------
BBN_Q = 1.293            # MeV, n-p mass difference
BBN_B_D = 2.225          # MeV, deuterium binding energy
BBN_M_N = 938.272        # MeV, nucleon mass
BBN_M_E = 0.511          # MeV, where g_star steps (e+e- annihilation)
BBN_M_PL = 1.2209e22     # MeV, Planck mass
BBN_ZETA3 = 1.2020569
BBN_N_EFF_SM = 3.044
BBN_TAU_N = 879.4 / 6.582119e-22   # neutron lifetime in MeV^-1
BBN_SIGMA_NP = 4.0e-11   # MeV^-2, <sigma v> of n + p -> D
BBN_SIGMA_DD = 3.4e-5    # MeV^-2, <sigma v> of D + D -> 4He
BBN_X_SPAN = np.array([BBN_Q / 10.0, BBN_Q / 0.01])   # T from 10 MeV to 10 keV
#: the consumer's prior box: log10(eta_10) in [0.5, 1.0], N_eff in [2, 4]
#: (main.py:284-286), its observed abundances and errors (main.py:274-279),
#: and its solver settings (main.py:176-196)
BBN_PRIOR_LO = np.array([0.5, 2.0])
BBN_PRIOR_HI = np.array([1.0, 4.0])
BBN_OBSERVED = np.array([0.106, 4.2e-6])     # Y_P, D/H
BBN_SIGMA = np.array([0.004, 0.3e-6])
BBN_SOLVER = dict(lu_precision="fp32", rtol=1e-4, atol=1e-9, first_step=0.1, max_steps=256)


def make_bbn(*, exp, sqrt):
    """The network's dY/dx as a factory over the two functions a backend spells
    its own way (the reference's dual-backend form, examples/dual_backend.py:
    1-60): ``make_bbn(exp=math.exp, sqrt=math.sqrt)`` is a callback ``solve``
    traces, ``make_bbn(exp=torch.exp, sqrt=torch.sqrt)`` evaluates it on
    tensors. State Y = (Y_n, Y_p, Y_d, Y_4He), parameters (log10 eta_10,
    N_eff). Y_n + Y_p + 2 Y_d + 4 Y_4He is conserved."""
    pi = math.pi

    def bbn(y, x, p):
        eta = 10.0 ** (p[0] - 10.0)
        temp = BBN_Q / x
        # relativistic degrees of freedom, stepping down through e+e- annihilation
        g_star = 3.91 + (10.75 - 3.91) * (temp > BBN_M_E) + (7.0 / 4.0) * (p[1] - BBN_N_EFF_SM)
        hubble = sqrt(4.0 * pi ** 3 * g_star / 45.0) * temp ** 2 / BBN_M_PL
        n_b = eta * (2.0 * BBN_ZETA3 / pi ** 2 * temp ** 3)
        gamma_np = (255.0 / BBN_TAU_N) * (12.0 + 6.0 * x + x ** 2) / x ** 5 + 1.0 / BBN_TAU_N
        gamma_pn = gamma_np * exp(-x)
        # Saha equilibrium ratio Y_d / (Y_n Y_p)
        k_d = n_b * (3.0 / 4.0) * (4.0 * pi / (BBN_M_N * temp)) ** 1.5 * exp(BBN_B_D / temp)
        r_np = n_b * BBN_SIGMA_NP * (y[0] * y[1] - y[2] / k_d)
        r_dd = n_b * BBN_SIGMA_DD * y[2] ** 2
        den = hubble * x
        return (
            (gamma_pn * y[1] - gamma_np * y[0] - r_np) / den,
            (gamma_np * y[0] - gamma_pn * y[1] - r_np) / den,
            (r_np - 2.0 * r_dd) / den,
            r_dd / den,
        )

    return bbn


bbn = make_bbn(exp=math.exp, sqrt=math.sqrt)


def bbn_initial_state():
    """Weak equilibrium at x0 = Q / 10 MeV: Y_n / Y_p = e^-x0 (main.py:220-225)."""
    x0 = BBN_X_SPAN[0]
    return np.array([math.exp(-x0) / (1.0 + math.exp(-x0)), 1.0 / (1.0 + math.exp(-x0)),
                     1e-20, 0.0])


def bbn_grid(n):
    """A near-square (log10 eta_10, N_eff) grid over the prior box, flattened
    (the consumer's chi^2-grid benchmark, main.py:240-247)."""
    side = int(math.ceil(math.sqrt(n)))
    ee, nn = np.meshgrid(np.linspace(BBN_PRIOR_LO[0], BBN_PRIOR_HI[0], side),
                         np.linspace(BBN_PRIOR_LO[1], BBN_PRIOR_HI[1], side))
    return np.column_stack([ee.ravel(), nn.ravel()])[:n]


def bbn_abundances(final):
    """(Y_P, D/H) from final states (..., 4): Y_P = 4 Y_4He, D/H = Y_d / Y_p."""
    return np.stack([4.0 * final[..., 3], final[..., 2] / final[..., 1]], axis=-1)
