"""Algorithmic work of a solve, for the VALU roofline.

The ensemble kernels are bound by vector ALU throughput: no two trajectories
share a contraction (so there is nothing for the matrix cores), and HBM only
carries the inputs once and the saves once. The work of a solve is therefore
counted in flops, split by precision, from the same structure the kernels
execute:

  - per attempted step (accepted or rejected -- a rejected step does all of it);
  - per save written by the continuous extension.

Conventions: +, -, *, / are one flop, a fused multiply-add two, a negation or
an absolute value zero (they are operand modifiers on this ISA), comparisons
and selects zero; square roots and the controller's three powers are counted
apart as special functions. ``F`` and ``Jf`` are the flops of the generated
right-hand side and of the generated Jacobian sweeps (including the primal
values those sweeps need), counted when the code is emitted.

Peaks are MI355X's published dense vector rates: FP64 78.6 TFLOP/s, FP32
157.3 TFLOP/s (the FP32 figure assumes packed dual-issue, which these kernels
are built without, so it is an upper bound they cannot reach).
"""

from __future__ import annotations

from dataclasses import dataclass

P64_TFLOPS = 78.6
P32_TFLOPS = 157.3
HBM_GBPS = 8000.0


@dataclass
class Work:
    fp64: float = 0.0
    fp32: float = 0.0
    special: float = 0.0

    def __add__(self, o):
        return Work(self.fp64 + o.fp64, self.fp32 + o.fp32, self.special + o.special)

    def scale(self, k):
        return Work(self.fp64 * k, self.fp32 * k, self.special * k)

    @property
    def total(self):
        return self.fp64 + self.fp32

    def roof_seconds(self):
        return self.fp64 / (P64_TFLOPS * 1e12) + self.fp32 / (P32_TFLOPS * 1e12)


def rodas5p_stage_flops(n: int) -> int:
    """Stage right-hand sides, all eight stages: stage 0 one fma per component;
    stage s >= 1 one multiply and s-1 fmas for the C-sum, then one fma (s >= 5)
    or two (s < 5) to add it and the dT term to the right-hand side."""
    per_j = 2
    for s in range(1, 8):
        per_j += 1 + 2 * (s - 1) + (4 if s < 5 else 2)
    return per_j * n


def rodas5p_step(built) -> Work:
    """Flops of one attempted Rodas5P step of ``built`` (a ``_kernels.Built``)."""
    n = built.n_vars
    F = built.rhs_flops.flops
    Jf = built.jac_flops.flops
    fp64 = 8 * F + Jf
    fp64 += rodas5p_stage_flops(n)
    # advances: 1 mul + s fmas + 1 add for s = 0..4, then two FSAL adds
    fp64 += sum(2 + 2 * s for s in range(5)) * n + 2 * n
    # error norm: y_new add, scale fma, weight mul, divide, square-accumulate fma
    fp64 += (1 + 2 + 1 + 1 + 2) * n + 2
    # step-size bookkeeping: dt_use, 1/dt, t_end, 1/(dt*gamma), the stage times,
    # the controller's products and clamp
    fp64 += 16
    special = 1 + 3  # sqrt, three powers
    lu = built.lu_flops
    solves = 8 * built.solve_flops
    diag = n
    if built.lu_precision == "fp32":
        if built.structure is not None and built.layout is None:
            # dense: the multiplier and every update are formed in double and
            # stored in float (the reference's typing); the solves are float
            fp64 += lu
            fp32 = solves + diag
        else:
            fp32 = lu + solves + diag
    else:
        fp64 += lu + solves + diag
        fp32 = 0
    # refined stage solves (lu_refine): per stage one Jacobian-vector product
    # in double, the residual (an fma and an add per component), a second FP32
    # solve and the correction's add
    jvp = built.extra.get("jvp_flops") if built.extra else None
    if jvp is not None:
        fp64 += 8 * (jvp.flops + 4 * n)
        fp32 += 8 * built.solve_flops
    # the accepted-step state update (counted on every attempt: an upper bound
    # of n flops on rejected steps)
    fp64 += n
    return Work(fp64=float(fp64), fp32=float(fp32), special=float(special))


def rodas5p_save(n: int) -> Work:
    """One dense-output save: h1..h3 (1 mul + 7 fmas each), y_new, 4 fmas/muls."""
    return Work(fp64=float((3 * (1 + 14) + 1 + 2 + 2 + 2 + 2 + 1) * n + 4))


def tsit5_step(built) -> Work:
    """One attempted Tsit5 step: six or seven right-hand sides (first same as
    last after an accept; counted as six -- the seventh is the next step's
    first), the stage combinations, the error estimate and the update."""
    n = built.n_vars
    F = built.rhs_flops.flops
    fp64 = 6 * F
    # u for stages 2..7: mul + (s-1) fmas + final fma
    fp64 += sum(1 + 2 * (s - 1) + 2 for s in range(1, 7)) * n
    # error: mul + 6 fmas, dt mul, scale fma, weight mul, divide, square fma
    fp64 += (1 + 12 + 1 + 2 + 1 + 1 + 2) * n + 2
    fp64 += 8  # stage times and the controller products
    return Work(fp64=float(fp64), special=4.0)


def tsit5_save(n: int) -> Work:
    """b1..b7 polynomials (about 35 flops) and per component 1 mul + 6 fmas + 1 fma."""
    return Work(fp64=float((1 + 12 + 2) * n + 35))


def solve_work(built, attempted_steps: int, saves: int, rejected_steps: int = 0) -> Work:
    """Total work of a solve from the kernel's own counters.

    ``attempted_steps`` and ``rejected_steps`` are summed over trajectories; a
    rejected Tsit5 step evaluates its first stage afresh, so it costs one extra
    right-hand side."""
    if built.method == "rodas5p":
        return rodas5p_step(built).scale(attempted_steps) + rodas5p_save(built.n_vars).scale(saves)
    w = tsit5_step(built).scale(attempted_steps) + tsit5_save(built.n_vars).scale(saves)
    w.fp64 += float(built.rhs_flops.flops) * rejected_steps
    return w


def hbm_bytes(n_traj: int, n_vars: int, n_params: int, n_save: int, weights: bool) -> int:
    """Algorithmic HBM traffic of a solve: every input read once, every output
    written once (y0, params, optional weights in; history and three counters out)."""
    read = 8 * n_traj * (n_vars + n_params + (n_vars if weights else 0))
    write = 8 * n_traj * n_save * n_vars + 12 * n_traj
    return read + write
