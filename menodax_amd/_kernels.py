"""Kernel construction: one generated source per (callback, shapes, options).

Plays the role of modax.rodas5P._make_kernel (rodas5P.py:306-1007) and
modax.tsit5._make_kernel (tsit5.py:409-602): the options are frozen into a
hashable key, the callback is traced and emitted as HIP C++
(``_codegen``), the method template is included, and the result is compiled
once through the code-object cache. Both launch shapes -- one lane per
trajectory and the persistent refill kernel -- come from the same code object.
"""

from __future__ import annotations

import math

import functools
from dataclasses import dataclass, field, replace

from . import _codegen, _runtime
from ._sparsity import CompressedJacobian, check_superset, colour_sparsity, dense_jacobian

#: one wavefront: the granularity at which divergence between trajectories is free
DEFAULT_BLOCK = 64
#: parameter rows up to this width are held in registers
PARAMS_IN_REGS_MAX = 16
#: systems up to this size run the fully unrolled, register-resident step
UNROLLED_MAX_VARS = 16
#: the dense LU is unrolled into registers up to this size (MDX_DENSE_UNROLL_MAX);
#: above it the matrix lives in per-lane scratch or LDS
DENSE_UNROLL_MAX = 16
#: the adjoint kernel holds eight stage vectors, their adjoints and the
#: factorisation in both precisions; unrolled only up to this many variables
ADJOINT_UNROLLED_MAX_VARS = 8
#: LDS per compute unit on gfx950
LDS_BYTES_PER_CU = 160 * 1024
PLACEMENTS = ("registers", "rolled", "lds")
#: one trajectory per 64-lane workgroup, the lanes sharing out the dense LU
#: (Rodas5P, a dense matrix above the unrolled size and at most one row per lane)
WAVE = "wave"
WAVE_MAX_VARS = 64

RODAS_EXPONENT = -1.0 / 6.0
TSIT_EXPONENT = -1.0 / 5.0


def trajectories_per_block_or_default(requested=None) -> int:
    """A wavefront by default; any positive count up to 1024 otherwise
    (modax/rodas5P.py:120-134)."""
    if requested is None:
        return DEFAULT_BLOCK
    requested = int(requested)
    if requested < 1:
        raise ValueError(f"trajectories_per_block must be positive, got {requested}")
    if requested > 1024:
        raise ValueError(f"trajectories_per_block is at most 1024 on this GPU, got {requested}")
    return requested


@dataclass(frozen=True)
class KernelOptions:
    """Everything besides the callback and its shapes that selects a kernel
    (modax/rodas5P.py:137-164). Hashable: it is the cache key."""

    pcoeff: float = 0.0
    icoeff: float = 1.0
    dcoeff: float = 0.0
    lu_precision: str = "fp32"
    trajectories_per_block: int = DEFAULT_BLOCK
    sparsity: tuple | None = None
    ordering: str = "amd"
    tf_index: int = -1
    max_registers: int | None = None
    unit_weights: bool = True
    params_in_regs: bool | None = None
    #: where the Rodas5P step keeps its vectors: "registers" (the step fully
    #: unrolled), "rolled" (one stage body run eight times, stage vectors and
    #: state in scratch), "lds" (rolled, with as many of the stage vectors, the
    #: stage argument and the state in LDS as fit), or None to choose by size
    placement: str | None = None
    #: keep the Rodas5P tableau in vector registers across the step loop (the
    #: unrolled step only): fewer instructions per step for more registers, or
    #: None for the size rule
    tableau_in_registers: bool | None = None
    #: forward sensitivities carried with the state (a ``SensitivitySpec``), or None
    sens: object = None
    #: an array-form right-hand side ``f(y, t, p, out)`` for the stage
    #: evaluations; the Jacobian still comes from ``ode_fn`` (rodas5P.py:382-392)
    array_rhs: object = None
    #: a save hook ``hook(save_idx, y, t, p_row, acc)`` (modax/rodas5P.py:157-164),
    #: the width of its accumulator row, whether the history is kept, and the
    #: number of save times the hook was traced for
    save_hook: object = None
    hook_size: int = 0
    save_history: bool = True
    hook_n_save: int = 0
    #: with an FP32 LU: refine each stage solve once against the FP64 iteration
    #: matrix and keep the stage vectors in FP64 (no counterpart in the
    #: reference, whose FP32 path stores the stage vectors in FP32)
    lu_refine: bool = False
    #: the LDS adjoint also keeps its stage arguments in LDS and pulls each
    #: stage's adjoint from the later ones when they fit at the workgroup size
    #: (the same bits, fewer instructions), or never
    adjoint_cache: bool = True


#: the unrolled Rodas5P step keeps its tableau in vector registers up to this
#: many variables: about 100 registers more (243 for Robertson), two waves per
#: SIMD; a 4-variable step would reach into the accumulation registers
TABLEAU_REGS_MAX_VARS = 3


def wave_placement(n_vars: int, n_traj: int, n_cu: int, placement=None, sparsity=None) -> object:
    """The size rule's placement for a plain Rodas5P solve: ``"wave"`` (one
    trajectory per 64-lane workgroup, the lanes sharing out the dense LU) for a
    dense matrix of ``DENSE_UNROLL_MAX + 1 .. WAVE_MAX_VARS`` variables when
    every trajectory can have a SIMD of its own (at most 4 per compute unit);
    otherwise ``placement`` unchanged. A wave steps one trajectory where a
    per-lane kernel steps up to 64, so it only wins while the ensemble leaves
    SIMDs idle: 1,000 64-variable Van der Pol rings 87.6 -> 17.4 ms (FP32 LU)."""
    if placement is not None or sparsity is not None:
        return placement
    if DENSE_UNROLL_MAX < n_vars <= WAVE_MAX_VARS and n_traj <= 4 * n_cu:
        return WAVE
    return placement


def tableau_in_registers(n_vars: int, n_traj: int, n_cu: int, placement=None,
                         max_registers=None) -> bool:
    """Whether a plain Rodas5P solve keeps its tableau in registers: for the
    unrolled step of a small system when every trajectory has a lane resident
    at two waves per SIMD, the occupancy the larger register file leaves. Then
    the step issues about 10% fewer instructions (Robertson: 1,032 against
    1,145, the scalar moves that rebuild each double constant at every use);
    beyond it the lost occupancy costs more than that (262,144 Robertson
    trajectories, FP64 LU: 1.55 ms against 1.73 ms)."""
    if placement not in (None, "registers") or max_registers is not None:
        return False
    if n_vars > TABLEAU_REGS_MAX_VARS:
        return False
    return n_traj <= n_cu * 4 * 2 * 64


def rodas5p_placement(n_vars: int, lu_precision: str, block: int, placement=None,
                      dense_matrix: int = 0, k_bytes: int | None = None) -> dict:
    """The placement macros for a Rodas5P kernel.

    Up to ``UNROLLED_MAX_VARS`` the whole step stays in registers. Above it,
    the stage loop is rolled and the per-lane arrays go to LDS, laid out
    ``[element][lane]``, in order of traffic per step: a dense iteration
    matrix of ``dense_matrix`` elements (its factorisation touches each
    element about n/3 times a step), the eight stage vectors (about 50 element
    reads per component per step), the stage argument and the state; whatever
    does not fit in one compute unit's LDS next to them stays in scratch.
    Outputs are bit-identical across placements. ``k_bytes`` is a stage
    vector element's size when it differs from the LU's (refined FP32 solves
    keep them in FP64).
    """
    if placement is None:
        placement = "registers" if n_vars <= UNROLLED_MAX_VARS else "lds"
    if placement == WAVE:
        n = math.isqrt(dense_matrix)
        if not DENSE_UNROLL_MAX < n <= WAVE_MAX_VARS:
            raise ValueError("the wave placement is for a dense matrix of "
                             f"{DENSE_UNROLL_MAX + 1}..{WAVE_MAX_VARS} variables (the forward solve)")
        lu_bytes = 4 if lu_precision == "fp32" else 8
        # stage vectors, stage argument and state (n_vars long: the joint
        # vector of a sensitivity solve), the padded matrix, its pivots
        lds = (8 * n_vars * lu_bytes + 16 * n_vars + n * (n + 1) * lu_bytes
               + n * (lu_bytes + 12))
        return {"MDX_ROLLED": 1, "MDX_LDS_LU": 1, "MDX_LDS_K": 1, "MDX_LDS_U": 1,
                "MDX_LDS_Y": 1, "MDX_COOP": 1, "lds_bytes": lds, "placement": WAVE}
    if placement not in PLACEMENTS:
        raise ValueError(f"placement must be one of {PLACEMENTS + (WAVE,)}, got {placement!r}")
    defs = {"MDX_ROLLED": 0 if placement == "registers" else 1,
            "MDX_LDS_LU": 0, "MDX_LDS_K": 0, "MDX_LDS_U": 0, "MDX_LDS_Y": 0, "MDX_COOP": 0}
    lds = 0
    if placement == "lds":
        lu_bytes = 4 if lu_precision == "fp32" else 8
        for macro, per_lane in (("MDX_LDS_LU", dense_matrix * lu_bytes),
                                # refined solves keep each stage's right-hand
                                # side in LDS beside the stage vectors
                                ("MDX_LDS_K", 8 * n_vars * (k_bytes or lu_bytes)
                                 + (8 * n_vars if k_bytes else 0)),
                                ("MDX_LDS_U", 8 * n_vars), ("MDX_LDS_Y", 8 * n_vars)):
            if per_lane and lds + per_lane * block <= LDS_BYTES_PER_CU:
                defs[macro] = 1
                lds += per_lane * block
    defs["lds_bytes"] = lds
    defs["placement"] = placement
    return defs


def adjoint_lds_bytes(n_vars: int, n_lu: int, lu_precision: str, cache: bool = False) -> int:
    """Per-lane LDS of the adjoint's LDS placement: the eight stage vectors (in
    the LU's type), their eight adjoints and the factors' double copy; with
    ``cache``, seven more double vectors (the stage arguments u_1..u_7, then
    their adjoints)."""
    lu_bytes = 4 if lu_precision == "fp32" else 8
    return 8 * n_vars * lu_bytes + 8 * n_vars * 8 + 8 * n_lu + (7 * n_vars * 8 if cache else 0)


def lu_size(n_vars: int, sparsity=None, ordering: str = "amd") -> int:
    """Stored elements of the iteration matrix: n^2 dense, else the sparse
    factorisation's L + U."""
    if sparsity is None:
        return n_vars * n_vars
    from . import _sparse_direct

    return _sparse_direct.sparse_direct_solver_for(sparsity, ordering).compressed.size


def adjoint_lds_supported(sparsity=None, ordering: str = "amd") -> bool:
    """Whether the adjoint's LDS placement serves this pattern: dense systems
    and sparse ones whose factorisation and solves are unrolled. With the
    table-driven sparse routines (patterns above the unrolling thresholds) the
    LDS adjoint kernel gave gradients different from the scratch kernel's, for
    a cause not found (the emitted routines are clean under AddressSanitizer on
    the host, the forward kernel with the same LDS layout is exact, and so is
    the adjoint with either routine alone table driven; zeroing the LDS stage
    vectors before every forward step made it exact, so the replay from a
    checkpoint saw the LDS's earlier contents), so those patterns run the
    scratch adjoint."""
    if sparsity is None:
        return True
    from . import _sparse_direct

    return all(_sparse_direct.sparse_direct_solver_for(sparsity, ordering).unrolled())


def adjoint_block(n_vars: int, n_lu: int, lu_precision: str, n_traj: int, n_cu: int):
    """The rolled Rodas5P adjoint's (placement, workgroup size, stage cache)
    for an ensemble of ``n_traj``: the LDS placement at the widest of 64, 32
    and 16 lanes whose LDS share holds the stage vectors, their adjoints and
    the factors -- with the stage cache (``KernelOptions.adjoint_cache``) where
    that still keeps the whole ensemble resident at once, else without it --
    and otherwise the scratch ("rolled") step at 64 lanes. Residency decides because the reverse step
    waits on every access to these arrays: a small ensemble leaves SIMDs idle
    and gains from LDS latency, while a large one at a narrow LDS workgroup
    keeps fewer lanes resident than 64-lane scratch workgroups. The 20-variable
    Van der Pol chain (67 factor elements), 8,192 trajectories: FP64 LU
    405 -> 182-192 ms (LDS at 32 lanes), 164 ms with the stage cache; FP32
    369 ms in scratch, 197-202 ms in LDS at 64 lanes without the cache,
    178-187 ms with it at 32 (whole gradient calls); 65,536 trajectories stay
    in scratch (907 / 786 ms against 908-1366 /
    794-813 ms for the LDS forms)."""
    for cache in (True, False):
        per_lane = adjoint_lds_bytes(n_vars, n_lu, lu_precision, cache)
        for block in (64, 32, 16):
            group = per_lane * block
            if group <= LDS_BYTES_PER_CU and -(-n_traj // block) <= resident_groups(group, n_cu):
                return "lds", block, cache
    return "rolled", DEFAULT_BLOCK, False


def resident_groups(lds_per_group: int, n_cu: int) -> int:
    """Workgroups the GPU keeps resident when each takes ``lds_per_group``
    bytes of LDS: per compute unit as many as its LDS holds, at most one per
    SIMD (these kernels' register use leaves a SIMD one wave)."""
    per_cu = 4 if lds_per_group <= 0 else min(4, LDS_BYTES_PER_CU // lds_per_group)
    return n_cu * per_cu


def rodas5p_block(requested, n_aug: int, lu_precision: str, n_traj: int, n_cu: int,
                  placement=None, sparsity=None, lu_refine=False) -> int:
    """The workgroup size of a plain or joint Rodas5P solve: the caller's, else
    a wavefront, except for a small ensemble of a sparse system whose rolled
    step (``n_aug``, the joint vector's length, above ``UNROLLED_MAX_VARS``;
    LDS placement) cannot keep its eight stage vectors in LDS at 64 lanes:
    then the larger of 32 and 16 lanes at which the stage vectors, the stage
    argument and the state all fit, provided the whole ensemble is resident at
    once at that size -- a narrow workgroup's LDS share limits how many a
    compute unit holds (128 FP32 joint rows at 16 lanes: 96 KiB, one per CU,
    4,096 lanes on 256 CUs, where 64-lane workgroups with the vectors in
    scratch keep 16,384), so a larger ensemble stays at 64 lanes. The
    reference's stiff Van der Pol rings, 1,000 trajectories, FP32 sparse LU:
    the 128-variable ring 5.26 -> 3.61 ms, the joint solve of the 64-variable
    ring with one parameter direction 7.82 -> 5.48 ms; the 64-variable ring's
    plain solve (stage vectors in LDS at 64 lanes) keeps 64 lanes: 1.31 ms
    there, 1.41-1.45 ms at 32 or 16.

    Refined FP32 solves (``lu_refine``) take the narrow workgroup whatever the
    ensemble size: their FP64 stage vectors and each stage's right-hand side,
    which the residual's Jacobian-vector product reads back, cost more in
    scratch than the lanes a narrow workgroup leaves unresident. The 100,000
    uneven stiff trajectories (50 variables): 1.80 s at 32 lanes (one
    workgroup a CU), 2.29 s at 16, 3.41 s at 64 with the stage vectors in
    scratch."""
    if requested is not None:
        return trajectories_per_block_or_default(requested)
    if placement not in (None, "lds") or sparsity is None or n_aug <= UNROLLED_MAX_VARS:
        return DEFAULT_BLOCK
    lu_bytes = 4 if lu_precision == "fp32" and not lu_refine else 8
    stage_vectors = 8 * n_aug * lu_bytes + (8 * n_aug if lu_refine else 0)
    if stage_vectors * DEFAULT_BLOCK <= LDS_BYTES_PER_CU:
        return DEFAULT_BLOCK
    for block in (32, 16):
        group = (stage_vectors + 16 * n_aug) * block
        if group <= LDS_BYTES_PER_CU and (
                lu_refine or -(-n_traj // block) <= resident_groups(group, n_cu)):
            return block
    return DEFAULT_BLOCK


def tsit5_block(requested, n_vars: int, n_traj: int, n_cu: int) -> int:
    """The workgroup size of a Tsit5 solve: the caller's, else a wavefront,
    except that a rolled step (more than ``UNROLLED_MAX_VARS`` variables) of
    an ensemble that is resident at once at the narrower size takes the
    largest of 64, 32 and 16 lanes at which its seven stage vectors fit in LDS
    (the residency bound as for ``rodas5p_block``: a 128-variable ring at 16
    lanes takes 144 KiB, one workgroup per CU). The reference's non-stiff Van
    der Pol rings, 1,000 trajectories: 64 variables 0.81 ms at 64 lanes (stage
    vectors in scratch), 0.55 ms at 32; 128 variables 1.60 ms at 64 lanes,
    1.05 ms at 16."""
    if requested is not None or n_vars <= UNROLLED_MAX_VARS:
        return trajectories_per_block_or_default(requested)
    for b in (64, 32, 16):
        place = tsit5_placement(n_vars, b)
        if not place["MDX_LDS_K"]:
            continue
        if b == DEFAULT_BLOCK or -(-n_traj // b) <= resident_groups(place["lds_bytes"], n_cu):
            return b
        break
    return DEFAULT_BLOCK


def tsit5_placement(n_vars: int, block: int, placement=None) -> dict:
    """The placement macros for a Tsit5 kernel: registers up to
    ``UNROLLED_MAX_VARS``, above it the rolled stage loop with the seven stage
    vectors, the stage argument and the state in LDS as far as they fit."""
    if placement is None:
        placement = "registers" if n_vars <= UNROLLED_MAX_VARS else "lds"
    if placement not in PLACEMENTS:
        raise ValueError(f"placement must be one of {PLACEMENTS}, got {placement!r}")
    defs = {"MDX_ROLLED": 0 if placement == "registers" else 1,
            "MDX_LDS_K": 0, "MDX_LDS_U": 0, "MDX_LDS_Y": 0}
    lds = 0
    if placement == "lds":
        for macro, per_lane in (("MDX_LDS_K", 7 * n_vars * 8), ("MDX_LDS_U", 8 * n_vars),
                                ("MDX_LDS_Y", 8 * n_vars)):
            if lds + per_lane * block <= LDS_BYTES_PER_CU:
                defs[macro] = 1
                lds += per_lane * block
    defs["lds_bytes"] = lds
    defs["placement"] = placement
    return defs


@dataclass
class Built:
    """A compiled kernel pair and what the roofline needs to know about it."""

    method: str
    plain: int
    refill: int
    block: int
    n_vars: int
    n_params: int
    unit_weights: bool
    source: str
    rhs_flops: _codegen.FlopCount
    jac_flops: _codegen.FlopCount
    structure: CompressedJacobian | None = None
    lu_flops: int = 0
    solve_flops: int = 0
    lu_precision: str = "fp64"
    layout: object = None
    extra: dict = field(default_factory=dict)


def _hex(v: float) -> str:
    return _codegen._lit(float(v))


def pid_exponents(exponent, pcoeff, icoeff, dcoeff):
    """modax/rodas5P.py:340-342 (tsit5.py:177-179)."""
    e1 = exponent * (icoeff + pcoeff + dcoeff)
    e2 = -exponent * (pcoeff + 2.0 * dcoeff)
    e3 = exponent * dcoeff
    return e1, e2, e3


def _prelude(defs: dict) -> list:
    lines = ["// generated by menodax_amd: do not edit"]
    for k, v in defs.items():
        lines.append(f"#define {k} {v}")
    lines.append('#include "mdx_common.hpp"')
    return lines


def _max_vgprs(max_registers):
    """The per-lane register cap (the reference's max_registers): 0 leaves it
    to the compiler, otherwise a VGPR budget of 24..512."""
    if max_registers is None:
        return 0
    v = int(max_registers)
    if not 24 <= v <= 512:
        raise ValueError(f"max_registers must be between 24 and 512, got {max_registers}")
    return v


def _params_in_regs(options, n_params):
    if options.params_in_regs is not None:
        return bool(options.params_in_regs)
    return n_params <= PARAMS_IN_REGS_MAX


def resume_rows(method: str, n_vars: int) -> int:
    """Doubles per trajectory of the step state a cost pilot leaves for the
    solve that resumes it: the state and the eight step scalars (Rodas5P),
    and for Tsit5 also its first-same-as-last stage and whether it holds."""
    return n_vars + 8 if method == "rodas5p" else 2 * n_vars + 9


def _register(code, block, unit_weights, plain_entry, refill_entry, hook=False, wave=False,
              resumable=False):
    """Handles of the one-lane-per-trajectory kernel and of the persistent
    refill kernel (0 when there is none: a wave-per-trajectory kernel);
    ``resumable`` kernels take a cost pilot's step state (``resume_rows``)."""
    flags = _runtime.KERNEL_UNIT_WEIGHTS if unit_weights else 0
    if hook:
        flags |= _runtime.KERNEL_HOOK
    if resumable and not hook and not wave:
        flags |= _runtime.KERNEL_RESUMABLE
    plain = _runtime.kernel_create(code, plain_entry, block,
                                   flags | (_runtime.KERNEL_WAVE if wave else 0))
    if refill_entry is None:
        return plain, 0
    refill = _runtime.kernel_create(code, refill_entry, block, flags | _runtime.KERNEL_PERSISTENT)
    return plain, refill


def dense_lu_flops(n):
    """Flops of the dense factorisation and of one solve (multiply-add = 2)."""
    factor = n + n * (n - 1) // 2 + (n - 1) * n * (2 * n - 1) // 3
    solve = 2 * n * n - n
    return factor, solve


def rodas5p_source(ode_fn, n_vars: int, n_params: int, options: KernelOptions, adjoint=False):
    """Generate the full Rodas5P kernel source; returns (source, details).
    ``adjoint=True`` generates the discrete-adjoint kernel instead (the
    reverse-mode contractions of the callback; registers up to
    ``ADJOINT_UNROLLED_MAX_VARS`` variables, the rolled step above; with
    ``placement="lds"`` the rolled adjoint keeps its stage vectors, their
    adjoints and the factors in LDS when they fit at the workgroup size, else
    in scratch)."""
    if options.lu_precision not in ("fp32", "fp64"):
        raise ValueError(f"lu_precision must be 'fp32' or 'fp64', got {options.lu_precision!r}")
    if options.lu_refine:
        if options.lu_precision != "fp32":
            raise ValueError("lu_refine refines FP32 solves; it needs lu_precision='fp32'")
        if adjoint or options.sens is not None:
            raise ValueError("lu_refine is for a plain solve: the adjoint and the forward "
                             "sensitivities differentiate the unrefined step")
        if options.placement == WAVE:
            raise ValueError("lu_refine runs one trajectory per lane, not the wave placement")
    traced = _codegen.trace(ode_fn, n_vars, n_params)
    layout = None
    if options.sparsity is None:
        structure = dense_jacobian(n_vars)
        sparse_src = ""
        lu_f, solve_f = dense_lu_flops(n_vars)
    else:
        from . import _sparse_direct

        check_superset(options.sparsity, _codegen.structural_pattern(traced))
        solver = _sparse_direct.sparse_direct_solver_for(options.sparsity, options.ordering)
        structure = solver.compressed
        layout = solver.layout
        sparse_src = solver.device_source()
        m, u, s = layout.flop_counts()
        lu_f, solve_f = m + 2 * u + n_vars, 2 * s + n_vars
    # the wave placement clears its shared matrix a column per lane, so the
    # Jacobian writes only its structurally nonzero slots
    wave = (options.placement == WAVE and not adjoint and options.sparsity is None)
    gen = _codegen.generate(
        traced, colour=structure.colour, n_colours=structure.n_colours,
        store=structure.stores(), skip_zero=wave,
    )
    if options.array_rhs is not None and options.sens is None:
        # the joint system's rows come from ode_fn's sweeps, as in the reference
        arr = _codegen.trace_array_rhs(options.array_rhs, n_vars, n_params)
        rhs_src, rhs_fc = _codegen.generate_rhs(arr)
        jac_src, _ = _codegen.generate_neg_jac(
            traced, colour=structure.colour, n_colours=structure.n_colours,
            store=structure.stores(), skip_zero=wave)
        gen = replace(gen, source=rhs_src + "\n\n" + jac_src + "\n", rhs_flops=rhs_fc)
    if options.tf_index >= n_params:
        raise ValueError(f"tf_index {options.tf_index} is not a column of the {n_params} parameters")
    e1, e2, e3 = pid_exponents(RODAS_EXPONENT, options.pcoeff, options.icoeff, options.dcoeff)
    block = trajectories_per_block_or_default(options.trajectories_per_block)
    spec = options.sens
    if spec is not None and (spec.n_vars, spec.n_params) != (n_vars, n_params):
        raise ValueError("the sensitivity spec was made for another system size")
    if options.placement == WAVE and not adjoint:
        if options.save_hook is not None:
            raise ValueError("the wave placement runs no save hook")
        block = 64  # lanes of the one workgroup a trajectory takes
    n_aug = n_vars if spec is None else spec.n_aug
    if adjoint and spec is not None:
        raise ValueError("the adjoint kernel differentiates the plain solve; drop the sensitivity spec")
    placement = options.placement
    adjoint_lds = adjoint_cache = False
    if adjoint:
        # the adjoint reads the state from per-lane arrays: registers for a few
        # variables (every stage loop unrolled), the rolled step above, its
        # stage-indexed arrays in scratch or, when they fit, in LDS
        if placement is None:
            placement = "registers" if n_aug <= ADJOINT_UNROLLED_MAX_VARS else "rolled"
        elif placement == WAVE:
            placement = "rolled"
        elif placement == "lds":
            per_lane = adjoint_lds_bytes(n_vars, structure.size, options.lu_precision)
            adjoint_lds = (per_lane * block <= LDS_BYTES_PER_CU
                           and adjoint_lds_supported(options.sparsity, options.ordering))
            adjoint_cache = adjoint_lds and options.adjoint_cache and adjoint_lds_bytes(
                n_vars, structure.size, options.lu_precision, cache=True) * block <= LDS_BYTES_PER_CU
            placement = "rolled"
    # a dense matrix above the unrolled size lives in per-lane memory: LDS
    # when it fits (the forward step only)
    dense_matrix = (structure.size if (options.sparsity is None and n_vars > DENSE_UNROLL_MAX
                                       and not adjoint) else 0)
    place = rodas5p_placement(n_aug, options.lu_precision, block, placement, dense_matrix,
                              k_bytes=8 if options.lu_refine else None)
    if adjoint_lds:
        place = dict(place, MDX_LDS_K=1, placement="lds",
                     lds_bytes=adjoint_lds_bytes(n_vars, structure.size, options.lu_precision,
                                                 cache=adjoint_cache) * block)
    defs = {
        "MDX_N": n_vars,
        "MDX_NP": n_params,
        "MDX_NLU": structure.size,
        "MDX_LU_FP32": 1 if options.lu_precision == "fp32" else 0,
        "MDX_SPARSE": 0 if options.sparsity is None else 1,
        "MDX_NEEDS_CLEAR": 1 if structure.needs_clear else 0,
        "MDX_N_ERROR": n_vars if spec is None else spec.n_error,
        "MDX_TF_INDEX": options.tf_index,
        "MDX_BLOCK": block,
        "MDX_PARAMS_IN_REGS": 1 if _params_in_regs(options, n_params) else 0,
        "MDX_UNIT_WEIGHTS": 1 if options.unit_weights else 0,
        "MDX_E1": _hex(e1),
        "MDX_E2": _hex(e2),
        "MDX_E3": _hex(e3),
        "MDX_DT_ZERO": 1 if gen.dt_zero else 0,
        "MDX_NSENS": 0 if spec is None else spec.n_sens,
        "MDX_ROLLED": place["MDX_ROLLED"],
        "MDX_LDS_LU": place["MDX_LDS_LU"],
        "MDX_COOP": place["MDX_COOP"],
        "MDX_LDS_K": place["MDX_LDS_K"],
        "MDX_LDS_U": place["MDX_LDS_U"],
        "MDX_LDS_Y": place["MDX_LDS_Y"],
        "MDX_MAX_VGPRS": _max_vgprs(options.max_registers),
        "MDX_TAB_REGS": 1 if (place["MDX_ROLLED"] == 0 and options.tableau_in_registers
                              and not adjoint and spec is None) else 0,
        "MDX_REFINE": 1 if options.lu_refine else 0,
        "MDX_ADJ_LDS": 1 if adjoint_lds else 0,
        "MDX_ADJ_CACHE": 1 if adjoint_cache else 0,
    }
    sens_src = ""
    if options.save_hook is not None or not options.save_history:
        if adjoint or spec is not None:
            raise ValueError("a solve with a save hook is not differentiated")
        if options.save_hook is None:
            raise ValueError("save_history=False needs a save_hook to consume the saves")
        if options.hook_size <= 0:
            raise ValueError("a save_hook needs a positive hook_size")
        th = _codegen.trace_hook(options.save_hook, n_vars, n_params, options.hook_n_save,
                                 options.hook_size)
        defs["MDX_HOOK"] = 1
        defs["MDX_HOOK_SIZE"] = int(options.hook_size)
        defs["MDX_SAVE_HISTORY"] = 1 if options.save_history else 0
        sens_src = _codegen.generate_hook(th)
    if adjoint:
        defs["MDX_ADJOINT"] = 1
        sens_src = _codegen.generate_adjoint(traced)
    if spec is not None:
        defs["MDX_SENS_PCOLS"] = ", ".join(str(c) for c in spec.direction_columns)
        defs["MDX_SENS_UNROLL"] = spec.n_sens if spec.n_sens <= 4 else 1
        sens_src = _codegen.generate_sensitivity(traced)
    jvp_src, jvp_fc = "", None
    if options.lu_refine:
        jvp_src, jvp_fc = _codegen.generate_jvp(traced)
    lines = _prelude(defs)
    lines += ["namespace sys {", gen.source, sens_src, jvp_src, sparse_src, "}  // namespace sys",
              '#include "rodas5p.hpp"']
    if adjoint:
        lines.append('#include "rodas5p_adjoint.hpp"')
    source = "\n".join(lines) + "\n"
    return source, dict(gen=gen, structure=structure, layout=layout, lu_f=lu_f, solve_f=solve_f,
                        block=block, placement=place, n_aug=n_aug, jvp_flops=jvp_fc)


@functools.lru_cache(maxsize=None)
def rodas5p_kernel(ode_fn, n_vars: int, n_params: int, options: KernelOptions) -> Built:
    source, d = rodas5p_source(ode_fn, n_vars, n_params, options)
    code = _runtime.code_object(source, f"rodas5p_{getattr(ode_fn, '__name__', 'rhs')}.hip")
    wave = d["placement"]["placement"] == WAVE
    plain, refill = _register(code, d["block"], options.unit_weights, "mdx_rodas5p",
                              None if wave else "mdx_rodas5p_refill",
                              hook=options.save_hook is not None, wave=wave, resumable=True)
    return Built(
        method="rodas5p", plain=plain, refill=refill, block=d["block"], n_vars=n_vars,
        n_params=n_params, unit_weights=options.unit_weights, source=source,
        rhs_flops=d["gen"].rhs_flops, jac_flops=d["gen"].jac_flops, structure=d["structure"],
        lu_flops=d["lu_f"], solve_flops=d["solve_f"], lu_precision=options.lu_precision,
        layout=d["layout"], extra={"placement": d["placement"], "n_aug": d["n_aug"],
                                   "sens": options.sens, "jvp_flops": d["jvp_flops"]},
    )


@functools.lru_cache(maxsize=None)
def rodas5p_adjoint_kernel(ode_fn, n_vars: int, n_params: int, options: KernelOptions) -> Built:
    """The discrete-adjoint kernel for the solve ``rodas5p_kernel`` builds with
    the same options; ``plain`` is its handle (``refill`` is 0)."""
    source, d = rodas5p_source(ode_fn, n_vars, n_params, options, adjoint=True)
    code = _runtime.code_object(source, f"rodas5p_adjoint_{getattr(ode_fn, '__name__', 'rhs')}.hip")
    flags = _runtime.KERNEL_ADJOINT | (_runtime.KERNEL_UNIT_WEIGHTS if options.unit_weights else 0)
    handle = _runtime.kernel_create(code, "mdx_rodas5p_adjoint", d["block"], flags)
    return Built(
        method="rodas5p_adjoint", plain=handle, refill=0, block=d["block"], n_vars=n_vars,
        n_params=n_params, unit_weights=options.unit_weights, source=source,
        rhs_flops=d["gen"].rhs_flops, jac_flops=d["gen"].jac_flops, structure=d["structure"],
        lu_flops=d["lu_f"], solve_flops=d["solve_f"], lu_precision=options.lu_precision,
        layout=d["layout"],
    )


@dataclass(frozen=True)
class Tsit5Options:
    pcoeff: float = 0.0
    icoeff: float = 1.0
    dcoeff: float = 0.0
    trajectories_per_block: int = DEFAULT_BLOCK
    unit_weights: bool = True
    params_in_regs: bool | None = None
    #: forward sensitivities carried with the state (a ``SensitivitySpec``), or None
    sens: object = None
    #: "registers", "rolled", "lds" or None (by size); see ``tsit5_placement``
    placement: str | None = None
    #: keep the coefficients in vector registers across the step loop (the
    #: unrolled step only), or None for the size rule
    tableau_in_registers: bool | None = None


def tsit5_source(ode_fn, n_vars: int, n_params: int, options: Tsit5Options, adjoint=False):
    """Generate the Tsit5 kernel source; returns (source, details).
    ``adjoint=True`` generates the discrete-adjoint kernel instead (the
    callback's reverse-mode contraction; registers up to
    ``ADJOINT_UNROLLED_MAX_VARS`` variables, the rolled scratch step above,
    never LDS)."""
    traced = _codegen.trace(ode_fn, n_vars, n_params)
    spec = options.sens
    n_error = n_vars
    if adjoint and spec is not None:
        raise ValueError("the adjoint kernel differentiates the plain solve; drop the sensitivity spec")
    if spec is not None:
        if (spec.n_vars, spec.n_params) != (n_vars, n_params):
            raise ValueError("the sensitivity spec was made for another system size")
        traced = _codegen.augmented_trace(traced, spec.direction_columns)
        n_vars, n_error = spec.n_aug, spec.n_error
    rhs_src, rhs_fc = _codegen.generate_rhs(traced)
    e1, e2, e3 = pid_exponents(TSIT_EXPONENT, options.pcoeff, options.icoeff, options.dcoeff)
    block = trajectories_per_block_or_default(options.trajectories_per_block)
    placement = options.placement
    if adjoint:
        # the adjoint reads the state from per-lane arrays
        if placement is None:
            placement = "registers" if n_vars <= ADJOINT_UNROLLED_MAX_VARS else "rolled"
        elif placement == "lds":
            placement = "rolled"
    place = tsit5_placement(n_vars, block, placement)
    defs = {
        "MDX_N": n_vars,
        "MDX_NP": n_params,
        "MDX_N_ERROR": n_error,
        "MDX_BLOCK": block,
        "MDX_PARAMS_IN_REGS": 1 if _params_in_regs(options, n_params) else 0,
        "MDX_UNIT_WEIGHTS": 1 if options.unit_weights else 0,
        "MDX_E1": _hex(e1),
        "MDX_E2": _hex(e2),
        "MDX_E3": _hex(e3),
        "MDX_ROLLED": place["MDX_ROLLED"],
        "MDX_LDS_K": place["MDX_LDS_K"],
        "MDX_LDS_U": place["MDX_LDS_U"],
        "MDX_LDS_Y": place["MDX_LDS_Y"],
        "MDX_TAB_REGS": 1 if (place["MDX_ROLLED"] == 0 and options.tableau_in_registers
                              and not adjoint and spec is None) else 0,
    }
    adj_src = ""
    if adjoint:
        defs["MDX_ADJOINT"] = 1
        adj_src = _codegen.generate_adjoint(traced)
    lines = _prelude(defs) + ["namespace sys {", rhs_src, adj_src, "}  // namespace sys",
                              '#include "tsit5.hpp"']
    if adjoint:
        lines.append('#include "tsit5_adjoint.hpp"')
    return "\n".join(lines) + "\n", dict(rhs_flops=rhs_fc, block=block, placement=place)


@functools.lru_cache(maxsize=None)
def tsit5_adjoint_kernel(ode_fn, n_vars: int, n_params: int, options: Tsit5Options) -> Built:
    """The discrete-adjoint kernel for the solve ``tsit5_kernel`` builds with
    the same options; ``plain`` is its handle (``refill`` is 0)."""
    source, d = tsit5_source(ode_fn, n_vars, n_params, options, adjoint=True)
    code = _runtime.code_object(source, f"tsit5_adjoint_{getattr(ode_fn, '__name__', 'rhs')}.hip")
    flags = _runtime.KERNEL_ADJOINT | (_runtime.KERNEL_UNIT_WEIGHTS if options.unit_weights else 0)
    handle = _runtime.kernel_create(code, "mdx_tsit5_adjoint", d["block"], flags)
    return Built(
        method="tsit5_adjoint", plain=handle, refill=0, block=d["block"], n_vars=n_vars,
        n_params=n_params, unit_weights=options.unit_weights, source=source,
        rhs_flops=d["rhs_flops"], jac_flops=_codegen.FlopCount(), lu_precision="fp64",
    )


@functools.lru_cache(maxsize=None)
def tsit5_kernel(ode_fn, n_vars: int, n_params: int, options: Tsit5Options) -> Built:
    source, d = tsit5_source(ode_fn, n_vars, n_params, options)
    code = _runtime.code_object(source, f"tsit5_{getattr(ode_fn, '__name__', 'rhs')}.hip")
    plain, refill = _register(code, d["block"], options.unit_weights, "mdx_tsit5", "mdx_tsit5_refill",
                              resumable=True)
    return Built(
        method="tsit5", plain=plain, refill=refill, block=d["block"], n_vars=n_vars,
        n_params=n_params, unit_weights=options.unit_weights, source=source,
        rhs_flops=d["rhs_flops"], jac_flops=_codegen.FlopCount(), lu_precision="fp64",
        extra={"sens": options.sens},
    )
