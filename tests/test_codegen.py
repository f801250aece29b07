"""The tracer, the forward-mode derivatives and the emitted HIP source.

The derivatives replace the reference's Enzyme sweeps
(modax/_sensitivity.py:148-179 via rodas5P.py:445-500), which its tests hold
against JAX's jacfwd (tests/test_enzyme_jacobian.py:109-138). Here they are held
against the oracle's hand-written closed-form Jacobians, bit for bit, through a
host interpreter that rounds exactly as the emitted code does; and against
central differences for functions the oracle does not carry.
"""

import math

import numpy as np
import pytest

from menodax_amd import _codegen as cg
from menodax_amd import _kernels, _runtime
from menodax_amd._sparsity import normalize_sparsity
from tests import systems as S
from oracle import oracle as O

CASES = [
    ("robertson", S.robertson, 3, 3, O.robertson),
    ("lorenz", S.lorenz, 3, 1, O.lorenz),
    ("forced_decay", S.forced_decay, 1, 2, O.forced_decay),
    ("chain4", S.chain4, 4, 3, O.chain4),
    ("ring6", S.ring6, 6, 2, O.ring6),
    ("damped", S.damped_oscillator, 2, 2, O.damped_osc),
    ("vdp_chain", S.vdp_lattice(5, mu=1.0, d=0.1, wrap=False)[0], 10, 1,
     lambda: O.vdp_lattice(5, mu=1.0, d=0.1, wrap=False)),
    ("vdp_ring", S.vdp_lattice(6)[0], 12, 1, lambda: O.vdp_lattice(6)),
    ("kaps", S.kaps(3, 1e-3)[0], 6, 1, lambda: O.kaps(3, 1e-3)),
]


@pytest.mark.parametrize("name, fn, n, n_p, osys", CASES, ids=[c[0] for c in CASES])
def test_rhs_and_jacobian_are_bit_exact_with_the_closed_forms(name, fn, n, n_p, osys):
    traced = cg.trace(fn, n, n_p)
    system = osys()
    rng = np.random.default_rng(0)
    for _ in range(25):
        y = rng.normal(size=n) * 3
        p = rng.uniform(0.1, 5, size=n_p)
        t = rng.uniform(0, 2)
        assert np.array_equal(cg.rhs_from_graph(traced, y, t, p), O.rhs(system, y, t, p))
        J, dT = cg.jacobian_from_graph(traced, y, t, p)
        Jo, dTo = O.jacobian(system, y, t, p)
        assert np.array_equal(J, Jo)
        assert np.array_equal(dT, dTo)


def test_block_arrow_rhs_and_jacobian_bit_exact():
    fn, mask, (rows, cols, vals, pidx) = S.arrow_system(12)
    n_p = S.arrow_n_params(12)
    traced = cg.trace(fn, 12, n_p)
    system = O.linquad(12, rows, cols, vals, 0.01, off_pidx=pidx, n_p=n_p)
    rng = np.random.default_rng(1)
    for _ in range(10):
        y, p = rng.normal(size=12), rng.normal(size=n_p) * 10
        assert np.array_equal(cg.rhs_from_graph(traced, y, 0.0, p), O.rhs(system, y, 0.0, p))
        J, _ = cg.jacobian_from_graph(traced, y, 0.0, p)
        assert np.array_equal(J, O.jacobian(system, y, 0.0, p)[0])
    assert (cg.structural_pattern(traced) <= mask).all()


def test_coloured_sweeps_recover_the_jacobian():
    """Seeding a whole colour group per sweep returns every column of the group."""
    from menodax_amd._sparse_direct import sparse_direct_solver_for
    from menodax_amd._sparsity import normalize_sparsity

    fn, _ = S.vdp_lattice(8, mu=3.0, d=0.1, wrap=False)
    mask = S.vdp_sparsity(8, wrap=False)
    compressed = sparse_direct_solver_for(normalize_sparsity(mask, 16), "amd").compressed
    assert compressed.n_colours < 16
    traced = cg.trace(fn, 16, 1)
    rng = np.random.default_rng(2)
    y, p = rng.normal(size=16), np.array([7.0])
    full, _ = cg.jacobian_from_graph(traced, y, 0.0, p)
    coloured, _ = cg.jacobian_from_graph(traced, y, 0.0, p, colour=compressed.colour,
                                         n_colours=compressed.n_colours)
    assert np.array_equal(full, coloured)


def _fd_jacobian(fn, y, t, p, h=1e-6):
    n = len(y)
    J = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = h
        J[:, c] = (np.array(fn(y + e, t, p)) - np.array(fn(y - e, t, p))) / (2 * h)
    dT = (np.array(fn(y, t + h, p)) - np.array(fn(y, t - h, p))) / (2 * h)
    return J, dT


def test_transcendental_callbacks_against_central_differences():
    def f(y, t, p):
        return (
            math.exp(-p[0] * y[0]) * math.sin(y[1] * t) + abs(y[2]) ** 1.5,
            math.log(1.0 + y[0] * y[0]) / math.sqrt(2.0 + y[1] ** 2) - max(y[0], y[2]),
            math.tanh(y[2]) * math.cos(p[1] * t) + min(y[0] * 3.0, 1.0) + y[1] ** -2,
        )

    traced = cg.trace(f, 3, 2)
    rng = np.random.default_rng(3)
    for _ in range(10):
        y = rng.uniform(0.3, 1.5, size=3) * np.array([1, 1, -1])
        p = rng.uniform(0.5, 2, size=2)
        t = rng.uniform(0.1, 2)
        np.testing.assert_allclose(cg.rhs_from_graph(traced, y, t, p), f(y, t, p), rtol=1e-14)
        J, dT = cg.jacobian_from_graph(traced, y, t, p)
        Jf, dTf = _fd_jacobian(f, y, t, p)
        np.testing.assert_allclose(J, Jf, rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(dT, dTf, rtol=1e-6, atol=1e-7)


def test_integer_powers_square_like_numba():
    """x**3 is x*(x*x) and x**4 is (x*x)*(x*x): numba's exponentiation by squaring."""
    g = cg.Graph()
    x = g.y(0)
    n3 = cg._int_power(g, x, 3)
    assert n3.op == "mul" and {n3.args[0].id, n3.args[1].id} == {x.id, g.mul(x, x).id}
    n4 = cg._int_power(g, x, 4)
    sq = g.mul(x, x)
    assert n4.op == "mul" and n4.args[0] is sq and n4.args[1] is sq


def test_where_selects_and_branching_raises():
    def ok(y, t, p):
        return (cg.where(y[0] > 0.0, y[0], -y[0]),)

    tr = cg.trace(ok, 1, 0)
    assert cg.rhs_from_graph(tr, np.array([-2.0]), 0.0, np.array([]))[0] == 2.0

    def bad(y, t, p):
        return (y[0] if y[0] > 0 else -y[0],)

    with pytest.raises(cg.TraceError, match="branches"):
        cg.trace(bad, 1, 0)


def test_wrong_component_count_is_an_error():
    with pytest.raises(ValueError, match="returned 2 components"):
        cg.trace(lambda y, t, p: (y[0], y[1]), 3, 0)


def test_math_module_is_restored_after_tracing():
    cg.trace(lambda y, t, p: (math.exp(y[0]),), 1, 0)
    assert math.exp(0.0) == 1.0 and math.exp.__module__ == "math"


def test_constants_are_emitted_exactly():
    for v in (0.1, 1.0 / 3.0, -2.5e-300, 6.02214076e23, -0.0):
        assert float.fromhex(cg._lit(v)) == v if v != -0.0 else cg._lit(v) == "-0.0"


def test_flop_counts_of_robertson():
    traced = cg.trace(S.robertson, 3, 3)
    _, fc = cg.generate_rhs(traced)
    # -p0*y0 + (p1*y1)*y2 ; p0*y0 - ... - p2*(y1*y1) ; p2*(y1*y1), shared products counted once
    assert 8 <= fc.flops <= 12 and fc.transcendentals == 0


# --- the generated kernels compile for gfx950 (hipRTC needs no GPU) ----------
@pytest.mark.parametrize("lu", ["fp32", "fp64"])
def test_rodas5p_dense_kernel_compiles(lu):
    src, _ = _kernels.rodas5p_source(S.robertson, 3, 3, _kernels.KernelOptions(lu_precision=lu))
    code = _runtime.rtc_compile(src, "t.hip")
    assert code[:4] == b"\x7fELF"


def test_rodas5p_sparse_kernel_compiles():
    from menodax_amd._sparsity import normalize_sparsity

    fn, mask, _ = S.arrow_system(50)
    opts = _kernels.KernelOptions(sparsity=normalize_sparsity(mask, 50), unit_weights=False)
    src, d = _kernels.rodas5p_source(fn, 50, S.arrow_n_params(50), opts)
    assert "sparse_factor" in src and d["structure"].size >= int(mask.sum())
    assert _runtime.rtc_compile(src, "arrow.hip")[:4] == b"\x7fELF"


def test_tsit5_kernel_compiles():
    src, _ = _kernels.tsit5_source(S.lorenz, 3, 1, _kernels.Tsit5Options(unit_weights=False))
    assert _runtime.rtc_compile(src, "lorenz.hip")[:4] == b"\x7fELF"


def test_a_compile_error_reports_the_source():
    with pytest.raises(_runtime.NativeError, match="hipRTC compilation"):
        _runtime.rtc_compile("this is not C++", "bad.hip")


def test_placement_fills_lds_in_order_of_traffic():
    P = _kernels.rodas5p_placement
    assert P(3, "fp32", 64)["placement"] == "registers"
    assert P(3, "fp32", 64)["MDX_ROLLED"] == 0
    d = P(50, "fp32", 64)
    assert (d["MDX_ROLLED"], d["MDX_LDS_K"], d["MDX_LDS_U"], d["MDX_LDS_Y"]) == (1, 1, 1, 1)
    assert d["lds_bytes"] == (8 * 50 * 4 + 2 * 50 * 8) * 64 <= _kernels.LDS_BYTES_PER_CU
    # double stage vectors of 50 components do not fit next to anything: state and argument do
    d = P(50, "fp64", 64)
    assert (d["MDX_LDS_K"], d["MDX_LDS_U"], d["MDX_LDS_Y"]) == (0, 1, 1)
    # 64 components: stage vectors and argument fill the 160 KiB exactly, the state stays out
    d = P(64, "fp32", 64)
    assert (d["MDX_LDS_K"], d["MDX_LDS_U"], d["MDX_LDS_Y"]) == (1, 1, 0)
    assert d["lds_bytes"] == _kernels.LDS_BYTES_PER_CU
    assert P(50, "fp32", 64, "rolled")["lds_bytes"] == 0
    with pytest.raises(ValueError, match="placement"):
        P(50, "fp32", 64, "shared")


def test_the_tableau_goes_to_registers_for_small_systems_and_ensembles():
    """The size rule: the unrolled step of up to three variables, when the
    ensemble fits two waves per SIMD (256 CUs: 131,072 trajectories), never
    under a register cap or another placement; only the plain forward kernel
    gets it, and the kernel still compiles."""
    T = _kernels.tableau_in_registers
    assert T(3, 65536, 256) and T(3, 131072, 256) and T(2, 1, 256)
    assert not T(3, 131073, 256) and not T(4, 1000, 256)
    assert not T(3, 1000, 256, "rolled") and not T(3, 1000, 256, max_registers=128)
    assert T(3, 1000, 256, "registers")
    on = _kernels.KernelOptions(tableau_in_registers=True)
    src, _ = _kernels.rodas5p_source(S.robertson, 3, 3, on)
    assert "#define MDX_TAB_REGS 1" in src
    assert _runtime.rtc_compile(src, "robertson_tab.hip")[:4] == b"\x7fELF"
    adj, _ = _kernels.rodas5p_source(S.robertson, 3, 3, on, adjoint=True)
    assert "#define MDX_TAB_REGS 0" in adj
    rolled, _ = _kernels.rodas5p_source(
        S.robertson, 3, 3, _kernels.KernelOptions(tableau_in_registers=True, placement="rolled"))
    assert "#define MDX_TAB_REGS 0" in rolled


def test_a_rolled_tsit5_step_takes_the_largest_block_its_stage_vectors_fit():
    B = _kernels.tsit5_block
    assert B(None, 3, 1000, 256) == 64 and B(None, 24, 1000, 256) == 64
    assert B(8, 64, 1000, 256) == 8
    assert B(None, 64, 1000, 256) == 32 and B(None, 128, 1000, 256) == 16
    assert B(None, 400, 1000, 256) == 64
    # a large ensemble is not resident at the narrow size: one 16-lane group of
    # the 128-variable ring per CU is 4,096 lanes, fewer than 64-lane groups keep
    assert B(None, 128, 100000, 256) == 64 and B(None, 64, 100000, 256) == 64
    assert B(None, 128, 4096, 256) == 16 and B(None, 128, 4097, 256) == 64
    assert _kernels.tsit5_placement(64, 32)["MDX_LDS_K"] == 1
    assert _kernels.tsit5_placement(64, 64)["MDX_LDS_K"] == 0


def test_a_small_sparse_ensemble_narrows_its_workgroup_to_keep_stage_vectors_in_lds():
    B = _kernels.rodas5p_block
    ring = ((0,),)
    # 128 joint rows, FP32: 4 KB of stage vectors a lane, not at 64 lanes
    assert B(None, 128, "fp32", 1000, 256, None, ring) == 16
    assert B(None, 128, "fp64", 1000, 256, None, ring) == 16
    assert B(None, 64, "fp64", 1000, 256, None, ring) == 32
    # stage vectors fit at 64 lanes; dense; a placement chosen; the caller's size
    assert B(None, 64, "fp32", 1000, 256, None, ring) == 64
    assert B(None, 128, "fp32", 1000, 256, None, None) == 64
    assert B(None, 128, "fp32", 1000, 256, "rolled", ring) == 64
    assert B(8, 128, "fp32", 1000, 256, None, ring) == 8
    # the ensemble not resident at once at the narrow size (a 16-lane group of
    # 128 FP32 joint rows takes 96 KiB: one per CU), or nothing fits at 16
    # lanes: a wavefront
    assert B(None, 128, "fp32", 10000, 256, None, ring) == 64
    assert B(None, 128, "fp32", 4096, 256, None, ring) == 16
    assert B(None, 128, "fp32", 4097, 256, None, ring) == 64
    assert B(None, 128, "fp32", 20000, 256, None, ring) == 64
    assert B(None, 440, "fp32", 1000, 256, None, ring) == 64
    # refined solves: FP64 stage vectors plus the stage right-hand side, 3,600 B
    # a lane at 50 variables, narrow whatever the ensemble size
    assert B(None, 50, "fp32", 100000, 256, None, ring, True) == 32
    assert B(None, 50, "fp32", 100000, 256, None, ring) == 64
    assert B(None, 30, "fp32", 100000, 256, None, ring, True) == 64
    assert B(None, 100, "fp32", 100000, 256, None, ring, True) == 16
    refined = _kernels.rodas5p_placement(50, "fp32", 32, "lds", k_bytes=8)
    assert refined["MDX_LDS_K"] == 1 and refined["lds_bytes"] == (3600 + 800) * 32
    place = _kernels.rodas5p_placement(128, "fp32", 16, "lds")
    assert place["MDX_LDS_K"] == place["MDX_LDS_U"] == place["MDX_LDS_Y"] == 1


def test_the_rolled_adjoint_keeps_its_stage_arrays_in_lds_while_the_ensemble_is_resident():
    A = _kernels.adjoint_block
    # 20-variable chain, 67 factor elements: 3,096 B a lane in FP64, 2,456 in
    # FP32, and 1,120 more with the stage cache (seven double vectors)
    assert _kernels.adjoint_lds_bytes(20, 67, "fp64") == 8 * 20 * 8 * 2 + 8 * 67
    assert _kernels.adjoint_lds_bytes(20, 67, "fp32") == 8 * 20 * 4 + 8 * 20 * 8 + 8 * 67
    assert _kernels.adjoint_lds_bytes(20, 67, "fp64", cache=True) == 3096 + 7 * 20 * 8
    assert A(20, 67, "fp64", 8192, 256) == ("lds", 32, True)
    # the cache at 32 lanes keeps 8,192 resident, where 64 lanes without it would too
    assert A(20, 67, "fp32", 8192, 256) == ("lds", 32, True)
    # 8,193 FP64 trajectories need 257 groups of 32 (one group a CU with or
    # without the cache) or 513 of 16 (two a CU with it): 16 lanes without it,
    # three groups a CU
    assert A(20, 67, "fp64", 8193, 256) == ("lds", 16, False)
    assert A(20, 67, "fp64", 12288, 256) == ("lds", 16, False)
    assert A(20, 67, "fp64", 12289, 256) == ("rolled", 64, False)
    assert A(20, 67, "fp64", 65536, 256) == ("rolled", 64, False)
    assert A(20, 67, "fp32", 65536, 256) == ("rolled", 64, False)
    # nothing fits at 16 lanes
    assert A(200, 2000, "fp64", 10, 256) == ("rolled", 64, False)
    fn, _ = S.vdp_lattice(10, mu=1.0, d=0.1, wrap=False)
    sp = normalize_sparsity(S.vdp_sparsity(10, wrap=False), 20)
    assert _kernels.lu_size(20, sp) == 67 and _kernels.lu_size(20) == 400
    lds, d = _kernels.rodas5p_source(
        fn, 20, 1, _kernels.KernelOptions(lu_precision="fp64", sparsity=sp, placement="lds",
                                          trajectories_per_block=32), adjoint=True)
    assert "#define MDX_ADJ_LDS 1" in lds and "#define MDX_LDS_K 1" in lds
    assert "#define MDX_ADJ_CACHE 1" in lds
    assert d["placement"]["lds_bytes"] == 4216 * 32
    assert _runtime.rtc_compile(lds, "chain_adjoint_lds.hip")[:4] == b"\x7fELF"
    plain, d = _kernels.rodas5p_source(
        fn, 20, 1, _kernels.KernelOptions(lu_precision="fp64", sparsity=sp, placement="lds",
                                          trajectories_per_block=32, adjoint_cache=False),
        adjoint=True)
    assert "#define MDX_ADJ_LDS 1" in plain and "#define MDX_ADJ_CACHE 0" in plain
    assert d["placement"]["lds_bytes"] == 3096 * 32
    # table-driven sparse routines keep the scratch adjoint
    assert _kernels.adjoint_lds_supported(None) and _kernels.adjoint_lds_supported(sp)
    from menodax_amd import _sparse_direct as SD

    saved = SD.MAX_UNROLLED_UPDATES, SD.MAX_UNROLLED_SUBSTITUTIONS
    SD.MAX_UNROLLED_UPDATES = SD.MAX_UNROLLED_SUBSTITUTIONS = 0
    try:
        assert not _kernels.adjoint_lds_supported(sp)
        tab, _ = _kernels.rodas5p_source(
            fn, 20, 1, _kernels.KernelOptions(lu_precision="fp64", sparsity=sp, placement="lds",
                                              trajectories_per_block=32), adjoint=True)
        assert "#define MDX_ADJ_LDS 0" in tab and "#define MDX_LDS_K 0" in tab
    finally:
        SD.MAX_UNROLLED_UPDATES, SD.MAX_UNROLLED_SUBSTITUTIONS = saved
    # at 64 lanes the FP64 arrays miss LDS: the scratch step
    wide, _ = _kernels.rodas5p_source(
        fn, 20, 1, _kernels.KernelOptions(lu_precision="fp64", sparsity=sp, placement="lds"),
        adjoint=True)
    assert "#define MDX_ADJ_LDS 0" in wide and "#define MDX_LDS_K 0" in wide


def test_the_wave_placement_is_for_small_ensembles_of_dense_systems():
    W = _kernels.wave_placement
    assert W(64, 1000, 256) == "wave" and W(24, 1024, 256) == "wave"
    assert W(64, 1025, 256) is None and W(16, 10, 256) is None and W(66, 10, 256) is None
    assert W(64, 10, 256, "lds") == "lds" and W(64, 10, 256, None, sparsity=((0,),)) is None
    fn, _ = S.vdp_lattice(12)
    src, d = _kernels.rodas5p_source(fn, 24, 1, _kernels.KernelOptions(placement="wave"))
    assert "#define MDX_COOP 1" in src and d["block"] == 64
    assert _runtime.rtc_compile(src, "ring_wave.hip")[:4] == b"\x7fELF"
    # the adjoint of a wave solve runs the rolled step
    adj, _ = _kernels.rodas5p_source(fn, 24, 1, _kernels.KernelOptions(placement="wave"),
                                     adjoint=True)
    assert "#define MDX_COOP 0" in adj
    with pytest.raises(ValueError, match="wave"):
        _kernels.rodas5p_placement(24, "fp32", 64, "wave", dense_matrix=0)
    with pytest.raises(ValueError, match="wave"):
        _kernels.rodas5p_placement(80, "fp32", 64, "wave", dense_matrix=6400)


def test_structurally_zero_time_derivative_is_flagged():
    g = cg.generate(cg.trace(S.robertson, 3, 3), colour=[0, 1, 2], n_colours=3,
                    store=[(r, c, r * 3 + c) for r in range(3) for c in range(3)])
    assert g.dt_zero
    g = cg.generate(cg.trace(S.forced_decay, 1, 2), colour=[0], n_colours=1, store=[(0, 0, 0)])
    assert not g.dt_zero


@pytest.mark.parametrize("placement", ["registers", "rolled", "lds"])
def test_every_placement_compiles(placement):
    from menodax_amd._sparsity import normalize_sparsity

    fn, mask, _ = S.arrow_system(12)
    opts = _kernels.KernelOptions(sparsity=normalize_sparsity(mask, 12), placement=placement,
                                  max_registers=128)
    src, d = _kernels.rodas5p_source(fn, 12, S.arrow_n_params(12), opts)
    assert d["placement"]["placement"] == placement
    assert _runtime.rtc_compile(src, f"arrow_{placement}.hip")[:4] == b"\x7fELF"


def test_save_history_false_needs_a_hook():
    from menodax_amd import rodas5P

    with pytest.raises(ValueError, match="save_history=False"):
        rodas5P.solve(S.damped_oscillator, [1.0, 0.0], [0.0, 1.0], [4.0, 0.2], save_history=False)
    with pytest.raises(ValueError, match="positive hook_size"):
        rodas5P.solve(S.damped_oscillator, [1.0, 0.0], [0.0, 1.0], [4.0, 0.2],
                      save_hook=S.energy_hook, hook_size=0)


def test_hook_trace_resolves_the_save_index_and_shares_identical_cases():
    th = cg.trace_hook(S.trapezoid_hook(np.linspace(0, 1, 5)), 2, 2, 5, 3)
    src = cg.generate_hook(th)
    # the two ends share one trapezoid weight and the interior saves another: two cases
    assert src.count("acc.put(2,") == 2 and "case 1: case 2: case 3:" in src
    th = cg.trace_hook(S.energy_hook, 2, 2, 4, 4)
    assert [k for k, _ in th.cases[3]] == [3]
    with pytest.raises(IndexError, match="hook_size"):
        cg.trace_hook(S.energy_hook, 2, 2, 4, 3)


def test_array_rhs_traces_to_the_same_expressions():
    tr = cg.trace(S.robertson, 3, 3)
    arr = cg.trace_array_rhs(S.robertson_array, 3, 3)
    rng = np.random.default_rng(7)
    for _ in range(10):
        y, p = rng.uniform(0, 1, 3), rng.uniform(0.5, 2, 3)
        assert np.array_equal(cg.rhs_from_graph(arr, y, 0.0, p), cg.rhs_from_graph(tr, y, 0.0, p))

    def partial(y, t, p, out):
        out[0] = y[0]

    with pytest.raises(ValueError, match="left out components"):
        cg.trace_array_rhs(partial, 2, 0)


def test_array_rhs_kernel_compiles_with_the_jacobian_of_ode_fn():
    opts = _kernels.KernelOptions(array_rhs=S.robertson_array)
    src, _ = _kernels.rodas5p_source(S.robertson, 3, 3, opts)
    assert "void rhs(" in src and "void neg_jac(" in src
    assert _runtime.rtc_compile(src, "arr.hip")[:4] == b"\x7fELF"


def test_vmap_batched_inputs_are_detected():
    import torch
    from torch.func import vmap

    from menodax_amd import _autodiff

    seen = []

    def f(x):
        seen.append(_autodiff.under_vmap(x))
        return x * 2.0

    vmap(f)(torch.ones(3, 2))
    assert seen == [True] and not _autodiff.under_vmap(torch.ones(2), np.ones(2))


def test_callbacks_built_by_a_factory_of_math_operations_trace():
    """A body written once and closed over its math operations -- the device
    form of examples/dual_backend.py -- traces like the hand-written callback."""
    def make_rhs(*, exp, sqrt, maximum, helper):
        def rhs(y, t, p):
            return (-exp(-y[0]) * p[0] + helper(y[1]), sqrt(maximum(y[1], 0.25)) - y[0])

        return rhs

    def make_helper(*, exp):
        def helper(x):
            return 0.5 * exp(-x * x)

        return helper

    device = make_rhs(exp=math.exp, sqrt=math.sqrt, maximum=max,
                      helper=make_helper(exp=math.exp))

    def direct(y, t, p):
        return (-math.exp(-y[0]) * p[0] + 0.5 * math.exp(-y[1] * y[1]),
                math.sqrt(max(y[1], 0.25)) - y[0])

    a, b = cg.trace(device, 2, 1), cg.trace(direct, 2, 1)
    rng = np.random.default_rng(0)
    for _ in range(10):
        y, p = rng.uniform(-1, 1, 2), rng.uniform(0.5, 2, 1)
        assert np.array_equal(cg.rhs_from_graph(a, y, 0.0, p), cg.rhs_from_graph(b, y, 0.0, p))
        np.testing.assert_allclose(cg.rhs_from_graph(a, y, 0.0, p), device(y, 0.0, p), rtol=1e-14)


def test_a_comparison_used_as_a_number():
    """``a + (b - a) * (x > c)``, the branchless select the reference's callbacks
    write (examples/bbn_estimation/main.py:74-84): traced as select(cond, 1.0,
    0.0) and then the same multiply and add as Python's bool arithmetic, so the
    values are bit-identical on both sides of the step, signed zeros included;
    its derivative is the other factor's."""
    import math

    from menodax_amd import _codegen as C

    def f(y, t, p):
        temp = 1.293 / t
        return (3.91 + (10.75 - 3.91) * (temp > 0.511) * y[0], -y[1] * (t > 2.0),
                (t < 1.0) - (t >= 1.0) + 0.5 * y[0])

    traced = C.trace(f, 3, 0)
    for t in (0.5, 1.0, 2.0, 2.5, 3.0):
        y = [2.0, 3.0, 0.0]
        got = C.rhs_from_graph(traced, y, t, [])
        want = np.array(f(y, t, []), dtype=float)
        assert np.array_equal(got, want) and all(
            math.copysign(1.0, a) == math.copysign(1.0, b) for a, b in zip(got, want))
        J, dT = C.jacobian_from_graph(traced, y, t, [])
        assert J[0, 0] == 3.91 * 0.0 + (10.75 - 3.91) * (1.293 / t > 0.511)
        assert J[1, 1] == -(1.0 if t > 2.0 else 0.0)
        assert J[2, 0] == 0.5 and np.all(dT[:2] == 0.0)


def test_the_bbn_network_traces_to_the_math_form():
    """The dual-form BBN factory traced through the math module equals its plain
    Python evaluation to rounding (integer powers are taken by squaring when
    traced, as numba takes them, and by pow in Python)."""
    from menodax_amd import _codegen as C

    traced = C.trace(S.bbn, 4, 2)
    y = S.bbn_initial_state()
    for x in (0.2, 2.0, 2.6, 40.0):
        for p in ([0.5, 2.0], [0.785, 3.044], [1.0, 4.0]):
            np.testing.assert_allclose(C.rhs_from_graph(traced, y, x, p),
                                       np.array(S.bbn(y, x, p)), rtol=1e-14, atol=0)
