"""Refined FP32 stage solves (``lu_refine=True``) on an MI355X.

The mode has no reference counterpart; what pins it:
  - the oracle's restatement of the refinement (``oracle.rodas5p(...,
    lu_refine=True)``), itself pinned on the CPU by converging to the FP64-LU
    solution (tests/test_oracle.py). One refinement step with FP32 factors
    contracts the error by about cond(M) * eps32, so where the iteration
    matrix is well conditioned the refined stages carry FP64 accuracy and
    where it is not (Robertson late in its solve, J singular by mass
    conservation; the stiffest block arrows, condition up to ~2e4) they keep
    an error of about (cond(M) * eps32)^2 relative. At that level the result
    depends on rounding details: the kernel forms J x with the generated
    forward sweep and the oracle with its closed-form Jacobian summed in
    column order, and the residual, a small difference of large terms, takes
    their last-bit differences into its float rounding. So the bar is the
    method's: the GPU's refined solve lies as close to the FP64-LU solution
    (the FP64 oracle, which the FP64 kernel matches bit for bit) as the
    oracle's refined solve does, within a factor of four; at least 95% of
    the trajectories take exactly the oracle's step counts (printed with the
    share of bit-identical ones);
  - bit-identity across the kernel's own placements and launch shapes (the
    same arithmetic in the same order);
  - the property the mode exists for: the history follows the FP64-LU
    kernel's to 1e-6, where the unrefined FP32 solve does not on the uneven
    stiff ensemble.
"""

import numpy as np
import pytest

from menodax_amd import rodas5P
from tests import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

def _close(tag, hist, st, ref, rst, fp64):
    """``hist`` (GPU, refined) against ``ref`` (oracle, refined) and ``fp64``
    (the FP64-LU oracle)."""
    steps = ((np.asarray(st["accepted_steps"]) == rst["accepted_steps"])
             & (np.asarray(st["rejected_steps"]) == rst["rejected_steps"]))
    bit = np.all(hist == ref, axis=tuple(range(1, hist.ndim)))
    gap_gpu = np.abs(hist - fp64).max()
    gap_oracle = np.abs(ref - fp64).max()
    print(f"{tag}: {steps.mean() * 100:.2f}% same step counts, {bit.mean() * 100:.1f}% "
          f"bit-identical to the oracle, max abs {np.abs(hist - ref).max():.2e}; from the "
          f"FP64 solve: GPU {gap_gpu:.2e}, oracle {gap_oracle:.2e}", flush=True)
    assert steps.mean() >= 0.95
    assert gap_gpu <= 4.0 * gap_oracle + 1e-12


@pytest.mark.parametrize("placement", [None, "rolled"])
def test_robertson_refined_against_the_oracle(placement):
    y0, p = S.robertson_lognormal_ensemble(2000, seed=1)
    t = S.robertson_benchmark_times()
    hist, st = rodas5P.solve(S.robertson, y0, t, p, rtol=1e-6, atol=1e-8, lu_precision="fp32",
                             lu_refine=True, placement=placement, return_stats=True)
    ref, rst = O.rodas5p(O.robertson(), y0, t, p, rtol=1e-6, atol=1e-8, lu_precision="fp32",
                         lu_refine=True)
    fp64, _ = O.rodas5p(O.robertson(), y0, t, p, rtol=1e-6, atol=1e-8, lu_precision="fp64")
    _close(f"robertson refined placement={placement}", hist, st, ref, rst, fp64)


def test_robertson_refined_placements_and_refill_bit_identical():
    y0, p = S.robertson_lognormal_ensemble(4096, seed=2)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8, lu_precision="fp32", lu_refine=True, return_stats=True)
    base, bst = rodas5P.solve(S.robertson, y0, t, p, **kw)
    for extra in (dict(placement="rolled"), dict(refill=True), dict(tableau_in_registers=False),
                  dict(tableau_in_registers=True)):
        h, s = rodas5P.solve(S.robertson, y0, t, p, **kw, **extra)
        np.testing.assert_array_equal(h, base, err_msg=str(extra))
        for k in bst:
            np.testing.assert_array_equal(s[k], bst[k], err_msg=f"{extra} {k}")


def test_robertson_refined_follows_the_fp64_kernel():
    y0, p = S.robertson_lognormal_ensemble(8192, seed=3)
    t = S.robertson_benchmark_times()
    kw = dict(rtol=1e-6, atol=1e-8, return_stats=True)
    h64, s64 = rodas5P.solve(S.robertson, y0, t, p, lu_precision="fp64", **kw)
    hr, sr = rodas5P.solve(S.robertson, y0, t, p, lu_precision="fp32", lu_refine=True, **kw)
    assert np.abs(hr - h64).max() < 1e-6
    assert np.mean(sr["accepted_steps"] == s64["accepted_steps"]) > 0.99


def _uneven(count):
    from menodax_amd import _sparse_direct as SD
    from menodax_amd._sparsity import normalize_sparsity

    fn, mask, (rows, cols, vals, pidx) = S.arrow_system(50)
    sp = normalize_sparsity(mask, 50)
    system = O.linquad(50, rows, cols, vals, 0.01, off_pidx=pidx, n_p=S.arrow_n_params(50))
    layout = O.sparse_structure(sp, SD.fill_reducing_order(sp, "amd"))
    y0, p = S.arrow_ensemble(100000, n=50)
    idx = np.linspace(0, 99999, count).astype(int)
    return fn, mask, system, layout, y0[idx], p[idx], np.linspace(0.0, 10.0, 20)


@pytest.mark.parametrize("placement,block", [("lds", None), ("lds", 32), ("rolled", None)])
def test_uneven_stiff_refined_against_the_oracle(placement, block):
    """The 50-variable block arrow (sparse LU, AMD) over its whole window on
    256 trajectories spread across the 100,000-trajectory ensemble."""
    fn, mask, system, layout, y0, p, t = _uneven(256)
    hist, st = rodas5P.solve(fn, y0, t, p, rtol=1e-6, atol=1e-8, lu_precision="fp32",
                             lu_refine=True, sparsity=mask, placement=placement,
                             trajectories_per_block=block, return_stats=True)
    ref, rst = O.rodas5p(system, y0, t, p, layout=layout, rtol=1e-6, atol=1e-8,
                         lu_precision="fp32", lu_refine=True)
    h64, _ = O.rodas5p(system, y0, t, p, layout=layout, rtol=1e-6, atol=1e-8,
                       lu_precision="fp64")
    _close(f"uneven stiff refined {placement}/{block}", hist, st, ref, rst, h64)
    assert np.abs(hist[:, -1] - h64[:, -1]).max() < 1e-6


def test_uneven_stiff_refined_launch_shapes_and_claim_orders_bit_identical():
    """2,048 spread trajectories of the block arrow: the refined kernel at the
    size rule's workgroup (32 lanes, stage vectors and right-hand sides in
    LDS) equals the 64-lane and scratch forms, plain and refill launches, and
    the refill launch with the pilot claim order."""
    fn, mask, _, _, y0, p, t = _uneven(2048)
    kw = dict(rtol=1e-6, atol=1e-8, lu_precision="fp32", lu_refine=True, sparsity=mask,
              return_stats=True)
    base, bst = rodas5P.solve(fn, y0, t, p, trajectories_per_block=32, refill=False, **kw)
    for extra in (dict(trajectories_per_block=32, refill=True),
                  dict(trajectories_per_block=32, refill=True, work_order="pilot"),
                  dict(trajectories_per_block=64, refill=True),
                  dict(placement="rolled", refill=False)):
        h, st = rodas5P.solve(fn, y0, t, p, **kw, **extra)
        np.testing.assert_array_equal(h, base, err_msg=str(extra))
        for k in bst:
            np.testing.assert_array_equal(st[k], bst[k], err_msg=f"{extra} {k}")


def test_refine_is_rejected_where_it_does_not_apply():
    y0, p = S.robertson_lognormal_ensemble(4, seed=1)
    t = S.robertson_benchmark_times()
    with pytest.raises(ValueError, match="fp32"):
        rodas5P.solve(S.robertson, y0, t, p, lu_precision="fp64", lu_refine=True)
    torch = pytest.importorskip("torch")
    y0_t = torch.tensor(y0, device="cuda", requires_grad=True)
    with pytest.raises(ValueError, match="plain solve"):
        rodas5P.solve(S.robertson, y0_t, torch.tensor(t), torch.tensor(p, device="cuda"),
                      lu_refine=True)
