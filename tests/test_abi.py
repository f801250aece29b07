"""The C ABI of libmenodax_hip.so, without a GPU.

Loads the library, checks it exports every function ``include/menodax_hip.h``
declares, and exercises the calls that need no device: compilation through
hipRTC, kernel registration, and the launcher's argument validation, which
rejects malformed launches before touching the GPU (the reference's handler
does the same with kInvalidArgument, modax/_jax_numba_custom_call.py:137-196).
"""

import ctypes
import os
import re

import pytest

from menodax_amd import _kernels, _runtime
from tests import systems as S

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "menodax_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|int64_t|void|const char\*)\s+(mdx_\w+)\(", text, re.M)))


def test_library_exports_every_declared_function():
    lib = _runtime.lib()
    names = _declared()
    assert len(names) >= 9
    for name in names:
        assert hasattr(lib, name), name
    assert lib.mdx_abi_version() == 1


def test_unknown_handle_is_reported():
    with pytest.raises(_runtime.NativeError, match="unknown kernel handle"):
        _runtime.solve_launch(999999, None, 0, 4, 3, 3, 5, 1, 1, 1, 1, 0.0, 1e-6, 1e-8, 10, 0,
                              1, 1, 1, 1)


@pytest.fixture(scope="module")
def handles():
    src, _ = _kernels.rodas5p_source(S.robertson, 3, 3, _kernels.KernelOptions(unit_weights=False))
    code = _runtime.rtc_compile(src, "abi.hip")
    plain = _runtime.kernel_create(code, "mdx_rodas5p", 64, 0)
    refill = _runtime.kernel_create(code, "mdx_rodas5p_refill", 64, _runtime.KERNEL_PERSISTENT)
    return plain, refill


def _launch(handle, **over):
    kw = dict(grid=0, n_traj=4, n_vars=3, n_params=3, n_save=5, layout=1, y0=8, times=8, params=8,
              dt0=0.0, rtol=1e-6, atol=1e-8, max_steps=10, weights=8, hist=8, accepted=8,
              rejected=8, loop_steps=8, hook_out=0, hook_size=0, workspace=0, workspace_bytes=0)
    kw.update(over)
    _runtime.solve_launch(handle, None, **kw)


@pytest.mark.parametrize(
    "override, message",
    [
        (dict(n_traj=-1), "bad extents"),
        (dict(n_vars=0), "bad extents"),
        (dict(n_save=0), "bad extents"),
        (dict(layout=7), "layout must be 0 or 1"),
        (dict(hist=0), "null buffer"),
        (dict(weights=0), "reads error weights"),
        (dict(max_steps=-3), "max_steps"),
        (dict(hook_size=2), "needs hook_out"),
        (dict(n_traj=200, grid=3), "grid >= ceil"),
    ],
)
def test_malformed_launches_are_rejected_before_the_gpu(handles, override, message):
    with pytest.raises(_runtime.NativeError, match=message):
        _launch(handles[0], **override)


def test_a_persistent_kernel_needs_its_workspace(handles):
    with pytest.raises(_runtime.NativeError, match="workspace"):
        _launch(handles[1])


def test_a_claim_order_needs_a_persistent_kernel(handles):
    with pytest.raises(_runtime.NativeError, match="needs a persistent"):
        _launch(handles[0], order=8)


def test_a_cost_pilot_runs_a_one_lane_kernel(handles):
    kw = dict(n_traj=4, n_vars=3, n_params=3, n_save=5, layout=1, y0=8, times=8, params=8,
              dt0=0.0, rtol=1e-6, atol=1e-8, steps=10, weights=8, hist=8, accepted=8,
              rejected=8, loop_steps=8, cost=8)
    with pytest.raises(_runtime.NativeError, match="one-lane-per-trajectory"):
        _runtime.cost_pilot_launch(handles[1], None, **kw)
    with pytest.raises(_runtime.NativeError, match="steps must be positive"):
        _runtime.cost_pilot_launch(handles[0], None, **dict(kw, steps=0))
    with pytest.raises(_runtime.NativeError, match="null cost"):
        _runtime.cost_pilot_launch(handles[0], None, **dict(kw, cost=0))


def test_a_step_state_buffer_needs_a_resumable_kernel(handles):
    """The pilot's step state goes only to kernels created with
    MDX_KERNEL_RESUMABLE (the fixture's are not), and both resumable entry
    points refuse a null buffer; all before touching the GPU."""
    kw = dict(n_traj=4, n_vars=3, n_params=3, n_save=5, layout=1, y0=8, times=8, params=8,
              dt0=0.0, rtol=1e-6, atol=1e-8, steps=10, weights=8, hist=8, accepted=8,
              rejected=8, loop_steps=8, cost=8)
    with pytest.raises(_runtime.NativeError, match="MDX_KERNEL_RESUMABLE"):
        _runtime.cost_pilot_launch(handles[0], None, **kw, resume=8)
    with pytest.raises(_runtime.NativeError, match="MDX_KERNEL_RESUMABLE"):
        _launch(handles[1], workspace=8, workspace_bytes=256, resume=8)
    lib = _runtime.lib()
    args = [handles[0], None, 4, 3, 3, 5, 1, 8, 8, 8, 0.0, 1e-6, 1e-8, 10, 8, 8, 8, 8, 8, 8]
    assert lib.mdx_cost_pilot_launch_resumable(*args, None) != 0
    assert b"null step-state" in lib.mdx_last_error()
    assert lib.mdx_solve_launch_resumed(handles[1], None, 0, 4, 3, 3, 5, 1, 8, 8, 8, 0.0, 1e-6,
                                        1e-8, 10, 8, 8, 8, 8, 8, 8, 256, None, None) != 0
    assert b"null step-state" in lib.mdx_last_error()


def test_the_step_state_rows_of_each_method():
    """What a resumable pilot leaves per trajectory, as the kernels lay it out:
    Rodas5P the state and eight step scalars, Tsit5 also its
    first-same-as-last stage and whether that stage holds."""
    from menodax_amd import _kernels

    assert _kernels.resume_rows("rodas5p", 3) == 3 + 8
    assert _kernels.resume_rows("rodas5p", 50) == 58
    assert _kernels.resume_rows("tsit5", 3) == 2 * 3 + 9
    assert _kernels.resume_rows("tsit5", 24) == 57


def test_work_order_must_be_a_permutation():
    import torch

    from menodax_amd import _common

    cpu = torch.device("cpu")
    assert _common.work_order([2, 0, 1], 3, cpu).tolist() == [2, 0, 1]
    for bad in ([0, 0, 1], [0, 1], [0, 1, 3]):
        with pytest.raises(ValueError, match="work_order"):
            _common.work_order(bad, 3, cpu)
    assert _common.cost_order([5, 9, 1, 9]).tolist() == [1, 3, 0, 2]
    assert _common.work_order_arg("pilot", 3, cpu) == "pilot"
    with pytest.raises(ValueError, match="pilot"):
        _common.work_order_arg("fastest", 3, cpu)


def test_an_empty_ensemble_launches_nothing(handles):
    _launch(handles[0], n_traj=0)  # returns before any device call


def test_an_empty_ensemble_may_pass_null_buffers(handles):
    """Zero-size device arrays have null data pointers; with N = 0 every entry
    point accepts them and returns before any device call, while N > 0 still
    rejects them."""
    nulls = dict(y0=0, times=0, params=0, weights=0, hist=0, accepted=0, rejected=0, loop_steps=0)
    _launch(handles[0], n_traj=0, **nulls)
    _launch(handles[1], n_traj=0, **nulls)  # persistent, and no workspace either
    with pytest.raises(_runtime.NativeError, match="null buffer"):
        _launch(handles[0], **nulls)
    lib = _runtime.lib()
    args = [handles[0], None, 0, 3, 3, 5, 1, 0, 0, 0, 0.0, 1e-6, 1e-8, 10, 0, 0, 0, 0, 0]
    assert lib.mdx_cost_pilot_launch(*args, None) == 0
    fn = S.vdp_lattice(1, mu=1.0, d=0.1, wrap=False)[0]
    adj = _kernels.rodas5p_adjoint_kernel(fn, 2, 1, _kernels.KernelOptions(lu_precision="fp64"))
    _runtime.adjoint_launch(adj.plain, None, 0, 2, 1, 20, 1, 0, 0, 0, 0.0, 1e-8, 1e-10, 100, 0,
                            0, 0, 0, 0, 1, 1, 0, 0)
    with pytest.raises(_runtime.NativeError, match="null buffer"):
        _runtime.adjoint_launch(adj.plain, None, 4, 2, 1, 20, 1, 0, 0, 0, 0.0, 1e-8, 1e-10, 100,
                                0, 0, 0, 0, 0, 1, 1, 0, 0)


def test_kernel_create_validates():
    with pytest.raises(_runtime.NativeError, match="block must be"):
        _runtime.kernel_create(b"\x7fELF....", "x", 0, 0)
    # a wave-per-trajectory kernel is a plain solve kernel
    for other in (_runtime.KERNEL_PERSISTENT, _runtime.KERNEL_ADJOINT):
        with pytest.raises(_runtime.NativeError, match="wave-per-trajectory"):
            _runtime.kernel_create(b"\x7fELF....", "x", 64, _runtime.KERNEL_WAVE | other)


def test_compile_returns_a_gfx950_code_object():
    src, _ = _kernels.tsit5_source(S.lorenz, 3, 1, _kernels.Tsit5Options())
    code = _runtime.rtc_compile(src, "lorenz.hip")
    assert code[:4] == b"\x7fELF"
    assert b"gfx950" in code


def test_last_error_is_thread_local_text():
    lib = _runtime.lib()
    assert isinstance(lib.mdx_last_error(), bytes)
    assert ctypes.c_int32(lib.mdx_abi_version()).value == 1


def test_the_static_adjoint_rejects_a_request_beyond_its_scratch():
    """mdx_adjoint_launch_static refuses, before touching the GPU, checkpoints
    that would not fit its static device scratch."""
    fn = S.vdp_lattice(1, mu=1.0, d=0.1, wrap=False)[0]
    adj = _kernels.rodas5p_adjoint_kernel(fn, 2, 1, _kernels.KernelOptions(lu_precision="fp64"))
    # one checkpoint (2 + 8 words) and one step record (2 + 4 words) per trajectory
    fits = _runtime.STATIC_SCRATCH_BYTES // (8 * (10 + 6))
    assert _runtime.adjoint_workspace_bytes(fits, 2, 1, 1) <= _runtime.STATIC_SCRATCH_BYTES
    assert _runtime.adjoint_workspace_bytes(fits + 1, 2, 1, 1) > _runtime.STATIC_SCRATCH_BYTES
    with pytest.raises(_runtime.NativeError, match="static scratch"):
        _runtime.adjoint_launch_static(adj.plain, None, fits + 1, 2, 1, 20, 1, 8, 8, 8, 0.0,
                                       1e-8, 1e-10, 100, 0, 8, 8, 8, 8, 1, 1)
    with pytest.raises(_runtime.NativeError, match="null buffer"):
        _runtime.adjoint_launch_static(adj.plain, None, 4, 2, 1, 20, 1, 8, 8, 8, 0.0,
                                       1e-8, 1e-10, 100, 0, 8, 0, 8, 8, 1, 1)


def test_the_static_scratch_size_matches_the_header():
    text = open(HEADER).read()
    m = re.search(r"#define MDX_STATIC_SCRATCH_BYTES \((\d+)ll << (\d+)\)", text)
    assert m and int(m.group(1)) << int(m.group(2)) == _runtime.STATIC_SCRATCH_BYTES
