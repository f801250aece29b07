"""The adjoint entry point that allocates nothing (mdx_adjoint_launch_static),
called through the C ABI on the GPU and held bit for bit against
mdx_adjoint_launch with a caller's workspace on the same inputs: two calls in a
row, the second with fewer trajectories (so it runs over what the first left
in the shared static scratch), and the rejection of a request larger than the
scratch.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from menodax_amd import _common, _kernels, _runtime, rodas5P  # noqa: E402
from tests import systems as S  # noqa: E402

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    mu = 10.0 ** rng.uniform(1.0, 2.0, size=(n, 1))
    y0 = rng.uniform(-2.0, 2.0, size=(n, 2))
    return y0, mu


def _both(n, seed, ckpt_every):
    """Gradients by both entry points for n trajectories of the stiff VdP."""
    dev = torch.device("cuda", 0)
    fn = S.vdp_lattice(1, mu=1.0, d=0.1, wrap=False)[0]
    options = _kernels.KernelOptions(lu_precision="fp64")
    built = _kernels.rodas5p_kernel(fn, 2, 1, options)
    adj = _kernels.rodas5p_adjoint_kernel(fn, 2, 1, options)
    y0, mu = _inputs(n, seed)
    times = torch.linspace(0.0, 3.0, 20, dtype=torch.float64, device=dev)
    y0_t = torch.tensor(y0, device=dev)
    p_t = torch.tensor(mu, device=dev)
    out = rodas5P.launch(built, y0_t, times, p_t, n, dt0=0.0, rtol=1e-8, atol=1e-10,
                         max_steps=100000)
    max_ckpt = int(out.accepted.max().item()) // ckpt_every + 1
    gen = torch.Generator(device="cpu").manual_seed(seed)
    # (n_save, n_vars, N): the structure-of-arrays layout of the history
    gbar = torch.randn((20, 2, n), generator=gen, dtype=torch.float64).to(dev)
    y0_s, p_s = _common.soa(y0_t), _common.soa(p_t)
    stream = torch.cuda.current_stream(dev).cuda_stream
    res = []
    for static in (False, True):
        gy = torch.full((2, n), np.nan, dtype=torch.float64, device=dev)
        gp = torch.full((1, n), np.nan, dtype=torch.float64, device=dev)
        status = torch.full((n,), -7, dtype=torch.int32, device=dev)
        common = (adj.plain, stream, n, 2, 1, 20, _common.LAYOUT_SOA, y0_s.data_ptr(),
                  times.data_ptr(), p_s.data_ptr(), 0.0, 1e-8, 1e-10, 100000, 0,
                  gbar.data_ptr(), gy.data_ptr(), gp.data_ptr(), status.data_ptr(),
                  ckpt_every, max_ckpt)
        if static:
            _runtime.adjoint_launch_static(*common)
        else:
            ws_bytes = _runtime.adjoint_workspace_bytes(n, 2, ckpt_every, max_ckpt)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
            _runtime.adjoint_launch(*common, ws.data_ptr(), ws_bytes)
        torch.cuda.synchronize()
        res.append((gy.cpu().numpy(), gp.cpu().numpy(), status.cpu().numpy()))
    return res


def test_static_scratch_matches_the_workspace_entry_point_twice():
    print("no-alloc adjoint: N=512 then N=96, ckpt_every 7", flush=True)
    for n, seed in ((512, 1), (96, 2)):
        (gy_w, gp_w, st_w), (gy_s, gp_s, st_s) = _both(n, seed, 7)
        assert np.all(st_w == 0) and np.array_equal(st_s, st_w)
        assert np.all(np.isfinite(gy_w)) and np.all(np.isfinite(gp_w))
        assert np.array_equal(gy_s, gy_w), n
        assert np.array_equal(gp_s, gp_w), n

