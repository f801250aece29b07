"""ctypes binding of ``libmenodax_hip.so`` and the code-object cache.

The library is the C ABI declared in ``include/menodax_hip.h``. It is loaded
after ``torch`` so that both share one HIP runtime (they link the same
``libamdhip64.so.7`` soname; the first one loaded serves the process), which is
what lets a launch run on a torch stream and write into torch-allocated memory.

A missing library is an error, never a reason to fall back: there is no CPU
path in the product.
"""

from __future__ import annotations

import ctypes
import hashlib
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmenodax_hip.so")
KERNEL_DIR = os.path.join(_HERE, "kernels")
ARCH = "gfx950"

KERNEL_PERSISTENT = 1
KERNEL_UNIT_WEIGHTS = 2
KERNEL_ADJOINT = 4
KERNEL_HOOK = 8
KERNEL_WAVE = 16
KERNEL_RESUMABLE = 32

# Every kernel is compiled with contraction off (the fused multiply-adds are
# explicit in the source) and without packed FP32 instructions.
COMPILE_OPTIONS = (
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
)

_lib = None
_lock = threading.Lock()


class NativeError(RuntimeError):
    """A call into libmenodax_hip.so failed."""


def lib():
    """The loaded library, built on demand if this tree has never built it."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        import torch  # noqa: F401  (first: its HIP runtime is the process's)

        if not os.path.exists(LIB_PATH):
            raise NativeError(
                f"{LIB_PATH} is not built; run `python -m menodax_amd.build` "
                "(or __graft_entry__.build()) first"
            )
        L = ctypes.CDLL(LIB_PATH)
        i32, i64, vp, cp, dbl, sz = (
            ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_char_p,
            ctypes.c_double, ctypes.c_size_t,
        )
        L.mdx_abi_version.restype = i32
        L.mdx_last_error.restype = cp
        L.mdx_free.argtypes = [vp]
        L.mdx_free.restype = None
        L.mdx_rtc_compile.argtypes = [
            cp, cp, i32, ctypes.POINTER(cp), ctypes.POINTER(cp), i32, ctypes.POINTER(cp),
            ctypes.POINTER(vp), ctypes.POINTER(sz), ctypes.POINTER(ctypes.c_char_p),
        ]
        L.mdx_rtc_compile.restype = i32
        L.mdx_kernel_create.argtypes = [vp, sz, cp, i32, i32, ctypes.POINTER(i64)]
        L.mdx_kernel_create.restype = i32
        L.mdx_kernel_info.argtypes = [i64] + [ctypes.POINTER(i32)] * 4
        L.mdx_kernel_info.restype = i32
        L.mdx_persistent_grid.argtypes = [i64, ctypes.POINTER(i32)]
        L.mdx_persistent_grid.restype = i32
        L.mdx_workspace_bytes.argtypes = [i64, i64]
        L.mdx_workspace_bytes.restype = i64
        L.mdx_solve_launch.argtypes = [
            i64, vp, i32, i64, i32, i32, i32, i32,
            vp, vp, vp, dbl, dbl, dbl, i32, vp,
            vp, vp, vp, vp, vp, i32, vp, i64,
        ]
        L.mdx_solve_launch.restype = i32
        L.mdx_solve_launch_ordered.argtypes = L.mdx_solve_launch.argtypes + [vp]
        L.mdx_solve_launch_ordered.restype = i32
        L.mdx_cost_pilot_launch.argtypes = [
            i64, vp, i64, i32, i32, i32, i32, vp, vp, vp, dbl, dbl, dbl, i32, vp,
            vp, vp, vp, vp, vp,
        ]
        L.mdx_cost_pilot_launch.restype = i32
        L.mdx_cost_pilot_launch_resumable.argtypes = L.mdx_cost_pilot_launch.argtypes + [vp]
        L.mdx_cost_pilot_launch_resumable.restype = i32
        L.mdx_solve_launch_resumed.argtypes = [
            i64, vp, i32, i64, i32, i32, i32, i32, vp, vp, vp, dbl, dbl, dbl, i32, vp,
            vp, vp, vp, vp, vp, i64, vp, vp,
        ]
        L.mdx_solve_launch_resumed.restype = i32
        L.mdx_adjoint_workspace_bytes.argtypes = [i64, i32, i32, i32]
        L.mdx_adjoint_workspace_bytes.restype = i64
        L.mdx_adjoint_launch.argtypes = [
            i64, vp, i64, i32, i32, i32, i32, vp, vp, vp, dbl, dbl, dbl, i32, vp,
            vp, vp, vp, vp, i32, i32, vp, i64,
        ]
        L.mdx_adjoint_launch.restype = i32
        L.mdx_adjoint_launch_static.argtypes = L.mdx_adjoint_launch.argtypes[:-2]
        L.mdx_adjoint_launch_static.restype = i32
        L.mdx_device_info.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(i32)]
        L.mdx_device_info.restype = i32
        if L.mdx_abi_version() != 1:
            raise NativeError("libmenodax_hip.so has an unexpected ABI version; rebuild it")
        _lib = L
        return L


def _check(rc, what):
    if rc != 0:
        msg = lib().mdx_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed ({rc}): {msg}")


def headers():
    """The kernel headers, as (name, text) pairs handed to hipRTC in memory."""
    out = []
    for name in sorted(os.listdir(KERNEL_DIR)):
        if name.endswith(".hpp"):
            with open(os.path.join(KERNEL_DIR, name)) as f:
                out.append((name, f.read()))
    return out


#: code objects compiled ahead of time by ``build`` travel with the tree here
PACKAGE_CACHE = os.path.join(_HERE, "_kcache")


def _cache_dirs():
    """Lookup order for compiled code objects; the first writable one is where
    a fresh compile is stored."""
    dirs = []
    env = os.environ.get("MENODAX_KERNEL_CACHE")
    if env:
        dirs.append(env)
    dirs.append(PACKAGE_CACHE)
    dirs.append(os.path.join(os.path.expanduser("~"), ".cache", "menodax_amd", "kernels"))
    return dirs


def cache_key(source: str, hdrs, options) -> str:
    h = hashlib.sha256()
    h.update(b"mdx-code-object-v1\0")
    h.update(source.encode())
    for name, text in hdrs:
        h.update(b"\0" + name.encode() + b"\0" + text.encode())
    for o in options:
        h.update(b"\0" + o.encode())
    return h.hexdigest()


def rtc_compile(source: str, name: str = "mdx_kernel.hip", hdrs=None, options=COMPILE_OPTIONS) -> bytes:
    """Compile with hipRTC (no cache)."""
    L = lib()
    hdrs = headers() if hdrs is None else hdrs
    n_h = len(hdrs)
    h_src = (ctypes.c_char_p * max(n_h, 1))(*[t.encode() for _, t in hdrs])
    h_nam = (ctypes.c_char_p * max(n_h, 1))(*[n.encode() for n, _ in hdrs])
    opts = (ctypes.c_char_p * len(options))(*[o.encode() for o in options])
    code = ctypes.c_void_p()
    size = ctypes.c_size_t()
    log = ctypes.c_char_p()
    rc = L.mdx_rtc_compile(
        source.encode(), name.encode(), n_h, h_src, h_nam, len(options), opts,
        ctypes.byref(code), ctypes.byref(size), ctypes.byref(log),
    )
    text = ""
    if log.value is not None:
        text = log.value.decode(errors="replace")
        L.mdx_free(ctypes.cast(log, ctypes.c_void_p))
    if rc != 0:
        raise NativeError(f"hipRTC compilation of {name} failed:\n{text}\n--- source ---\n{source}")
    try:
        return ctypes.string_at(code, size.value)
    finally:
        L.mdx_free(code)


_memory_cache: dict[str, bytes] = {}


def code_object(source: str, name: str = "mdx_kernel.hip", store_in: str | None = None) -> bytes:
    """Compile through the cache: memory, then disk (keyed on source, headers
    and options), then hipRTC. ``store_in`` forces a copy into that directory."""
    hdrs = headers()
    key = cache_key(source, hdrs, COMPILE_OPTIONS)
    if store_in is not None:
        data = code_object(source, name)
        os.makedirs(store_in, exist_ok=True)
        path = os.path.join(store_in, key + ".co")
        if not os.path.exists(path):
            tmp = f"{path}.{os.getpid()}.tmp"
            with open(tmp, "wb") as f:
                f.write(data)
            os.replace(tmp, path)
        return data
    hit = _memory_cache.get(key)
    if hit is not None:
        return hit
    for d in _cache_dirs():
        path = os.path.join(d, key + ".co")
        if os.path.exists(path):
            with open(path, "rb") as f:
                data = f.read()
            _memory_cache[key] = data
            return data
    data = rtc_compile(source, name, hdrs)
    _memory_cache[key] = data
    for d in _cache_dirs():
        try:
            os.makedirs(d, exist_ok=True)
            tmp = os.path.join(d, f"{key}.{os.getpid()}.tmp")
            with open(tmp, "wb") as f:
                f.write(data)
            os.replace(tmp, os.path.join(d, key + ".co"))
            break
        except OSError:
            continue
    return data


def kernel_create(code: bytes, entry: str, block: int, flags: int) -> int:
    handle = ctypes.c_int64()
    buf = ctypes.create_string_buffer(code, len(code))
    _check(
        lib().mdx_kernel_create(buf, len(code), entry.encode(), int(block), int(flags), ctypes.byref(handle)),
        f"mdx_kernel_create({entry})",
    )
    return int(handle.value)


def kernel_info(handle: int) -> dict:
    v, s, l, b = (ctypes.c_int32() for _ in range(4))
    _check(
        lib().mdx_kernel_info(handle, ctypes.byref(v), ctypes.byref(s), ctypes.byref(l), ctypes.byref(b)),
        "mdx_kernel_info",
    )
    return dict(vgprs=v.value, scratch_bytes=s.value, lds_bytes=l.value, blocks_per_cu=b.value)


def persistent_grid(handle: int) -> int:
    g = ctypes.c_int32()
    _check(lib().mdx_persistent_grid(handle, ctypes.byref(g)), "mdx_persistent_grid")
    return int(g.value)


def workspace_bytes(handle: int, n_traj: int) -> int:
    v = lib().mdx_workspace_bytes(handle, n_traj)
    if v < 0:
        _check(1, "mdx_workspace_bytes")
    return int(v)


def device_info():
    buf = ctypes.create_string_buffer(128)
    n_cu = ctypes.c_int32()
    _check(lib().mdx_device_info(buf, 128, ctypes.byref(n_cu)), "mdx_device_info")
    return buf.value.decode(), int(n_cu.value)


def adjoint_workspace_bytes(n_traj, n_vars, ckpt_every, max_ckpt) -> int:
    v = lib().mdx_adjoint_workspace_bytes(int(n_traj), int(n_vars), int(ckpt_every), int(max_ckpt))
    if v < 0:
        raise ValueError("mdx_adjoint_workspace_bytes: bad extents")
    return int(v)


def adjoint_launch(handle, stream, n_traj, n_vars, n_params, n_save, layout, y0, times, params,
                   dt0, rtol, atol, max_steps, weights, gbar, grad_y0, grad_p, status,
                   ckpt_every, max_ckpt, workspace, workspace_bytes):
    """Launch through ``mdx_adjoint_launch``; pointers are plain integers."""
    _check(
        lib().mdx_adjoint_launch(
            handle, stream, int(n_traj), int(n_vars), int(n_params), int(n_save), int(layout),
            y0, times, params or None, float(dt0), float(rtol), float(atol), int(max_steps),
            weights or None, gbar, grad_y0, grad_p or None, status, int(ckpt_every),
            int(max_ckpt), workspace, int(workspace_bytes),
        ),
        "mdx_adjoint_launch",
    )


#: bytes of the static device scratch behind ``adjoint_launch_static``
#: (``MDX_STATIC_SCRATCH_BYTES`` in include/menodax_hip.h)
STATIC_SCRATCH_BYTES = 64 << 20


def adjoint_launch_static(handle, stream, n_traj, n_vars, n_params, n_save, layout, y0, times,
                          params, dt0, rtol, atol, max_steps, weights, gbar, grad_y0, grad_p,
                          status, ckpt_every, max_ckpt):
    """Launch through ``mdx_adjoint_launch_static``: as ``adjoint_launch``, with
    the checkpoints in the library's static device scratch instead of a
    workspace (nothing is allocated)."""
    _check(
        lib().mdx_adjoint_launch_static(
            handle, stream, int(n_traj), int(n_vars), int(n_params), int(n_save), int(layout),
            y0, times, params or None, float(dt0), float(rtol), float(atol), int(max_steps),
            weights or None, gbar, grad_y0, grad_p or None, status, int(ckpt_every),
            int(max_ckpt),
        ),
        "mdx_adjoint_launch_static",
    )


def solve_launch(handle, stream, grid, n_traj, n_vars, n_params, n_save, layout, y0, times, params,
                 dt0, rtol, atol, max_steps, weights, hist, accepted, rejected, loop_steps,
                 hook_out=0, hook_size=0, workspace=0, workspace_bytes=0, order=0, resume=0):
    """Launch through ``mdx_solve_launch`` (``mdx_solve_launch_ordered`` when a
    claim ``order`` pointer is given, ``mdx_solve_launch_resumed`` when a
    pilot's step-state buffer ``resume`` is); pointers are plain integers."""
    if resume:
        if hook_out or hook_size:
            raise ValueError("a resumed solve runs no save hook")
        _check(lib().mdx_solve_launch_resumed(
            handle, stream, int(grid), int(n_traj), int(n_vars), int(n_params), int(n_save),
            int(layout), y0, times, params, float(dt0), float(rtol), float(atol),
            int(max_steps), weights or None, hist, accepted, rejected, loop_steps,
            workspace or None, int(workspace_bytes), order or None, resume),
            "mdx_solve_launch_resumed")
        return
    args = (handle, stream, int(grid), int(n_traj), int(n_vars), int(n_params), int(n_save),
            int(layout), y0, times, params, float(dt0), float(rtol), float(atol), int(max_steps),
            weights or None, hist, accepted, rejected, loop_steps, hook_out or None,
            int(hook_size), workspace or None, int(workspace_bytes))
    if order:
        _check(lib().mdx_solve_launch_ordered(*args, order), "mdx_solve_launch_ordered")
    else:
        _check(lib().mdx_solve_launch(*args), "mdx_solve_launch")


def cost_pilot_launch(handle, stream, n_traj, n_vars, n_params, n_save, layout, y0, times, params,
                      dt0, rtol, atol, steps, weights, hist, accepted, rejected, loop_steps, cost,
                      resume=0):
    """Launch through ``mdx_cost_pilot_launch`` (``_resumable`` when a
    step-state buffer ``resume`` is given); pointers are plain integers."""
    args = (handle, stream, int(n_traj), int(n_vars), int(n_params), int(n_save), int(layout),
            y0, times, params, float(dt0), float(rtol), float(atol), int(steps),
            weights or None, hist, accepted, rejected, loop_steps, cost)
    if resume:
        _check(lib().mdx_cost_pilot_launch_resumable(*args, resume),
               "mdx_cost_pilot_launch_resumable")
    else:
        _check(lib().mdx_cost_pilot_launch(*args), "mdx_cost_pilot_launch")
