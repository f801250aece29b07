/*
 * menodax_hip.h -- C ABI of the MI355X ensemble ODE engine (libmenodax_hip.so).
 *
 * This is the drop-in boundary. The reference (lawrenceberry/menodax, package
 * `modax`) reaches its GPU through three things in
 * modax/_jax_numba_custom_call.py: a compile-and-load step that turns a numba
 * kernel into a CUfunction (compile_kernel, :286-314), a launcher that takes
 * that function, a stream, the launch shape and the solver's buffers
 * (LaunchNumbaCudaAbi, :145-213), and the XLA FFI symbol that exposes the
 * launcher (modax_numba_cuda_abi_launch, :215-231). The entry points below
 * replace them one for one. An XLA FFI handler is a thin wrapper over
 * mdx_solve_launch that reads its buffers and attributes and calls it; it is not
 * shipped, because no jaxlib with the FFI headers exists on the machines this
 * library is built and tested on.
 *
 * Conventions: every function returns 0 on success and a nonzero code on
 * failure, with the message available from mdx_last_error() on the calling
 * thread. Kernel handles are int64 keys into a process-lifetime registry, like
 * the reference's `function` attribute (:348-362). Nothing here allocates
 * device memory on a launch, and nothing synchronises the stream. A launch
 * of zero trajectories validates its extents and returns; its buffers may be
 * null (the data pointer of a zero-size allocation).
 */
#ifndef MENODAX_HIP_H
#define MENODAX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Version of this ABI; bumped on any signature change. */
int32_t mdx_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* mdx_last_error(void);

/* Compile HIP source to a gfx950 code object with hipRTC.
 * Replaces the numba-cuda-mlir compile inside compile_kernel
 * (modax/_jax_numba_custom_call.py:295-297). `headers`/`header_names` are
 * in-memory include files; `options` are compiler options. On success
 * *code and *log are malloc'd (free with mdx_free); on failure *log holds the
 * compiler's diagnostics. */
int32_t mdx_rtc_compile(const char* source, const char* name, int32_t n_headers,
                        const char* const* headers, const char* const* header_names,
                        int32_t n_options, const char* const* options, void** code,
                        size_t* code_size, char** log);

void mdx_free(void* p);

/* Kernel flags for mdx_kernel_create. */
#define MDX_KERNEL_PERSISTENT 1  /* refill kernel; module also holds mdx_reset_queue */
#define MDX_KERNEL_UNIT_WEIGHTS 2 /* compiled with unit error weights: `weights` unread */
#define MDX_KERNEL_ADJOINT 4      /* a discrete-adjoint kernel, launched by mdx_adjoint_launch */
#define MDX_KERNEL_HOOK 8         /* runs a save hook: hook_out (N, hook_size) required */
#define MDX_KERNEL_WAVE 16        /* one trajectory per workgroup of `block` lanes: grid = N */
#define MDX_KERNEL_RESUMABLE 32   /* reads / writes a step-state buffer (the *_resum* entries) */

/* Register a kernel from a code object: `entry` is its symbol, `block` the
 * workgroup size it was compiled for, `flags` a combination of the
 * MDX_KERNEL_* bits above. The code object is loaded on each device when first
 * launched there; handles live as long as the process.
 * Replaces cuModuleLoadData/cuModuleGetFunction in compile_kernel
 * (modax/_jax_numba_custom_call.py:299-314). */
int32_t mdx_kernel_create(const void* code, size_t code_size, const char* entry, int32_t block,
                          int32_t flags, int64_t* handle);

/* Resource use of a registered kernel on the current device: vector registers
 * per lane, scratch (private) bytes per lane, static LDS bytes, and the
 * workgroups per compute unit it can keep resident at its block size. No
 * counterpart in the reference (numba's occupancy is not exposed there). */
int32_t mdx_kernel_info(int64_t handle, int32_t* vgprs, int32_t* scratch_bytes,
                        int32_t* lds_bytes, int32_t* blocks_per_cu);

/* Workgroups a persistent launch should use on the current device: compute
 * units times the resident workgroups per unit. */
int32_t mdx_persistent_grid(int64_t handle, int32_t* grid);

/* Device memory a launch of `handle` needs as `workspace` (the refill
 * counter), in bytes. */
int64_t mdx_workspace_bytes(int64_t handle, int64_t n_traj);

/* Launch one ensemble solve on `stream` (a hipStream_t; NULL is the default
 * stream). Replaces LaunchNumbaCudaAbi (modax/_jax_numba_custom_call.py:145-213)
 * with the buffers in the reference kernel's argument order
 * (modax/_numba_common.py:27-52, 84-130, rodas5P.py:680-694):
 *   inputs  y0 (N, n_vars), times (n_save,), params (N, n_params),
 *           scalars dt0, rtol, atol, max_steps, weights (N, n_vars);
 *   outputs hist (N, n_save, n_vars), accepted, rejected, loop_steps (N,) int32,
 *           hook_out (N, hook_size) or NULL.
 * `layout` 0: every array trajectory-major and contiguous, as the reference
 * passes them. `layout` 1: structure of arrays -- y0/weights (n_vars, N),
 * params (n_params, N), hist (n_save, n_vars, N) -- the coalesced form.
 * `grid` 0 sizes a one-lane-per-trajectory launch from N, or a persistent
 * kernel's resident grid; a positive `grid` sets the workgroup count (any
 * count for a persistent kernel, whose lanes claim trajectories from the
 * queue; at least ceil(N / block) for a one-lane-per-trajectory kernel, which
 * would otherwise leave trajectories unsolved, so a smaller one is an error). `weights` may be NULL only for a kernel registered
 * with MDX_KERNEL_UNIT_WEIGHTS. `workspace` must hold mdx_workspace_bytes; it
 * is reset on the stream before the solve. A kernel registered with
 * MDX_KERNEL_HOOK zeroes and accumulates hook_out (N, hook_size), laid out like
 * params; one compiled without the history writes only hist's first save row
 * (the final state), so hist may hold that one row. */
int32_t mdx_solve_launch(int64_t handle, void* stream, int32_t grid, int64_t n_traj,
                         int32_t n_vars, int32_t n_params, int32_t n_save, int32_t layout,
                         const double* y0, const double* times, const double* params, double dt0,
                         double rtol, double atol, int32_t max_steps, const double* weights,
                         double* hist, int32_t* accepted, int32_t* rejected,
                         int32_t* loop_steps, double* hook_out, int32_t hook_size,
                         void* workspace, int64_t workspace_bytes);

/* mdx_solve_launch with a claim order for a persistent (refill) kernel: its
 * lanes take trajectories order[0], order[1], ... as lanes free up, instead of
 * 0, 1, .... `order` is a device array of N int64, a permutation of 0..N-1
 * (the host checks that; an entry outside [0, N) is skipped on the device
 * rather than read past the arrays); each trajectory's results do not depend
 * on it. A caller that can rank its trajectories by cost -- the step counts of
 * an earlier solve of the same or a similar ensemble -- claims the expensive
 * ones first and shortens the ensemble's tail. NULL is mdx_solve_launch; a
 * one-lane-per-trajectory kernel takes no order. No counterpart in the
 * reference, whose launch gives every trajectory a thread of its own
 * (modax/rodas5P.py:1010-1021). */
int32_t mdx_solve_launch_ordered(int64_t handle, void* stream, int32_t grid, int64_t n_traj,
                                 int32_t n_vars, int32_t n_params, int32_t n_save, int32_t layout,
                                 const double* y0, const double* times, const double* params,
                                 double dt0, double rtol, double atol, int32_t max_steps,
                                 const double* weights, double* hist, int32_t* accepted,
                                 int32_t* rejected, int32_t* loop_steps, double* hook_out,
                                 int32_t hook_size, void* workspace, int64_t workspace_bytes,
                                 const int64_t* order);

/* A cost pilot for a claim order: the one-lane-per-trajectory kernel runs each
 * trajectory for at most `steps` steps and writes cost[i], the steps left at
 * its current step size ((t_end - t) / dt; 0 when it finished), instead of its
 * remaining saves and counters. Trajectories by decreasing cost are a claim
 * order for mdx_solve_launch_ordered (a first solve's estimate of what an
 * earlier solve's step counts give exactly). hist and the counters receive the
 * pilot's partial results and are scratch to the caller; the other buffers as
 * mdx_solve_launch. The kernel must not be persistent or run a save hook. No
 * counterpart in the reference. */
int32_t mdx_cost_pilot_launch(int64_t handle, void* stream, int64_t n_traj, int32_t n_vars,
                              int32_t n_params, int32_t n_save, int32_t layout, const double* y0,
                              const double* times, const double* params, double dt0, double rtol,
                              double atol, int32_t steps, const double* weights, double* hist,
                              int32_t* accepted, int32_t* rejected, int32_t* loop_steps,
                              double* cost);

/* mdx_cost_pilot_launch that also leaves each trajectory's step state in
 * `resume`, with the trajectory index fastest: for Rodas5P (n_vars + 8) x
 * n_traj doubles, the state vector, then t, dt, the controller's two previous
 * error norms, the next save index and the step, accepted and rejected counts;
 * for Tsit5 (2 n_vars + 9) x n_traj, the state vector, its first-same-as-last
 * stage, the same eight scalars and 1.0 when that stage holds (else 0.0). The
 * saves it passed are already in `hist`. For kernels created with
 * MDX_KERNEL_RESUMABLE (the one-lane Rodas5P and Tsit5 kernels). No
 * counterpart in the reference. */
int32_t mdx_cost_pilot_launch_resumable(int64_t handle, void* stream, int64_t n_traj,
                                        int32_t n_vars, int32_t n_params, int32_t n_save,
                                        int32_t layout, const double* y0, const double* times,
                                        const double* params, double dt0, double rtol,
                                        double atol, int32_t steps, const double* weights,
                                        double* hist, int32_t* accepted, int32_t* rejected,
                                        int32_t* loop_steps, double* cost, double* resume);

/* mdx_solve_launch_ordered (order may be NULL) whose trajectories start from
 * the step state mdx_cost_pilot_launch_resumable left in `resume`, into the
 * same `hist`, so the pilot's steps are not taken again; the results are the
 * bits of a solve from y0. For kernels created with MDX_KERNEL_RESUMABLE
 * (one-lane or refill Rodas5P or Tsit5, no save hook). No counterpart in the
 * reference. */
int32_t mdx_solve_launch_resumed(int64_t handle, void* stream, int32_t grid, int64_t n_traj,
                                 int32_t n_vars, int32_t n_params, int32_t n_save, int32_t layout,
                                 const double* y0, const double* times, const double* params,
                                 double dt0, double rtol, double atol, int32_t max_steps,
                                 const double* weights, double* hist, int32_t* accepted,
                                 int32_t* rejected, int32_t* loop_steps, void* workspace,
                                 int64_t workspace_bytes, const int64_t* order,
                                 const double* resume);

/* Device memory mdx_adjoint_launch needs as `workspace`: per trajectory,
 * max_ckpt checkpoints of n_vars + 8 doubles and ckpt_every step records of
 * n_vars + 4 doubles. */
int64_t mdx_adjoint_workspace_bytes(int64_t n_traj, int32_t n_vars, int32_t ckpt_every,
                                    int32_t max_ckpt);

/* The discrete adjoint of one ensemble solve, Rodas5P or Tsit5 (the handle's
 * kernel decides): given gbar = dL/dhist (N, n_save, n_vars) for the solve
 * that mdx_solve_launch runs with the same inputs, write dL/dy0 (N, n_vars)
 * and dL/dparams (N, n_params). The kernel
 * re-runs the solve storing a checkpoint before every ckpt_every-th accepted
 * step (at most max_ckpt per trajectory), then goes back through the accepted
 * steps segment by segment, recomputing each segment from its checkpoint; the
 * step sizes are held fixed. status[i] is 0, or 1 when trajectory i needed
 * more than max_ckpt checkpoints (its gradients are then NaN). `layout` as for
 * mdx_solve_launch (gbar like hist, the gradients like y0 and params).
 * Replaces the reverse-mode use of the reference's forward-sensitivity JVP rule
 * (modax/_sensitivity.py:331-413, transposed by JAX): the same gradients for
 * O(1) memory per step instead of n_vars + n_params sensitivity columns. */
int32_t mdx_adjoint_launch(int64_t handle, void* stream, int64_t n_traj, int32_t n_vars,
                           int32_t n_params, int32_t n_save, int32_t layout, const double* y0,
                           const double* times, const double* params, double dt0, double rtol,
                           double atol, int32_t max_steps, const double* weights,
                           const double* gbar, double* grad_y0, double* grad_p, int32_t* status,
                           int32_t ckpt_every, int32_t max_ckpt, void* workspace,
                           int64_t workspace_bytes);

/* Bytes of device scratch behind mdx_adjoint_launch_static (per device). */
#define MDX_STATIC_SCRATCH_BYTES (64ll << 20)

/* mdx_adjoint_launch without a workspace: the checkpoints and step records go
 * to a static device array of MDX_STATIC_SCRATCH_BYTES that the library holds
 * on each device it runs on until the process ends. No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that forbid allocation). A call whose
 * mdx_adjoint_workspace_bytes(n_traj, n_vars, ckpt_every, max_ckpt) exceeds
 * MDX_STATIC_SCRATCH_BYTES is rejected before anything runs. It calls no
 * allocator: its first call on a device resolves the array's address there,
 * and (as any launch) loads the kernel's code object if no earlier call on
 * that device has, so a caller that must not allocate at all makes that first
 * call (or mdx_kernel_info) before it starts. Every call on a device shares
 * that device's one array, which keeps the last call's contents (the kernel
 * writes each checkpoint before it reads it, so nothing needs clearing):
 * calls that overlap on one device race, so issue them on one stream or wait
 * for one before the next. Arguments and results otherwise as
 * mdx_adjoint_launch, bit for bit. */
int32_t mdx_adjoint_launch_static(int64_t handle, void* stream, int64_t n_traj, int32_t n_vars,
                                  int32_t n_params, int32_t n_save, int32_t layout,
                                  const double* y0, const double* times, const double* params,
                                  double dt0, double rtol, double atol, int32_t max_steps,
                                  const double* weights, const double* gbar, double* grad_y0,
                                  double* grad_p, int32_t* status, int32_t ckpt_every,
                                  int32_t max_ckpt);

/* Device facts used for code generation and the roofline: architecture name
 * (e.g. "gfx950") and compute-unit count of the current device. */
int32_t mdx_device_info(char* arch, size_t arch_cap, int32_t* n_cu);

#ifdef __cplusplus
}
#endif
#endif /* MENODAX_HIP_H */
