// runtime.cpp -- host side of libmenodax_hip.so: hipRTC compilation, a
// registry of loaded kernels, and the ensemble launch (ABI in
// include/menodax_hip.h).
//
// The launch mirrors the reference's LaunchNumbaCudaAbi
// (modax/_jax_numba_custom_call.py:145-213): it validates what it is given,
// marshals the buffers into the kernel's argument block and launches on the
// caller's stream without synchronising. It differs where the hardware does:
// the argument block is one struct of pointers and element strides rather
// than MLIR memref descriptors, a code object is loaded lazily per device, and
// a persistent refill kernel gets its work counter reset by a one-lane kernel
// on the same stream first (a kernel, not a memset, so the sequence can be
// captured into a HIP graph).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "menodax_hip.h"

// Device scratch of mdx_adjoint_launch_static, the entry point that allocates
// nothing: the adjoint's checkpoints and step records live here instead of in a
// caller's workspace. Each device that runs the library holds its own copy
// until the process ends. In host code `scratch` names a host-side stand-in,
// never the device array: kernels get the address hipGetSymbolAddress resolves
// for the current device (static_scratch below), and only that address.
static __device__ __align__(256) unsigned char scratch[MDX_STATIC_SCRATCH_BYTES];

namespace {

// Must match mdx::SolveArgs in menodax_amd/kernels/mdx_common.hpp member for
// member: both sides are 64-bit with natural alignment.
struct SolveArgs {
  const double* y0;
  const double* times;
  const double* params;
  const double* weights;
  double* hist;
  int32_t* accepted;
  int32_t* rejected;
  int32_t* loop;
  double* hook_out;
  unsigned long long* queue;
  const int64_t* order;
  double* cost;
  double* resume;
  double dt0, rtol, atol;
  int64_t n_traj;
  int64_t y0_ts, y0_vs;
  int64_t p_ts, p_vs;
  int64_t w_ts, w_vs;
  int64_t h_ts, h_ss, h_vs;
  int64_t ho_ts, ho_vs;
  int32_t max_steps;
  int32_t n_save;
};
static_assert(sizeof(SolveArgs) == 13 * 8 + 3 * 8 + 8 + 11 * 8 + 8, "SolveArgs layout");

// Mirrors mdx::AdjointArgs in kernels/mdx_common.hpp (both adjoint kernels).
struct AdjointArgs {
  SolveArgs s;
  const double* gbar;
  double* grad_y0;
  double* grad_p;
  int32_t* status;
  double* ws;
  int64_t g_ts, g_ss, g_vs;
  int64_t gy_ts, gy_vs, gp_ts, gp_vs;
  int32_t ckpt_every, max_ckpt;
};
static_assert(sizeof(AdjointArgs) == sizeof(SolveArgs) + 5 * 8 + 7 * 8 + 8, "AdjointArgs layout");

thread_local std::string g_error;

int32_t fail(int32_t code, const std::string& msg) {
  g_error = msg;
  return code;
}

int32_t hip_fail(hipError_t e, const char* what) {
  g_error = std::string(what) + ": " + hipGetErrorString(e) + " (" + std::to_string((int)e) + ")";
  return (int32_t)e;
}

struct Loaded {
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;
  hipFunction_t reset = nullptr;
  int32_t persistent_grid = 0;  // cached resident grid on this device
};

struct Kernel {
  std::vector<char> code;
  std::string entry;
  int32_t block = 64;
  bool persistent = false;
  bool unit_weights = false;
  bool adjoint = false;
  bool hook = false;
  bool wave = false;  // one trajectory per workgroup
  bool resumable = false;  // reads and writes a step-state buffer (SolveArgs::resume)
  std::map<int, Loaded> per_device;
};

std::mutex g_mutex;
std::vector<std::unique_ptr<Kernel>> g_kernels;  // handle = index + 1; never freed

std::map<int, void*> g_scratch;  // device -> resolved address of `scratch`

int32_t lookup(int64_t handle, Kernel** out) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (handle < 1 || handle > (int64_t)g_kernels.size())
    return fail(1, "unknown kernel handle " + std::to_string(handle));
  *out = g_kernels[(size_t)(handle - 1)].get();
  return 0;
}

// The kernel's module on the current device, loaded on first use.
int32_t loaded_on_current_device(Kernel* k, Loaded** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_mutex);
  auto it = k->per_device.find(dev);
  if (it != k->per_device.end()) {
    *out = &it->second;
    return 0;
  }
  Loaded l;
  e = hipModuleLoadData(&l.module, k->code.data());
  if (e != hipSuccess) return hip_fail(e, "hipModuleLoadData");
  e = hipModuleGetFunction(&l.fn, l.module, k->entry.c_str());
  if (e != hipSuccess) return hip_fail(e, ("hipModuleGetFunction " + k->entry).c_str());
  if (k->persistent) {
    e = hipModuleGetFunction(&l.reset, l.module, "mdx_reset_queue");
    if (e != hipSuccess) return hip_fail(e, "hipModuleGetFunction mdx_reset_queue");
  }
  auto res = k->per_device.emplace(dev, l);
  *out = &res.first->second;
  return 0;
}

// The address of `scratch` on the current device, resolved on the first call
// there (with that device current) and kept.
int32_t static_scratch(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_mutex);
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) {
    *out = it->second;
    return 0;
  }
  void* ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!ptr) return fail(1, "hipGetSymbolAddress(scratch) returned a null address");
  g_scratch.emplace(dev, ptr);
  *out = ptr;
  return 0;
}

int32_t adjoint_impl(int64_t handle, void* stream, int64_t n_traj, int32_t n_vars,
                     int32_t n_params, int32_t n_save, int32_t layout, const double* y0,
                     const double* times, const double* params, double dt0, double rtol,
                     double atol, int32_t max_steps, const double* weights, const double* gbar,
                     double* grad_y0, double* grad_p, int32_t* status, int32_t ckpt_every,
                     int32_t max_ckpt, void* workspace, int64_t workspace_bytes,
                     bool use_static, const char* who);

}  // namespace

extern "C" {

int32_t mdx_abi_version(void) { return 1; }

const char* mdx_last_error(void) { return g_error.c_str(); }

void mdx_free(void* p) { std::free(p); }

int32_t mdx_rtc_compile(const char* source, const char* name, int32_t n_headers,
                        const char* const* headers, const char* const* header_names,
                        int32_t n_options, const char* const* options, void** code,
                        size_t* code_size, char** log) {
  if (!source || !code || !code_size || !log) return fail(1, "mdx_rtc_compile: null argument");
  *code = nullptr;
  *code_size = 0;
  *log = nullptr;
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, source, name ? name : "mdx_kernel.hip", n_headers,
                                       headers, header_names);
  if (r != HIPRTC_SUCCESS)
    return fail(2, std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));
  r = hiprtcCompileProgram(prog, n_options, options);
  size_t log_size = 0;
  hiprtcGetProgramLogSize(prog, &log_size);
  char* text = (char*)std::malloc(log_size + 1);
  if (text) {
    if (log_size) hiprtcGetProgramLog(prog, text);
    text[log_size] = '\0';
    *log = text;
  }
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail(3, std::string("hiprtcCompileProgram: ") + hiprtcGetErrorString(r));
  }
  size_t size = 0;
  r = hiprtcGetCodeSize(prog, &size);
  if (r != HIPRTC_SUCCESS || size == 0) {
    hiprtcDestroyProgram(&prog);
    return fail(4, "hiprtcGetCodeSize failed");
  }
  void* buf = std::malloc(size);
  if (!buf) {
    hiprtcDestroyProgram(&prog);
    return fail(5, "out of host memory");
  }
  r = hiprtcGetCode(prog, (char*)buf);
  hiprtcDestroyProgram(&prog);
  if (r != HIPRTC_SUCCESS) {
    std::free(buf);
    return fail(4, std::string("hiprtcGetCode: ") + hiprtcGetErrorString(r));
  }
  *code = buf;
  *code_size = size;
  return 0;
}

int32_t mdx_kernel_create(const void* code, size_t code_size, const char* entry, int32_t block,
                          int32_t flags, int64_t* handle) {
  if (!code || code_size == 0 || !entry || !handle) return fail(1, "mdx_kernel_create: null argument");
  if (block < 1 || block > 1024) return fail(1, "mdx_kernel_create: block must be in [1, 1024]");
  auto k = std::make_unique<Kernel>();
  k->code.assign((const char*)code, (const char*)code + code_size);
  k->entry = entry;
  k->block = block;
  k->persistent = (flags & MDX_KERNEL_PERSISTENT) != 0;
  k->unit_weights = (flags & MDX_KERNEL_UNIT_WEIGHTS) != 0;
  k->adjoint = (flags & MDX_KERNEL_ADJOINT) != 0;
  k->hook = (flags & MDX_KERNEL_HOOK) != 0;
  k->wave = (flags & MDX_KERNEL_WAVE) != 0;
  k->resumable = (flags & MDX_KERNEL_RESUMABLE) != 0;
  if (k->adjoint && k->persistent)
    return fail(1, "mdx_kernel_create: an adjoint kernel is not a persistent kernel");
  if (k->wave && (k->persistent || k->adjoint))
    return fail(1, "mdx_kernel_create: a wave-per-trajectory kernel is a plain solve kernel");
  std::lock_guard<std::mutex> lock(g_mutex);
  g_kernels.push_back(std::move(k));
  *handle = (int64_t)g_kernels.size();
  return 0;
}

int32_t mdx_kernel_info(int64_t handle, int32_t* vgprs, int32_t* scratch_bytes, int32_t* lds_bytes,
                        int32_t* blocks_per_cu) {
  Kernel* k = nullptr;
  if (int32_t rc = lookup(handle, &k)) return rc;
  Loaded* l = nullptr;
  if (int32_t rc = loaded_on_current_device(k, &l)) return rc;
  int v = 0;
  hipError_t e;
  if (vgprs) {
    e = hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, l->fn);
    if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttribute(NUM_REGS)");
    *vgprs = v;
  }
  if (scratch_bytes) {
    e = hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, l->fn);
    if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttribute(LOCAL_SIZE_BYTES)");
    *scratch_bytes = v;
  }
  if (lds_bytes) {
    e = hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, l->fn);
    if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttribute(SHARED_SIZE_BYTES)");
    *lds_bytes = v;
  }
  if (blocks_per_cu) {
    int nb = 0;
    e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, l->fn, k->block, 0);
    if (e != hipSuccess) return hip_fail(e, "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
    *blocks_per_cu = nb;
  }
  return 0;
}

int32_t mdx_persistent_grid(int64_t handle, int32_t* grid) {
  if (!grid) return fail(1, "mdx_persistent_grid: null argument");
  int32_t per_cu = 0;
  if (int32_t rc = mdx_kernel_info(handle, nullptr, nullptr, nullptr, &per_cu)) return rc;
  int dev = 0, n_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute(MultiprocessorCount)");
  *grid = (per_cu < 1 ? 1 : per_cu) * n_cu;
  return 0;
}

int64_t mdx_workspace_bytes(int64_t handle, int64_t n_traj) {
  (void)n_traj;
  Kernel* k = nullptr;
  if (lookup(handle, &k)) return -1;
  return k->persistent ? 256 : 0;
}

int32_t mdx_solve_launch(int64_t handle, void* stream, int32_t grid, int64_t n_traj,
                         int32_t n_vars, int32_t n_params, int32_t n_save, int32_t layout,
                         const double* y0, const double* times, const double* params, double dt0,
                         double rtol, double atol, int32_t max_steps, const double* weights,
                         double* hist, int32_t* accepted, int32_t* rejected,
                         int32_t* loop_steps, double* hook_out, int32_t hook_size,
                         void* workspace, int64_t workspace_bytes) {
  return mdx_solve_launch_ordered(handle, stream, grid, n_traj, n_vars, n_params, n_save, layout,
                                  y0, times, params, dt0, rtol, atol, max_steps, weights, hist,
                                  accepted, rejected, loop_steps, hook_out, hook_size, workspace,
                                  workspace_bytes, nullptr);
}

namespace {
int32_t launch_impl(int64_t handle, void* stream, int32_t grid, int64_t n_traj, int32_t n_vars,
                    int32_t n_params, int32_t n_save, int32_t layout, const double* y0,
                    const double* times, const double* params, double dt0, double rtol,
                    double atol, int32_t max_steps, const double* weights, double* hist,
                    int32_t* accepted, int32_t* rejected, int32_t* loop_steps, double* hook_out,
                    int32_t hook_size, void* workspace, int64_t workspace_bytes,
                    const int64_t* order, double* cost, double* resume);
}  // namespace

int32_t mdx_solve_launch_ordered(int64_t handle, void* stream, int32_t grid, int64_t n_traj,
                                 int32_t n_vars, int32_t n_params, int32_t n_save, int32_t layout,
                                 const double* y0, const double* times, const double* params,
                                 double dt0, double rtol, double atol, int32_t max_steps,
                                 const double* weights, double* hist, int32_t* accepted,
                                 int32_t* rejected, int32_t* loop_steps, double* hook_out,
                                 int32_t hook_size, void* workspace, int64_t workspace_bytes,
                                 const int64_t* order) {
  return launch_impl(handle, stream, grid, n_traj, n_vars, n_params, n_save, layout, y0, times,
                     params, dt0, rtol, atol, max_steps, weights, hist, accepted, rejected,
                     loop_steps, hook_out, hook_size, workspace, workspace_bytes, order, nullptr,
                     nullptr);
}

int32_t mdx_solve_launch_resumed(int64_t handle, void* stream, int32_t grid, int64_t n_traj,
                                 int32_t n_vars, int32_t n_params, int32_t n_save, int32_t layout,
                                 const double* y0, const double* times, const double* params,
                                 double dt0, double rtol, double atol, int32_t max_steps,
                                 const double* weights, double* hist, int32_t* accepted,
                                 int32_t* rejected, int32_t* loop_steps, void* workspace,
                                 int64_t workspace_bytes, const int64_t* order,
                                 const double* resume) {
  if (!resume && n_traj > 0) return fail(1, "mdx_solve_launch_resumed: null step-state buffer");
  return launch_impl(handle, stream, grid, n_traj, n_vars, n_params, n_save, layout, y0, times,
                     params, dt0, rtol, atol, max_steps, weights, hist, accepted, rejected,
                     loop_steps, nullptr, 0, workspace, workspace_bytes, order, nullptr,
                     const_cast<double*>(resume));
}

int32_t mdx_cost_pilot_launch(int64_t handle, void* stream, int64_t n_traj, int32_t n_vars,
                              int32_t n_params, int32_t n_save, int32_t layout, const double* y0,
                              const double* times, const double* params, double dt0, double rtol,
                              double atol, int32_t steps, const double* weights, double* hist,
                              int32_t* accepted, int32_t* rejected, int32_t* loop_steps,
                              double* cost) {
  if (!cost && n_traj > 0) return fail(1, "mdx_cost_pilot_launch: null cost buffer");
  if (steps < 1) return fail(1, "mdx_cost_pilot_launch: steps must be positive");
  return launch_impl(handle, stream, 0, n_traj, n_vars, n_params, n_save, layout, y0, times,
                     params, dt0, rtol, atol, steps, weights, hist, accepted, rejected,
                     loop_steps, nullptr, 0, nullptr, 0, nullptr, cost, nullptr);
}

int32_t mdx_cost_pilot_launch_resumable(int64_t handle, void* stream, int64_t n_traj,
                                        int32_t n_vars, int32_t n_params, int32_t n_save,
                                        int32_t layout, const double* y0, const double* times,
                                        const double* params, double dt0, double rtol,
                                        double atol, int32_t steps, const double* weights,
                                        double* hist, int32_t* accepted, int32_t* rejected,
                                        int32_t* loop_steps, double* cost, double* resume) {
  if (!cost && n_traj > 0) return fail(1, "mdx_cost_pilot_launch_resumable: null cost buffer");
  if (!resume && n_traj > 0)
    return fail(1, "mdx_cost_pilot_launch_resumable: null step-state buffer");
  if (steps < 1) return fail(1, "mdx_cost_pilot_launch_resumable: steps must be positive");
  return launch_impl(handle, stream, 0, n_traj, n_vars, n_params, n_save, layout, y0, times,
                     params, dt0, rtol, atol, steps, weights, hist, accepted, rejected,
                     loop_steps, nullptr, 0, nullptr, 0, nullptr, cost, resume);
}

namespace {
int32_t launch_impl(int64_t handle, void* stream, int32_t grid, int64_t n_traj, int32_t n_vars,
                    int32_t n_params, int32_t n_save, int32_t layout, const double* y0,
                    const double* times, const double* params, double dt0, double rtol,
                    double atol, int32_t max_steps, const double* weights, double* hist,
                    int32_t* accepted, int32_t* rejected, int32_t* loop_steps, double* hook_out,
                    int32_t hook_size, void* workspace, int64_t workspace_bytes,
                    const int64_t* order, double* cost, double* resume) {
  Kernel* k = nullptr;
  if (int32_t rc = lookup(handle, &k)) return rc;
  if (n_traj < 0 || n_vars < 1 || n_params < 0 || n_save < 1)
    return fail(1, "mdx_solve_launch: bad extents (N=" + std::to_string(n_traj) +
                       ", n_vars=" + std::to_string(n_vars) + ", n_params=" +
                       std::to_string(n_params) + ", n_save=" + std::to_string(n_save) + ")");
  if (layout != 0 && layout != 1) return fail(1, "mdx_solve_launch: layout must be 0 or 1");
  // an empty ensemble reads and writes nothing: its buffers may be null (the
  // data pointer of a zero-size allocation)
  if (n_traj > 0 && (!y0 || !times || (!params && n_params > 0) || !hist || !accepted ||
                     !rejected || !loop_steps))
    return fail(1, "mdx_solve_launch: null buffer");
  if (n_traj > 0 && !weights && !k->unit_weights)
    return fail(1, "mdx_solve_launch: this kernel reads error weights; weights is NULL");
  if (n_traj > 0 && hook_size > 0 && !hook_out)
    return fail(1, "mdx_solve_launch: hook_size > 0 needs hook_out");
  if (k->hook && (hook_size <= 0 || (n_traj > 0 && !hook_out)))
    return fail(1, "mdx_solve_launch: this kernel runs a save hook; hook_out and hook_size are required");
  if (max_steps < 0) return fail(1, "mdx_solve_launch: max_steps must be non-negative");
  if (k->adjoint) return fail(1, "mdx_solve_launch: an adjoint kernel is launched by mdx_adjoint_launch");
  const int64_t per_group = k->wave ? 1 : k->block;  // trajectories per workgroup
  if (grid > 0 && !k->persistent && (int64_t)grid * per_group < n_traj)
    return fail(1, "mdx_solve_launch: a one-lane-per-trajectory launch needs grid >= "
                   "ceil(n_traj / block) (n_traj for a wave-per-trajectory kernel)");
  if (order && !k->persistent)
    return fail(1, "mdx_solve_launch_ordered: a claim order needs a persistent (refill) kernel");
  if (cost && (k->persistent || k->hook))
    return fail(1, "mdx_cost_pilot_launch: a cost pilot runs a one-lane-per-trajectory kernel "
                   "without a save hook");
  if (resume && !k->resumable)
    return fail(1, "mdx_solve_launch: a step-state buffer is for kernels created with "
                   "MDX_KERNEL_RESUMABLE (the one-lane and refill Rodas5P and Tsit5 kernels)");
  if (k->persistent && n_traj > 0 && (!workspace || workspace_bytes < 8))
    return fail(1, "mdx_solve_launch: a persistent kernel needs a workspace of at least 8 bytes");
  if (n_traj == 0) return 0;

  SolveArgs a;
  std::memset(&a, 0, sizeof(a));
  a.y0 = y0;
  a.times = times;
  a.params = params;
  a.weights = weights;
  a.hist = hist;
  a.accepted = accepted;
  a.rejected = rejected;
  a.loop = loop_steps;
  a.hook_out = hook_out;
  a.queue = (unsigned long long*)workspace;
  a.order = order;
  a.cost = cost;
  a.resume = resume;
  a.dt0 = dt0;
  a.rtol = rtol;
  a.atol = atol;
  a.n_traj = n_traj;
  a.max_steps = max_steps;
  a.n_save = n_save;
  if (layout == 0) {  // trajectory-major, contiguous
    a.y0_ts = n_vars, a.y0_vs = 1;
    a.p_ts = n_params, a.p_vs = 1;
    a.w_ts = n_vars, a.w_vs = 1;
    a.h_ts = (int64_t)n_save * n_vars, a.h_ss = n_vars, a.h_vs = 1;
    a.ho_ts = hook_size, a.ho_vs = 1;
  } else {  // structure of arrays: trajectory index fastest
    a.y0_ts = 1, a.y0_vs = n_traj;
    a.p_ts = 1, a.p_vs = n_traj;
    a.w_ts = 1, a.w_vs = n_traj;
    a.h_ts = 1, a.h_ss = (int64_t)n_vars * n_traj, a.h_vs = n_traj;
    a.ho_ts = 1, a.ho_vs = n_traj;
  }

  Loaded* l = nullptr;
  if (int32_t rc = loaded_on_current_device(k, &l)) return rc;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (k->persistent) {
    void* qargs[] = {&a.queue};
    e = hipModuleLaunchKernel(l->reset, 1, 1, 1, 1, 1, 1, 0, s, qargs, nullptr);
    if (e != hipSuccess) return hip_fail(e, "hipModuleLaunchKernel(mdx_reset_queue)");
  }
  int64_t blocks = grid;
  if (blocks <= 0) {
    if (k->persistent) {
      if (l->persistent_grid <= 0) {
        int32_t g = 0;
        if (int32_t rc = mdx_persistent_grid(handle, &g)) return rc;
        l->persistent_grid = g;
      }
      blocks = l->persistent_grid;
      // no more resident workgroups than there are trajectories to start
      int64_t needed = (n_traj + k->block - 1) / k->block;
      if (blocks > needed) blocks = needed;
    } else {
      blocks = (n_traj + per_group - 1) / per_group;
    }
  }
  if (blocks > 0x7fffffff) return fail(1, "mdx_solve_launch: grid too large");
  void* kargs[] = {&a};
  e = hipModuleLaunchKernel(l->fn, (unsigned)blocks, 1, 1, (unsigned)k->block, 1, 1, 0, s, kargs,
                            nullptr);
  if (e != hipSuccess) return hip_fail(e, ("hipModuleLaunchKernel(" + k->entry + ")").c_str());
  return 0;
}
}  // namespace

int64_t mdx_adjoint_workspace_bytes(int64_t n_traj, int32_t n_vars, int32_t ckpt_every,
                                    int32_t max_ckpt) {
  if (n_traj < 0 || n_vars < 1 || ckpt_every < 1 || max_ckpt < 1) return -1;
  return n_traj * 8 * ((int64_t)max_ckpt * (n_vars + 8) + (int64_t)ckpt_every * (n_vars + 4));
}

int32_t mdx_adjoint_launch(int64_t handle, void* stream, int64_t n_traj, int32_t n_vars,
                           int32_t n_params, int32_t n_save, int32_t layout, const double* y0,
                           const double* times, const double* params, double dt0, double rtol,
                           double atol, int32_t max_steps, const double* weights,
                           const double* gbar, double* grad_y0, double* grad_p, int32_t* status,
                           int32_t ckpt_every, int32_t max_ckpt, void* workspace,
                           int64_t workspace_bytes) {
  return adjoint_impl(handle, stream, n_traj, n_vars, n_params, n_save, layout, y0, times, params,
                      dt0, rtol, atol, max_steps, weights, gbar, grad_y0, grad_p, status,
                      ckpt_every, max_ckpt, workspace, workspace_bytes, false,
                      "mdx_adjoint_launch");
}

int32_t mdx_adjoint_launch_static(int64_t handle, void* stream, int64_t n_traj, int32_t n_vars,
                                  int32_t n_params, int32_t n_save, int32_t layout,
                                  const double* y0, const double* times, const double* params,
                                  double dt0, double rtol, double atol, int32_t max_steps,
                                  const double* weights, const double* gbar, double* grad_y0,
                                  double* grad_p, int32_t* status, int32_t ckpt_every,
                                  int32_t max_ckpt) {
  return adjoint_impl(handle, stream, n_traj, n_vars, n_params, n_save, layout, y0, times, params,
                      dt0, rtol, atol, max_steps, weights, gbar, grad_y0, grad_p, status,
                      ckpt_every, max_ckpt, nullptr, 0, true, "mdx_adjoint_launch_static");
}

}  // extern "C"

namespace {
int32_t adjoint_impl(int64_t handle, void* stream, int64_t n_traj, int32_t n_vars,
                     int32_t n_params, int32_t n_save, int32_t layout, const double* y0,
                     const double* times, const double* params, double dt0, double rtol,
                     double atol, int32_t max_steps, const double* weights, const double* gbar,
                     double* grad_y0, double* grad_p, int32_t* status, int32_t ckpt_every,
                     int32_t max_ckpt, void* workspace, int64_t workspace_bytes,
                     bool use_static, const char* who) {
  const std::string w(who);
  Kernel* k = nullptr;
  if (int32_t rc = lookup(handle, &k)) return rc;
  if (!k->adjoint) return fail(1, w + ": not an adjoint kernel");
  if (n_traj < 0 || n_vars < 1 || n_params < 0 || n_save < 1)
    return fail(1, w + ": bad extents");
  if (layout != 0 && layout != 1) return fail(1, w + ": layout must be 0 or 1");
  if (ckpt_every < 1 || max_ckpt < 1)
    return fail(1, w + ": ckpt_every and max_ckpt must be positive");
  if (n_traj > 0 && (!y0 || !times || (!params && n_params > 0) || !gbar || !grad_y0 ||
                     (!grad_p && n_params > 0) || !status || (!use_static && !workspace)))
    return fail(1, w + ": null buffer");
  if (n_traj > 0 && !weights && !k->unit_weights)
    return fail(1, w + ": this kernel reads error weights; weights is NULL");
  if (max_steps < 0) return fail(1, w + ": max_steps must be non-negative");
  const int64_t need = mdx_adjoint_workspace_bytes(n_traj, n_vars, ckpt_every, max_ckpt);
  if (use_static) {
    if (need > (int64_t)MDX_STATIC_SCRATCH_BYTES)
      return fail(1, w + ": the checkpoints need " + std::to_string(need) +
                         " bytes of scratch, more than the static scratch's " +
                         std::to_string((long long)MDX_STATIC_SCRATCH_BYTES) +
                         " (fewer trajectories per call, a larger ckpt_every, or "
                         "mdx_adjoint_launch with a workspace)");
  } else if (workspace_bytes < need) {
    return fail(1, w + ": workspace holds " + std::to_string(workspace_bytes) +
                       " bytes, needs " + std::to_string(need));
  }
  if (n_traj == 0) return 0;
  if (use_static) {
    if (int32_t rc = static_scratch(&workspace)) return rc;
  }

  AdjointArgs g;
  std::memset(&g, 0, sizeof(g));
  SolveArgs& a = g.s;
  a.y0 = y0;
  a.times = times;
  a.params = params;
  a.weights = weights;
  a.dt0 = dt0;
  a.rtol = rtol;
  a.atol = atol;
  a.n_traj = n_traj;
  a.max_steps = max_steps;
  a.n_save = n_save;
  g.gbar = gbar;
  g.grad_y0 = grad_y0;
  g.grad_p = grad_p;
  g.status = status;
  g.ws = (double*)workspace;
  g.ckpt_every = ckpt_every;
  g.max_ckpt = max_ckpt;
  if (layout == 0) {
    a.y0_ts = n_vars, a.y0_vs = 1;
    a.p_ts = n_params, a.p_vs = 1;
    a.w_ts = n_vars, a.w_vs = 1;
    g.g_ts = (int64_t)n_save * n_vars, g.g_ss = n_vars, g.g_vs = 1;
    g.gy_ts = n_vars, g.gy_vs = 1;
    g.gp_ts = n_params, g.gp_vs = 1;
  } else {
    a.y0_ts = 1, a.y0_vs = n_traj;
    a.p_ts = 1, a.p_vs = n_traj;
    a.w_ts = 1, a.w_vs = n_traj;
    g.g_ts = 1, g.g_ss = (int64_t)n_vars * n_traj, g.g_vs = n_traj;
    g.gy_ts = 1, g.gy_vs = n_traj;
    g.gp_ts = 1, g.gp_vs = n_traj;
  }
  Loaded* l = nullptr;
  if (int32_t rc = loaded_on_current_device(k, &l)) return rc;
  const int64_t blocks = (n_traj + k->block - 1) / k->block;
  if (blocks > 0x7fffffff) return fail(1, w + ": grid too large");
  void* kargs[] = {&g};
  hipError_t e = hipModuleLaunchKernel(l->fn, (unsigned)blocks, 1, 1, (unsigned)k->block, 1, 1, 0,
                                       (hipStream_t)stream, kargs, nullptr);
  if (e != hipSuccess) return hip_fail(e, ("hipModuleLaunchKernel(" + k->entry + ")").c_str());
  return 0;
}
}  // namespace

extern "C" {

int32_t mdx_device_info(char* arch, size_t arch_cap, int32_t* n_cu) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  if (arch && arch_cap) {
    std::snprintf(arch, arch_cap, "%s", prop.gcnArchName);
  }
  if (n_cu) *n_cu = prop.multiProcessorCount;
  return 0;
}

}  // extern "C"
