// mdx_common.hpp -- shared pieces of the ensemble kernels (gfx950, wave64).
//
// Compiled by hipRTC together with the code generated for one right-hand side.
// Everything here runs one trajectory per lane: a lane owns its state outright,
// nothing is shared inside a step and nothing synchronises, so a wave's lanes
// diverge freely (modax/rodas5P.py:306-324 describes the same ownership model).
//
// Rounding: every fused multiply-add is an explicit fma() and the file is
// compiled with -ffp-contract=off, so no other operation is fused. The CPU
// oracle places the same fma()s, which makes the two agree bit for bit.
#pragma once

namespace mdx {

// Element type of an LU buffer: a pointer, or an accessor that names it.
template <class L>
struct lu_elem {
  typedef typename L::value_type type;
};
template <class T>
struct lu_elem<T*> {
  typedef T type;
};


typedef long long i64;
typedef unsigned long long u64;

// Kernel arguments, one struct passed by value. Arrays are addressed through
// element strides so the same kernel reads the reference API's
// trajectory-major layout or the structure-of-arrays layout the host prefers
// (stride 1 over trajectories, which makes every lane-parallel access
// coalesced).
struct SolveArgs {
  const double* y0;       // (traj, var)
  const double* times;    // (save,)
  const double* params;   // (traj, param)
  const double* weights;  // (traj, var)
  double* hist;           // (traj, save, var)
  int* accepted;          // (traj,)
  int* rejected;
  int* loop;
  double* hook_out;       // (traj, hook_size): save-hook accumulators
  u64* queue;             // persistent refill: next unclaimed trajectory
  const i64* order;       // persistent refill: claim order (a permutation), or null
  double* cost;           // cost pilot: work left per trajectory, written instead of saves
  double* resume;         // step state per trajectory, (n_vars + 8, traj) for Rodas5P,
                          // (2 n_vars + 9, traj) for Tsit5: a cost pilot writes it,
                          // a solve given it starts each trajectory there
  double dt0, rtol, atol;
  i64 n_traj;
  i64 y0_ts, y0_vs;
  i64 p_ts, p_vs;
  i64 w_ts, w_vs;
  i64 h_ts, h_ss, h_vs;
  i64 ho_ts, ho_vs;       // hook_out strides
  int max_steps;
  int n_save;
};

// The discrete-adjoint launch (both methods): the forward solve's inputs, the
// cotangent of its history, the gradient outputs and the checkpoint workspace.
struct AdjointArgs {
  SolveArgs s;             // the forward solve's inputs; hist/counters unused
  const double* gbar;      // dL/dhist, strides g_ts, g_ss, g_vs
  double* grad_y0;         // (traj, var), strides gy_ts, gy_vs
  double* grad_p;          // (traj, param), strides gp_ts, gp_vs
  int* status;             // (traj,): 0 ok, 1 checkpoint space exhausted
  double* ws;              // workspace
  i64 g_ts, g_ss, g_vs;
  i64 gy_ts, gy_vs, gp_ts, gp_vs;
  int ckpt_every, max_ckpt;
};

// A trajectory's save-hook accumulator row in global memory.
struct HookRow {
  double* base;
  i64 stride;
  __device__ __forceinline__ double operator[](int k) const { return base[k * stride]; }
  __device__ __forceinline__ void put(int k, double v) const { base[k * stride] = v; }
};

__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double absT(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float absT(float a) { return __builtin_fabsf(a); }

// ---------------------------------------------------------------------------
// Deterministic power for the step-size controller: x**y for finite x > 0 as
// exp(y*log(x)); log from the atanh series of s = (m-1)/(m+1) on
// m in [sqrt(1/2), sqrt(2)), exp from a Cody-Waite reduction by ln 2 and a
// degree-13 Taylor polynomial. Only correctly rounded operations, so host and
// device produce the same bits. pow(x, +-0) is exactly 1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double det_log(double x) {
  int e;
  double m = __builtin_frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double q = 1.0 / 23.0;
  q = fmaT(q, s2, 1.0 / 21.0);
  q = fmaT(q, s2, 1.0 / 19.0);
  q = fmaT(q, s2, 1.0 / 17.0);
  q = fmaT(q, s2, 1.0 / 15.0);
  q = fmaT(q, s2, 1.0 / 13.0);
  q = fmaT(q, s2, 1.0 / 11.0);
  q = fmaT(q, s2, 1.0 / 9.0);
  q = fmaT(q, s2, 1.0 / 7.0);
  q = fmaT(q, s2, 1.0 / 5.0);
  q = fmaT(q, s2, 1.0 / 3.0);
  const double two_s = 2.0 * s;
  const double log_m = fmaT(two_s * s2, q, two_s);
  const double de = (double)e;
  return fmaT(de, 6.93147180369123816490e-01, fmaT(de, 1.90821492927058770002e-10, log_m));
}

__device__ __forceinline__ double det_exp(double z) {
  const double k = __builtin_rint(z * 1.44269504088896338700e+00);
  double r = fmaT(-k, 6.93147180369123816490e-01, z);
  r = fmaT(-k, 1.90821492927058770002e-10, r);
  double q = 1.0 / 6227020800.0;
  q = fmaT(q, r, 1.0 / 479001600.0);
  q = fmaT(q, r, 1.0 / 39916800.0);
  q = fmaT(q, r, 1.0 / 3628800.0);
  q = fmaT(q, r, 1.0 / 362880.0);
  q = fmaT(q, r, 1.0 / 40320.0);
  q = fmaT(q, r, 1.0 / 5040.0);
  q = fmaT(q, r, 1.0 / 720.0);
  q = fmaT(q, r, 1.0 / 120.0);
  q = fmaT(q, r, 1.0 / 24.0);
  q = fmaT(q, r, 1.0 / 6.0);
  q = fmaT(q, r, 0.5);
  q = fmaT(q, r, 1.0);
  q = fmaT(q, r, 1.0);
  return __builtin_ldexp(q, (int)k);
}

__device__ __forceinline__ double det_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  return det_exp(y * det_log(x));
}

// Error weight of component j. A kernel built for unit weights (the default
// when the caller passes none) never reads the array: 1.0 * x is x exactly, so
// the result is the same bits either way.
__device__ __forceinline__ double weight(const SolveArgs& a, i64 i, int j) {
#if MDX_UNIT_WEIGHTS
  (void)a;
  (void)i;
  (void)j;
  return 1.0;
#else
  return a.weights[i * a.w_ts + (i64)j * a.w_vs];
#endif
}

// Step-size factor shared by both methods (rodas5P.py:913-928, tsit5.py:383-398):
// the error is clamped into [1e-18, 1e18] (NaN counts as huge), the PID factor
// formed from the error history, the history advanced only on an accepted
// step, and the factor clamped.
__device__ __forceinline__ double pid_factor(double err_norm, bool accept, double& err_prev,
                                             double& err_prev2, double e1, double e2, double e3,
                                             double fmin, double fmax) {
  double safe;
  if (__builtin_isnan(err_norm) || err_norm > 1e18)
    safe = 1e18;
  else if (err_norm == 0.0)
    safe = 1e-18;
  else
    safe = err_norm;
  double factor = 0.9 * det_pow(safe, e1) * det_pow(err_prev, e2) * det_pow(err_prev2, e3);
  if (accept) {
    err_prev2 = err_prev;
    err_prev = safe;
  }
  if (factor < fmin)
    factor = fmin;
  else if (factor > fmax)
    factor = fmax;
  return factor;
}

// Parameters of one trajectory: a register copy when the row is short, read
// through the strided global pointer otherwise (the loads are coalesced and
// stay in L1/L2, and the registers go to the stage vectors instead).
template <int NP>
struct RegParams {
  double v[NP > 0 ? NP : 1];
  __device__ __forceinline__ double operator[](int j) const { return v[j]; }
  __device__ __forceinline__ void load(const SolveArgs& a, i64 i) {
#pragma unroll
    for (int j = 0; j < NP; ++j) v[j] = a.params[i * a.p_ts + (i64)j * a.p_vs];
  }
};

struct GlobalParams {
  const double* __restrict__ row;
  i64 stride;
  __device__ __forceinline__ double operator[](int j) const { return row[(i64)j * stride]; }
  __device__ __forceinline__ void load(const SolveArgs& a, i64 i) {
    row = a.params + i * a.p_ts;
    stride = a.p_vs;
  }
};

// ---------------------------------------------------------------------------
// Persistent refill: a lane whose trajectory has finished claims the next one
// from a device-wide counter. The claim is wave-aggregated -- one atomic per
// wave per round, the indices handed out by rank among the requesting lanes --
// so the counter sees a few thousand atomics rather than one per trajectory.
// Called in wave-uniform control flow.
// ---------------------------------------------------------------------------
__device__ __forceinline__ i64 wave_claim(u64* queue, bool want) {
  const u64 mask = __ballot(want);
  if (mask == 0) return -1;
  const int lane = __lane_id();
  const int leader = __builtin_ctzll(mask);
  u64 base = 0;
  if (lane == leader) base = atomicAdd(queue, (u64)__builtin_popcountll(mask));
  base = __shfl(base, leader);
  const u64 below = mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
  return want ? (i64)(base + (u64)__builtin_popcountll(below)) : -1;
}

}  // namespace mdx

// Zeroes the refill counter on the solve's own stream (a kernel rather than a
// memset, so the whole solve can be captured into a graph and replayed).
extern "C" __global__ void mdx_reset_queue(unsigned long long* q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) q[0] = 0ull;
}
