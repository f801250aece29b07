// tsit5_adjoint.hpp -- discrete adjoint of the Tsit5 ensemble solve, one
// trajectory per lane.
//
// Included after tsit5.hpp in a source generated with MDX_ADJOINT=1, whose sys
// namespace also holds f_vjp (the callback's reverse-mode contraction
// w^T [df/dy, df/dp], accumulated into ybar and pbar).
//
// Given dL/dhist for a loss of the saved history, each lane returns dL/dy0 and
// dL/dparams of its trajectory, differentiating the solver's own arithmetic:
// every accepted step with its step size held fixed (the controller is not
// differentiated), its seven stages and the continuous extension that
// produced each save. Phases as in the Rodas5P adjoint:
//   phase A  the forward solve again, storing the step state (state vector,
//            time, step size, controller memory, counters) before every
//            ckpt_every-th accepted step;
//   phase B  checkpoints last to first: replay the segment from its
//            checkpoint -- the same step code, so the same bits -- recording
//            each accepted step's (y_n, t_n, h_n, saves it produced), then run
//            back through the recorded steps.
// A checkpoint does not hold the first-same-as-last carry: the previous step's
// k7 was evaluated at exactly (y_n, t_n) (u7 is y_{n+1} and t + 1.0 h is
// t_{n+1}), so the replayed step recomputes the same k1 bits, and the reverse
// of each step accounts for its own k1.
//
// Reverse of one step (u_1 = y, u_s = y + h sum_{q<s} a_sq k_q,
// k_s = f(u_s, t + c_s h), y_{n+1} = u_7, a save = y + h sum_q b_q(theta) k_q):
//   ybar += g, kbar_q += h b_q(theta) g          (each save of the step)
//   ubar_7 = lam
//   s = 7..2:  ubar_s += J(u_s)^T kbar_s ; pbar += J_p(u_s)^T kbar_s
//              ybar += ubar_s ; kbar_q += h a_sq ubar_s  (q < s)
//   ybar += J(y)^T kbar_1 ; pbar += J_p(y)^T kbar_1 ; lam = ybar
// Workspace, per lane, [slot][lane] (coalesced): max_ckpt checkpoints of
// N + 8 doubles, then ckpt_every step records of N + 4 doubles (the layout the
// host sizes for either method).
#pragma once

namespace mdx {
namespace t5 {

#if MDX_ROLLED
#define MDX_ADJ_STAGES _Pragma("unroll 1")
#else
#define MDX_ADJ_STAGES _Pragma("unroll")
#endif

constexpr int CK = N + 8;   // checkpoint words
constexpr int RC = N + 4;   // step record words
constexpr int NPB = MDX_NP > 0 ? MDX_NP : 1;

__device__ __forceinline__ double& wsat(const AdjointArgs& g, i64 i, i64 slot) {
  return g.ws[slot * g.s.n_traj + i];
}

__device__ __forceinline__ void store_ckpt(const AdjointArgs& g, i64 i, int c, const State& s) {
  const i64 b = (i64)c * CK;
  for (int j = 0; j < N; ++j) wsat(g, i, b + j) = s.y[j];
  wsat(g, i, b + N + 0) = s.t;
  wsat(g, i, b + N + 1) = s.dt;
  wsat(g, i, b + N + 2) = s.err_prev;
  wsat(g, i, b + N + 3) = s.err_prev2;
  wsat(g, i, b + N + 4) = (double)s.save_idx;
  wsat(g, i, b + N + 5) = (double)s.n_steps;
  wsat(g, i, b + N + 6) = (double)s.accepted;
  wsat(g, i, b + N + 7) = (double)s.rejected;
}

__device__ __forceinline__ void load_ckpt(const AdjointArgs& g, i64 i, int c, State& s) {
  const i64 b = (i64)c * CK;
  for (int j = 0; j < N; ++j) s.y[j] = wsat(g, i, b + j);
  s.t = wsat(g, i, b + N + 0);
  s.dt = wsat(g, i, b + N + 1);
  s.err_prev = wsat(g, i, b + N + 2);
  s.err_prev2 = wsat(g, i, b + N + 3);
  s.save_idx = (int)wsat(g, i, b + N + 4);
  s.n_steps = (int)wsat(g, i, b + N + 5);
  s.accepted = (int)wsat(g, i, b + N + 6);
  s.rejected = (int)wsat(g, i, b + N + 7);
  s.has_fsal = false;  // k1 recomputed at (y_n, t_n): the carry's own bits
}

__device__ __forceinline__ double gbar_at(const AdjointArgs& g, i64 i, int sv, int j) {
  return g.gbar[i * g.g_ts + (i64)sv * g.g_ss + (i64)j * g.g_vs];
}

// The step size the forward step takes from state s (its first lines).
__device__ __forceinline__ double step_size(const State& s) {
  double dt_use = s.dt;
  if (dt_use > s.tf - s.t) dt_use = s.tf - s.t;
  if (dt_use < 1e-30) dt_use = 1e-30;
  return dt_use;
}

// u_S = y + h sum_{q<S} a_Sq k_q (0-based S = 1..6 for stages 2..7) in the
// forward step's order: the first term a product, one fma per further stage
// vector, then the fma with h and y.
__device__ __forceinline__ void stage_arg(int S, const double* __restrict__ y, double h,
                                          const double (*__restrict__ k)[N],
                                          double* __restrict__ u) {
  const double* __restrict__ ar = AT[S];
  for (int j = 0; j < N; ++j) {
    double q = ar[0] * k[0][j];
    for (int r = 1; r < S; ++r) q = fmaT(ar[r], k[r][j], q);
    u[j] = fmaT(h, q, y[j]);
  }
}

// The step's stage vectors at (y, t, h), by the forward step's arithmetic.
__device__ __forceinline__ void recompute(const double* __restrict__ y, double t, double h,
                                          PA& p, double (*__restrict__ k)[N]) {
  sys::rhs(y, t, p, k[0]);
  MDX_ADJ_STAGES
  for (int S = 1; S < 7; ++S) {
    double u[N];
    stage_arg(S, y, h, k, u);
    sys::rhs(u, t + CS[S] * h, p, k[S]);
  }
}

// lam: dL/dy_{n+1} in, dL/dy_n out; pbar accumulates dL/dp.
__device__ void reverse_step(const AdjointArgs& g, i64 i, PA& p, const double* __restrict__ y,
                             double t, double h, int save_lo, int save_hi,
                             double* __restrict__ lam, double* __restrict__ pbar) {
  const SolveArgs& a = g.s;
  double k[7][N];
  recompute(y, t, h, p, k);
  double kbar[7][N], ybar[N], ub[N];
  for (int q = 0; q < 7; ++q)
    for (int j = 0; j < N; ++j) kbar[q][j] = 0.0;
  for (int j = 0; j < N; ++j) {
    ybar[j] = 0.0;
    ub[j] = lam[j];  // y_{n+1} = u_7
  }
  // saves the step produced: y + h sum_q b_q(theta) k_q
  for (int sv = save_lo; sv < save_hi; ++sv) {
    double b[7];
    dense_weights((a.times[sv] - t) / h, b);
    for (int j = 0; j < N; ++j) {
      const double gj = gbar_at(g, i, sv, j);
      ybar[j] += gj;
      for (int q = 0; q < 7; ++q) kbar[q][j] += h * b[q] * gj;
    }
  }
  MDX_ADJ_STAGES
  for (int S = 6; S >= 1; --S) {
    double u[N];
    stage_arg(S, y, h, k, u);
    sys::f_vjp(u, t + CS[S] * h, p, kbar[S], ub, pbar);
    const double* __restrict__ ar = AT[S];
    for (int j = 0; j < N; ++j) ybar[j] += ub[j];
    for (int q = 0; q < S; ++q) {
      const double c = h * ar[q];
      for (int j = 0; j < N; ++j) kbar[q][j] += c * ub[j];
    }
    for (int j = 0; j < N; ++j) ub[j] = 0.0;
  }
  sys::f_vjp(y, t, p, kbar[0], ybar, pbar);
  for (int j = 0; j < N; ++j) lam[j] = ybar[j];
}

}  // namespace t5
}  // namespace mdx

extern "C" __global__ __launch_bounds__(MDX_BLOCK) void mdx_tsit5_adjoint(mdx::AdjointArgs g) {
  using namespace mdx::t5;
  const mdx::SolveArgs& a = g.s;
  const mdx::i64 i = (mdx::i64)blockIdx.x * MDX_BLOCK + threadIdx.x;
  if (i >= a.n_traj) return;
  PA p;
  p.load(a, i);
  State s;
  init(s, a, i);
  const Tab tb;
  const int C = g.ckpt_every;

  // phase A: the forward solve, checkpointing before every C-th accepted step
  int n_ck = 0, last = -1, status = 0;
  while (active(s, a)) {
    if (s.accepted % C == 0 && s.accepted != last) {
      if (n_ck < g.max_ckpt) store_ckpt(g, i, n_ck, s);
      else status = 1;
      ++n_ck;
      last = s.accepted;
    }
    step(s, p, a, i, tb);
  }
  if (status != 0) n_ck = 0;  // not enough room: report, leave the gradients NaN

  // saves the integration never reached are NaN in the history and carry no
  // dependence on the inputs
  double lam[N], pbar[NPB];
  for (int j = 0; j < N; ++j) lam[j] = 0.0;
  for (int m = 0; m < NPB; ++m) pbar[m] = 0.0;

  // phase B: segments last to first
  const mdx::i64 rec0 = (mdx::i64)g.max_ckpt * CK;
#pragma unroll 1
  for (int c = n_ck - 1; c >= 0; --c) {
    load_ckpt(g, i, c, s);
    const int target = s.accepted + C;
    int n_rec = 0;
    while (active(s, a) && s.accepted < target) {
      const mdx::i64 b = rec0 + (mdx::i64)n_rec * RC;
      for (int j = 0; j < N; ++j) wsat(g, i, b + j) = s.y[j];
      wsat(g, i, b + N + 0) = s.t;
      wsat(g, i, b + N + 1) = step_size(s);
      wsat(g, i, b + N + 2) = (double)s.save_idx;
      const int acc0 = s.accepted;
      step(s, p, a, i, tb);
      if (s.accepted > acc0) {
        wsat(g, i, b + N + 3) = (double)s.save_idx;
        ++n_rec;
      }
    }
#pragma unroll 1
    for (int r = n_rec - 1; r >= 0; --r) {
      const mdx::i64 b = rec0 + (mdx::i64)r * RC;
      double y[N];
      for (int j = 0; j < N; ++j) y[j] = wsat(g, i, b + j);
      reverse_step(g, i, p, y, wsat(g, i, b + N + 0), wsat(g, i, b + N + 1),
                   (int)wsat(g, i, b + N + 2), (int)wsat(g, i, b + N + 3), lam, pbar);
    }
  }
  // hist[0] is y0 itself
  const double bad = status != 0 ? __builtin_nan("") : 0.0;
  for (int j = 0; j < N; ++j)
    g.grad_y0[i * g.gy_ts + (mdx::i64)j * g.gy_vs] = lam[j] + gbar_at(g, i, 0, j) + bad;
  for (int m = 0; m < MDX_NP; ++m) g.grad_p[i * g.gp_ts + (mdx::i64)m * g.gp_vs] = pbar[m] + bad;
  g.status[i] = status;
}
