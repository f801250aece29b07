// rodas5p_adjoint.hpp -- discrete adjoint of the Rodas5P ensemble solve, one
// trajectory per lane.
//
// Included after rodas5p.hpp in a source generated with MDX_ADJOINT=1 (the
// forward step then writes no history), whose sys namespace also holds f_vjp /
// jac_vjp / dt_vjp (reverse-mode contractions of the callback) and, for a
// sparse system, sparse_solve_t. With MDX_ROLLED=1 (larger systems) the stage
// loops stay loops and the per-stage arrays live in scratch; with MDX_ROLLED=0
// (a few variables) every stage loop is unrolled, so the stage vectors and
// their adjoints are registers.
//
// Given dL/dhist for a loss of the saved history, each lane returns dL/dy0 and
// dL/dparams of its trajectory, differentiating the solver's own arithmetic:
// the accepted steps with their step sizes held fixed (the step-size
// controller is not differentiated), every stage, the iteration matrix's
// dependence on the state and parameters through J(y, p), the df/dt term, and
// the dense output that produced each save.
//
//   phase A  the forward solve again, storing the whole step state (state
//            vector, time, step size, controller memory, counters) before
//            every ckpt_every-th accepted step;
//   phase B  checkpoints last to first: replay the segment from its
//            checkpoint -- the same code, so the same bits -- recording each
//            accepted step's (y_n, t_n, h_n, saves it produced), then run back
//            through the recorded steps.
//
// Reverse of one step (y_{n+1} = u7 + k7, k_s = M^{-1} r_s, M = I/(h gamma) - J):
//   r_bar_s  = M^{-T} k_bar_s                    (s = 7 .. 0)
//   J_bar   += r_bar_s k_s^T   -> d/d(y,p) of r_bar_s . J(y, p) k_s
//   dT_bar  += h d_s r_bar_s ;  k_bar_q += C_sq / h r_bar_s
//   u_bar_s += J(u_s)^T r_bar_s ; p_bar += J_p(u_s)^T r_bar_s
//   u_bar_s flows to y (and to k_q through the A rows, or to u_{s-1}, k_{s-1}
//   for the two end stages)
// Workspace, per lane, [slot][lane] (coalesced): max_ckpt checkpoints of
// N + 8 doubles, then ckpt_every step records of N + 4 doubles.
//
// With MDX_ADJ_LDS=1 (rolled step only) the per-lane arrays the reverse step
// indexes by stage -- the stage vectors k (the forward step's own LDS stage
// vectors, free while a step is reversed), their adjoints k_bar and the double
// copy of the factors -- live in LDS, [element][lane], instead of scratch: a
// lane whose ensemble leaves most SIMDs idle waits on each access, and LDS
// answers in a fraction of scratch's time. Same arithmetic, same bits.
#pragma once

namespace mdx {
namespace r5 {


#if MDX_ROLLED
#define MDX_ADJ_STAGES _Pragma("unroll 1")
#else
#define MDX_ADJ_STAGES _Pragma("unroll")
#endif

#ifndef MDX_ADJ_LDS
#define MDX_ADJ_LDS 0
#endif
#if MDX_ADJ_LDS && !(MDX_ROLLED && MDX_LDS_K && !MDX_LDS_U && !MDX_LDS_Y && !MDX_LDS_LU)
#error "the adjoint's LDS arrays go with the rolled step's LDS stage vectors only"
#endif
// MDX_ADJ_CACHE=1 (LDS placement only): seven more vectors a lane in LDS. The
// step's recomputation leaves there each stage argument u_S (S = 1..7), which
// the reverse sweep reads instead of re-forming it from the stage vectors, and
// the sweep then stores each stage's u_bar_S in the same slot. Stage S's
// k_bar_S is pulled when the sweep reaches it, from the r_bar_R and u_bar_R of
// the stages R > S done before, in the order the pushing form adds them
// (R = 7 first; its C term, then its u term), so the sums and their bits are
// the same; the pushing form's read-modify-writes of every earlier k_bar at
// every stage go.
#ifndef MDX_ADJ_CACHE
#define MDX_ADJ_CACHE 0
#endif
#if MDX_ADJ_CACHE && !MDX_ADJ_LDS
#error "the adjoint's stage cache is part of its LDS placement"
#endif

#if MDX_ADJ_LDS
__shared__ double lds_kb[8 * N * MDX_BLOCK];
__shared__ double lds_lud[MDX_NLU * MDX_BLOCK];
#if MDX_ADJ_CACHE
// u_S of the step, then u_bar_S once stage S is reversed (S = 1..7)
__shared__ double lds_us[7 * N * MDX_BLOCK];
#define USV(S, j) lds_us[(((S) - 1) * N + (j)) * MDX_BLOCK + lane_]
#endif
// element e of this lane's row in an [element][lane] LDS array
struct LdsRows {
  double* base;
  __device__ __forceinline__ double& operator[](int e) const { return base[e * MDX_BLOCK]; }
};
// Stage vectors k[q][j] and adjoints k_bar[q][j]; each user re-reads the lane
// index (MDX_LANE) where its loop starts, as the forward step does.
struct KStore {};
#define KV(q, j) KQ(q, j)
#define KBV(q, j) lds_kb[((q) * N + (j)) * MDX_BLOCK + lane_]
#else
struct KStore {
  LUT v[8][N];
};
#define KV(q, j) ks.v[q][j]
#define KBV(q, j) kbar[q][j]
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

// The step's factorisation and stage vectors at (y, t, h), by the forward
// step's arithmetic.
__device__ void recompute(const double* __restrict__ y, double t, double h, PA& p,
                          LUT* __restrict__ lu, Pivots& ipiv, KStore& ks) {
  (void)ks;
  const double inv_dt = 1.0 / h;
  const double t_end = t + h;
  const double dtgamma_inv = 1.0 / (h * GAMMA);
#if MDX_NEEDS_CLEAR
  for (int q = 0; q < MDX_NLU; ++q) lu[q] = (LUT)0;
#endif
  double dT[N];
  sys::neg_jac(y, t, p, lu, dT);
  for (int d = 0; d < N; ++d) lu[diag_slot(d)] = lu[diag_slot(d)] + (LUT)dtgamma_inv;
  factor(lu, ipiv);
  double u[N];
  MDX_ADJ_STAGES
  for (int S = 0; S < 8; ++S) {
    MDX_LANE();
    double ts;
    if (S == 0) {
      for (int j = 0; j < N; ++j) u[j] = y[j];
      ts = t;
    } else if (S < 6) {
      for (int j = 0; j < N; ++j) {
        double acc = AM[S - 1][0] * (double)KV(0, j);
        for (int q = 1; q < S; ++q) acc = fmaT(AM[S - 1][q], (double)KV(q, j), acc);
        u[j] = y[j] + acc;
      }
      ts = S < 5 ? t + CT[S - 1] * h : t_end;
    } else {
      for (int j = 0; j < N; ++j) u[j] = u[j] + (double)KV(S - 1, j);
      ts = t_end;
    }
#if MDX_ADJ_CACHE
    if (S > 0)
      for (int j = 0; j < N; ++j) USV(S, j) = u[j];
#endif
    double work[N];
    sys::rhs(u, ts, p, work);
    LUT r[N];
    for (int j = 0; j < N; ++j) {
      double v;
      if (S == 0) {
        v = fmaT(h * DV[0], DTJ(j), work[j]);
      } else {
        double acc = CM[S][0] * (double)KV(0, j);
        for (int q = 1; q < S; ++q) acc = fmaT(CM[S][q], (double)KV(q, j), acc);
        v = S < 5 ? fmaT(acc, inv_dt, fmaT(h * DV[S], DTJ(j), work[j])) : fmaT(acc, inv_dt, work[j]);
      }
      r[j] = (LUT)v;
    }
    solve(lu, ipiv, r);
    for (int j = 0; j < N; ++j) KV(S, j) = r[j];
  }
}

// u_S of the step (the argument of stage S's right-hand side) and its time.
__device__ __forceinline__ double stage_point(int S, const double* __restrict__ y, double t,
                                              double h, KStore& ks, double* __restrict__ u) {
  (void)ks;
  MDX_LANE();
  if (S == 0) {
    for (int j = 0; j < N; ++j) u[j] = y[j];
    return t;
  }
  const int A = S < 6 ? S - 1 : 4;
  for (int j = 0; j < N; ++j) {
    double acc = AM[A][0] * (double)KV(0, j);
    for (int q = 1; q <= A; ++q) acc = fmaT(AM[A][q], (double)KV(q, j), acc);
    u[j] = y[j] + acc;
  }
  if (S >= 6) for (int j = 0; j < N; ++j) u[j] = u[j] + (double)KV(5, j);
  if (S >= 7) for (int j = 0; j < N; ++j) u[j] = u[j] + (double)KV(6, j);
  return S == 0 ? t : (S < 5 ? t + CT[S - 1] * h : t + h);
}

// lam: dL/dy_{n+1} in, dL/dy_n out; pbar accumulates dL/dp.
__device__ void reverse_step(const AdjointArgs& g, i64 i, PA& p, const double* __restrict__ y,
                             double t, double h, int save_lo, int save_hi,
                             double* __restrict__ lam, double* __restrict__ pbar) {
  const SolveArgs& a = g.s;
  LUT lu[MDX_NLU];
  Pivots ipiv;
  KStore ks;
  recompute(y, t, h, p, lu, ipiv, ks);
  MDX_LANE();
#if MDX_ADJ_LDS
  const LdsRows lud{&lds_lud[lane_]};
#else
  double lud[MDX_NLU];
  double kbar[8][N];
#endif
  for (int q = 0; q < MDX_NLU; ++q) lud[q] = (double)lu[q];
  const double inv_dt = 1.0 / h;

  double ybar[N], ucarry[N], dTbar[N];
  for (int q = 0; q < 8; ++q)
    for (int j = 0; j < N; ++j) KBV(q, j) = 0.0;
  for (int j = 0; j < N; ++j) {
    ybar[j] = 0.0;
    dTbar[j] = 0.0;
    ucarry[j] = lam[j];   // y_{n+1} = u7 + k7
    KBV(7, j) = lam[j];
  }
  // saves the step produced: theta1 y + theta (y_new + theta1 (h1 + theta (h2 + theta h3)))
  for (int sv = save_lo; sv < save_hi; ++sv) {
    const double theta = (a.times[sv] - t) / h;
    const double theta1 = 1.0 - theta;
    const double c1 = theta * theta1, c2 = c1 * theta, c3 = c2 * theta;
    for (int j = 0; j < N; ++j) {
      const double gj = gbar_at(g, i, sv, j);
      ybar[j] += theta1 * gj;
      ucarry[j] += theta * gj;
      KBV(7, j) += theta * gj;
      for (int q = 0; q < 8; ++q) KBV(q, j) += (H1[q] * c1 + H2[q] * c2 + H3[q] * c3) * gj;
    }
  }
  MDX_ADJ_STAGES
  for (int S = 7; S >= 0; --S) {
    MDX_LANE();
    double rb[N], kd[N], u[N], ub[N];
    for (int j = 0; j < N; ++j) {
      rb[j] = KBV(S, j);
      kd[j] = (double)KV(S, j);
      ub[j] = 0.0;
    }
#if MDX_ADJ_CACHE
    // k_bar_S: what the step's end and its saves gave it, then every later
    // stage's share, stage 7 first (a stage's loads across the elements are
    // independent, so they are in flight together)
    for (int R = 7; R > S; --R) {
      const double c = CM[R][S] * inv_dt;
      for (int j = 0; j < N; ++j) rb[j] += c * KBV(R, j);  // KBV(R) holds r_bar_R
      if (R >= 6) {
        if (S == R - 1)
          for (int j = 0; j < N; ++j) rb[j] += USV(R, j);  // USV(R) holds u_bar_R
      } else {
        const double a = AM[R - 1][S];
        for (int j = 0; j < N; ++j) rb[j] += a * USV(R, j);
      }
    }
#endif
    solve_t(lud, ipiv, rb);
    sys::jac_vjp(y, t, p, rb, kd, ybar, pbar);
    if (S < 5) {
      const double dd = h * DV[S];
      for (int j = 0; j < N; ++j) dTbar[j] += dd * rb[j];
    }
#if MDX_ADJ_CACHE
    for (int j = 0; j < N; ++j) KBV(S, j) = rb[j];
    double ts;
    if (S == 0) {
      for (int j = 0; j < N; ++j) u[j] = y[j];
      ts = t;
    } else {
      for (int j = 0; j < N; ++j) u[j] = USV(S, j);
      ts = S < 5 ? t + CT[S - 1] * h : t + h;
    }
#else
    for (int q = 0; q < S; ++q) {
      const double c = CM[S][q] * inv_dt;
      for (int j = 0; j < N; ++j) KBV(q, j) += c * rb[j];
    }
    const double ts = stage_point(S, y, t, h, ks, u);
#endif
    sys::f_vjp(u, ts, p, rb, ub, pbar);
    if (S >= 5) {
      for (int j = 0; j < N; ++j) ub[j] += ucarry[j];
    }
    if (S == 7 || S == 6) {
      // u_S = u_{S-1} + k_{S-1}
      for (int j = 0; j < N; ++j) {
        ucarry[j] = ub[j];
#if MDX_ADJ_CACHE
        USV(S, j) = ub[j];
#else
        KBV(S - 1, j) += ub[j];
#endif
      }
    } else if (S > 0) {
      // u_S = y + sum_q A_{S,q} k_q
      for (int j = 0; j < N; ++j) ybar[j] += ub[j];
#if MDX_ADJ_CACHE
      for (int j = 0; j < N; ++j) USV(S, j) = ub[j];
#else
      for (int q = 0; q < S; ++q)
        for (int j = 0; j < N; ++j) KBV(q, j) += AM[S - 1][q] * ub[j];
#endif
    } else {
      for (int j = 0; j < N; ++j) ybar[j] += ub[j];
    }
  }
#if !MDX_DT_ZERO
  sys::dt_vjp(y, t, p, dTbar, ybar, pbar);
#endif
  for (int j = 0; j < N; ++j) lam[j] = ybar[j];
}

}  // namespace r5
}  // namespace mdx

extern "C" __global__ __launch_bounds__(MDX_BLOCK) void mdx_rodas5p_adjoint(
    mdx::AdjointArgs g) {
  using namespace mdx::r5;
  const mdx::SolveArgs& a = g.s;
  const mdx::i64 i = (mdx::i64)blockIdx.x * MDX_BLOCK + threadIdx.x;
  if (i >= a.n_traj) return;
  PA p;
  p.load(a, i);
  State s;
  init(s, a, i, p);
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

  // saves the loop never reached hold the final state
  double lam[N], pbar[NPB];
  for (int j = 0; j < N; ++j) lam[j] = 0.0;
  for (int m = 0; m < NPB; ++m) pbar[m] = 0.0;
  for (int sv = s.save_idx; sv < a.n_save; ++sv)
    for (int j = 0; j < N; ++j) lam[j] += gbar_at(g, i, sv, j);

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
