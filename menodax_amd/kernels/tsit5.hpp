// tsit5.hpp -- Tsit5 ensemble kernels for gfx950, one trajectory per lane.
//
// Included by the source generated for one right-hand side, after it has
// defined namespace sys { rhs } and MDX_N, MDX_NP, MDX_N_ERROR, MDX_BLOCK,
// MDX_PARAMS_IN_REGS, MDX_E1/E2/E3.
//
// The method is the reference kernel's (modax/tsit5.py: tableau 34-81,
// controller 85-90, loop 238-400): seven stages with first-same-as-last reuse
// of k7 after an accepted step, the embedded error estimate, the PID
// controller with factors clamped to [0.2, 10], and the continuous extension
// at every save time a step passes. The state and its nine vectors live in
// registers. A save the integration never reaches (max_steps exhausted) is
// written as NaN; the reference leaves that memory as it found it.
#pragma once
#include "mdx_common.hpp"

namespace mdx {
namespace t5 {

constexpr int N = MDX_N;

#if MDX_PARAMS_IN_REGS
typedef RegParams<MDX_NP> PA;
#else
typedef GlobalParams PA;
#endif

constexpr double C2 = 161.0 / 1000.0, C3 = 327.0 / 1000.0, C4 = 9.0 / 10.0,
                 C5 = 0.9800255409045097, C6 = 1.0, C7 = 1.0;
constexpr double A21 = 161.0 / 1000.0;
constexpr double A31 = -0.008480655492356989, A32 = 0.335480655492357;
constexpr double A41 = 2.8971530571054935, A42 = -6.359448489975075, A43 = 4.3622954328695815;
constexpr double A51 = 5.325864828439257, A52 = -11.748883564062828, A53 = 7.4955393428898365,
                 A54 = -0.09249506636175525;
constexpr double A61 = 5.86145544294642, A62 = -12.92096931784711, A63 = 8.159367898576159,
                 A64 = -0.071584973281401, A65 = -0.028269050394068383;
constexpr double B1 = 0.09646076681806523, B2 = 0.01, B3 = 0.4798896504144996,
                 B4 = 1.379008574103742, B5 = -3.2900695154360807, B6 = 2.324710524099774;
constexpr double E1 = 0.0017800110522257773, E2 = 0.0008164344596567463,
                 E3 = -0.007880878010261994, E4 = 0.1447110071732629, E5 = -0.5823571654525552,
                 E6 = 0.45808210592918686, E7 = -1.0 / 66.0;

struct State {
  double y[N];
  double k7[N];  // first-same-as-last carry
  double t, dt, tf, err_prev, err_prev2;
  int save_idx, n_steps, accepted, rejected;
  bool has_fsal;
};

__device__ __forceinline__ void init(State& s, const SolveArgs& a, i64 i) {
  const double* __restrict__ times = a.times;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    s.y[j] = a.y0[i * a.y0_ts + (i64)j * a.y0_vs];
    a.hist[i * a.h_ts + (i64)j * a.h_vs] = s.y[j];
    s.k7[j] = 0.0;
  }
  s.t = times[0];
  s.tf = times[a.n_save - 1];
  s.dt = a.dt0 > 0.0 ? a.dt0 : (s.tf - s.t) * 1e-6;
  s.err_prev = 1.0;
  s.err_prev2 = 1.0;
  s.save_idx = 1;
  s.n_steps = 0;
  s.accepted = 0;
  s.rejected = 0;
  s.has_fsal = false;
}

__device__ __forceinline__ bool active(const State& s, const SolveArgs& a) {
  return s.save_idx < a.n_save && s.t < s.tf && s.n_steps < a.max_steps;
}

__device__ __forceinline__ void step(State& s, const PA& p, const SolveArgs& a, i64 i) {
  double dt_use = s.dt;
  if (dt_use > s.tf - s.t) dt_use = s.tf - s.t;
  if (dt_use < 1e-30) dt_use = 1e-30;
  const double t = s.t;
  double k1[N], k2[N], k3[N], k4[N], k5[N], k6[N], u[N];
  if (s.has_fsal) {
#pragma unroll
    for (int j = 0; j < N; ++j) k1[j] = s.k7[j];
  } else {
    sys::rhs(s.y, t, p, k1);
  }
#pragma unroll
  for (int j = 0; j < N; ++j) u[j] = fmaT(dt_use, A21 * k1[j], s.y[j]);
  sys::rhs(u, t + C2 * dt_use, p, k2);
#pragma unroll
  for (int j = 0; j < N; ++j) u[j] = fmaT(dt_use, fmaT(A32, k2[j], A31 * k1[j]), s.y[j]);
  sys::rhs(u, t + C3 * dt_use, p, k3);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double q = A41 * k1[j];
    q = fmaT(A42, k2[j], q);
    q = fmaT(A43, k3[j], q);
    u[j] = fmaT(dt_use, q, s.y[j]);
  }
  sys::rhs(u, t + C4 * dt_use, p, k4);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double q = A51 * k1[j];
    q = fmaT(A52, k2[j], q);
    q = fmaT(A53, k3[j], q);
    q = fmaT(A54, k4[j], q);
    u[j] = fmaT(dt_use, q, s.y[j]);
  }
  sys::rhs(u, t + C5 * dt_use, p, k5);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double q = A61 * k1[j];
    q = fmaT(A62, k2[j], q);
    q = fmaT(A63, k3[j], q);
    q = fmaT(A64, k4[j], q);
    q = fmaT(A65, k5[j], q);
    u[j] = fmaT(dt_use, q, s.y[j]);
  }
  sys::rhs(u, t + C6 * dt_use, p, k6);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double q = B1 * k1[j];
    q = fmaT(B2, k2[j], q);
    q = fmaT(B3, k3[j], q);
    q = fmaT(B4, k4[j], q);
    q = fmaT(B5, k5[j], q);
    q = fmaT(B6, k6[j], q);
    u[j] = fmaT(dt_use, q, s.y[j]);
  }
  sys::rhs(u, t + C7 * dt_use, p, s.k7);

  double err = 0.0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double e = E1 * k1[j];
    e = fmaT(E2, k2[j], e);
    e = fmaT(E3, k3[j], e);
    e = fmaT(E4, k4[j], e);
    e = fmaT(E5, k5[j], e);
    e = fmaT(E6, k6[j], e);
    e = fmaT(E7, s.k7[j], e);
    const double err_est = dt_use * e;
    const double ya = absT(s.y[j]), yb = absT(u[j]);
    const double scale = fmaT(a.rtol, ya > yb ? ya : yb, a.atol);
    const double r = weight(a, i, j) * err_est / scale;
    err = fmaT(r, r, err);
  }
  const double err_norm = __builtin_sqrt(err / (double)MDX_N_ERROR);
  const bool accept = err_norm <= 1.0 && !__builtin_isnan(err_norm);

  double t_new = t;
  if (accept) {
    t_new = t + dt_use;
    const double* __restrict__ times = a.times;
    while (s.save_idx < a.n_save) {
      const double ts = times[s.save_idx];
      const double at = absT(ts);
      if (!(ts <= t_new + 1e-12 * (1.0 > at ? 1.0 : at))) break;
      const double th = (ts - t) / dt_use;
      // tsit5.py:309-361
      const double b1 = -1.0530884977290216 * th * (th - 1.3299890189751412) *
                        (th * th - 1.4364028541716351 * th + 0.7139816917074209);
      const double b2 =
          0.1017 * th * th * (th * th - 2.1966568338249754 * th + 1.2949852507374631);
      const double b3 = 2.490627285651252793 * th * th *
                        (th * th - 2.38535645472061657 * th + 1.57803468208092486);
      const double b4 = -16.54810288924490272 * (th - 1.21712927295533244) *
                        (th - 0.61620406037800089) * th * th;
      const double b5 = 47.37952196281928122 * (th - 1.203071208372362603) *
                        (th - 0.658047292653547382) * th * th;
      const double b6 = -34.87065786149660974 * (th - 1.2) * (th - 0.666666666666666667) * th * th;
      const double b7 = 2.5 * (th - 1.0) * (th - 0.6) * th * th;
      double* __restrict__ row = a.hist + i * a.h_ts + (i64)s.save_idx * a.h_ss;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double q = b1 * k1[j];
        q = fmaT(b2, k2[j], q);
        q = fmaT(b3, k3[j], q);
        q = fmaT(b4, k4[j], q);
        q = fmaT(b5, k5[j], q);
        q = fmaT(b6, k6[j], q);
        q = fmaT(b7, s.k7[j], q);
        row[(i64)j * a.h_vs] = fmaT(dt_use, q, s.y[j]);
      }
      s.save_idx++;
    }
#pragma unroll
    for (int j = 0; j < N; ++j) s.y[j] = u[j];
    s.accepted++;
    s.has_fsal = true;
  } else {
    s.rejected++;
#pragma unroll
    for (int j = 0; j < N; ++j) s.k7[j] = 0.0;
    s.has_fsal = false;
  }
  const double factor_ =
      pid_factor(err_norm, accept, s.err_prev, s.err_prev2, MDX_E1, MDX_E2, MDX_E3, 0.2, 10.0);
  s.dt = dt_use * factor_;
  s.t = t_new;
  s.n_steps++;
}

__device__ __forceinline__ void finish(State& s, const SolveArgs& a, i64 i) {
  const double nan = __builtin_nan("");
  while (s.save_idx < a.n_save) {
    double* __restrict__ row = a.hist + i * a.h_ts + (i64)s.save_idx * a.h_ss;
#pragma unroll
    for (int j = 0; j < N; ++j) row[(i64)j * a.h_vs] = nan;
    s.save_idx++;
  }
  a.accepted[i] = s.accepted;
  a.rejected[i] = s.rejected;
  a.loop[i] = s.n_steps;
}

}  // namespace t5
}  // namespace mdx

extern "C" __global__ __launch_bounds__(MDX_BLOCK) void mdx_tsit5(mdx::SolveArgs a) {
  using namespace mdx::t5;
  const mdx::i64 i = (mdx::i64)blockIdx.x * MDX_BLOCK + threadIdx.x;
  if (i >= a.n_traj) return;
  PA p;
  p.load(a, i);
  State s;
  init(s, a, i);
  while (active(s, a)) step(s, p, a, i);
  finish(s, a, i);
}

extern "C" __global__ __launch_bounds__(MDX_BLOCK) void mdx_tsit5_refill(mdx::SolveArgs a) {
  using namespace mdx::t5;
  PA p;
  State s;
  mdx::i64 i = -1;
  bool exhausted = false;
  for (;;) {
    const bool want = (i < 0) && !exhausted;
    const mdx::i64 got = mdx::wave_claim(a.queue, want);
    if (want) {
      if (got < a.n_traj) {
        i = got;
        p.load(a, i);
        init(s, a, i);
      } else {
        exhausted = true;
      }
    }
    if (!__any(i >= 0)) break;
    if (i >= 0) {
      if (active(s, a)) step(s, p, a, i);
      if (!active(s, a)) {
        finish(s, a, i);
        i = -1;
      }
    }
  }
}
