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
// at every save time a step passes. A save the integration never reaches
// (max_steps exhausted) is written as NaN; the reference leaves that memory as
// it found it.
//
// Placement (the reference's "shared"/"local" backends, tsit5.py:92-139):
//   MDX_ROLLED 0   the step unrolled, the state and its nine vectors in
//                  registers (small systems);
//   MDX_ROLLED 1   one stage body run six times; the seven stage vectors (k7
//                  carried between steps), the stage argument and the state
//                  in LDS [element][lane] as far as MDX_LDS_K/U/Y say, else in
//                  scratch. Same arithmetic, same bits.
//
// MDX_ADJOINT 1 builds the source of the discrete-adjoint kernel
// (tsit5_adjoint.hpp): the step writes no history and the solve kernels are
// left out.
#pragma once
#include "mdx_common.hpp"

namespace mdx {
namespace t5 {

constexpr int N = MDX_N;

#ifndef MDX_ROLLED
#define MDX_ROLLED 0
#endif
#ifndef MDX_LDS_K
#define MDX_LDS_K 0
#endif
// Elements per chunk where a loop over the per-lane vectors of the rolled step
// loads a chunk before computing with it.
#ifndef MDX_CHUNK
#define MDX_CHUNK 16
#endif
#ifndef MDX_LDS_U
#define MDX_LDS_U 0
#endif
#ifndef MDX_LDS_Y
#define MDX_LDS_Y 0
#endif
#if (MDX_LDS_K || MDX_LDS_U || MDX_LDS_Y) && !MDX_ROLLED
#error "LDS placement needs the rolled stage loop"
#endif

#if MDX_PARAMS_IN_REGS || MDX_ROLLED
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

#ifndef MDX_ADJOINT
#define MDX_ADJOINT 0
#endif
// Hold the tableau in vector registers across the step loop (unrolled step).
// A double that is not an inline constant costs two scalar moves at every
// use, and when a SIMD runs a single wave every instruction takes an issue
// slot; loaded once into registers, the step issues none of them.
#ifndef MDX_TAB_REGS
#define MDX_TAB_REGS 0
#endif

#define MDX_T5_COEFS(X)                                                                       \
  X(C2) X(C3) X(C4) X(C5) X(A21) X(A31) X(A32) X(A41) X(A42) X(A43) X(A51) X(A52) X(A53)      \
  X(A54) X(A61) X(A62) X(A63) X(A64) X(A65) X(B1) X(B2) X(B3) X(B4) X(B5) X(B6) X(E1) X(E2) \
  X(E3) X(E4) X(E5) X(E6) X(E7)

// The coefficients as the unrolled step reads them: compile-time constants,
// or (MDX_TAB_REGS) copies made once per kernel in registers the compiler
// cannot fold back into constants.
struct Tab {
#if MDX_TAB_REGS
#define MDX_T5_FIELD(c) double c;
#define MDX_T5_LOAD(c) c = reg(::mdx::t5::c);
  MDX_T5_COEFS(MDX_T5_FIELD)
  static __device__ __forceinline__ double reg(double x) {
    asm("" : "+v"(x));
    return x;
  }
  __device__ __forceinline__ Tab() { MDX_T5_COEFS(MDX_T5_LOAD) }
#undef MDX_T5_LOAD
#else
#define MDX_T5_FIELD(c) static constexpr double c = ::mdx::t5::c;
  MDX_T5_COEFS(MDX_T5_FIELD)
#endif
#undef MDX_T5_FIELD
};

// Stage rows 2..7 (the last is the solution weights B) and abscissae, for the
// rolled step and the adjoint.
__device__ constexpr double AT[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {A21, 0, 0, 0, 0, 0},
    {A31, A32, 0, 0, 0, 0},
    {A41, A42, A43, 0, 0, 0},
    {A51, A52, A53, A54, 0, 0},
    {A61, A62, A63, A64, A65, 0},
    {B1, B2, B3, B4, B5, B6},
};
__device__ constexpr double CS[7] = {0.0, C2, C3, C4, C5, C6, C7};

// Weights b_1..b_7 of the continuous extension at theta (tsit5.py:309-361).
__device__ __forceinline__ void dense_weights(double th, double* __restrict__ b) {
  b[0] = -1.0530884977290216 * th * (th - 1.3299890189751412) *
         (th * th - 1.4364028541716351 * th + 0.7139816917074209);
  b[1] = 0.1017 * th * th * (th * th - 2.1966568338249754 * th + 1.2949852507374631);
  b[2] = 2.490627285651252793 * th * th *
         (th * th - 2.38535645472061657 * th + 1.57803468208092486);
  b[3] = -16.54810288924490272 * (th - 1.21712927295533244) * (th - 0.61620406037800089) * th *
         th;
  b[4] = 47.37952196281928122 * (th - 1.203071208372362603) * (th - 0.658047292653547382) * th *
         th;
  b[5] = -34.87065786149660974 * (th - 1.2) * (th - 0.666666666666666667) * th * th;
  b[6] = 2.5 * (th - 1.0) * (th - 0.6) * th * th;
}

#if MDX_ROLLED
__device__ __forceinline__ unsigned opaque_lane() {
  unsigned v = threadIdx.x;
  asm volatile("; lane" : "+v"(v));
  return v;
}
#define MDX_LANE() const unsigned lane_ = opaque_lane()
// k1..k7 are rows 0..6
#if MDX_LDS_K
__shared__ double lds_k[7 * N * MDX_BLOCK];
#define KQ(st, q, j) lds_k[((q) * N + (j)) * MDX_BLOCK + lane_]
#else
#define KQ(st, q, j) (st).kk[q][j]
#endif
#if MDX_LDS_Y
__shared__ double lds_y[N * MDX_BLOCK];
#define YJ(st, j) lds_y[(j) * MDX_BLOCK + lane_]
#else
#define YJ(st, j) (st).y[j]
#endif
#if MDX_LDS_U
__shared__ double lds_u[N * MDX_BLOCK];
#define UJ(j) lds_u[(j) * MDX_BLOCK + lane_]
#else
#define UJ(j) u[j]
#endif
struct LdsVec {
  const double* base;
  __device__ __forceinline__ double operator[](int j) const { return base[j * MDX_BLOCK]; }
};
#endif

struct State {
#if !MDX_ROLLED
  double y[N];
  double k7[N];  // first-same-as-last carry
#else
#if !MDX_LDS_Y
  double y[N];
#endif
#if !MDX_LDS_K
  double kk[7][N];  // stage vectors; row 6 (k7) is the first-same-as-last carry
#endif
#endif
  double t, dt, tf, err_prev, err_prev2;
  int save_idx, n_steps, accepted, rejected;
  bool has_fsal;
};

__device__ __forceinline__ void init(State& s, const SolveArgs& a, i64 i) {
  const double* __restrict__ times = a.times;
#if MDX_ROLLED
  MDX_LANE();
#pragma unroll
  for (int j = 0; j < N; ++j) {
    YJ(s, j) = a.y0[i * a.y0_ts + (i64)j * a.y0_vs];
#if !MDX_ADJOINT
    a.hist[i * a.h_ts + (i64)j * a.h_vs] = YJ(s, j);
#endif
    KQ(s, 6, j) = 0.0;
  }
#else
#pragma unroll
  for (int j = 0; j < N; ++j) {
    s.y[j] = a.y0[i * a.y0_ts + (i64)j * a.y0_vs];
#if !MDX_ADJOINT
    a.hist[i * a.h_ts + (i64)j * a.h_vs] = s.y[j];
#endif
    s.k7[j] = 0.0;
  }
#endif
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
#if !MDX_ADJOINT
  if (a.resume && !a.cost) {
    // where a cost pilot left the trajectory: the state, the first-same-as-last
    // stage and the controller's memory, so the solve continues with the bits
    // an unbroken one has (the pilot wrote the saves it passed). Layout
    // [y | k7 | t dt err_prev err_prev2 save_idx n_steps accepted rejected
    // has_fsal], a row of N trajectories per element.
    const double* __restrict__ r = a.resume + i;
    const i64 st = a.n_traj;
#if MDX_ROLLED
#pragma unroll 1
    for (int j = 0; j < N; ++j) {
      YJ(s, j) = r[j * st];
      KQ(s, 6, j) = r[(N + j) * st];
    }
#else
#pragma unroll
    for (int j = 0; j < N; ++j) {
      s.y[j] = r[j * st];
      s.k7[j] = r[(N + j) * st];
    }
#endif
    s.t = r[(2 * N + 0) * st];
    s.dt = r[(2 * N + 1) * st];
    s.err_prev = r[(2 * N + 2) * st];
    s.err_prev2 = r[(2 * N + 3) * st];
    s.save_idx = (int)r[(2 * N + 4) * st];
    s.n_steps = (int)r[(2 * N + 5) * st];
    s.accepted = (int)r[(2 * N + 6) * st];
    s.rejected = (int)r[(2 * N + 7) * st];
    s.has_fsal = r[(2 * N + 8) * st] != 0.0;
  }
#endif
}

__device__ __forceinline__ bool active(const State& s, const SolveArgs& a) {
  return s.save_idx < a.n_save && s.t < s.tf && s.n_steps < a.max_steps;
}

#if MDX_ROLLED
// The Tsit5 step with the stage loop kept as a loop (large systems). Each stage
// forms u = y + dt * sum_q a_Sq k_q in the order the unrolled step uses (first
// term a product, then one fma per stage vector, then the fma with dt and y),
// so the bits are the same. For scratch placement, the never-executed stores
// (max_steps >= 0 is checked before launch) keep the arrays in memory.
__device__ __forceinline__ void step(State& s, PA& p, const SolveArgs& a, i64 i,
                                     const Tab& tb) {
  double dt_use = s.dt;
  if (dt_use > s.tf - s.t) dt_use = s.tf - s.t;
  if (dt_use < 1e-30) dt_use = 1e-30;
  const double t = s.t;
  MDX_LANE();
#if !MDX_LDS_U
  double u[N];
#endif
  {
    double out[N];
    if (s.has_fsal) {
#pragma unroll
      for (int j = 0; j < N; ++j) out[j] = KQ(s, 6, j);
    } else {
#if MDX_LDS_Y
      sys::rhs(LdsVec{&lds_y[lane_]}, t, p, out);
#else
      sys::rhs(s.y, t, p, out);
#endif
    }
#pragma unroll
    for (int j = 0; j < N; ++j) KQ(s, 0, j) = out[j];
  }
#define MDX_PIN(arr, len) \
  if (a.max_steps < 0) arr[(a.n_save + S) % (len)] = 0
#pragma unroll 1
  for (int S = 1; S < 7; ++S) {
    MDX_LANE();
#if !MDX_LDS_K
    MDX_PIN(s.kk[S], N);
#endif
#if !MDX_LDS_Y
    MDX_PIN(s.y, N);
#endif
#if !MDX_PARAMS_IN_REGS && MDX_NP > 0
    MDX_PIN(p.v, MDX_NP);
#endif
    // a chunk of elements at a time, each chunk's loads ahead of its
    // arithmetic (the stage sums over every element at once would keep one
    // accumulator per element live and leave the loads no registers)
    const double* __restrict__ ar = AT[S];
#pragma unroll
    for (int j0 = 0; j0 < N; j0 += MDX_CHUNK) {
      double q[MDX_CHUNK];
#pragma unroll
      for (int jj = 0; jj < MDX_CHUNK; ++jj)
        if (j0 + jj < N) q[jj] = ar[0] * KQ(s, 0, j0 + jj);
#pragma unroll 1
      for (int r = 1; r < S; ++r) {
        const double c = ar[r];
        double kv[MDX_CHUNK];
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj)
          if (j0 + jj < N) kv[jj] = KQ(s, r, j0 + jj);
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj)
          if (j0 + jj < N) q[jj] = fmaT(c, kv[jj], q[jj]);
      }
      double yv[MDX_CHUNK];
#pragma unroll
      for (int jj = 0; jj < MDX_CHUNK; ++jj)
        if (j0 + jj < N) yv[jj] = YJ(s, j0 + jj);
#pragma unroll
      for (int jj = 0; jj < MDX_CHUNK; ++jj)
        if (j0 + jj < N) UJ(j0 + jj) = fmaT(dt_use, q[jj], yv[jj]);
    }
    double out[N];
#if MDX_LDS_U
    sys::rhs(LdsVec{&lds_u[lane_]}, t + CS[S] * dt_use, p, out);
#else
    sys::rhs(u, t + CS[S] * dt_use, p, out);
#endif
#pragma unroll
    for (int j = 0; j < N; ++j) KQ(s, S, j) = out[j];
  }
#undef MDX_PIN

  // the embedded error, a chunk of elements at a time with each stage's
  // chunk loaded ahead of its multiply-adds; the norm still sums the elements
  // in ascending order
  double err = 0.0;
#pragma unroll
  for (int j0 = 0; j0 < N; j0 += MDX_CHUNK) {
    constexpr double EW[7] = {E1, E2, E3, E4, E5, E6, E7};
    double e[MDX_CHUNK];
#pragma unroll
    for (int jj = 0; jj < MDX_CHUNK; ++jj)
      if (j0 + jj < N) e[jj] = EW[0] * KQ(s, 0, j0 + jj);
#pragma unroll
    for (int r = 1; r < 7; ++r) {
      double kv[MDX_CHUNK];
#pragma unroll
      for (int jj = 0; jj < MDX_CHUNK; ++jj)
        if (j0 + jj < N) kv[jj] = KQ(s, r, j0 + jj);
#pragma unroll
      for (int jj = 0; jj < MDX_CHUNK; ++jj)
        if (j0 + jj < N) e[jj] = fmaT(EW[r], kv[jj], e[jj]);
    }
    double ya[MDX_CHUNK], yb[MDX_CHUNK];
#pragma unroll
    for (int jj = 0; jj < MDX_CHUNK; ++jj)
      if (j0 + jj < N) {
        ya[jj] = absT(YJ(s, j0 + jj));
        yb[jj] = absT(UJ(j0 + jj));
      }
#pragma unroll
    for (int jj = 0; jj < MDX_CHUNK; ++jj)
      if (j0 + jj < N) {
        const double err_est = dt_use * e[jj];
        const double scale = fmaT(a.rtol, ya[jj] > yb[jj] ? ya[jj] : yb[jj], a.atol);
        const double r = weight(a, i, j0 + jj) * err_est / scale;
        err = fmaT(r, r, err);
      }
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
#if !MDX_ADJOINT
      double b[7];
      dense_weights(th, b);
      double* __restrict__ row = a.hist + i * a.h_ts + (i64)s.save_idx * a.h_ss;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double q = b[0] * KQ(s, 0, j);
        q = fmaT(b[1], KQ(s, 1, j), q);
        q = fmaT(b[2], KQ(s, 2, j), q);
        q = fmaT(b[3], KQ(s, 3, j), q);
        q = fmaT(b[4], KQ(s, 4, j), q);
        q = fmaT(b[5], KQ(s, 5, j), q);
        q = fmaT(b[6], KQ(s, 6, j), q);
        row[(i64)j * a.h_vs] = fmaT(dt_use, q, YJ(s, j));
      }
#else
      (void)th;
#endif
      s.save_idx++;
    }
#pragma unroll
    for (int j = 0; j < N; ++j) YJ(s, j) = UJ(j);
    s.accepted++;
    s.has_fsal = true;
  } else {
    s.rejected++;
#pragma unroll
    for (int j = 0; j < N; ++j) KQ(s, 6, j) = 0.0;
    s.has_fsal = false;
  }
  const double factor_ =
      pid_factor(err_norm, accept, s.err_prev, s.err_prev2, MDX_E1, MDX_E2, MDX_E3, 0.2, 10.0);
  s.dt = dt_use * factor_;
  s.t = t_new;
  s.n_steps++;
}
#else
__device__ __forceinline__ void step(State& s, PA& p, const SolveArgs& a, i64 i,
                                     const Tab& tb) {
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
  for (int j = 0; j < N; ++j) u[j] = fmaT(dt_use, tb.A21 * k1[j], s.y[j]);
  sys::rhs(u, t + tb.C2 * dt_use, p, k2);
#pragma unroll
  for (int j = 0; j < N; ++j) u[j] = fmaT(dt_use, fmaT(tb.A32, k2[j], tb.A31 * k1[j]), s.y[j]);
  sys::rhs(u, t + tb.C3 * dt_use, p, k3);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double q = tb.A41 * k1[j];
    q = fmaT(tb.A42, k2[j], q);
    q = fmaT(tb.A43, k3[j], q);
    u[j] = fmaT(dt_use, q, s.y[j]);
  }
  sys::rhs(u, t + tb.C4 * dt_use, p, k4);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double q = tb.A51 * k1[j];
    q = fmaT(tb.A52, k2[j], q);
    q = fmaT(tb.A53, k3[j], q);
    q = fmaT(tb.A54, k4[j], q);
    u[j] = fmaT(dt_use, q, s.y[j]);
  }
  sys::rhs(u, t + tb.C5 * dt_use, p, k5);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double q = tb.A61 * k1[j];
    q = fmaT(tb.A62, k2[j], q);
    q = fmaT(tb.A63, k3[j], q);
    q = fmaT(tb.A64, k4[j], q);
    q = fmaT(tb.A65, k5[j], q);
    u[j] = fmaT(dt_use, q, s.y[j]);
  }
  sys::rhs(u, t + C6 * dt_use, p, k6);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double q = tb.B1 * k1[j];
    q = fmaT(tb.B2, k2[j], q);
    q = fmaT(tb.B3, k3[j], q);
    q = fmaT(tb.B4, k4[j], q);
    q = fmaT(tb.B5, k5[j], q);
    q = fmaT(tb.B6, k6[j], q);
    u[j] = fmaT(dt_use, q, s.y[j]);
  }
  sys::rhs(u, t + C7 * dt_use, p, s.k7);

  double err = 0.0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double e = tb.E1 * k1[j];
    e = fmaT(tb.E2, k2[j], e);
    e = fmaT(tb.E3, k3[j], e);
    e = fmaT(tb.E4, k4[j], e);
    e = fmaT(tb.E5, k5[j], e);
    e = fmaT(tb.E6, k6[j], e);
    e = fmaT(tb.E7, s.k7[j], e);
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
#if !MDX_ADJOINT
      double b[7];
      dense_weights(th, b);
      double* __restrict__ row = a.hist + i * a.h_ts + (i64)s.save_idx * a.h_ss;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double q = b[0] * k1[j];
        q = fmaT(b[1], k2[j], q);
        q = fmaT(b[2], k3[j], q);
        q = fmaT(b[3], k4[j], q);
        q = fmaT(b[4], k5[j], q);
        q = fmaT(b[5], k6[j], q);
        q = fmaT(b[6], s.k7[j], q);
        row[(i64)j * a.h_vs] = fmaT(dt_use, q, s.y[j]);
      }
#else
      (void)th;
#endif
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
#endif

__device__ __forceinline__ void finish(State& s, const SolveArgs& a, i64 i) {
  if (a.cost) {
    // a cost pilot (max_steps is its step budget): the steps left at the
    // current step size instead of the saves and counters, 0 once finished
    a.cost[i] = (s.save_idx < a.n_save && s.t < s.tf) ? (s.tf - s.t) / s.dt : 0.0;
    if (a.resume) {
      // the step state the resumed solve starts from (layout as init reads it)
      double* __restrict__ r = a.resume + i;
      const i64 st = a.n_traj;
#if MDX_ROLLED
      MDX_LANE();
#pragma unroll 1
      for (int j = 0; j < N; ++j) {
        r[j * st] = YJ(s, j);
        r[(N + j) * st] = KQ(s, 6, j);
      }
#else
#pragma unroll
      for (int j = 0; j < N; ++j) {
        r[j * st] = s.y[j];
        r[(N + j) * st] = s.k7[j];
      }
#endif
      r[(2 * N + 0) * st] = s.t;
      r[(2 * N + 1) * st] = s.dt;
      r[(2 * N + 2) * st] = s.err_prev;
      r[(2 * N + 3) * st] = s.err_prev2;
      r[(2 * N + 4) * st] = (double)s.save_idx;
      r[(2 * N + 5) * st] = (double)s.n_steps;
      r[(2 * N + 6) * st] = (double)s.accepted;
      r[(2 * N + 7) * st] = (double)s.rejected;
      r[(2 * N + 8) * st] = s.has_fsal ? 1.0 : 0.0;
    }
    return;
  }
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

#if !MDX_ADJOINT
extern "C" __global__ __launch_bounds__(MDX_BLOCK) void mdx_tsit5(mdx::SolveArgs a) {
  using namespace mdx::t5;
  const mdx::i64 i = (mdx::i64)blockIdx.x * MDX_BLOCK + threadIdx.x;
  if (i >= a.n_traj) return;
  PA p;
  p.load(a, i);
  State s;
  init(s, a, i);
  const Tab tb;
  while (active(s, a)) step(s, p, a, i, tb);
  finish(s, a, i);
}

extern "C" __global__ __launch_bounds__(MDX_BLOCK) void mdx_tsit5_refill(mdx::SolveArgs a) {
  using namespace mdx::t5;
  PA p;
  State s;
  const Tab tb;
  mdx::i64 i = -1;
  bool exhausted = false;
  for (;;) {
    const bool want = (i < 0) && !exhausted;
    const mdx::i64 got = mdx::wave_claim(a.queue, want);
    if (want) {
      if (got < a.n_traj) {
        // the claim order, when given, maps the claim to a trajectory; an
        // entry out of range is skipped and the lane claims again
        const mdx::i64 j = a.order ? a.order[got] : got;
        if (j >= 0 && j < a.n_traj) {
          i = j;
          p.load(a, i);
          init(s, a, i);
        }
      } else {
        exhausted = true;
      }
    }
    if (!__any(i >= 0)) break;
    if (i >= 0) {
      if (active(s, a)) step(s, p, a, i, tb);
      if (!active(s, a)) {
        finish(s, a, i);
        i = -1;
      }
    }
  }
}
#endif  // !MDX_ADJOINT
