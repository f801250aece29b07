/*
 * oracle.c -- CPU restatement of modax's Tsit5 and Rodas5P ensemble solvers.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is held
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and never as the thing measured or shipped. The product path in
 * menodax_amd/ neither links nor calls it.
 *
 * What it restates (file:line in the reference, modax package):
 *   - Rodas5P tableau, d-coefficients and controller constants: rodas5P.py:44-113
 *   - dense LU with partial pivoting (first max wins, whole-row swap, zero pivot
 *     skipped, f64 reciprocal of an f32 pivot):             rodas5P.py:187-216
 *   - dense forward/back substitution applying pivots in order: rodas5P.py:218-236
 *   - negated Jacobian written into the LU buffer, dT = df/dt: rodas5P.py:251-303,
 *     diagonal 1/(h*gamma) added in LU precision, then factorise: rodas5P.py:770-775
 *   - eight stages (f64 accumulation, LU-precision right-hand side and storage):
 *     rodas5P.py:592-668, 780-901
 *   - weighted RMS error, PID controller, dense output, saves, tail fill and
 *     counters: rodas5P.py:903-1005
 *   - up-looking unpivoted sparse LU on the CSR image of L+U with the inverse
 *     diagonal stored, and the two sparse sweeps reading the caller's indices:
 *     _sparse_direct.py:472-622
 *   - Tsit5 tableau, controller and loop incl. FSAL, dense output, no tail fill:
 *     tsit5.py:34-90, 218-404
 *
 * The Jacobian is NOT taken by automatic differentiation here: every system in
 * this file carries a hand-written closed-form Jacobian and df/dt, so the
 * oracle is independent of the product's code generator.
 *
 * Rounding. The reference runs on a GPU whose compiler contracts a multiply
 * feeding an add into one fused multiply-add; its published Robertson step
 * counts (157 attempted, 10 rejected with the FP32 LU) are not reproduced
 * without fusing the f32 triangular solves that way (and are reproduced exactly
 * with the fusions below). So this file spells out every fusion as an explicit
 * fma() and is built with -ffp-contract=off: the product's
 * HIP kernels place the same fma()s in the same spots under the same flag, so
 * the two agree bit for bit wherever the right-hand side is polynomial. The
 * step-size controller's power is the deterministic det_pow() below (built from
 * correctly rounded IEEE operations only), written out identically on the GPU.
 *
 * Precision semantics in LU-fp32 mode follow numba's typing of the reference
 * kernel (a Python float literal is float64):
 *   stage right-hand side accumulated in f64 and cast to f32; stage vectors
 *   stored f32; -J cast to f32; 1/(h*gamma) cast to f32 then added in f32; dense
 *   factorisation forms inv_piv and the multiplier in f64 and stores f32; dense
 *   and sparse solves pure f32; sparse pivot reciprocal formed in f64, stored f32.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Deterministic power for the step-size controller                           */
/* ------------------------------------------------------------------------- */
/* x**y for finite x > 0 as exp(y * log(x)), with log from the atanh series of
 * s = (m-1)/(m+1) on m in [sqrt(1/2), sqrt(2)) and exp from a Cody-Waite
 * reduction by ln 2 and a degree-13 Taylor polynomial. Every operation is a
 * correctly rounded IEEE one (fma, division, frexp/ldexp, rint), so the CPU and
 * the GPU evaluate it to the same bits. Accuracy is a few ulp, ample for a
 * step-size factor; determinism is the point. det_pow(x, +-0) is exactly 1. */
static const double DP_LN2_HI = 6.93147180369123816490e-01;
static const double DP_LN2_LO = 1.90821492927058770002e-10;
static const double DP_INV_LN2 = 1.44269504088896338700e+00;

static double det_log(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double q = 1.0 / 23.0;
  q = fma(q, s2, 1.0 / 21.0);
  q = fma(q, s2, 1.0 / 19.0);
  q = fma(q, s2, 1.0 / 17.0);
  q = fma(q, s2, 1.0 / 15.0);
  q = fma(q, s2, 1.0 / 13.0);
  q = fma(q, s2, 1.0 / 11.0);
  q = fma(q, s2, 1.0 / 9.0);
  q = fma(q, s2, 1.0 / 7.0);
  q = fma(q, s2, 1.0 / 5.0);
  q = fma(q, s2, 1.0 / 3.0);
  double two_s = 2.0 * s;
  double log_m = fma(two_s * s2, q, two_s);
  double de = (double)e;
  return fma(de, DP_LN2_HI, fma(de, DP_LN2_LO, log_m));
}

static double det_exp(double z) {
  double k = rint(z * DP_INV_LN2);
  double r = fma(-k, DP_LN2_HI, z);
  r = fma(-k, DP_LN2_LO, r);
  double q = 1.0 / 6227020800.0;
  q = fma(q, r, 1.0 / 479001600.0);
  q = fma(q, r, 1.0 / 39916800.0);
  q = fma(q, r, 1.0 / 3628800.0);
  q = fma(q, r, 1.0 / 362880.0);
  q = fma(q, r, 1.0 / 40320.0);
  q = fma(q, r, 1.0 / 5040.0);
  q = fma(q, r, 1.0 / 720.0);
  q = fma(q, r, 1.0 / 120.0);
  q = fma(q, r, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  q = fma(q, r, 1.0);
  q = fma(q, r, 1.0);
  return ldexp(q, (int)k);
}

double det_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  return det_exp(y * det_log(x));
}

/* ------------------------------------------------------------------------- */
/* Systems                                                                    */
/* ------------------------------------------------------------------------- */
enum {
  SYS_LORENZ = 0,       /* reference/systems/python/lorenz.py:36-43           */
  SYS_ROBERTSON = 1,    /* reference/systems/python/robertson.py:17-22        */
  SYS_FORCED_DECAY = 2, /* tests/test_solver_options.py:142-143               */
  SYS_DECAY = 3,        /* tests/test_solver_vmap.py decay: -p0*y0            */
  SYS_CHAIN4 = 4,       /* tests/test_sparsity.py:40-46                       */
  SYS_RING6 = 5,        /* tests/test_sparse_direct.py:333-342                */
  SYS_VDP_LATTICE = 6,  /* reference/systems/python/vdp.py:253-290 (+ open chain) */
  SYS_LINQUAD = 7,      /* f = diag(p) y + off-diagonal terms + q y*y (uneven stiff ensemble) */
  SYS_DAMPED_OSC = 8,   /* tests/test_save_hook.py:23-24                       */
  SYS_KAPS = 9,         /* reference/systems/python/kaps.py:26-48             */
  SYS_CALLBACK = 10,    /* any system: f, dense J and df/dt from host callbacks */
  SYS_BBN = 11,         /* examples/bbn_estimation/main.py:74-161, x = Q/T the time */
};

/* The four-species BBN network (n, p, D, 4He) with x = Q/T as the independent
 * variable and parameters (log10 eta_10, N_eff): constants and rates as in
 * examples/bbn_estimation/main.py:44-60, 74-161. Written once in complex
 * arithmetic: with real arguments it is f itself; with a small imaginary
 * perturbation of one input it gives that input's derivative column by the
 * complex step, Im f(z + ih) / h, with no cancellation (an independent
 * derivative method from the product's forward-mode sweeps). The g_star step
 * at T = m_e is taken on the real part. */
static void bbn_eval(const double complex* y, double complex x, const double* p,
                     double complex* f) {
  const double Q = 1.293, B_D = 2.225, M_N = 938.272, M_E = 0.511, M_PL = 1.2209e22;
  const double ZETA3 = 1.2020569, N_EFF_SM = 3.044, TAU = 879.4 / 6.582119e-22;
  const double S_NP = 4.0e-11, S_DD = 3.4e-5, PI = 3.14159265358979323846;
  const double eta = pow(10.0, p[0] - 10.0);
  const double complex T = Q / x;
  const double g_star = 3.91 + (creal(T) > M_E ? 10.75 - 3.91 : 0.0) + 1.75 * (p[1] - N_EFF_SM);
  const double complex H = sqrt(4.0 * PI * PI * PI * g_star / 45.0) * (T * T) / M_PL;
  const double complex n_b = eta * (2.0 * ZETA3 / (PI * PI)) * (T * T * T);
  const double complex x2 = x * x;
  const double complex g_np = (255.0 / TAU) * (12.0 + 6.0 * x + x2) / (x2 * x2 * x) + 1.0 / TAU;
  const double complex g_pn = g_np * cexp(-x);
  const double complex v = 4.0 * PI / (M_N * T);
  const double complex k_d = n_b * 0.75 * (v * csqrt(v)) * cexp(B_D / T);
  const double complex r_np = n_b * S_NP * (y[0] * y[1] - y[2] / k_d);
  const double complex r_dd = n_b * S_DD * (y[2] * y[2]);
  const double complex den = H * x;
  f[0] = (g_pn * y[1] - g_np * y[0] - r_np) / den;
  f[1] = (g_np * y[0] - g_pn * y[1] - r_np) / den;
  f[2] = (r_np - 2.0 * r_dd) / den;
  f[3] = r_dd / den;
}

/* complex-step derivatives of the BBN network: J[r*4+c] and df/dx */
static void bbn_jac(const double* y, double x, const double* p, double* J, double* dT) {
  const double h = 1e-30;
  double complex yc[4], f[4];
  for (int c = 0; c <= 4; ++c) {
    for (int j = 0; j < 4; ++j) yc[j] = y[j];
    double complex xc = x;
    if (c < 4) yc[c] += h * I; else xc += h * I;
    bbn_eval(yc, xc, p, f);
    for (int r = 0; r < 4; ++r) {
      if (c < 4) J[r * 4 + c] = cimag(f[r]) / h; else dT[r] = cimag(f[r]) / h;
    }
  }
}

/* Host callbacks of a SYS_CALLBACK system, and the forward-sensitivity
 * derivatives of any system a joint solve carries: with a state direction s and
 * a parameter direction pd,
 *   sens_f   = J_y s + J_p pd            (the sensitivity rows of the joint rhs)
 *   sens_dt  = d/dt of sens_f            (their df/dt)
 *   sens_cpl = D_y(sens_f)[v]            (the block-triangular stage coupling)
 * modax/_sensitivity.py:148-243, rodas5P.py:400-436, 512-561. */
typedef void (*cb_f_t)(const double* y, double t, const double* p, double* out);
typedef void (*cb_jac_t)(const double* y, double t, const double* p, double* J, double* dT);
typedef void (*cb_sens_t)(const double* y, double t, const double* p, const double* s,
                          const double* pd, double* out);
typedef void (*cb_cpl_t)(const double* y, double t, const double* p, const double* s,
                         const double* pd, const double* v, double* out);

typedef struct {
  int kind;
  int n, n_p;
  /* lattice */
  int n_osc, wrap, coupling_range;
  double mu, d, omega;
  /* linquad: off-diagonal terms and quadratic coefficient; a term is
   * val * y[col], or (val * p[pidx]) * y[col] when pidx >= 0 */
  int n_off;
  const int* off_row;
  const int* off_col;
  const int* off_pidx;
  const double* off_val;
  double quad;
  /* kaps: per-pair inverse epsilon */
  const double* inv_eps;
  /* callbacks */
  cb_f_t cb_f;
  cb_jac_t cb_jac;
  cb_sens_t cb_sens_f, cb_sens_dt;
  cb_cpl_t cb_sens_cpl;
} osys;

/* neighbours along the ring or chain, one entry per place (vdp.py:235-250);
 * an open chain drops the neighbours past either end. */
static int lattice_neighbours(const osys* s, int osc, int* out) {
  int cnt = 0;
  for (int off = 1; off <= s->coupling_range; ++off) {
    int cand[2] = {osc - off, osc + off};
    for (int c = 0; c < 2; ++c) {
      int o = cand[c];
      if (s->wrap) {
        o = ((o % s->n_osc) + s->n_osc) % s->n_osc;
      } else if (o < 0 || o >= s->n_osc) {
        continue;
      }
      if (o != osc) out[cnt++] = o;
    }
  }
  return cnt;
}

static void sys_f(const osys* s, const double* y, double t, const double* p, double* f) {
  switch (s->kind) {
    case SYS_CALLBACK:
      s->cb_f(y, t, p, f);
      break;
    case SYS_LORENZ: {
      const double sigma = 10.0, beta = 8.0 / 3.0;
      f[0] = sigma * (y[1] - y[0]);
      f[1] = y[0] * (p[0] - y[2]) - y[1];
      f[2] = y[0] * y[1] - beta * y[2];
      break;
    }
    case SYS_ROBERTSON:
      f[0] = -p[0] * y[0] + p[1] * y[1] * y[2];
      f[1] = p[0] * y[0] - p[1] * y[1] * y[2] - p[2] * (y[1] * y[1]);
      f[2] = p[2] * (y[1] * y[1]);
      break;
    case SYS_FORCED_DECAY:
      f[0] = -p[0] * y[0] + p[1] * t;
      break;
    case SYS_DECAY:
      f[0] = -p[0] * y[0];
      break;
    case SYS_CHAIN4:
      f[0] = -p[0] * y[0];
      f[1] = p[0] * y[0] - p[1] * y[1];
      f[2] = p[1] * y[1] - p[2] * y[2];
      f[3] = p[2] * y[2];
      break;
    case SYS_RING6:
      for (int i = 0; i < 6; ++i) {
        int nx = (i + 1) % 6, pv = (i + 5) % 6;
        f[i] = p[0] * (y[nx] - 2.0 * y[i] + y[pv]) - p[1] * y[i];
      }
      break;
    case SYS_VDP_LATTICE: {
      int nb[64];
      double strength = s->d / s->coupling_range;
      for (int osc = 0; osc < s->n_osc; ++osc) {
        int x = 2 * osc, v = x + 1;
        int cnt = lattice_neighbours(s, osc, nb);
        double lap = 0.0;
        int first = 1;
        for (int k = 0; k < cnt; ++k) {
          lap = first ? y[2 * nb[k]] : lap + y[2 * nb[k]];
          first = 0;
        }
        double self = -(double)cnt * y[x];
        lap = first ? self : lap + self;
        f[x] = y[v] + 0.0;
        f[v] = p[0] * s->mu * (1.0 - y[x] * y[x]) * y[v] + -(s->omega * s->omega) * y[x] +
               strength * lap + 0.0;
      }
      break;
    }
    case SYS_LINQUAD: {
      /* p[j]*y[j] + terms in order + quad*(y[j]*y[j]), then + 0.0 as the
       * tuple-callback form writes it */
      for (int j = 0; j < s->n; ++j) f[j] = p[j] * y[j];
      for (int k = 0; k < s->n_off; ++k) {
        double c = s->off_pidx[k] >= 0 ? s->off_val[k] * p[s->off_pidx[k]] : s->off_val[k];
        f[s->off_row[k]] += c * y[s->off_col[k]];
      }
      for (int j = 0; j < s->n; ++j) f[j] = (f[j] + s->quad * (y[j] * y[j])) + 0.0;
      break;
    }
    case SYS_DAMPED_OSC:
      f[0] = y[1];
      f[1] = -p[0] * y[0] - p[1] * y[1];
      break;
    case SYS_BBN: {
      double complex yc[4], fc[4];
      for (int j = 0; j < 4; ++j) yc[j] = y[j];
      bbn_eval(yc, t, p, fc);
      for (int j = 0; j < 4; ++j) f[j] = creal(fc[j]);
      break;
    }
    case SYS_KAPS:
      for (int pr = 0; pr < s->n / 2; ++pr) {
        double ie = s->inv_eps[pr];
        double y1 = y[2 * pr], y2 = y[2 * pr + 1];
        f[2 * pr] = p[0] * (-(ie + 2.0) * y1 + ie * (y2 * y2));
        f[2 * pr + 1] = p[0] * ((y1 - y2) - y2 * y2);
      }
      break;
    default:
      break;
  }
}

/* Closed-form Jacobian, dense row-major J[r*n+c], and df/dt. */
static void sys_jac(const osys* s, const double* y, double t, const double* p, double* J,
                    double* dT) {
  const int n = s->n;
  (void)t;
  memset(J, 0, sizeof(double) * (size_t)n * (size_t)n);
  for (int i = 0; i < n; ++i) dT[i] = 0.0;
  switch (s->kind) {
    case SYS_CALLBACK:
      s->cb_jac(y, t, p, J, dT);
      break;
    case SYS_LORENZ: {
      const double sigma = 10.0, beta = 8.0 / 3.0;
      J[0] = -sigma; J[1] = sigma; J[2] = 0.0;
      J[3] = p[0] - y[2]; J[4] = -1.0; J[5] = -y[0];
      J[6] = y[1]; J[7] = y[0]; J[8] = -beta;
      break;
    }
    case SYS_ROBERTSON:
      J[0] = -p[0]; J[1] = p[1] * y[2]; J[2] = p[1] * y[1];
      J[3] = p[0]; J[4] = -p[1] * y[2] - 2.0 * p[2] * y[1]; J[5] = -p[1] * y[1];
      J[6] = 0.0; J[7] = 2.0 * p[2] * y[1]; J[8] = 0.0;
      break;
    case SYS_FORCED_DECAY:
      J[0] = -p[0];
      dT[0] = p[1];
      break;
    case SYS_DECAY:
      J[0] = -p[0];
      break;
    case SYS_CHAIN4:
      J[0 * 4 + 0] = -p[0];
      J[1 * 4 + 0] = p[0]; J[1 * 4 + 1] = -p[1];
      J[2 * 4 + 1] = p[1]; J[2 * 4 + 2] = -p[2];
      J[3 * 4 + 2] = p[2];
      break;
    case SYS_RING6:
      for (int i = 0; i < 6; ++i) {
        int nx = (i + 1) % 6, pv = (i + 5) % 6;
        J[i * 6 + i] = -2.0 * p[0] - p[1];
        J[i * 6 + nx] += p[0];
        J[i * 6 + pv] += p[0];
      }
      break;
    case SYS_VDP_LATTICE: {
      int nb[64];
      double strength = s->d / s->coupling_range;
      for (int osc = 0; osc < s->n_osc; ++osc) {
        int x = 2 * osc, v = x + 1;
        int cnt = lattice_neighbours(s, osc, nb);
        J[x * n + v] = 1.0;
        J[v * n + x] = p[0] * s->mu * (-2.0 * y[x]) * y[v] - s->omega * s->omega -
                       strength * (double)cnt;
        J[v * n + v] = p[0] * s->mu * (1.0 - y[x] * y[x]);
        for (int k = 0; k < cnt; ++k) J[v * n + 2 * nb[k]] += strength;
      }
      break;
    }
    case SYS_LINQUAD:
      for (int j = 0; j < n; ++j) J[j * n + j] = p[j] + s->quad * (2.0 * y[j]);
      for (int k = 0; k < s->n_off; ++k) {
        double c = s->off_pidx[k] >= 0 ? s->off_val[k] * p[s->off_pidx[k]] : s->off_val[k];
        J[s->off_row[k] * n + s->off_col[k]] += c;
      }
      break;
    case SYS_DAMPED_OSC:
      J[1] = 1.0;
      J[2] = -p[0]; J[3] = -p[1];
      break;
    case SYS_BBN:
      bbn_jac(y, t, p, J, dT);
      break;
    case SYS_KAPS:
      for (int pr = 0; pr < n / 2; ++pr) {
        double ie = s->inv_eps[pr];
        int a = 2 * pr, b = a + 1;
        J[a * n + a] = p[0] * (-(ie + 2.0));
        J[a * n + b] = p[0] * (ie * (2.0 * y[b]));
        J[b * n + a] = p[0];
        J[b * n + b] = p[0] * (-1.0 - 2.0 * y[b]);
      }
      break;
    default:
      break;
  }
}

/* Directional parameter derivative (df/dp_k) for the forward-sensitivity and
 * finite-difference checks: central differences are taken by the caller, this
 * is only the primal.  (Kept here so the ABI below exposes f itself.) */

/* ------------------------------------------------------------------------- */
/* Rodas5P constants: rodas5P.py:44-113                                       */
/* ------------------------------------------------------------------------- */
static const double GAMMA = 0.21193756319429014;
static const double A21 = 3.0;
static const double A31 = 2.849394379747939, A32 = 0.45842242204463923;
static const double A41 = -6.954028509809101, A42 = 2.489845061869568,
                    A43 = -10.358996098473584;
static const double A51 = 2.8029986275628964, A52 = 0.5072464736228206,
                    A53 = -0.3988312541770524, A54 = -0.04721187230404641;
static const double A61 = -7.502846399306121, A62 = 2.561846144803919,
                    A63 = -11.627539656261098, A64 = -0.18268767659942256,
                    A65 = 0.030198172008377946;
static const double C21 = -14.155112264123755;
static const double C31 = -17.97296035885952, C32 = -2.859693295451294;
static const double C41 = 147.12150275711716, C42 = -1.41221402718213,
                    C43 = 71.68940251302358;
static const double C51 = 165.43517024871676, C52 = -0.4592823456491126,
                    C53 = 42.90938336958603, C54 = -5.961986721573306;
static const double C61 = 24.854864614690072, C62 = -3.0009227002832186,
                    C63 = 47.4931110020768, C64 = 5.5814197821558125,
                    C65 = -0.6610691825249471;
static const double C71 = 30.91273214028599, C72 = -3.1208243349937974,
                    C73 = 77.79954646070892, C74 = 34.28646028294783,
                    C75 = -19.097331116725623, C76 = -28.087943162872662;
static const double C81 = 37.80277123390563, C82 = -3.2571969029072276,
                    C83 = 112.26918849496327, C84 = 66.9347231244047,
                    C85 = -40.06618937091002, C86 = -54.66780262877968,
                    C87 = -9.48861652309627;
static const double RC2 = 0.6358126895828704, RC3 = 0.4095798393397535,
                    RC4 = 0.9769306725060716, RC5 = 0.4288403609558664;
static const double D1 = 0.21193756319429014, D2 = -0.42387512638858027,
                    D3 = -0.3384627126235924, D4 = 1.8046452872882734,
                    D5 = 2.325825639765069;
static const double R_SAFETY = 0.9, R_FACTOR_MIN = 0.2, R_FACTOR_MAX = 6.0;
static const double R_EXPONENT = -1.0 / 6.0;

/* Dense-output coefficients, rodas5P.py:939-968 */
static const double H1[8] = {25.948786856663858, -2.5579724845846235, 10.433815404888879,
                             -2.3679251022685204, 0.524948541321073, 1.1241088310450404,
                             0.4272876194431874, -0.17202221070155493};
static const double H2[8] = {-9.91568850695171, -0.9689944594115154, 3.0438037242978453,
                             -24.495224566215796, 20.176138334709044, 15.98066361424651,
                             -6.789040303419874, -6.710236069923372};
static const double H3[8] = {11.419903575922262, 2.8879645146136994, 72.92137995996029,
                             80.12511834622643, -52.072871366152654, -59.78993625266729,
                             -0.15582684282751913, 4.883087185713722};

/* ------------------------------------------------------------------------- */
/* Linear algebra in both precisions                                           */
/* ------------------------------------------------------------------------- */
typedef struct {
  int n;
  int sparse;
  /* sparse CSR image of L+U in elimination order */
  int nnz;
  const int* row_ptr;
  const int* col_ind;
  const int* diag_ptr;
  const int* order;    /* caller index of each elimination row */
  const int* slot_of;  /* (n*n) caller (row,col) -> slot or -1 */
  const int* pattern;  /* (n*n) 1 where the caller's pattern has an entry */
  int* elim_slot;      /* (n*n) elimination (i,c) -> slot, built once */
} olin;

/* rodas5P.py:187-216 in f64 */
static void dense_factor_d(int n, double* lu, int* ipiv) {
  for (int i = 0; i < n; ++i) {
    double max_val = fabs(lu[i * n + i]);
    int pivot = i;
    for (int k = i + 1; k < n; ++k) {
      double v = fabs(lu[k * n + i]);
      if (v > max_val) { max_val = v; pivot = k; }
    }
    ipiv[i] = pivot;
    if (pivot != i)
      for (int j = 0; j < n; ++j) {
        double tmp = lu[i * n + j]; lu[i * n + j] = lu[pivot * n + j]; lu[pivot * n + j] = tmp;
      }
    double piv = lu[i * n + i];
    if (piv != 0.0) {
      double inv_piv = 1.0 / piv;
      for (int k = i + 1; k < n; ++k) {
        double factor = lu[k * n + i] * inv_piv;
        lu[k * n + i] = factor;
        for (int j = i + 1; j < n; ++j) lu[k * n + j] = fma(-factor, lu[i * n + j], lu[k * n + j]);
      }
    }
  }
}

/* rodas5P.py:187-216 with f32 storage: inv_piv and factor are f64 (numba typing
 * of `1.0 / piv`), the update is evaluated in f64 and stored f32. */
static void dense_factor_s(int n, float* lu, int* ipiv) {
  for (int i = 0; i < n; ++i) {
    float max_val = fabsf(lu[i * n + i]);
    int pivot = i;
    for (int k = i + 1; k < n; ++k) {
      float v = fabsf(lu[k * n + i]);
      if (v > max_val) { max_val = v; pivot = k; }
    }
    ipiv[i] = pivot;
    if (pivot != i)
      for (int j = 0; j < n; ++j) {
        float tmp = lu[i * n + j]; lu[i * n + j] = lu[pivot * n + j]; lu[pivot * n + j] = tmp;
      }
    float piv = lu[i * n + i];
    if (piv != 0.0f) {
      double inv_piv = 1.0 / (double)piv;
      for (int k = i + 1; k < n; ++k) {
        double factor = (double)lu[k * n + i] * inv_piv;
        lu[k * n + i] = (float)factor;
        for (int j = i + 1; j < n; ++j)
          lu[k * n + j] = (float)fma(-factor, (double)lu[i * n + j], (double)lu[k * n + j]);
      }
    }
  }
}

/* rodas5P.py:218-236 */
static void dense_solve_d(int n, const double* lu, const int* ipiv, double* rhs) {
  for (int i = 0; i < n; ++i) {
    int pivot = ipiv[i];
    if (pivot != i) { double tmp = rhs[i]; rhs[i] = rhs[pivot]; rhs[pivot] = tmp; }
    double acc = rhs[i];
    for (int j = 0; j < i; ++j) acc = fma(-lu[i * n + j], rhs[j], acc);
    rhs[i] = acc;
  }
  for (int i = n - 1; i >= 0; --i) {
    double acc = rhs[i];
    for (int j = i + 1; j < n; ++j) acc = fma(-lu[i * n + j], rhs[j], acc);
    rhs[i] = acc / lu[i * n + i];
  }
}

static void dense_solve_s(int n, const float* lu, const int* ipiv, float* rhs) {
  for (int i = 0; i < n; ++i) {
    int pivot = ipiv[i];
    if (pivot != i) { float tmp = rhs[i]; rhs[i] = rhs[pivot]; rhs[pivot] = tmp; }
    float acc = rhs[i];
    for (int j = 0; j < i; ++j) acc = fmaf(-lu[i * n + j], rhs[j], acc);
    rhs[i] = acc;
  }
  for (int i = n - 1; i >= 0; --i) {
    float acc = rhs[i];
    for (int j = i + 1; j < n; ++j) acc = fmaf(-lu[i * n + j], rhs[j], acc);
    rhs[i] = acc / lu[i * n + i];
  }
}

/* Up-looking unpivoted LU on the CSR image, _sparse_direct.py:472-546. The
 * destination of each update is looked up in a dense (row, column) table built
 * here from the CSR image itself -- not from the product's precomputed
 * destination lists -- so a mistake in those shows up as a mismatch. */
#define SPARSE_FACTOR(T, FMA, RECIP)                                                  \
  for (int i = 0; i < L->n; ++i) {                                               \
    int pivot = L->diag_ptr[i];                                                  \
    for (int p = L->row_ptr[i]; p < pivot; ++p) {                                \
      int k = L->col_ind[p];                                                     \
      int above = L->diag_ptr[k];                                                \
      T f = lu[p] * lu[above];                                                   \
      lu[p] = f;                                                                 \
      for (int s = above + 1; s < L->row_ptr[k + 1]; ++s) {                      \
        int dst = L->elim_slot[i * L->n + L->col_ind[s]];                        \
        lu[dst] = FMA(-f, lu[s], lu[dst]);                                       \
      }                                                                          \
    }                                                                            \
    lu[pivot] = RECIP;                                                           \
  }

static void sparse_factor_d(const olin* L, double* lu) { SPARSE_FACTOR(double, fma, 1.0 / lu[pivot]) }
static void sparse_factor_s(const olin* L, float* lu) {
  SPARSE_FACTOR(float, fmaf, (float)(1.0 / (double)lu[pivot]))
}

/* _sparse_direct.py:549-622 */
#define SPARSE_SOLVE(T, FMA)                                                          \
  for (int i = 0; i < L->n; ++i) {                                               \
    int row = L->order[i];                                                       \
    T acc = rhs[row];                                                            \
    for (int p = L->row_ptr[i]; p < L->diag_ptr[i]; ++p)                         \
      acc = FMA(-lu[p], rhs[L->order[L->col_ind[p]]], acc);                      \
    rhs[row] = acc;                                                              \
  }                                                                              \
  for (int i = L->n - 1; i >= 0; --i) {                                          \
    int row = L->order[i], pivot = L->diag_ptr[i];                               \
    T acc = rhs[row];                                                            \
    for (int p = pivot + 1; p < L->row_ptr[i + 1]; ++p)                          \
      acc = FMA(-lu[p], rhs[L->order[L->col_ind[p]]], acc);                      \
    rhs[row] = acc * lu[pivot];                                                  \
  }

static void sparse_solve_d(const olin* L, const double* lu, double* rhs) { SPARSE_SOLVE(double, fma) }
static void sparse_solve_s(const olin* L, const float* lu, float* rhs) { SPARSE_SOLVE(float, fmaf) }

/* ------------------------------------------------------------------------- */
/* Rodas5P, one trajectory (rodas5P.py:698-1005)                              */
/* ------------------------------------------------------------------------- */
typedef struct {
  int n, n_p, n_save;
  const double* times;
  double dt0, rtol, atol;
  int max_steps;
  int lu_fp32;
  double e1, e2, e3;
  int tf_index;
  /* forward sensitivities: directions, vector width n*(1+n_sens), error divisor,
   * and per direction the parameter it seeds (-1: an initial-state direction) */
  int n_sens, na, n_error;
  const int* pcol;
} orun;

typedef struct {
  double *y, *u, *work, *dT, *J;
  double *pd, *ky, *cpl; /* sensitivity scratch: n_p, n, n */
  double* kd;   /* 8*na, f64 storage */
  float* ks;    /* 8*na, f32 storage */
  double* lud;
  float* lus;
  double* rhsd;
  float* rhss;
  int* ipiv;
} owork;

/* lu_fp32 == 2: FP32 factorisation and solves, each stage solve refined once
 * against the double iteration matrix, stage vectors stored in double */
static int refined(const orun* R) { return R->lu_fp32 == 2; }

static double kget(const orun* R, const owork* W, int s, int j) {
  if (refined(R)) return W->kd[s * R->na + j];
  return R->lu_fp32 ? (double)W->ks[s * R->na + j] : W->kd[s * R->na + j];
}

/* unit parameter direction of sensitivity direction d (zeros for -1) */
static void direction_seed(const orun* R, int d, double* pd) {
  for (int m = 0; m < R->n_p; ++m) pd[m] = 0.0;
  if (R->pcol[d] >= 0) pd[R->pcol[d]] = 1.0;
}

/* the joint right-hand side: state rows, then J_y S_d + J_p e_d per direction */
static void joint_f(const orun* R, const osys* S, const double* z, double t, const double* p,
                    double* out, double* pd) {
  sys_f(S, z, t, p, out);
  const int n = R->n;
  for (int d = 0; d < R->n_sens; ++d) {
    direction_seed(R, d, pd);
    S->cb_sens_f(z, t, p, z + n + d * n, pd, out + n + d * n);
  }
}

static void assemble_and_factor(const orun* R, const olin* L, const osys* S, owork* W,
                                const double* p, double t, double dt_use) {
  const int n = R->n;
  double dtgamma_inv = 1.0 / (dt_use * GAMMA);
  sys_jac(S, W->y, t, p, W->J, W->dT);
  for (int d = 0; d < R->n_sens; ++d) {
    direction_seed(R, d, W->pd);
    S->cb_sens_dt(W->y, t, p, W->y + n + d * n, W->pd, W->dT + n + d * n);
  }
  if (!L->sparse) {
    if (R->lu_fp32) {
      for (int i = 0; i < n * n; ++i) W->lus[i] = (float)(-W->J[i]);
      for (int d = 0; d < n; ++d) W->lus[d * n + d] += (float)dtgamma_inv;
      dense_factor_s(n, W->lus, W->ipiv);
    } else {
      for (int i = 0; i < n * n; ++i) W->lud[i] = -W->J[i];
      for (int d = 0; d < n; ++d) W->lud[d * n + d] += dtgamma_inv;
      dense_factor_d(n, W->lud, W->ipiv);
    }
  } else {
    /* clear, then one slot per declared pattern entry (rodas5P.py:465-476,
     * 251-303); entries of J outside the pattern are not written. */
    if (R->lu_fp32) {
      for (int i = 0; i < L->nnz; ++i) W->lus[i] = 0.0f;
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c)
          if (L->pattern[r * n + c]) W->lus[L->slot_of[r * n + c]] = (float)(-W->J[r * n + c]);
      for (int d = 0; d < n; ++d) W->lus[L->slot_of[d * n + d]] += (float)dtgamma_inv;
      sparse_factor_s(L, W->lus);
    } else {
      for (int i = 0; i < L->nnz; ++i) W->lud[i] = 0.0;
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c)
          if (L->pattern[r * n + c]) W->lud[L->slot_of[r * n + c]] = -W->J[r * n + c];
      for (int d = 0; d < n; ++d) W->lud[L->slot_of[d * n + d]] += dtgamma_inv;
      sparse_factor_d(L, W->lud);
    }
  }
}

/* stage s: rhs_j = LU(work_j + dt*d*dT_j + (sum_k C_sk k_kj) * inv_dt), solve, store k_s.
 * With sensitivities the rows past n keep their value in work (rodas5P.py:567-583),
 * and after the state rows are solved each direction solves against the same
 * factorisation with the coupling L_d k_y added, evaluated at the step's base
 * point (rodas5P.py:512-561). */
static void stage(const orun* R, const olin* L, const osys* S, owork* W, int s,
                  const double* crow, double d, double dt_use, double inv_dt, const double* p,
                  double t) {
  const int n = R->n, na = R->na;
  for (int j = 0; j < na; ++j) {
    double value;
    if (s == 0) {
      value = fma(dt_use * d, W->dT[j], W->work[j]);
    } else {
      double acc = crow[0] * kget(R, W, 0, j);
      for (int k = 1; k < s; ++k) acc = fma(crow[k], kget(R, W, k, j), acc);
      if (d != 0.0)
        value = fma(acc, inv_dt, fma(dt_use * d, W->dT[j], W->work[j]));
      else
        value = fma(acc, inv_dt, W->work[j]);
    }
    if (j >= n) W->work[j] = value;
    else if (R->lu_fp32) { W->rhss[j] = (float)value; W->rhsd[j] = value; }
    else W->rhsd[j] = value;
  }
  for (int blk = 0; blk <= R->n_sens; ++blk) {
    const int base = blk * n;
    if (blk > 0) {
      const int dir = blk - 1;
      for (int j = 0; j < n; ++j) W->ky[j] = kget(R, W, s, j);
      direction_seed(R, dir, W->pd);
      S->cb_sens_cpl(W->y, t, p, W->y + base, W->pd, W->ky, W->cpl);
      for (int j = 0; j < n; ++j) {
        double v = W->work[base + j] + W->cpl[j];
        if (R->lu_fp32) W->rhss[j] = (float)v; else W->rhsd[j] = v;
      }
    }
    if (refined(R)) {
      /* x solves the rounded system in FP32; the residual of the stage's
       * double right-hand side b against the double matrix
       * M = I/(h gamma) - J(y_n, t_n), r = fma(-x, 1/(h gamma), b) + J x
       * (J the closed-form Jacobian the step evaluated, summed in column
       * order), is solved with the same FP32 factors; k = x + dx in double */
      const double gi = 1.0 / (dt_use * GAMMA);
      double* x = W->kd + s * na + base;
      if (L->sparse) sparse_solve_s(L, W->lus, W->rhss); else dense_solve_s(n, W->lus, W->ipiv, W->rhss);
      for (int j = 0; j < n; ++j) x[j] = (double)W->rhss[j];
      for (int j = 0; j < n; ++j) {
        double jx = 0.0;
        for (int c = 0; c < n; ++c) jx = jx + W->J[j * n + c] * x[c];
        W->rhss[j] = (float)(fma(-x[j], gi, W->rhsd[j]) + jx);
      }
      if (L->sparse) sparse_solve_s(L, W->lus, W->rhss); else dense_solve_s(n, W->lus, W->ipiv, W->rhss);
      for (int j = 0; j < n; ++j) x[j] = x[j] + (double)W->rhss[j];
    } else if (R->lu_fp32) {
      if (L->sparse) sparse_solve_s(L, W->lus, W->rhss); else dense_solve_s(n, W->lus, W->ipiv, W->rhss);
      for (int j = 0; j < n; ++j) W->ks[s * na + base + j] = W->rhss[j];
    } else {
      if (L->sparse) sparse_solve_d(L, W->lud, W->rhsd); else dense_solve_d(n, W->lud, W->ipiv, W->rhsd);
      for (int j = 0; j < n; ++j) W->kd[s * na + base + j] = W->rhsd[j];
    }
  }
}

static void advance(const orun* R, owork* W, const double* arow, int n_prev) {
  for (int j = 0; j < R->na; ++j) {
    double acc = arow[0] * kget(R, W, 0, j);
    for (int k = 1; k < n_prev; ++k) acc = fma(arow[k], kget(R, W, k, j), acc);
    W->u[j] = W->y[j] + acc;
  }
}

static void rodas5p_one(const orun* R, const olin* L, const osys* S, owork* W, const double* y0,
                        const double* p, const double* w, double* hist, int* acc_out,
                        int* rej_out, int* loop_out) {
  const int na = R->na, n_save = R->n_save;
  const double* times = R->times;
  const double tf = times[n_save - 1];
  const double dt_init = R->dt0 > 0.0 ? R->dt0 : (tf - times[0]) * 1e-6;
  static const double c1[1] = {C21};
  static const double c2[2] = {C31, C32};
  static const double c3[3] = {C41, C42, C43};
  static const double c4[4] = {C51, C52, C53, C54};
  static const double c5[5] = {C61, C62, C63, C64, C65};
  static const double c6[6] = {C71, C72, C73, C74, C75, C76};
  static const double c7[7] = {C81, C82, C83, C84, C85, C86, C87};
  static const double a0[1] = {A21};
  static const double a1[2] = {A31, A32};
  static const double a2[3] = {A41, A42, A43};
  static const double a3[4] = {A51, A52, A53, A54};
  static const double a4[5] = {A61, A62, A63, A64, A65};

  for (int j = 0; j < na; ++j) W->y[j] = y0[j];
  for (int j = 0; j < na; ++j) hist[j] = W->y[j];
  double t = times[0];
  int save_idx = 1, n_steps = 0, accepted = 0, rejected = 0;
  double dt = dt_init, err_prev = 1.0, err_prev2 = 1.0;
  double tf_local = tf;
  if (R->tf_index >= 0) tf_local = p[R->tf_index];

  while (save_idx < n_save && t < tf_local && n_steps < R->max_steps) {
    double dt_use = dt;
    if (dt_use > tf_local - t) dt_use = tf_local - t;
    if (dt_use < 1e-30) dt_use = 1e-30;
    double inv_dt = 1.0 / dt_use;
    double t_end = t + dt_use;
    assemble_and_factor(R, L, S, W, p, t, dt_use);

    joint_f(R, S, W->y, t, p, W->work, W->pd);
    stage(R, L, S, W, 0, NULL, D1, dt_use, inv_dt, p, t);
    advance(R, W, a0, 1);
    joint_f(R, S, W->u, t + RC2 * dt_use, p, W->work, W->pd);
    stage(R, L, S, W, 1, c1, D2, dt_use, inv_dt, p, t);
    advance(R, W, a1, 2);
    joint_f(R, S, W->u, t + RC3 * dt_use, p, W->work, W->pd);
    stage(R, L, S, W, 2, c2, D3, dt_use, inv_dt, p, t);
    advance(R, W, a2, 3);
    joint_f(R, S, W->u, t + RC4 * dt_use, p, W->work, W->pd);
    stage(R, L, S, W, 3, c3, D4, dt_use, inv_dt, p, t);
    advance(R, W, a3, 4);
    joint_f(R, S, W->u, t + RC5 * dt_use, p, W->work, W->pd);
    stage(R, L, S, W, 4, c4, D5, dt_use, inv_dt, p, t);
    advance(R, W, a4, 5);
    joint_f(R, S, W->u, t_end, p, W->work, W->pd);
    stage(R, L, S, W, 5, c5, 0.0, dt_use, inv_dt, p, t);
    for (int j = 0; j < na; ++j) W->u[j] += kget(R, W, 5, j);
    joint_f(R, S, W->u, t_end, p, W->work, W->pd);
    stage(R, L, S, W, 6, c6, 0.0, dt_use, inv_dt, p, t);
    for (int j = 0; j < na; ++j) W->u[j] += kget(R, W, 6, j);
    joint_f(R, S, W->u, t_end, p, W->work, W->pd);
    stage(R, L, S, W, 7, c7, 0.0, dt_use, inv_dt, p, t);

    double err_local = 0.0;
    for (int j = 0; j < na; ++j) {
      double k7 = kget(R, W, 7, j);
      double y_new_j = W->u[j] + k7;
      double a = fabs(W->y[j]), b = fabs(y_new_j);
      double scale = fma(R->rtol, a > b ? a : b, R->atol);
      double r = w[j] * k7 / scale;
      err_local = fma(r, r, err_local);
    }
    double err_norm = sqrt(err_local / R->n_error);
    int accept = err_norm <= 1.0 && !isnan(err_norm);
    double safe_err;
    if (isnan(err_norm) || err_norm > 1e18) safe_err = 1e18;
    else if (err_norm == 0.0) safe_err = 1e-18;
    else safe_err = err_norm;
    double factor = R_SAFETY * det_pow(safe_err, R->e1) * det_pow(err_prev, R->e2) *
                    det_pow(err_prev2, R->e3);
    if (accept) { err_prev2 = err_prev; err_prev = safe_err; }
    if (factor < R_FACTOR_MIN) factor = R_FACTOR_MIN;
    else if (factor > R_FACTOR_MAX) factor = R_FACTOR_MAX;
    dt = dt_use * factor;

    if (accept) {
      double t_old = t, t_new = t_old + dt_use;
      while (save_idx < n_save) {
        double ts = times[save_idx];
        double at = fabs(ts);
        if (!(ts <= t_new + 1e-12 * (1.0 > at ? 1.0 : at))) break;
        double theta = (ts - t_old) / dt_use;
        double theta1 = 1.0 - theta;
        for (int j = 0; j < na; ++j) {
          double h1 = H1[0] * kget(R, W, 0, j), h2 = H2[0] * kget(R, W, 0, j),
                 h3 = H3[0] * kget(R, W, 0, j);
          for (int s = 1; s < 8; ++s) {
            double kv = kget(R, W, s, j);
            h1 = fma(H1[s], kv, h1);
            h2 = fma(H2[s], kv, h2);
            h3 = fma(H3[s], kv, h3);
          }
          double y_new_j = W->u[j] + kget(R, W, 7, j);
          /* theta1*y + theta*(y_new + theta1*(h1 + theta*(h2 + theta*h3))) */
          double v = fma(theta, h3, h2);
          v = fma(theta, v, h1);
          v = fma(theta1, v, y_new_j);
          hist[(size_t)save_idx * na + j] = fma(theta, v, theta1 * W->y[j]);
        }
        save_idx++;
      }
      for (int j = 0; j < na; ++j) W->y[j] = W->u[j] + kget(R, W, 7, j);
      t += dt_use;
      accepted++;
    } else {
      rejected++;
    }
    n_steps++;
  }
  while (save_idx < n_save) {
    for (int j = 0; j < na; ++j) hist[(size_t)save_idx * na + j] = W->y[j];
    save_idx++;
  }
  *acc_out = accepted;
  *rej_out = rejected;
  *loop_out = n_steps;
}

static int build_elim_slot(olin* L) {
  int n = L->n;
  L->elim_slot = (int*)malloc(sizeof(int) * (size_t)n * (size_t)n);
  if (!L->elim_slot) return -1;
  for (int i = 0; i < n * n; ++i) L->elim_slot[i] = -1;
  for (int i = 0; i < n; ++i)
    for (int p = L->row_ptr[i]; p < L->row_ptr[i + 1]; ++p) L->elim_slot[i * n + L->col_ind[p]] = p;
  /* every update destination must exist in the structure */
  for (int i = 0; i < n; ++i)
    for (int p = L->row_ptr[i]; p < L->diag_ptr[i]; ++p) {
      int k = L->col_ind[p];
      for (int s = L->diag_ptr[k] + 1; s < L->row_ptr[k + 1]; ++s)
        if (L->elim_slot[i * n + L->col_ind[s]] < 0) return -2;
    }
  return 0;
}

static void owork_alloc_aug(owork* W, int n, int na, int n_p, int lu_n) {
  W->y = (double*)calloc((size_t)na, sizeof(double));
  W->u = (double*)calloc((size_t)na, sizeof(double));
  W->work = (double*)calloc((size_t)na, sizeof(double));
  W->dT = (double*)calloc((size_t)na, sizeof(double));
  W->J = (double*)calloc((size_t)n * (size_t)n, sizeof(double));
  W->pd = (double*)calloc((size_t)(n_p > 0 ? n_p : 1), sizeof(double));
  W->ky = (double*)calloc((size_t)n, sizeof(double));
  W->cpl = (double*)calloc((size_t)n, sizeof(double));
  W->kd = (double*)calloc(8 * (size_t)na, sizeof(double));
  W->ks = (float*)calloc(8 * (size_t)na, sizeof(float));
  W->lud = (double*)calloc((size_t)lu_n, sizeof(double));
  W->lus = (float*)calloc((size_t)lu_n, sizeof(float));
  W->rhsd = (double*)calloc((size_t)n, sizeof(double));
  W->rhss = (float*)calloc((size_t)n, sizeof(float));
  W->ipiv = (int*)calloc((size_t)n, sizeof(int));
}

static void owork_free(owork* W) {
  free(W->y); free(W->u); free(W->work); free(W->dT); free(W->J); free(W->kd); free(W->ks);
  free(W->pd); free(W->ky); free(W->cpl);
  free(W->lud); free(W->lus); free(W->rhsd); free(W->rhss); free(W->ipiv);
}

/* System description passed across the ABI: ints[] and dbls[] as documented in
 * oracle.py (kind, n, n_p, n_osc, wrap, coupling_range, n_off / mu, d, omega,
 * quad), plus the off-diagonal terms (rows, cols, parameter indices) and
 * the Kaps inverse epsilons. */
static void unpack_sys(osys* S, const int* si, const double* sd, const int* off_rc,
                       const double* off_val, const double* inv_eps) {
  memset(S, 0, sizeof(*S));
  S->kind = si[0]; S->n = si[1]; S->n_p = si[2];
  S->n_osc = si[3]; S->wrap = si[4]; S->coupling_range = si[5] > 0 ? si[5] : 1;
  S->n_off = si[6];
  S->mu = sd[0]; S->d = sd[1]; S->omega = sd[2]; S->quad = sd[3];
  S->off_row = off_rc;
  S->off_col = off_rc ? off_rc + S->n_off : NULL;
  S->off_pidx = off_rc ? off_rc + 2 * S->n_off : NULL;
  S->off_val = off_val;
  S->inv_eps = inv_eps;
}

static int run_rodas5p(const osys* S, const orun* R, int N, const double* y0,
                       const double* params, const double* weights, int nnz,
                       const int* row_ptr, const int* col_ind, const int* diag_ptr,
                       const int* order, const int* slot_of, const int* pattern, double* hist,
                       int* accepted, int* rejected, int* loop, int n_threads) {
  const int n = R->n, na = R->na, n_p = R->n_p, n_save = R->n_save;
  olin L;
  memset(&L, 0, sizeof(L));
  L.n = n;
  L.sparse = nnz > 0;
  if (L.sparse) {
    L.nnz = nnz; L.row_ptr = row_ptr; L.col_ind = col_ind; L.diag_ptr = diag_ptr;
    L.order = order; L.slot_of = slot_of; L.pattern = pattern;
    int rc = build_elim_slot(&L);
    if (rc != 0) { free(L.elim_slot); return rc; }
  }
  int lu_n = L.sparse ? nnz : n * n;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
  {
    owork W;
    owork_alloc_aug(&W, n, na, n_p, lu_n);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 8)
#endif
    for (int i = 0; i < N; ++i) {
      rodas5p_one(R, &L, S, &W, y0 + (size_t)i * na, params + (size_t)i * n_p,
                  weights + (size_t)i * na, hist + (size_t)i * n_save * na, accepted + i,
                  rejected + i, loop + i);
    }
    owork_free(&W);
  }
  free(L.elim_slot);
  return 0;
}

/* Ensemble Rodas5P. Arrays are trajectory-major (the reference API's layout):
 * y0 (N,n), params (N,n_p), weights (N,n), hist (N,n_save,n). Returns 0 on
 * success, <0 on a malformed sparse structure. */
int oracle_rodas5p(const int* sys_i, const double* sys_d, const int* off_rc,
                   const double* off_val, const double* inv_eps, int N, const double* y0,
                   const double* times, int n_save, const double* params,
                   const double* weights, double dt0, double rtol, double atol, int max_steps,
                   int lu_fp32, double pcoeff, double icoeff, double dcoeff, int tf_index,
                   /* sparse structure, or nnz <= 0 for dense */
                   int nnz, const int* row_ptr, const int* col_ind, const int* diag_ptr,
                   const int* order, const int* slot_of, const int* pattern, double* hist,
                   int* accepted, int* rejected, int* loop, int n_threads) {
  osys S;
  unpack_sys(&S, sys_i, sys_d, off_rc, off_val, inv_eps);
  const int n = S.n, n_p = S.n_p;
  orun R = {n, n_p, n_save, times, dt0, rtol, atol, max_steps, lu_fp32,
            R_EXPONENT * (icoeff + pcoeff + dcoeff), -R_EXPONENT * (pcoeff + 2.0 * dcoeff),
            R_EXPONENT * dcoeff, tf_index, 0, n, n, NULL};
  return run_rodas5p(&S, &R, N, y0, params, weights, nnz, row_ptr, col_ind, diag_ptr, order,
                     slot_of, pattern, hist, accepted, rejected, loop, n_threads);
}

/* Ensemble Rodas5P of a callback system, optionally carrying n_sens forward
 * sensitivity directions (the joint [y, S] system of modax/_sensitivity.py):
 * y0, weights and hist rows are n*(1+n_sens) wide, pcol[d] is the parameter
 * direction d seeds (-1 for an initial-state direction) and n_error the error
 * norm's divisor. The callbacks must be reentrant (they run on every thread). */
int oracle_rodas5p_cb(void* cb_f, void* cb_jac, void* cb_sens_f, void* cb_sens_dt,
                      void* cb_sens_cpl, int n, int n_p, int n_sens, const int* pcol,
                      int n_error, int N, const double* y0, const double* times, int n_save,
                      const double* params, const double* weights, double dt0, double rtol,
                      double atol, int max_steps, int lu_fp32, double pcoeff, double icoeff,
                      double dcoeff, int tf_index, int nnz, const int* row_ptr,
                      const int* col_ind, const int* diag_ptr, const int* order,
                      const int* slot_of, const int* pattern, double* hist, int* accepted,
                      int* rejected, int* loop, int n_threads) {
  osys S;
  memset(&S, 0, sizeof(S));
  S.kind = SYS_CALLBACK;
  S.n = n;
  S.n_p = n_p;
  S.cb_f = (cb_f_t)cb_f;
  S.cb_jac = (cb_jac_t)cb_jac;
  S.cb_sens_f = (cb_sens_t)cb_sens_f;
  S.cb_sens_dt = (cb_sens_t)cb_sens_dt;
  S.cb_sens_cpl = (cb_cpl_t)cb_sens_cpl;
  if (n_sens > 0 && (!cb_sens_f || !cb_sens_dt || !cb_sens_cpl)) return -10;
  if (n_sens > 0 && lu_fp32 == 2) return -11; /* refinement: plain solves only */
  orun R = {n, n_p, n_save, times, dt0, rtol, atol, max_steps, lu_fp32,
            R_EXPONENT * (icoeff + pcoeff + dcoeff), -R_EXPONENT * (pcoeff + 2.0 * dcoeff),
            R_EXPONENT * dcoeff, tf_index, n_sens, n * (1 + n_sens), n_error, pcol};
  return run_rodas5p(&S, &R, N, y0, params, weights, nnz, row_ptr, col_ind, diag_ptr, order,
                     slot_of, pattern, hist, accepted, rejected, loop, n_threads);
}

/* ------------------------------------------------------------------------- */
/* Tsit5 (tsit5.py:34-90, 218-404)                                            */
/* ------------------------------------------------------------------------- */
static const double T_C2 = 161.0 / 1000.0, T_C3 = 327.0 / 1000.0, T_C4 = 9.0 / 10.0,
                    T_C5 = 0.9800255409045097, T_C6 = 1.0, T_C7 = 1.0;
static const double T_A21 = 161.0 / 1000.0;
static const double T_A31 = -0.008480655492356989, T_A32 = 0.335480655492357;
static const double T_A41 = 2.8971530571054935, T_A42 = -6.359448489975075,
                    T_A43 = 4.3622954328695815;
static const double T_A51 = 5.325864828439257, T_A52 = -11.748883564062828,
                    T_A53 = 7.4955393428898365, T_A54 = -0.09249506636175525;
static const double T_A61 = 5.86145544294642, T_A62 = -12.92096931784711,
                    T_A63 = 8.159367898576159, T_A64 = -0.071584973281401,
                    T_A65 = -0.028269050394068383;
static const double T_B1 = 0.09646076681806523, T_B2 = 0.01, T_B3 = 0.4798896504144996,
                    T_B4 = 1.379008574103742, T_B5 = -3.2900695154360807,
                    T_B6 = 2.324710524099774;
static const double T_E1 = 0.0017800110522257773, T_E2 = 0.0008164344596567463,
                    T_E3 = -0.007880878010261994, T_E4 = 0.1447110071732629,
                    T_E5 = -0.5823571654525552, T_E6 = 0.45808210592918686,
                    T_E7 = -1.0 / 66.0;
static const double T_SAFETY = 0.9, T_FACTOR_MIN = 0.2, T_FACTOR_MAX = 10.0;
static const double T_EXPONENT = -1.0 / 5.0;

static void tsit5_one(const osys* S, int n_save, const double* times, double dt0, double rtol,
                      double atol, int max_steps, double e1, double e2, double e3,
                      const double* y0, const double* p, const double* w, double* hist,
                      int* acc_out, int* rej_out, int* loop_out, double* buf) {
  const int n = S->n;
  double *y = buf, *u = buf + n, *k1 = buf + 2 * n, *k2 = buf + 3 * n, *k3 = buf + 4 * n,
         *k4 = buf + 5 * n, *k5 = buf + 6 * n, *k6 = buf + 7 * n, *k7 = buf + 8 * n;
  for (int j = 0; j < n; ++j) { y[j] = y0[j]; hist[j] = y0[j]; k7[j] = 0.0; }
  double t = times[0];
  const double tf = times[n_save - 1];
  double dt = dt0 > 0.0 ? dt0 : (tf - t) * 1e-6;
  int save_idx = 1, n_steps = 0, accepted = 0, rejected = 0, has_fsal = 0;
  double err_prev = 1.0, err_prev2 = 1.0;
  while (save_idx < n_save && t < tf && n_steps < max_steps) {
    double dt_use = dt;
    if (dt_use > tf - t) dt_use = tf - t;
    if (dt_use < 1e-30) dt_use = 1e-30;
    if (has_fsal) for (int j = 0; j < n; ++j) k1[j] = k7[j];
    else sys_f(S, y, t, p, k1);
    /* u = y + dt * (sum_k a_k k_k): the sum accumulates with fma, then one
     * fma adds dt times it to y (tsit5.py:251-284). */
    for (int j = 0; j < n; ++j) u[j] = fma(dt_use, T_A21 * k1[j], y[j]);
    sys_f(S, u, t + T_C2 * dt_use, p, k2);
    for (int j = 0; j < n; ++j) u[j] = fma(dt_use, fma(T_A32, k2[j], T_A31 * k1[j]), y[j]);
    sys_f(S, u, t + T_C3 * dt_use, p, k3);
    for (int j = 0; j < n; ++j) {
      double a = T_A41 * k1[j];
      a = fma(T_A42, k2[j], a);
      a = fma(T_A43, k3[j], a);
      u[j] = fma(dt_use, a, y[j]);
    }
    sys_f(S, u, t + T_C4 * dt_use, p, k4);
    for (int j = 0; j < n; ++j) {
      double a = T_A51 * k1[j];
      a = fma(T_A52, k2[j], a);
      a = fma(T_A53, k3[j], a);
      a = fma(T_A54, k4[j], a);
      u[j] = fma(dt_use, a, y[j]);
    }
    sys_f(S, u, t + T_C5 * dt_use, p, k5);
    for (int j = 0; j < n; ++j) {
      double a = T_A61 * k1[j];
      a = fma(T_A62, k2[j], a);
      a = fma(T_A63, k3[j], a);
      a = fma(T_A64, k4[j], a);
      a = fma(T_A65, k5[j], a);
      u[j] = fma(dt_use, a, y[j]);
    }
    sys_f(S, u, t + T_C6 * dt_use, p, k6);
    for (int j = 0; j < n; ++j) {
      double a = T_B1 * k1[j];
      a = fma(T_B2, k2[j], a);
      a = fma(T_B3, k3[j], a);
      a = fma(T_B4, k4[j], a);
      a = fma(T_B5, k5[j], a);
      a = fma(T_B6, k6[j], a);
      u[j] = fma(dt_use, a, y[j]);
    }
    sys_f(S, u, t + T_C7 * dt_use, p, k7);
    double err_sum = 0.0;
    for (int j = 0; j < n; ++j) {
      double e = T_E1 * k1[j];
      e = fma(T_E2, k2[j], e);
      e = fma(T_E3, k3[j], e);
      e = fma(T_E4, k4[j], e);
      e = fma(T_E5, k5[j], e);
      e = fma(T_E6, k6[j], e);
      e = fma(T_E7, k7[j], e);
      double err_est = dt_use * e;
      double a = fabs(y[j]), b = fabs(u[j]);
      double scale = fma(rtol, a > b ? a : b, atol);
      double r = w[j] * err_est / scale;
      err_sum = fma(r, r, err_sum);
    }
    double err_norm = sqrt(err_sum / n);
    int accept = err_norm <= 1.0 && !isnan(err_norm);
    double t_new = t;
    if (accept) {
      t_new = t + dt_use;
      while (save_idx < n_save) {
        double ts = times[save_idx];
        double at = fabs(ts);
        if (!(ts <= t_new + 1e-12 * (1.0 > at ? 1.0 : at))) break;
        double th = (ts - t) / dt_use;
        double b1 = -1.0530884977290216 * th * (th - 1.3299890189751412) *
                    (th * th - 1.4364028541716351 * th + 0.7139816917074209);
        double b2 = 0.1017 * th * th * (th * th - 2.1966568338249754 * th + 1.2949852507374631);
        double b3 = 2.490627285651252793 * th * th *
                    (th * th - 2.38535645472061657 * th + 1.57803468208092486);
        double b4 = -16.54810288924490272 * (th - 1.21712927295533244) *
                    (th - 0.61620406037800089) * th * th;
        double b5 = 47.37952196281928122 * (th - 1.203071208372362603) *
                    (th - 0.658047292653547382) * th * th;
        double b6 = -34.87065786149660974 * (th - 1.2) * (th - 0.666666666666666667) * th * th;
        double b7 = 2.5 * (th - 1.0) * (th - 0.6) * th * th;
        for (int j = 0; j < n; ++j) {
          double a = b1 * k1[j];
          a = fma(b2, k2[j], a);
          a = fma(b3, k3[j], a);
          a = fma(b4, k4[j], a);
          a = fma(b5, k5[j], a);
          a = fma(b6, k6[j], a);
          a = fma(b7, k7[j], a);
          hist[(size_t)save_idx * n + j] = fma(dt_use, a, y[j]);
        }
        save_idx++;
      }
      for (int j = 0; j < n; ++j) y[j] = u[j];
      accepted++;
      has_fsal = 1;
    } else {
      rejected++;
      for (int j = 0; j < n; ++j) k7[j] = 0.0;
      has_fsal = 0;
    }
    double safe_err;
    if (isnan(err_norm) || err_norm > 1e18) safe_err = 1e18;
    else if (err_norm == 0.0) safe_err = 1e-18;
    else safe_err = err_norm;
    double factor = T_SAFETY * det_pow(safe_err, e1) * det_pow(err_prev, e2) * det_pow(err_prev2, e3);
    if (accept) { err_prev2 = err_prev; err_prev = safe_err; }
    if (factor < T_FACTOR_MIN) factor = T_FACTOR_MIN;
    else if (factor > T_FACTOR_MAX) factor = T_FACTOR_MAX;
    dt = dt_use * factor;
    t = t_new;
    n_steps++;
  }
  *acc_out = accepted;
  *rej_out = rejected;
  *loop_out = n_steps;
}

/* Ensemble Tsit5, trajectory-major arrays as oracle_rodas5p. Saves the kernel
 * never reaches are left as the caller initialised them (tsit5.py:402-404). */
int oracle_tsit5(const int* sys_i, const double* sys_d, const int* off_rc, const double* off_val,
                 const double* inv_eps, int N, const double* y0, const double* times, int n_save,
                 const double* params, const double* weights, double dt0, double rtol,
                 double atol, int max_steps, double pcoeff, double icoeff, double dcoeff,
                 double* hist, int* accepted, int* rejected, int* loop, int n_threads) {
  osys S;
  unpack_sys(&S, sys_i, sys_d, off_rc, off_val, inv_eps);
  const int n = S.n, n_p = S.n_p;
  double e1 = T_EXPONENT * (icoeff + pcoeff + dcoeff);
  double e2 = -T_EXPONENT * (pcoeff + 2.0 * dcoeff);
  double e3 = T_EXPONENT * dcoeff;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
  {
    double* buf = (double*)calloc(9 * (size_t)n, sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int i = 0; i < N; ++i)
      tsit5_one(&S, n_save, times, dt0, rtol, atol, max_steps, e1, e2, e3, y0 + (size_t)i * n,
                params + (size_t)i * n_p, weights + (size_t)i * n,
                hist + (size_t)i * n_save * n, accepted + i, rejected + i, loop + i, buf);
    free(buf);
  }
  return 0;
}

/* The primal right-hand side and closed-form Jacobian, exposed so tests can
 * hold the product's generated derivatives against them. */
int oracle_rhs(const int* sys_i, const double* sys_d, const int* off_rc, const double* off_val,
               const double* inv_eps, const double* y, double t, const double* p, double* f) {
  osys S;
  unpack_sys(&S, sys_i, sys_d, off_rc, off_val, inv_eps);
  sys_f(&S, y, t, p, f);
  return 0;
}

int oracle_jacobian(const int* sys_i, const double* sys_d, const int* off_rc,
                    const double* off_val, const double* inv_eps, const double* y, double t,
                    const double* p, double* J, double* dT) {
  osys S;
  unpack_sys(&S, sys_i, sys_d, off_rc, off_val, inv_eps);
  sys_jac(&S, y, t, p, J, dT);
  return 0;
}

/* Closed-form forward-sensitivity callbacks, derived by hand, for two systems
 * the joint-solve tests use: Robertson (modax's stiff kinetics, three rates)
 * and one Van der Pol oscillator y' = (v, mu p0 (1 - x^2) v - x) (vdp.py:
 * 253-290 with one oscillator, omega 1, no coupling). They have the signatures
 * of the callback solver's hooks (cb_f_t, cb_jac_t, cb_sens_t, cb_cpl_t), so
 * the oracle's joint solve can run on them instead of on the derivative graphs
 * the product's tracer emits -- an independent check of those graphs' second
 * derivatives through a whole integration.
 *
 * Robertson: f = (-a y0 + b y1 y2, a y0 - b y1 y2 - c y1^2, c y1^2), p = (a, b, c)
 *   J_y = [[-a, b y2, b y1], [a, -b y2 - 2c y1, -b y1], [0, 2c y1, 0]]
 *   J_p = [[-y0, y1 y2, 0], [y0, -y1 y2, -y1^2], [0, 0, y1^2]]
 *   D_y(J_y s)[v] = (b (v2 s1 + v1 s2), -b (v2 s1 + v1 s2) - 2c v1 s1, 2c v1 s1)
 *   D_y(J_p q)[v] = (-v0 q0 + (v1 y2 + y1 v2) q1,
 *                    v0 q0 - (v1 y2 + y1 v2) q1 - 2 y1 v1 q2, 2 y1 v1 q2) */
void oracle_cf_robertson_f(const double* y, double t, const double* p, double* out) {
  (void)t;
  out[0] = -p[0] * y[0] + p[1] * y[1] * y[2];
  out[1] = p[0] * y[0] - p[1] * y[1] * y[2] - p[2] * y[1] * y[1];
  out[2] = p[2] * y[1] * y[1];
}

void oracle_cf_robertson_jac(const double* y, double t, const double* p, double* J, double* dT) {
  (void)t;
  const double a = p[0], b = p[1], c = p[2];
  J[0] = -a, J[1] = b * y[2], J[2] = b * y[1];
  J[3] = a, J[4] = -b * y[2] - 2.0 * c * y[1], J[5] = -b * y[1];
  J[6] = 0.0, J[7] = 2.0 * c * y[1], J[8] = 0.0;
  dT[0] = dT[1] = dT[2] = 0.0;
}

void oracle_cf_robertson_sens_f(const double* y, double t, const double* p, const double* s,
                                const double* q, double* out) {
  double J[9], dT[3];
  oracle_cf_robertson_jac(y, t, p, J, dT);
  const double y12 = y[1] * y[2], y11 = y[1] * y[1];
  out[0] = J[0] * s[0] + J[1] * s[1] + J[2] * s[2] - y[0] * q[0] + y12 * q[1];
  out[1] = J[3] * s[0] + J[4] * s[1] + J[5] * s[2] + y[0] * q[0] - y12 * q[1] - y11 * q[2];
  out[2] = J[7] * s[1] + y11 * q[2];
}

void oracle_cf_zero_sens_dt(const double* y, double t, const double* p, const double* s,
                            const double* q, double* out) {
  (void)y, (void)t, (void)p, (void)s, (void)q;
  out[0] = out[1] = out[2] = 0.0;
}

void oracle_cf_robertson_sens_cpl(const double* y, double t, const double* p, const double* s,
                                  const double* q, const double* v, double* out) {
  (void)t;
  const double b = p[1], c = p[2];
  const double bs = b * (v[2] * s[1] + v[1] * s[2]);
  const double dy12 = v[1] * y[2] + y[1] * v[2];
  out[0] = bs - v[0] * q[0] + dy12 * q[1];
  out[1] = -bs - 2.0 * c * v[1] * s[1] + v[0] * q[0] - dy12 * q[1] - 2.0 * y[1] * v[1] * q[2];
  out[2] = 2.0 * c * v[1] * s[1] + 2.0 * y[1] * v[1] * q[2];
}

/* Van der Pol, y = (x, v), p = (m): f = (v, mu m (1 - x^2) v - x)
 *   J_y = [[0, 1], [-2 mu m x v - 1, mu m (1 - x^2)]],  J_p = (0, mu (1 - x^2) v)
 *   D_y(J_y s)[w] = (0, -2 mu m (w_x v + x w_v) s_x - 2 mu m x w_x s_v)
 *   D_y(J_p q)[w] = (0, mu (-2 x w_x v + (1 - x^2) w_v) q) */
static const double CF_VDP_MU = 1.0;

void oracle_cf_vdp_f(const double* y, double t, const double* p, double* out) {
  (void)t;
  const double x = y[0], v = y[1], m = CF_VDP_MU * p[0];
  out[0] = v;
  out[1] = m * (1.0 - x * x) * v - x;
}

void oracle_cf_vdp_jac(const double* y, double t, const double* p, double* J, double* dT) {
  (void)t;
  const double x = y[0], v = y[1], m = CF_VDP_MU * p[0];
  J[0] = 0.0, J[1] = 1.0;
  J[2] = -2.0 * m * x * v - 1.0, J[3] = m * (1.0 - x * x);
  dT[0] = dT[1] = 0.0;
}

void oracle_cf_vdp_sens_f(const double* y, double t, const double* p, const double* s,
                          const double* q, double* out) {
  (void)t;
  const double x = y[0], v = y[1], m = CF_VDP_MU * p[0];
  out[0] = s[1];
  out[1] = (-2.0 * m * x * v - 1.0) * s[0] + m * (1.0 - x * x) * s[1]
           + CF_VDP_MU * (1.0 - x * x) * v * q[0];
}

void oracle_cf_vdp_sens_dt(const double* y, double t, const double* p, const double* s,
                           const double* q, double* out) {
  (void)y, (void)t, (void)p, (void)s, (void)q;
  out[0] = out[1] = 0.0;
}

void oracle_cf_vdp_sens_cpl(const double* y, double t, const double* p, const double* s,
                            const double* q, const double* w, double* out) {
  (void)t;
  const double x = y[0], v = y[1], m = CF_VDP_MU * p[0];
  out[0] = 0.0;
  out[1] = -2.0 * m * (w[0] * v + x * w[1]) * s[0] - 2.0 * m * x * w[0] * s[1]
           + CF_VDP_MU * (-2.0 * x * w[0] * v + (1.0 - x * x) * w[1]) * q[0];
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
