// rodas5p.hpp -- Rodas5P ensemble kernels for gfx950, one trajectory per lane.
//
// Included by the source generated for one right-hand side, after it has
// defined namespace sys { rhs, neg_jac, [sparse_factor, sparse_solve, DIAG] }
// and these macros:
//   MDX_N, MDX_NP        state and parameter widths
//   MDX_NLU              elements of the iteration-matrix buffer
//   MDX_LU_FP32          1: iteration matrix, its LU and the stage vectors in
//                        float; 0: double
//   MDX_SPARSE           1: pattern-compiled sparse LU from sys; 0: dense LU
//   MDX_NEEDS_CLEAR      1 when some buffer slot is written by no sweep
//   MDX_N_ERROR          components the error norm divides by
//   MDX_TF_INDEX         parameter column holding the end time, or -1
//   MDX_BLOCK            lanes per workgroup
//   MDX_PARAMS_IN_REGS   keep the parameter row in registers
//   MDX_E1/E2/E3         PID exponents
//   MDX_ROLLED           1: the stage loop is kept as a loop, the stage
//                        vectors and state live in memory (large n); 0: the
//                        step is unrolled into registers
//   MDX_LDS_K, MDX_LDS_Y, MDX_LDS_U   with MDX_ROLLED: the stage vectors, the
//                        state, the stage argument live in LDS, laid out
//                        [element][lane]; otherwise k and y are in scratch
//                        and u in registers
//   MDX_DT_ZERO          1 when df/dt is identically zero
//   MDX_NSENS            forward-sensitivity directions carried with the state
//                        (0: a plain solve); the vectors are then
//                        MDX_N * (1 + MDX_NSENS) long, the matrix stays n x n
//   MDX_SENS_PCOLS       per direction, the parameter it seeds (-1: a
//                        direction of the initial state)
//   MDX_HOOK, MDX_HOOK_SIZE  call sys::save_hook at every save with the saved
//                        state, accumulating into the trajectory's hook_out row
//   MDX_SAVE_HISTORY     0: write no history, only the final state in row 0
//   MDX_REFINE           1 (with MDX_LU_FP32): each stage solve takes one step
//                        of iterative refinement against the double iteration
//                        matrix (sys::jvp forms J x) and the stage vectors are
//                        stored in double
//
// The method, stage by stage, is the reference kernel's (modax/rodas5P.py:
// tableau 44-105, controller 108-113, 340-342, loop 761-989, tail 991-1005);
// the precision of each value follows the reference's numba typing (stage
// right-hand sides accumulated in double and rounded to the LU type, stage
// vectors stored in the LU type, state/error/controller in double).
#pragma once
#include "mdx_common.hpp"

namespace mdx {
namespace r5 {

#if MDX_LU_FP32
typedef float LUT;
#else
typedef double LUT;
#endif
#ifndef MDX_REFINE
#define MDX_REFINE 0
#endif
// The stage vectors' storage: the LU type (the reference's typing), or double
// when the FP32 solves are refined.
#if MDX_REFINE
typedef double KT;
#else
typedef LUT KT;
#endif

constexpr int N = MDX_N;
constexpr double GAMMA = 0.21193756319429014;

// W-transformed Rodas5P coefficients (Steinebach 2023), rows padded.
__device__ constexpr double CM[8][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {-14.155112264123755, 0, 0, 0, 0, 0, 0},
    {-17.97296035885952, -2.859693295451294, 0, 0, 0, 0, 0},
    {147.12150275711716, -1.41221402718213, 71.68940251302358, 0, 0, 0, 0},
    {165.43517024871676, -0.4592823456491126, 42.90938336958603, -5.961986721573306, 0, 0, 0},
    {24.854864614690072, -3.0009227002832186, 47.4931110020768, 5.5814197821558125,
     -0.6610691825249471, 0, 0},
    {30.91273214028599, -3.1208243349937974, 77.79954646070892, 34.28646028294783,
     -19.097331116725623, -28.087943162872662, 0},
    {37.80277123390563, -3.2571969029072276, 112.26918849496327, 66.9347231244047,
     -40.06618937091002, -54.66780262877968, -9.48861652309627},
};
__device__ constexpr double DV[8] = {0.21193756319429014, -0.42387512638858027,
                                     -0.3384627126235924, 1.8046452872882734,
                                     2.325825639765069,   0.0, 0.0, 0.0};
__device__ constexpr double AM[5][5] = {
    {3.0, 0, 0, 0, 0},
    {2.849394379747939, 0.45842242204463923, 0, 0, 0},
    {-6.954028509809101, 2.489845061869568, -10.358996098473584, 0, 0},
    {2.8029986275628964, 0.5072464736228206, -0.3988312541770524, -0.04721187230404641, 0},
    {-7.502846399306121, 2.561846144803919, -11.627539656261098, -0.18268767659942256,
     0.030198172008377946},
};
// Stage abscissae c2..c5 (c6 = c7 = c8 = 1 use t_end).
__device__ constexpr double CT[4] = {0.6358126895828704, 0.4095798393397535,
                                     0.9769306725060716, 0.4288403609558664};
// Dense-output coefficients (rodas5P.py:939-968).
__device__ constexpr double H1[8] = {25.948786856663858, -2.5579724845846235, 10.433815404888879,
                                     -2.3679251022685204, 0.524948541321073, 1.1241088310450404,
                                     0.4272876194431874, -0.17202221070155493};
__device__ constexpr double H2[8] = {-9.91568850695171, -0.9689944594115154, 3.0438037242978453,
                                     -24.495224566215796, 20.176138334709044, 15.98066361424651,
                                     -6.789040303419874, -6.710236069923372};
__device__ constexpr double H3[8] = {11.419903575922262, 2.8879645146136994, 72.92137995996029,
                                     80.12511834622643, -52.072871366152654, -59.78993625266729,
                                     -0.15582684282751913, 4.883087185713722};

#ifndef MDX_ROLLED
#define MDX_ROLLED 0
#endif
// Hold the tableau in vector registers across the step loop (unrolled step).
// A double that is not an inline constant costs two scalar moves at every
// use, and when a SIMD runs a single wave every instruction takes an issue
// slot; loaded once into registers, the step issues none of them.
#ifndef MDX_TAB_REGS
#define MDX_TAB_REGS 0
#endif
#ifndef MDX_ADJOINT
#define MDX_ADJOINT 0
#endif
#ifndef MDX_HOOK
#define MDX_HOOK 0
#endif
#ifndef MDX_SAVE_HISTORY
#define MDX_SAVE_HISTORY 1
#endif
#if MDX_HOOK && MDX_ADJOINT
#error "a save hook is not differentiated"
#endif
// The rolled kernel copies the parameter row into per-lane scratch: each read
// is then one load at a constant offset, with no address kept in registers.
#if MDX_PARAMS_IN_REGS || MDX_ROLLED
typedef RegParams<MDX_NP> PA;
#else
typedef GlobalParams PA;
#endif
#ifndef MDX_LDS_K
#define MDX_LDS_K 0
#endif
#ifndef MDX_LDS_Y
#define MDX_LDS_Y 0
#endif
#ifndef MDX_LDS_U
#define MDX_LDS_U 0
#endif
#ifndef MDX_DT_ZERO
#define MDX_DT_ZERO 0
#endif
// Elements per chunk where a loop over the per-lane vectors loads a chunk
// before computing with it.
#ifndef MDX_CHUNK
#define MDX_CHUNK 16
#endif
// Columns per chunk in the scratch-resident dense factorisation's row update.
#ifndef MDX_LU_CHUNK
#define MDX_LU_CHUNK MDX_CHUNK
#endif
#ifndef MDX_LDS_LU
#define MDX_LDS_LU 0
#endif
#if (MDX_LDS_K || MDX_LDS_Y || MDX_LDS_U || MDX_LDS_LU) && !MDX_ROLLED
#error "LDS placement needs the rolled stage loop"
#endif
#if MDX_LDS_LU && (MDX_SPARSE || MDX_ADJOINT)
#error "the LDS iteration matrix is for the dense forward step"
#endif
#ifndef MDX_NSENS
#define MDX_NSENS 0
#endif
#ifndef MDX_SENS_UNROLL
#define MDX_SENS_UNROLL 1
#endif
// Length of every per-trajectory vector: the state, then one block of N per
// sensitivity direction (modax/_sensitivity.py:24-40).
constexpr int NS = MDX_NSENS;
constexpr int NA = N * (1 + NS);

#if MDX_NSENS > 0
// The parameter each direction seeds, -1 for an initial-state direction.
__device__ constexpr int PCOL[NS] = {MDX_SENS_PCOLS};
// A unit parameter direction (or none) as the generated code reads it.
struct OneHot {
  int col;
  __device__ __forceinline__ double operator[](int m) const { return m == col ? 1.0 : 0.0; }
};
// Elements off..off+N of a vector, through whatever accessor holds it.
template <class V>
struct Shifted {
  const V& v;
  int off;
  __device__ __forceinline__ double operator[](int j) const { return v[off + j]; }
};
#endif

// Where the stage vectors and the state live. LDS arrays are [element][lane]:
// a wave's access to one element is 64 consecutive words, conflict-free. The
// lane index is re-read through an opaque move at the top of each stage, so
// every LDS address is that one register plus an immediate offset; otherwise
// the compiler hoists hundreds of per-element addresses out of the step loop
// and spills them.
// One trajectory per workgroup (MDX_COOP, the "wave" placement): every lane
// runs the step on the same values, all per-trajectory arrays live once in
// LDS, and the lanes share out the dense factorisation and the forward
// substitution (coop_factor, coop_solve). Otherwise each lane is a trajectory
// and the LDS arrays interleave the lanes' copies.
#ifndef MDX_COOP
#define MDX_COOP 0
#endif
#if MDX_COOP && !(MDX_ROLLED && MDX_LDS_K && MDX_LDS_U && MDX_LDS_Y && MDX_LDS_LU)
#error "a wave-per-trajectory kernel keeps every vector and the matrix in LDS"
#endif
#if MDX_COOP && (MDX_HOOK || MDX_ADJOINT || MDX_BLOCK != 64)
#error "a wave-per-trajectory kernel is a forward solve (plain or joint) on one 64-lane wave"
#endif
#if MDX_REFINE && !(MDX_LU_FP32 && MDX_NSENS == 0 && !MDX_COOP && !MDX_ADJOINT)
#error "refined stage solves are for a plain forward solve with an FP32 LU on one lane"
#endif
#if MDX_COOP
#define MDX_LSTRIDE 1
// row stride of the shared matrix: one padding element keeps a column's
// elements in different banks
#define MDX_LD (N + 1)
#else
#define MDX_LSTRIDE MDX_BLOCK
#endif
#if MDX_COOP
#define MDX_LANE() constexpr unsigned lane_ = 0
#elif MDX_LDS_K || MDX_LDS_Y || MDX_LDS_U || MDX_LDS_LU
__device__ __forceinline__ unsigned opaque_lane() {
  unsigned v = threadIdx.x;
  asm volatile("; lane" : "+v"(v));
  return v;
}
#define MDX_LANE() const unsigned lane_ = opaque_lane()
#else
#define MDX_LANE() (void)0
#endif
#if MDX_LDS_K
__shared__ KT lds_k[8 * NA * MDX_LSTRIDE];
#define KQ(q, j) lds_k[((q) * NA + (j)) * MDX_LSTRIDE + lane_]
#else
#define KQ(q, j) k[q][j]
#endif
#if MDX_REFINE && MDX_LDS_K
// a refined stage's right-hand side in double, [element][lane]
__shared__ double lds_b[N * MDX_LSTRIDE];
#endif
#if MDX_LDS_Y
__shared__ double lds_y[NA * MDX_LSTRIDE];
#define YJ(st, j) lds_y[(j) * MDX_LSTRIDE + lane_]
#else
#define YJ(st, j) (st).y[j]
#endif
#if MDX_LDS_U
__shared__ double lds_u[NA * MDX_LSTRIDE];
#define UJ(j) lds_u[(j) * MDX_LSTRIDE + lane_]
#else
#define UJ(j) u[j]
#endif
#if MDX_LDS_LU && MDX_COOP
// The trajectory's dense iteration matrix, row-major with padded rows;
// element e = row * N + col as the generated code numbers it.
__shared__ LUT lds_lu[N * MDX_LD];
__shared__ LUT lds_r[N];
__shared__ int lds_piv[N];
__shared__ int lds_perm[N];
__shared__ double lds_rd[N];
struct LdsMat {
  typedef LUT value_type;
  LUT* base;
  __device__ __forceinline__ LUT& operator[](int e) const {
    return base[(e / N) * MDX_LD + e % N];
  }
};
#elif MDX_LDS_LU
// The dense iteration matrix in LDS, element e of this lane at [e][lane]: the
// factorisation's loads then take LDS latency instead of scratch latency.
__shared__ LUT lds_lu[MDX_NLU * MDX_BLOCK];
struct LdsMat {
  typedef LUT value_type;
  LUT* base;
  __device__ __forceinline__ LUT& operator[](int e) const { return base[e * MDX_BLOCK]; }
};
#endif
// A vector in LDS as the generated right-hand side reads it.
struct LdsVec {
  const double* base;
  __device__ __forceinline__ double operator[](int j) const { return base[j * MDX_LSTRIDE]; }
};
#if MDX_DT_ZERO
#define DTJ(j) 0.0
#else
#define DTJ(j) dT[j]
#endif

// ---------------------------------------------------------------------------
// Dense LU with partial pivoting, in registers (rodas5P.py:187-236). The pivot
// row is chosen first max wins, the whole row is swapped, and a zero pivot
// leaves its column alone. The swap is a select against every candidate row so
// all indices stay compile-time constants and nothing leaves the register file.
// In float storage the reciprocal and multiplier are double and each update is
// one double fma rounded to float, as the reference's typing has it.
// ---------------------------------------------------------------------------
// What the dense factorisation leaves beside the matrix: the row swaps, and
// for each pivot u_ii the double reciprocal 1/u_ii the elimination formed
// anyway. With a float matrix the back substitution multiplies by it instead
// of dividing: for float a and u with a normal float quotient,
// float(double(a) * (1/u)) is the correctly rounded a/u -- the double product
// is within 2^-52 of a/u relative, while a quotient of two 24-bit significands
// that is not exactly a float midpoint lies at least 2^-49 from every midpoint
// (and an exact midpoint of the normal range needs u a power of two, where the
// product is exact), so both round to the same float.
struct Pivots {
  int p[N];
  double rd[N];
};

// Up to this size the dense LU is unrolled with every index a constant, so the
// matrix stays in registers; above it the loops are kept and the matrix lives
// in scratch (a dense solve at that size is the uninformative end of the
// sparse path, which is where larger systems belong).
#ifndef MDX_DENSE_UNROLL_MAX
#define MDX_DENSE_UNROLL_MAX 16
#endif
constexpr bool DENSE_UNROLLED = N <= MDX_DENSE_UNROLL_MAX;
static_assert(!(MDX_LDS_LU && DENSE_UNROLLED), "an unrolled dense matrix stays in registers");

// U x = z, the back substitution of the dense solve. With double storage each
// row divides by its pivot. With float storage each row multiplies by the
// pivot's double reciprocal (see Pivots); a quotient in the float subnormal
// range can sit exactly on a midpoint of that coarser grid with u not a power
// of two, so when any row's double product is nonzero and below FLT_MIN the
// whole substitution is redone with divisions -- rare, and one branch per
// solve. The test reads the products' exponents: |x| < 2^-126 exactly when
// frexp's exponent is -126 or less, and zero, inf and NaN give 0, so the rows
// combine with one integer minimum and one compare. It flags every case in
// which the rounded float would be subnormal or a nonzero quotient would round
// to zero, plus the products just below FLT_MIN that round up to it, for which
// the divisions give the same bits.
template <bool UNROLLED>
__device__ __forceinline__ void back_substitute(const double* __restrict__ lu, const Pivots& ipiv,
                                                double* __restrict__ r) {
  (void)ipiv;
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    double acc = r[i];
#pragma unroll
    for (int j = i + 1; j < N; ++j) acc = fmaT(-lu[i * N + j], r[j], acc);
    r[i] = acc / lu[i * N + i];
  }
}

template <bool UNROLLED>
__device__ __forceinline__ void back_substitute(const float* __restrict__ lu, const Pivots& ipiv,
                                                float* __restrict__ r) {
  float z[N];
  int e_min = 0x7fffffff;
#pragma unroll
  for (int i = 0; i < N; ++i) z[i] = r[i];
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    float acc = r[i];
#pragma unroll
    for (int j = i + 1; j < N; ++j) acc = fmaT(-lu[i * N + j], r[j], acc);
    const double qd = (double)acc * ipiv.rd[i];
    const int e = __builtin_amdgcn_frexp_exp(qd);
    e_min = e < e_min ? e : e_min;
    r[i] = (float)qd;
  }
  if (e_min <= -126) {
    // the first row's divisor passes through an opaque move so the divisions
    // stay on this branch instead of being computed on every solve
    float u0 = lu[0];
    asm volatile("; subnormal quotient" : "+v"(u0));
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
      float acc = z[i];
#pragma unroll
      for (int j = i + 1; j < N; ++j) acc = fmaT(-lu[i * N + j], z[j], acc);
      z[i] = acc / (i == 0 ? u0 : lu[i * N + i]);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] = z[i];
  }
}

// An index into a chunk clamped to the last element (a load that stays in
// bounds; its value is not used).
__device__ __forceinline__ int clampN(int x) { return x < N ? x : N - 1; }

template <class L>
__device__ __forceinline__ void dense_factor(L lu, Pivots& ipiv) {
  typedef typename lu_elem<L>::type T;
  if constexpr (DENSE_UNROLLED) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T maxv = absT(lu[i * N + i]);
      int piv = i;
#pragma unroll
      for (int k = i + 1; k < N; ++k) {
        const T v = absT(lu[k * N + i]);
        if (v > maxv) {
          maxv = v;
          piv = k;
        }
      }
      ipiv.p[i] = piv;
#pragma unroll
      for (int k = i + 1; k < N; ++k) {
        const bool sw = (piv == k);
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const T a = lu[i * N + j], b = lu[k * N + j];
          lu[i * N + j] = sw ? b : a;
          lu[k * N + j] = sw ? a : b;
        }
      }
      const T pv = lu[i * N + i];
      const double inv = 1.0 / (double)pv;
      ipiv.rd[i] = inv;
      if (pv != (T)0) {
#pragma unroll
        for (int k = i + 1; k < N; ++k) {
          const double f = (double)lu[k * N + i] * inv;
          lu[k * N + i] = (T)f;
#pragma unroll
          for (int j = i + 1; j < N; ++j)
            lu[k * N + j] = (T)fmaT(-f, (double)lu[i * N + j], (double)lu[k * N + j]);
        }
      }
    }
  } else {
    // Above the unrolled size the matrix lives in per-lane scratch. Every loop
    // loads a chunk of MDX_CHUNK elements before computing with it, so that
    // many loads are in flight per wait; each element still sees the same
    // operations in the same order. Chunk indices past the end are clamped
    // to the last element for the loads and skipped for the stores.
#pragma unroll 1
    for (int i = 0; i < N; ++i) {
      T maxv = absT(lu[i * N + i]);
      int piv = i;
#pragma unroll 1
      for (int k0 = i + 1; k0 < N; k0 += MDX_CHUNK) {
        T v[MDX_CHUNK];
#pragma unroll
        for (int c = 0; c < MDX_CHUNK; ++c) v[c] = absT(lu[clampN(k0 + c) * N + i]);
#pragma unroll
        for (int c = 0; c < MDX_CHUNK; ++c)
          if (k0 + c < N && v[c] > maxv) {
            maxv = v[c];
            piv = k0 + c;
          }
      }
      ipiv.p[i] = piv;
      if (piv != i) {
#pragma unroll 1
        for (int j0 = 0; j0 < N; j0 += MDX_CHUNK) {
          T a[MDX_CHUNK], b[MDX_CHUNK];
#pragma unroll
          for (int c = 0; c < MDX_CHUNK; ++c) {
            a[c] = lu[i * N + clampN(j0 + c)];
            b[c] = lu[piv * N + clampN(j0 + c)];
          }
#pragma unroll
          for (int c = 0; c < MDX_CHUNK; ++c)
            if (j0 + c < N) {
              lu[i * N + j0 + c] = b[c];
              lu[piv * N + j0 + c] = a[c];
            }
        }
      }
      const T pv = lu[i * N + i];
      const double inv = 1.0 / (double)pv;
      ipiv.rd[i] = inv;
      if (pv != (T)0) {
        // the pivot row's chunk stays in registers while every row below
        // takes its update; the multiplier f = a_ki / u_ii (in double) is
        // formed again for each chunk from a_ki, which is replaced by the
        // stored L entry only after the last chunk
#pragma unroll 1
        for (int j0 = i + 1; j0 < N; j0 += MDX_LU_CHUNK) {
          double prow[MDX_LU_CHUNK];
#pragma unroll
          for (int c = 0; c < MDX_LU_CHUNK; ++c) prow[c] = (double)lu[i * N + clampN(j0 + c)];
#pragma unroll 1
          for (int k = i + 1; k < N; ++k) {
            const double f = (double)lu[k * N + i] * inv;
            T kv[MDX_LU_CHUNK];
#pragma unroll
            for (int c = 0; c < MDX_LU_CHUNK; ++c) kv[c] = lu[k * N + clampN(j0 + c)];
#pragma unroll
            for (int c = 0; c < MDX_LU_CHUNK; ++c)
              if (j0 + c < N) lu[k * N + j0 + c] = (T)fmaT(-f, prow[c], (double)kv[c]);
          }
        }
#pragma unroll 1
        for (int k0 = i + 1; k0 < N; k0 += MDX_CHUNK) {
          T l[MDX_CHUNK];
#pragma unroll
          for (int c = 0; c < MDX_CHUNK; ++c) l[c] = lu[clampN(k0 + c) * N + i];
#pragma unroll
          for (int c = 0; c < MDX_CHUNK; ++c)
            if (k0 + c < N) lu[(k0 + c) * N + i] = (T)((double)l[c] * inv);
        }
      }
    }
  }
}

template <class L, class T>
__device__ __forceinline__ void dense_solve(const L& lu, const Pivots& ipiv,
                                            T* __restrict__ r) {
  if constexpr (DENSE_UNROLLED) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int piv = ipiv.p[i];
#pragma unroll
      for (int k = i + 1; k < N; ++k) {
        const bool sw = (piv == k);
        const T a = r[i], b = r[k];
        r[i] = sw ? b : a;
        r[k] = sw ? a : b;
      }
      T acc = r[i];
#pragma unroll
      for (int j = 0; j < i; ++j) acc = fmaT(-lu[i * N + j], r[j], acc);
      r[i] = acc;
    }
    back_substitute<true>(lu, ipiv, r);
  } else {
    // scratch-resident matrix and vector: each row's sum loads a chunk of the
    // row and of the vector before its multiply-adds, which stay in order
#pragma unroll 1
    for (int i = 0; i < N; ++i) {
      const int piv = ipiv.p[i];
      if (piv != i) {
        const T a = r[i];
        r[i] = r[piv];
        r[piv] = a;
      }
      T acc = r[i];
#pragma unroll 1
      for (int j0 = 0; j0 < i; j0 += MDX_CHUNK) {
        T lv[MDX_CHUNK], rv[MDX_CHUNK];
#pragma unroll
        for (int c = 0; c < MDX_CHUNK; ++c) {
          lv[c] = lu[i * N + clampN(j0 + c)];
          rv[c] = r[clampN(j0 + c)];
        }
#pragma unroll
        for (int c = 0; c < MDX_CHUNK; ++c)
          if (j0 + c < i) acc = fmaT(-lv[c], rv[c], acc);
      }
      r[i] = acc;
    }
#pragma unroll 1
    for (int i = N - 1; i >= 0; --i) {
      T acc = r[i];
#pragma unroll 1
      for (int j0 = i + 1; j0 < N; j0 += MDX_CHUNK) {
        T uv[MDX_CHUNK], rv[MDX_CHUNK];
#pragma unroll
        for (int c = 0; c < MDX_CHUNK; ++c) {
          uv[c] = lu[i * N + clampN(j0 + c)];
          rv[c] = r[clampN(j0 + c)];
        }
#pragma unroll
        for (int c = 0; c < MDX_CHUNK; ++c)
          if (j0 + c < N) acc = fmaT(-uv[c], rv[c], acc);
      }
      r[i] = acc / lu[i * N + i];
    }
  }
}

// M^T x = b for the factorisation above (PM = LU, swaps in ipiv): U^T z = b,
// L^T w = z, then the swaps undone in reverse order. Used by the adjoint.
template <class T, class L>
__device__ __forceinline__ void dense_solve_t(const L& lu, const Pivots& ipiv,
                                              T* __restrict__ r) {
#pragma unroll 1
  for (int i = 0; i < N; ++i) {
    T acc = r[i];
    for (int j = 0; j < i; ++j) acc = fmaT(-lu[j * N + i], r[j], acc);
    r[i] = acc / lu[i * N + i];
  }
#pragma unroll 1
  for (int i = N - 1; i >= 0; --i) {
    T acc = r[i];
    for (int j = i + 1; j < N; ++j) acc = fmaT(-lu[j * N + i], r[j], acc);
    r[i] = acc;
  }
#pragma unroll 1
  for (int i = N - 1; i >= 0; --i) {
    const int piv = ipiv.p[i];
    if (piv != i) {
      const T a = r[i];
      r[i] = r[piv];
      r[piv] = a;
    }
  }
}

__device__ __forceinline__ int diag_slot(int d) {
#if MDX_SPARSE
  return sys::DIAG[d];
#else
  return d * N + d;
#endif
}

#if MDX_COOP
// The dense factorisation shared out over the trajectory's 64 lanes, with the
// matrix in LDS (rows of MDX_LD elements). Per pivot i:
// - the pivot is the first maximum of |a_ki| over k >= i, as the sequential
//   scan finds it (a NaN below the diagonal never wins, a NaN diagonal keeps
//   row i): each lane offers its row, and a butterfly over the wave keeps the
//   larger value, or on a tie the smaller row;
// - the two rows are swapped a column per lane;
// - for each row k below, every lane forms the same multiplier
//   f = a_ki * (1/u_ii) in double and updates its own column j > i with
//   a_kj = (T)fma(-f, u_ij, a_kj), the sequential update of that element;
// - column i then takes the multipliers, a row per lane.
// Every element sees the operations of the sequential factorisation in the
// same order, so the bits are the same.
// Lane l's value of v, read by every lane (l uniform): a VALU read across the
// wave, with no LDS round trip on the path.
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ void coop_clear(LUT* __restrict__ a) {
  const int c = (int)threadIdx.x;
  __syncthreads();
  if (c < N) {
#pragma unroll 1
    for (int r = 0; r < N; ++r) a[r * MDX_LD + c] = (LUT)-0.0;
  }
  __syncthreads();
}

__device__ __forceinline__ void coop_factor(LUT* __restrict__ a) {
  typedef LUT T;
  const int c = (int)threadIdx.x;
#pragma unroll 1
  for (int i = 0; i < N; ++i) {
    int piv = i;
    const T dii = absT(a[i * MDX_LD + i]);
    if (dii == dii) {
      T v = (T)-1;
      if (c >= i && c < N) {
        v = absT(a[c * MDX_LD + i]);
        if (v != v) v = (T)-1;
      }
      int idx = c;
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        const T ov = __shfl_xor(v, m);
        const int oi = __shfl_xor(idx, m);
        if (ov > v || (ov == v && oi < idx)) {
          v = ov;
          idx = oi;
        }
      }
      piv = __builtin_amdgcn_readfirstlane(idx);
    }
    if (c == 0) lds_piv[i] = piv;
    if (piv != i && c < N) {
      const T x = a[i * MDX_LD + c];
      a[i * MDX_LD + c] = a[piv * MDX_LD + c];
      a[piv * MDX_LD + c] = x;
    }
    __syncthreads();
    const T pv = a[i * MDX_LD + i];
    const double inv = 1.0 / (double)pv;
    if (c == 0) lds_rd[i] = inv;
    if (pv != (T)0) {
      const bool mine = c > i && c < N;
      // this lane's row multiplier (lane = row), handed to the column lanes
      // by lane_value as each row's turn comes
      const double fk = mine ? (double)a[c * MDX_LD + i] * inv : 0.0;
      const double ui = mine ? (double)a[i * MDX_LD + c] : 0.0;
      // eight rows at a time, their elements loaded before any is stored
#pragma unroll 1
      for (int k0 = i + 1; k0 < N; k0 += 8) {
        T av[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) av[q] = mine ? a[clampN(k0 + q) * MDX_LD + c] : (T)0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const double f = lane_value(fk, clampN(k0 + q));
          if (mine && k0 + q < N) a[(k0 + q) * MDX_LD + c] = (T)fmaT(-f, ui, (double)av[q]);
        }
      }
      __syncthreads();
      if (mine) a[c * MDX_LD + i] = (T)fk;
    }
    __syncthreads();
  }
  // the row swaps composed into one permutation of the right-hand side,
  // applied by coop_solve as a gather: position q takes r[lds_perm[q]]
  if (c < N) lds_perm[c] = c;
  __syncthreads();
#pragma unroll 1
  for (int i = 0; i < N; ++i) {
    const int piv = lds_piv[i];
    if (piv != i) {
      const int x = lds_perm[i];
      const int y = lds_perm[piv];
      lds_perm[i] = y;
      lds_perm[piv] = x;
    }
  }
  __syncthreads();
}

// Solve with coop_factor's factorisation, the right-hand side in registers
// (r, the same on every lane): the row swaps as the gather coop_factor
// composed (the same values the swaps in order leave), the forward
// substitution a row per lane (row c takes column j's term once row j is
// final, so each row sums in ascending column order, as sequentially), then
// the back substitution as one chain, every lane computing the same values.
__device__ __forceinline__ void coop_solve(const LUT* __restrict__ a, LUT* __restrict__ r) {
  typedef LUT T;
  const int c = (int)threadIdx.x;
#pragma unroll
  for (int j = 0; j < N; ++j) lds_r[j] = r[j];
  __syncthreads();
  T acc = c < N ? lds_r[lds_perm[c]] : (T)0;
#pragma unroll 1
  for (int j0 = 0; j0 < N - 1; j0 += MDX_CHUNK) {
    T l[MDX_CHUNK];
#pragma unroll
    for (int q = 0; q < MDX_CHUNK; ++q)
      l[q] = c < N ? a[c * MDX_LD + clampN(j0 + q)] : (T)0;
#pragma unroll
    for (int q = 0; q < MDX_CHUNK; ++q) {
      const int j = j0 + q;
      if (j < N - 1) {
        const T rj = lane_value(acc, j);
        if (c > j && c < N) acc = fmaT(-l[q], rj, acc);
      }
    }
  }
  // back substitution: lane j keeps x_j once it is known, and row i reads
  // each x_j it needs from lane j; the row's elements come from LDS
  T xr = (T)0;
#pragma unroll 1
  for (int i = N - 1; i >= 0; --i) {
    T x = lane_value(acc, i);
    int j = i + 1;
    // sixteen terms at a time, the row's sixteen elements loaded first
#pragma unroll 1
    for (; j + 16 <= N; j += 16) {
      T u[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) u[q] = a[i * MDX_LD + j + q];
#pragma unroll
      for (int q = 0; q < 16; ++q) x = fmaT(-u[q], lane_value(xr, j + q), x);
    }
#pragma unroll 1
    for (; j + 4 <= N; j += 4) {
      T u[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) u[q] = a[i * MDX_LD + j + q];
#pragma unroll
      for (int q = 0; q < 4; ++q) x = fmaT(-u[q], lane_value(xr, j + q), x);
    }
#pragma unroll 1
    for (; j < N; ++j) x = fmaT(-a[i * MDX_LD + j], lane_value(xr, j), x);
    x = x / a[i * MDX_LD + i];
    if (c == i) xr = x;
  }
  __syncthreads();
  if (c < N) lds_r[c] = xr;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < N; ++j) r[j] = lds_r[j];
  __syncthreads();
}
#endif

template <class L>
__device__ __forceinline__ void factor(L lu, Pivots& ipiv) {
#if MDX_SPARSE
  (void)ipiv;
  sys::sparse_factor(lu);
#elif MDX_COOP
  (void)ipiv;
  __syncthreads();
  coop_factor(lu.base);
#else
  dense_factor(lu, ipiv);
#endif
}

template <class L>
__device__ __forceinline__ void solve(const L& lu, const Pivots& ipiv, LUT* __restrict__ r) {
#if MDX_SPARSE
  (void)ipiv;
  sys::sparse_solve(lu, r);
#elif MDX_COOP
  (void)ipiv;
  coop_solve(lu.base, r);
#else
  dense_solve(lu, ipiv, r);
#endif
}

#if MDX_ADJOINT
// M^T x = b against the step's factorisation, in double (the adjoint); the
// factors through any indexable type holding doubles.
template <class L>
__device__ __forceinline__ void solve_t(const L& lu, const Pivots& ipiv, double* __restrict__ r) {
#if MDX_SPARSE
  (void)ipiv;
  sys::sparse_solve_t(lu, r);
#else
  dense_solve_t(lu, ipiv, r);
#endif
}
#endif

// The joint right-hand side at z: the state block, then each direction's
// J_y s + J_p e (modax/_sensitivity.py:198-243).
template <class Z>
__device__ __forceinline__ void rhs_aug(const Z& z, double ts, const PA& p,
                                        double* __restrict__ work) {
  sys::rhs(z, ts, p, work);
#if MDX_NSENS > 0
#pragma unroll MDX_SENS_UNROLL
  for (int d = 0; d < NS; ++d)
    sys::sens_rhs(z, ts, p, Shifted<Z>{z, N + d * N}, OneHot{PCOL[d]}, work + N + d * N);
#endif
}

// The sensitivity rows of the joint system's df/dt (rodas5P.py:400-436).
template <class Z>
__device__ __forceinline__ void sens_time_derivative(const Z& y, double t, const PA& p,
                                                     double* __restrict__ dT) {
#if MDX_NSENS > 0 && !MDX_DT_ZERO
#pragma unroll MDX_SENS_UNROLL
  for (int d = 0; d < NS; ++d)
    sys::sens_dt(y, t, p, Shifted<Z>{y, N + d * N}, OneHot{PCOL[d]}, dT + N + d * N);
#else
  (void)y, (void)t, (void)p, (void)dT;
#endif
}

// Stage S's state rows as doubles, read where the coupling term uses them.
template <class KS>
struct StageRows {
  KS ks;
  int S;
  __device__ __forceinline__ double operator[](int j) const { return (double)ks(S, j); }
};

// A direction's right-hand side as the coupling term is produced: element j of
// the term c lands in r[j] = (LUT)(sv[j] + c), so neither c nor a copy of the
// stage's state rows is held whole next to r.
struct CoupledRhs {
  LUT* r;
  const double* sv;
  struct Elem {
    LUT* r;
    double s;
    __device__ __forceinline__ void operator=(double c) const { *r = (LUT)(s + c); }
  };
  __device__ __forceinline__ Elem operator[](int j) const { return Elem{r + j, sv[j]}; }
};

// The sensitivity rows of stage S, after its state rows are solved: the joint
// iteration matrix is block lower triangular with M on every diagonal block, so
// each direction solves against the same factorisation with its right-hand side
// corrected by L_d k_y, one second-order sweep at the step's base point
// (rodas5P.py:512-561). ``sv`` holds the directions' stage values (double),
// ``ks`` stores into the stage vectors.
#if MDX_NSENS > 0
template <class Z, class KS, class L>
__device__ __forceinline__ void sens_direction(int S, int d, const Z& ybase, double t,
                                               const PA& p, const double* __restrict__ sv,
                                               const L& lu, const Pivots& ipiv, KS ks) {
  LUT r[N];
  sys::sens_coupling(ybase, t, p, Shifted<Z>{ybase, N + d * N}, OneHot{PCOL[d]},
                     StageRows<KS>{ks, S}, CoupledRhs{r, sv});
  solve(lu, ipiv, r);
#pragma unroll
  for (int j = 0; j < N; ++j) ks(S, N + d * N + j) = r[j];
}
#endif

template <class Z, class KS, class L>
__device__ __forceinline__ void sens_stage(int S, const Z& ybase, double t, const PA& p,
                                           const double* __restrict__ sv, const L& lu,
                                           const Pivots& ipiv, KS ks) {
#if MDX_NSENS > 0
#pragma unroll MDX_SENS_UNROLL
  for (int d = 0; d < NS; ++d) sens_direction(S, d, ybase, t, p, sv + d * N, lu, ipiv, ks);
#else
  (void)S, (void)ybase, (void)t, (void)p, (void)sv, (void)lu, (void)ipiv, (void)ks;
#endif
}

#if MDX_REFINE
// Stage S's stored vector as a Jacobian-vector product's direction.
template <class KS>
struct StageVec {
  KS ks;
  int S;
  __device__ __forceinline__ double operator[](int j) const { return ks(S, j); }
};

// The refinement's right-hand side as J x is produced: element j of J x lands
// in r[j] = (LUT)(fma(-x_j, 1/(h gamma), b_j) + (J x)_j), the residual
// b - M x against M = I/(h gamma) - J in double, x_j = ks(S, j).
template <class KS, class B>
struct Residual {
  LUT* r;
  B b;
  KS ks;
  int S;
  double gi;
  struct Elem {
    LUT* r;
    double b, x, gi;
    __device__ __forceinline__ void operator=(double c) const {
      *r = (LUT)(fmaT(-x, gi, b) + c);
    }
  };
  __device__ __forceinline__ Elem operator[](int j) const {
    return Elem{r + j, b[j], ks(S, j), gi};
  }
};

// One step of mixed-precision iterative refinement of stage S's solve. On
// entry ks(S, .) holds x, the FP32 solution of the rounded system; the
// residual of b (the stage's right-hand side in double) against the double
// iteration matrix at the step's base point is solved with the same FP32
// factors and added, so the stored stage vector is x + dx in double.
template <class YB, class B, class KS, class L>
__device__ __forceinline__ void refine(int S, const YB& ybase, double t, const PA& p, double gi,
                                       B b, KS ks, const L& lu, const Pivots& ipiv) {
  LUT r[N];
  sys::jvp(ybase, t, p, StageVec<KS>{ks, S}, Residual<KS, B>{r, b, ks, S, gi});
  solve(lu, ipiv, r);
#pragma unroll
  for (int j = 0; j < N; ++j) ks(S, j) = ks(S, j) + (double)r[j];
}
#endif

// The tableau as the unrolled step reads it: compile-time constants, or
// (MDX_TAB_REGS) copies made once per kernel in registers the compiler cannot
// fold back into constants. Only the entries a step reads are kept live.
struct Tab {
#if MDX_TAB_REGS
  double cm[8][7], am[5][5], dv[8], ct[4];
  static __device__ __forceinline__ double reg(double x) {
    asm("" : "+v"(x));
    return x;
  }
  __device__ __forceinline__ Tab() {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
#pragma unroll
      for (int q = 0; q < 7; ++q) cm[r][q] = reg(CM[r][q]);
      dv[r] = reg(DV[r]);
    }
#pragma unroll
    for (int r = 0; r < 5; ++r)
#pragma unroll
      for (int q = 0; q < 5; ++q) am[r][q] = reg(AM[r][q]);
#pragma unroll
    for (int r = 0; r < 4; ++r) ct[r] = reg(CT[r]);
  }
  __device__ __forceinline__ double c(int r, int q) const { return cm[r][q]; }
  __device__ __forceinline__ double a(int r, int q) const { return am[r][q]; }
  __device__ __forceinline__ double d(int r) const { return dv[r]; }
  __device__ __forceinline__ double t(int r) const { return ct[r]; }
#else
  __device__ __forceinline__ double c(int r, int q) const { return CM[r][q]; }
  __device__ __forceinline__ double a(int r, int q) const { return AM[r][q]; }
  __device__ __forceinline__ double d(int r) const { return DV[r]; }
  __device__ __forceinline__ double t(int r) const { return CT[r]; }
#endif
};

// Stage S: rhs_j = LU(work_j + dt*d_S*dT_j + (sum_q C_Sq k_qj) / dt), solved
// against the step's factorisation into k_S (rodas5P.py:611-647). Rows past N
// (sensitivity blocks) keep their value in double for sens_stage.
template <int S>
__device__ __forceinline__ void stage(const double* __restrict__ work,
                                      const double* __restrict__ dT, KT (*__restrict__ k)[NA],
                                      const LUT* __restrict__ lu, const Pivots& ipiv,
                                      double dt_use, double inv_dt, double* __restrict__ sv,
                                      const Tab& tb, const double* __restrict__ ybase, double t,
                                      const PA& p, double gi) {
  (void)ybase, (void)t, (void)p, (void)gi;
  LUT r[N];
#if MDX_REFINE
  double b[N];
#endif
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    double v;
    if constexpr (S == 0) {
      v = fmaT(dt_use * tb.d(0), DTJ(j), work[j]);
    } else {
      double acc = tb.c(S, 0) * (double)k[0][j];
#pragma unroll
      for (int q = 1; q < S; ++q) acc = fmaT(tb.c(S, q), (double)k[q][j], acc);
      if constexpr (S < 5)
        v = fmaT(acc, inv_dt, fmaT(dt_use * tb.d(S), DTJ(j), work[j]));
      else
        v = fmaT(acc, inv_dt, work[j]);
    }
    if (j < N) {
      r[j] = (LUT)v;
#if MDX_REFINE
      b[j] = v;
#endif
    } else {
      sv[j - N] = v;
    }
  }
  solve(lu, ipiv, r);
#pragma unroll
  for (int j = 0; j < N; ++j) k[S][j] = r[j];
#if MDX_REFINE
  refine(S, ybase, t, p, gi, (const double*)b, [&](int q, int j) -> KT& { return k[q][j]; }, lu,
         ipiv);
#endif
}

// u = y + sum_q A_{S+1,q} k_q (rodas5P.py:649-661).
template <int S>
__device__ __forceinline__ void advance(double* __restrict__ u, const double* __restrict__ y,
                                        KT (*__restrict__ k)[NA], const Tab& tb) {
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    double acc = tb.a(S, 0) * (double)k[0][j];
#pragma unroll
    for (int q = 1; q <= S; ++q) acc = fmaT(tb.a(S, q), (double)k[q][j], acc);
    u[j] = y[j] + acc;
  }
}

// Everything one trajectory carries from step to step.
struct State {
#if !MDX_LDS_Y
  double y[NA];
#endif
  double t, dt, tf, err_prev, err_prev2;
  int save_idx, n_steps, accepted, rejected;
};

__device__ __forceinline__ void init(State& s, const SolveArgs& a, i64 i, const PA& p) {
  MDX_LANE();
  const double* __restrict__ times = a.times;
  const double tf = times[a.n_save - 1];
#pragma unroll
  for (int j = 0; j < NA; ++j) YJ(s, j) = a.y0[i * a.y0_ts + (i64)j * a.y0_vs];
#if !MDX_ADJOINT && MDX_SAVE_HISTORY
#pragma unroll
  for (int j = 0; j < NA; ++j) a.hist[i * a.h_ts + (i64)j * a.h_vs] = YJ(s, j);
#endif
#if MDX_HOOK
  {
    const HookRow acc{a.hook_out + i * a.ho_ts, a.ho_vs};
    for (int k = 0; k < MDX_HOOK_SIZE; ++k) acc.put(k, 0.0);
#if MDX_LDS_Y
    sys::save_hook(0, LdsVec{&lds_y[lane_]}, times[0], p, acc);
#else
    sys::save_hook(0, s.y, times[0], p, acc);
#endif
  }
#endif
  s.t = times[0];
  s.dt = a.dt0 > 0.0 ? a.dt0 : (tf - times[0]) * 1e-6;
#if MDX_TF_INDEX >= 0
  s.tf = p[MDX_TF_INDEX];
#else
  (void)p;
  s.tf = tf;
#endif
  s.err_prev = 1.0;
  s.err_prev2 = 1.0;
  s.save_idx = 1;
  s.n_steps = 0;
  s.accepted = 0;
  s.rejected = 0;
#if !MDX_ADJOINT
  if (a.resume && !a.cost) {
    // where a cost pilot left the trajectory: everything a step carries to the
    // next, so the solve continues with the bits an unbroken one has (the
    // pilot wrote the saves it passed)
    const double* __restrict__ r = a.resume + i;
    const i64 st = a.n_traj;
    // a loop in the rolled kernel (its state lies in LDS or scratch, and an
    // unrolled copy would hold every element in flight beside the step's
    // registers); unrolled where the state is registers
#if MDX_ROLLED
#pragma unroll 1
#else
#pragma unroll
#endif
    for (int j = 0; j < NA; ++j) YJ(s, j) = r[j * st];
    s.t = r[(NA + 0) * st];
    s.dt = r[(NA + 1) * st];
    s.err_prev = r[(NA + 2) * st];
    s.err_prev2 = r[(NA + 3) * st];
    s.save_idx = (int)r[(NA + 4) * st];
    s.n_steps = (int)r[(NA + 5) * st];
    s.accepted = (int)r[(NA + 6) * st];
    s.rejected = (int)r[(NA + 7) * st];
  }
#endif
}

__device__ __forceinline__ bool active(const State& s, const SolveArgs& a) {
  return s.save_idx < a.n_save && s.t < s.tf && s.n_steps < a.max_steps;
}

// acc[jj] = sum_{q<S} c_q k_q[j0+jj] for the first ``count`` elements of one
// chunk, q ascending (the rolled step's stage sums). A chunk's loads are independent, so they are
// all in flight before the first is waited for; summing every element at once
// would keep one accumulator per element live and leave the loads no registers.
template <class KS>
__device__ __forceinline__ void chunk_sum(const double* __restrict__ c, int S, int j0, KS ks,
                                          double* __restrict__ acc, int count) {
#pragma unroll
  for (int jj = 0; jj < MDX_CHUNK; ++jj)
    if (jj < count) acc[jj] = c[0] * (double)ks(0, j0 + jj);
#pragma unroll 1
  for (int q = 1; q < S; ++q) {
    const double cq = c[q];
    KT kv[MDX_CHUNK];
#pragma unroll
    for (int jj = 0; jj < MDX_CHUNK; ++jj)
      if (jj < count) kv[jj] = ks(q, j0 + jj);
#pragma unroll
    for (int jj = 0; jj < MDX_CHUNK; ++jj)
      if (jj < count) acc[jj] = fmaT(cq, (double)kv[jj], acc[jj]);
  }
}

#if MDX_ROLLED
// Rows [j_base, j_base + N) of stage S's right-hand side, rhs_j = work_j +
// dt*d_S*dT_j + (sum_q C_Sq k_qj) / dt (rodas5P.py:611-647), with work holding
// those rows' f values; the sum is formed a chunk at a time (see chunk_sum),
// one branch per stage and chunk outside the element loop, and each row handed
// to out(row - j_base, value).
template <class KST, class OUT>
__device__ __forceinline__ void stage_rows(int S, int j_base, const double* __restrict__ work,
                                           const double* __restrict__ dT, KST kst,
                                           double dt_use, double inv_dt, OUT out) {
  (void)dT;
  const double* __restrict__ cm = CM[S];
#pragma unroll
  for (int i0 = 0; i0 < N; i0 += MDX_CHUNK) {
    const int j0 = j_base + i0;
    if (S == 0) {
      const double dd = dt_use * DV[0];
#pragma unroll
      for (int jj = 0; jj < MDX_CHUNK; ++jj) {
        if (i0 + jj >= N) break;
        out(i0 + jj, fmaT(dd, DTJ(j0 + jj), work[i0 + jj]));
      }
    } else {
      double acc[MDX_CHUNK];
      chunk_sum(cm, S, j0, kst, acc, N - i0);
      if (S < 5) {
        const double dd = dt_use * DV[S];
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj) {
          if (i0 + jj >= N) break;
          out(i0 + jj, fmaT(acc[jj], inv_dt, fmaT(dd, DTJ(j0 + jj), work[i0 + jj])));
        }
      } else {
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj) {
          if (i0 + jj >= N) break;
          out(i0 + jj, fmaT(acc[jj], inv_dt, work[i0 + jj]));
        }
      }
    }
  }
}

// Stage S of the rolled step at the stage argument z: the state rows' f,
// right-hand side and solve, then each sensitivity direction's rows the same
// way (sens_direction), so that one block of n_vars rows is held at a time
// rather than the whole joint vector.
template <class Z, class YB, class KS, class KST, class L>
__device__ __forceinline__ void rolled_stage(int S, const Z& z, double ts, const YB& ybase,
                                             double t, const PA& p,
                                             const double* __restrict__ dT, KS ks, KST kst,
                                             const L& lu, const Pivots& ipiv, double dt_use,
                                             double inv_dt, double gi) {
  (void)ybase, (void)t, (void)gi;
  double work[N];
  sys::rhs(z, ts, p, work);
  LUT r[N];
#if MDX_REFINE && MDX_LDS_K
  // the right-hand side b, kept in LDS for the residual while the solve and
  // the Jacobian-vector product hold the registers
  MDX_LANE();
  stage_rows(S, 0, work, dT, kst, dt_use, inv_dt,
             [&](int i, double v) { r[i] = (LUT)v; lds_b[i * MDX_LSTRIDE + lane_] = v; });
#elif MDX_REFINE
  // the right-hand side replaces f in work, element by element (each row's f
  // is read before its right-hand side is written), for the residual
  stage_rows(S, 0, work, dT, kst, dt_use, inv_dt,
             [&](int i, double v) { r[i] = (LUT)v; work[i] = v; });
#else
  stage_rows(S, 0, work, dT, kst, dt_use, inv_dt,
             [&](int i, double v) { r[i] = (LUT)v; });
#endif
  solve(lu, ipiv, r);
#pragma unroll
  for (int j = 0; j < N; ++j) ks(S, j) = r[j];
#if MDX_REFINE && MDX_LDS_K
  refine(S, ybase, t, p, gi, LdsVec{&lds_b[lane_]}, ks, lu, ipiv);
#elif MDX_REFINE
  refine(S, ybase, t, p, gi, (const double*)work, ks, lu, ipiv);
#endif
#if MDX_NSENS > 0
#pragma unroll MDX_SENS_UNROLL
  for (int d = 0; d < NS; ++d) {
    sys::sens_rhs(z, ts, p, Shifted<Z>{z, N + d * N}, OneHot{PCOL[d]}, work);
    double sv[N];
    stage_rows(S, N + d * N, work, dT, kst, dt_use, inv_dt,
               [&](int i, double v) { sv[i] = v; });
    sens_direction(S, d, ybase, t, p, sv, lu, ipiv, ks);
  }
#endif
}
#endif

// One attempted step, accepted or rejected, with the saves it passes.
__device__ __forceinline__ void step(State& s, PA& p, const SolveArgs& a, i64 i,
                                     const Tab& tb) {
#if MDX_ROLLED && !MDX_PARAMS_IN_REGS && MDX_NP > 0
  if (a.max_steps < 0) p.v[a.n_save % MDX_NP] = 0;  // keeps the row in memory (see below)
#endif
  double dt_use = s.dt;
  if (dt_use > s.tf - s.t) dt_use = s.tf - s.t;
  if (dt_use < 1e-30) dt_use = 1e-30;
  MDX_LANE();
  const double inv_dt = 1.0 / dt_use;
  const double t = s.t;
  const double t_end = t + dt_use;
  const double dtgamma_inv = 1.0 / (dt_use * GAMMA);
#if MDX_LDS_Y
  const LdsVec ybase{&lds_y[lane_]};
#else
  const double (&ybase)[NA] = s.y;
#endif

  // M = I/(h gamma) - J straight into the buffer the factorisation works in.
#if MDX_LDS_LU
  const LdsMat lu{&lds_lu[lane_]};
#else
  LUT lu[MDX_NLU];
#endif
#if MDX_COOP
  // every slot to -0.0, a column per lane; the Jacobian then writes only its
  // structurally nonzero slots (the others would have stored -0.0)
  coop_clear(lu.base);
#elif MDX_NEEDS_CLEAR
#pragma unroll
  for (int q = 0; q < MDX_NLU; ++q) lu[q] = (LUT)0;
#endif
  double dT[NA];
  sys::neg_jac(ybase, t, p, lu, dT);
  sens_time_derivative(ybase, t, p, dT);
#pragma unroll
  for (int d = 0; d < N; ++d) lu[diag_slot(d)] = lu[diag_slot(d)] + (LUT)dtgamma_inv;
  Pivots ipiv;
  factor(lu, ipiv);

#if !MDX_LDS_K
  KT k[8][NA];
  auto ks = [&](int q, int j) -> KT& { return k[q][j]; };
#else
  auto ks = [&](int q, int j) -> KT& { return KQ(q, j); };
#endif
#if !MDX_LDS_U
  double u[NA];
#endif
  double sv[NS > 0 ? NA - N : 1];
#if MDX_ROLLED
  // Large systems: one copy of the stage code, run eight times, so the loop
  // body fits the instruction cache. The stage vectors and the state live in
  // LDS or in per-lane scratch (swizzled: one element of a wave is one
  // contiguous 256-byte access) and the registers hold the matrix and the
  // stage in flight. For scratch, the stores below are never executed
  // (max_steps >= 0 is checked before launch); their unresolvable index keeps
  // each array in memory, and repeating them inside the loop keeps the loads
  // where they are used rather than hoisted out of the loop.
#define MDX_PIN(arr, len) \
  if (a.max_steps < 0) arr[(a.n_save + S) % (len)] = 0
#pragma unroll 1
  for (int S = 0; S < 8; ++S) {
    MDX_LANE();
    // the stage vectors through this stage's lane register
    auto kst = [&](int q, int j) -> KT { return KQ(q, j); };
#if !MDX_LDS_K
    MDX_PIN(k[S & 7], NA);
#endif
#if !MDX_LDS_Y
    MDX_PIN(s.y, NA);
#endif
#if !MDX_DT_ZERO
    MDX_PIN(dT, NA);
#endif
#if !MDX_PARAMS_IN_REGS && MDX_NP > 0
    MDX_PIN(p.v, MDX_NP);
#endif
    double ts;
    if (S == 0) {
#if !(MDX_LDS_U && MDX_LDS_Y)
#pragma unroll
      for (int j0 = 0; j0 < NA; j0 += MDX_CHUNK) {
        double yv[MDX_CHUNK];
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj)
          if (j0 + jj < NA) yv[jj] = YJ(s, j0 + jj);
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj)
          if (j0 + jj < NA) UJ(j0 + jj) = yv[jj];
      }
#endif
      ts = t;
    } else if (S < 6) {
      // u = y + sum_q A_{S,q} k_q (rodas5P.py:649-661), a chunk of elements at
      // a time (see chunk_sum)
      const double* __restrict__ am = AM[S - 1];
#pragma unroll
      for (int j0 = 0; j0 < NA; j0 += MDX_CHUNK) {
        double acc[MDX_CHUNK], yv[MDX_CHUNK];
        chunk_sum(am, S, j0, kst, acc, NA - j0);
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj)
          if (j0 + jj < NA) yv[jj] = YJ(s, j0 + jj);
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj)
          if (j0 + jj < NA) UJ(j0 + jj) = yv[jj] + acc[jj];
      }
      ts = S < 5 ? t + CT[S - 1] * dt_use : t_end;
    } else {
#pragma unroll
      for (int j0 = 0; j0 < NA; j0 += MDX_CHUNK) {
        double uv[MDX_CHUNK];
        KT kv[MDX_CHUNK];
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj)
          if (j0 + jj < NA) uv[jj] = UJ(j0 + jj), kv[jj] = KQ(S - 1, j0 + jj);
#pragma unroll
        for (int jj = 0; jj < MDX_CHUNK; ++jj)
          if (j0 + jj < NA) UJ(j0 + jj) = uv[jj] + (double)kv[jj];
      }
      ts = t_end;
    }
#if MDX_LDS_U && MDX_LDS_Y
    // the first stage evaluates at the state itself, read where it lies
    rolled_stage(S, LdsVec{S == 0 ? &lds_y[lane_] : &lds_u[lane_]}, ts, ybase, t, p, dT, ks,
                 kst, lu, ipiv, dt_use, inv_dt, dtgamma_inv);
#elif MDX_LDS_U
    rolled_stage(S, LdsVec{&lds_u[lane_]}, ts, ybase, t, p, dT, ks, kst, lu, ipiv, dt_use,
                 inv_dt, dtgamma_inv);
#else
    rolled_stage(S, u, ts, ybase, t, p, dT, ks, kst, lu, ipiv, dt_use, inv_dt, dtgamma_inv);
#endif
  }
#undef MDX_PIN
#else
  double work[NA];
  rhs_aug(s.y, t, p, work);
  stage<0>(work, dT, k, lu, ipiv, dt_use, inv_dt, sv, tb, s.y, t, p, dtgamma_inv);
  sens_stage(0, ybase, t, p, sv, lu, ipiv, ks);
  advance<0>(u, s.y, k, tb);
  rhs_aug(u, t + tb.t(0) * dt_use, p, work);
  stage<1>(work, dT, k, lu, ipiv, dt_use, inv_dt, sv, tb, s.y, t, p, dtgamma_inv);
  sens_stage(1, ybase, t, p, sv, lu, ipiv, ks);
  advance<1>(u, s.y, k, tb);
  rhs_aug(u, t + tb.t(1) * dt_use, p, work);
  stage<2>(work, dT, k, lu, ipiv, dt_use, inv_dt, sv, tb, s.y, t, p, dtgamma_inv);
  sens_stage(2, ybase, t, p, sv, lu, ipiv, ks);
  advance<2>(u, s.y, k, tb);
  rhs_aug(u, t + tb.t(2) * dt_use, p, work);
  stage<3>(work, dT, k, lu, ipiv, dt_use, inv_dt, sv, tb, s.y, t, p, dtgamma_inv);
  sens_stage(3, ybase, t, p, sv, lu, ipiv, ks);
  advance<3>(u, s.y, k, tb);
  rhs_aug(u, t + tb.t(3) * dt_use, p, work);
  stage<4>(work, dT, k, lu, ipiv, dt_use, inv_dt, sv, tb, s.y, t, p, dtgamma_inv);
  sens_stage(4, ybase, t, p, sv, lu, ipiv, ks);
  advance<4>(u, s.y, k, tb);
  rhs_aug(u, t_end, p, work);
  stage<5>(work, dT, k, lu, ipiv, dt_use, inv_dt, sv, tb, s.y, t, p, dtgamma_inv);
  sens_stage(5, ybase, t, p, sv, lu, ipiv, ks);
#pragma unroll
  for (int j = 0; j < NA; ++j) u[j] = u[j] + (double)k[5][j];
  rhs_aug(u, t_end, p, work);
  stage<6>(work, dT, k, lu, ipiv, dt_use, inv_dt, sv, tb, s.y, t, p, dtgamma_inv);
  sens_stage(6, ybase, t, p, sv, lu, ipiv, ks);
#pragma unroll
  for (int j = 0; j < NA; ++j) u[j] = u[j] + (double)k[6][j];
  rhs_aug(u, t_end, p, work);
  stage<7>(work, dT, k, lu, ipiv, dt_use, inv_dt, sv, tb, s.y, t, p, dtgamma_inv);
  sens_stage(7, ybase, t, p, sv, lu, ipiv, ks);
#endif

  // Weighted RMS error of the embedded estimate k7 (rodas5P.py:903-911).
  // Loads a chunk at a time ahead of the arithmetic, so that several are in
  // flight when the vectors live in LDS or scratch.
  double err = 0.0;
#pragma unroll
  for (int j0 = 0; j0 < NA; j0 += MDX_CHUNK) {
    double k7v[MDX_CHUNK], uv[MDX_CHUNK], yv[MDX_CHUNK];
#pragma unroll
    for (int jj = 0; jj < MDX_CHUNK; ++jj)
      if (j0 + jj < NA) {
        k7v[jj] = (double)KQ(7, j0 + jj);
        uv[jj] = UJ(j0 + jj);
        yv[jj] = YJ(s, j0 + jj);
      }
#pragma unroll
    for (int jj = 0; jj < MDX_CHUNK; ++jj) {
      const int j = j0 + jj;
      if (j >= NA) break;
      const double k7 = k7v[jj];
      const double y_new = uv[jj] + k7;
      const double ya = absT(yv[jj]), yb = absT(y_new);
      const double scale = fmaT(a.rtol, ya > yb ? ya : yb, a.atol);
      const double r = weight(a, i, j) * k7 / scale;
      err = fmaT(r, r, err);
    }
  }
  const double err_norm = __builtin_sqrt(err / (double)MDX_N_ERROR);
  const bool accept = err_norm <= 1.0 && !__builtin_isnan(err_norm);
  const double factor_ =
      pid_factor(err_norm, accept, s.err_prev, s.err_prev2, MDX_E1, MDX_E2, MDX_E3, 0.2, 6.0);
  s.dt = dt_use * factor_;

  if (accept) {
    const double t_new = t + dt_use;
    const double* __restrict__ times = a.times;
    while (s.save_idx < a.n_save) {
      const double ts = times[s.save_idx];
      const double at = absT(ts);
      if (!(ts <= t_new + 1e-12 * (1.0 > at ? 1.0 : at))) break;
#if !MDX_ADJOINT
      const double theta = (ts - t) / dt_use;
      const double theta1 = 1.0 - theta;
      double* __restrict__ row = a.hist + i * a.h_ts + (i64)s.save_idx * a.h_ss;
#if MDX_HOOK
      double ysave[NA];
#endif
#pragma unroll
      for (int j = 0; j < NA; ++j, row += a.h_vs) {
        const double k0 = (double)KQ(0, j);
        double h1 = H1[0] * k0;
        double h2 = H2[0] * k0;
        double h3 = H3[0] * k0;
#pragma unroll
        for (int q = 1; q < 8; ++q) {
          const double kv = (double)KQ(q, j);
          h1 = fmaT(H1[q], kv, h1);
          h2 = fmaT(H2[q], kv, h2);
          h3 = fmaT(H3[q], kv, h3);
        }
        const double y_new = UJ(j) + (double)KQ(7, j);
        double v = fmaT(theta, h3, h2);
        v = fmaT(theta, v, h1);
        v = fmaT(theta1, v, y_new);
        const double val = fmaT(theta, v, theta1 * YJ(s, j));
#if MDX_SAVE_HISTORY
        *row = val;
#endif
#if MDX_HOOK
        ysave[j] = val;
#endif
      }
#if MDX_HOOK
      sys::save_hook(s.save_idx, ysave, ts, p, HookRow{a.hook_out + i * a.ho_ts, a.ho_vs});
#endif
#endif
      s.save_idx++;
    }
#pragma unroll
    for (int j0 = 0; j0 < NA; j0 += MDX_CHUNK) {
      double uv[MDX_CHUNK];
      KT kv[MDX_CHUNK];
#pragma unroll
      for (int jj = 0; jj < MDX_CHUNK; ++jj)
        if (j0 + jj < NA) uv[jj] = UJ(j0 + jj), kv[jj] = KQ(7, j0 + jj);
#pragma unroll
      for (int jj = 0; jj < MDX_CHUNK; ++jj)
        if (j0 + jj < NA) YJ(s, j0 + jj) = uv[jj] + (double)kv[jj];
    }
    s.t = t + dt_use;
    s.accepted++;
  } else {
    s.rejected++;
  }
  s.n_steps++;
}

// Saves past the trajectory's end hold its final state; counters out.
__device__ __forceinline__ void finish(State& s, const PA& p, const SolveArgs& a, i64 i) {
  MDX_LANE();
  (void)p;
  if (a.cost) {
    // a cost pilot (max_steps is its step budget): the steps left at the
    // current step size instead of the saves and counters, 0 once finished
    a.cost[i] = (s.save_idx < a.n_save && s.t < s.tf) ? (s.tf - s.t) / s.dt : 0.0;
    if (a.resume) {
      double* __restrict__ r = a.resume + i;
      const i64 st = a.n_traj;
#if MDX_ROLLED
#pragma unroll 1
#else
#pragma unroll
#endif
      for (int j = 0; j < NA; ++j) r[j * st] = YJ(s, j);
      r[(NA + 0) * st] = s.t;
      r[(NA + 1) * st] = s.dt;
      r[(NA + 2) * st] = s.err_prev;
      r[(NA + 3) * st] = s.err_prev2;
      r[(NA + 4) * st] = (double)s.save_idx;
      r[(NA + 5) * st] = (double)s.n_steps;
      r[(NA + 6) * st] = (double)s.accepted;
      r[(NA + 7) * st] = (double)s.rejected;
    }
    return;
  }
  while (s.save_idx < a.n_save) {
#if MDX_SAVE_HISTORY
    double* __restrict__ row = a.hist + i * a.h_ts + (i64)s.save_idx * a.h_ss;
#pragma unroll
    for (int j = 0; j < NA; ++j) row[(i64)j * a.h_vs] = YJ(s, j);
#endif
#if MDX_HOOK
    const HookRow acc{a.hook_out + i * a.ho_ts, a.ho_vs};
#if MDX_LDS_Y
    sys::save_hook(s.save_idx, LdsVec{&lds_y[lane_]}, a.times[s.save_idx], p, acc);
#else
    sys::save_hook(s.save_idx, s.y, a.times[s.save_idx], p, acc);
#endif
#endif
    s.save_idx++;
  }
#if !MDX_SAVE_HISTORY
  // without the history the one row holds the final state (rodas5P.py:999-1002)
#pragma unroll
  for (int j = 0; j < NA; ++j) a.hist[i * a.h_ts + (i64)j * a.h_vs] = YJ(s, j);
#endif
  a.accepted[i] = s.accepted;
  a.rejected[i] = s.rejected;
  a.loop[i] = s.n_steps;
}

}  // namespace r5
}  // namespace mdx

#if !MDX_ADJOINT
// A register cap, when the caller set one (the reference's max_registers).
#if defined(MDX_MAX_VGPRS) && MDX_MAX_VGPRS > 0
#define MDX_REG_CAP __attribute__((amdgpu_num_vgpr(MDX_MAX_VGPRS > 256 ? 256 : MDX_MAX_VGPRS)))
#else
#define MDX_REG_CAP
#endif

extern "C" __global__ __launch_bounds__(MDX_BLOCK) MDX_REG_CAP void mdx_rodas5p(mdx::SolveArgs a) {
  using namespace mdx::r5;
#if MDX_COOP
  // one trajectory per workgroup: every lane steps it, sharing the LU
  const mdx::i64 i = (mdx::i64)blockIdx.x;
#else
  const mdx::i64 i = (mdx::i64)blockIdx.x * MDX_BLOCK + threadIdx.x;
#endif
  if (i >= a.n_traj) return;
  PA p;
  p.load(a, i);
  State s;
  init(s, a, i, p);
  const Tab tb;
  while (active(s, a)) step(s, p, a, i, tb);
  finish(s, p, a, i);
}

#if !MDX_COOP
// Persistent refill: a resident grid whose lanes claim trajectories from a
// device counter and take the next one as soon as theirs ends, so a wave keeps
// every lane busy while step counts across the ensemble differ by orders of
// magnitude. Per trajectory it runs the same init/step/finish as mdx_rodas5p,
// hence the same bits. Every lane leaves once the counter has passed n_traj
// and its last trajectory is finished.
extern "C" __global__ __launch_bounds__(MDX_BLOCK) MDX_REG_CAP void mdx_rodas5p_refill(
    mdx::SolveArgs a) {
  using namespace mdx::r5;
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
          init(s, a, i, p);
        }
      } else {
        exhausted = true;
      }
    }
    if (!__any(i >= 0)) break;
    if (i >= 0) {
      if (active(s, a)) step(s, p, a, i, tb);
      if (!active(s, a)) {
        finish(s, p, a, i);
        i = -1;
      }
    }
  }
}
#endif  // !MDX_COOP
#endif  // !MDX_ADJOINT
