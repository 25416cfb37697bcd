// oj_oracle.cpp -- CPU oracle for the single-spin-flip annealing path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (openjij_amd/, include/)
// links, loads or calls this file.  Only tests/, the smoke check in
// __graft_entry__.py and the cpu_baseline leg of bench.py use it, and only as
// the checker / the timed CPU baseline, never as the thing measured or shipped.
//
// It holds two independent CPU restatements:
//
//  (1) "ref_*": the reference algorithm restated from its sources, with the
//      reference's own random stream (Marsaglia xorshift128 feeding libstdc++
//      std::uniform_real_distribution<double> / uniform_int_distribution<>),
//      FP64 state, the (N+1) dummy-spin layout and the sequential sweep order.
//        include/openjij/utility/random.hpp:39-87          (Xorshift)
//        include/openjij/graph/graph.hpp:54-65             (gen_spin)
//        include/openjij/system/classical_ising.hpp:60-90  (dE = -2 s (J s))
//        include/openjij/updater/single_spin_flip.hpp:67-95   (classical sweep)
//        include/openjij/updater/single_spin_flip.hpp:128-251 (Trotter sweep)
//        include/openjij/system/transverse_ising.hpp:165-186 (Trotter dE)
//        include/openjij/utility/schedule_list.hpp:165-207  (schedule builders)
//        include/openjij/result/get_solution.hpp:39-84      (solutions)
//        include/openjij/graph/dense.hpp:176-192            (energy)
//
//  (2) "replay_*": a CPU replay of the GPU kernels' rule: counter-based
//      Philox4x32-10 uniforms, the threshold acceptance test
//      dE < theta = log(U) * (-1/beta), the same field-update arithmetic and the
//      same sweep / slice order.  For integer couplings (and, with the
//      lane-sequential field initialisation, for any couplings in the fp64 and
//      fp32 modes) it reproduces GPU spin trajectories bit for bit.
//
// Plus exact (double-double compensated) energies and a Gray-code brute-force
// ground-state search for n <= 24.
//
// Parity pins for each function (see tests/test_oracle.py):
//   philox             -> Random123 known-answer vectors + rocRAND source header
//   det_log_*          -> libm log within 4 ulp-ish bounds, monotonicity, x->1
//   ref_* schedules    -> closed forms (geometric / linear), reference tests
//   ref_sa_* / ref_sqa -> the 5 ground-state fixtures of tests/cxxtest/testcase.hpp,
//                         brute force, invariant dE = -2 s (J s) after every run
//   replay_*           -> brute force, invariants, KS vs ref_* (statistical)
//   theta_hb_* (heat-bath rule) -> the reference's form 1/(1+exp(beta dE)) > U
//                         evaluated directly, the closed-form acceptance
//                         probability, Boltzmann stationarity by enumeration
//   energy_exact       -> closed forms of tests/cxxtest/graph/quadratic.hpp
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off -fopenmp).

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;
typedef uint64_t u64;
typedef uint32_t u32;
typedef int8_t i8;

namespace {

// ---------------------------------------------------------------------------
// Reference random engine restated: Marsaglia xorshift128 with the fixed x,y,z
// state and w = seed (include/openjij/utility/random.hpp:39-87).  result_type
// is uint_fast32_t (64-bit on Linux) while max() is UINT_MAX, which makes
// libstdc++'s generate_canonical<double,53> draw twice per uniform.
struct RefXorshift {
  using result_type = uint_fast32_t;
  static constexpr unsigned min() { return 0u; }
  static constexpr unsigned max() { return UINT_MAX; }
  explicit RefXorshift(unsigned s) : w(s) {}
  unsigned operator()() {
    unsigned t = x ^ (x << 11);
    x = y;
    y = z;
    z = w;
    return w = (w ^ (w >> 19)) ^ (t ^ (t >> 8));
  }
  unsigned x = 123456789u, y = 362436069u, z = 521288629u, w;
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), written from the paper.
inline void philox(const u32 ctr[4], const u32 key[2], u32 out[4]) {
  u32 c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  u32 k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    u64 p0 = (u64)0xD2511F53u * c0;
    u64 p1 = (u64)0xCD9E8D57u * c2;
    u32 hi0 = (u32)(p0 >> 32), lo0 = (u32)p0;
    u32 hi1 = (u32)(p1 >> 32), lo1 = (u32)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// The uniform used by spin i of replica r, Trotter slice m, global sweep t.
inline u32 philox_word(u64 seed, u64 t, u64 r, u64 m, u64 i) {
  u32 ctr[4] = {(u32)(i >> 2), (u32)t, (u32)r, (u32)m};
  u32 key[2] = {(u32)seed, (u32)(seed >> 32)};
  u32 o[4];
  philox(ctr, key, o);
  return o[i & 3];
}

// ---------------------------------------------------------------------------
// Deterministic natural log on (0,1): exact exponent/mantissa split, then
// ln m = 2 atanh(t), t = (m-1)/(m+1), as an odd series with coefficients
// 1/(2k+1).  Only IEEE +,-,*,/ and fma are used, so the CPU and the GPU give
// the same bits.
inline double det_log_d(double x) {
  u64 b;
  std::memcpy(&b, &x, 8);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m;
  std::memcpy(&m, &b, 8);
  if (m > 0x1.6a09e667f3bcdp+0) {
    m = m * 0.5;
    e += 1;
  }
  double f = m - 1.0;
  double t = f / (2.0 + f);
  double t2 = t * t;
  double p = 0x1.642c8590b2164p-5;            // 1/23
  p = std::fma(t2, p, 0x1.8618618618618p-5);  // 1/21
  p = std::fma(t2, p, 0x1.af286bca1af28p-5);  // 1/19
  p = std::fma(t2, p, 0x1.e1e1e1e1e1e1ep-5);  // 1/17
  p = std::fma(t2, p, 0x1.1111111111111p-4);  // 1/15
  p = std::fma(t2, p, 0x1.3b13b13b13b14p-4);  // 1/13
  p = std::fma(t2, p, 0x1.745d1745d1746p-4);  // 1/11
  p = std::fma(t2, p, 0x1.c71c71c71c71cp-4);  // 1/9
  p = std::fma(t2, p, 0x1.2492492492492p-3);  // 1/7
  p = std::fma(t2, p, 0x1.999999999999ap-3);  // 1/5
  p = std::fma(t2, p, 0x1.5555555555555p-2);  // 1/3
  double s = t2 * p;
  double tt = t + t;
  double lm = std::fma(tt, s, tt);
  return std::fma((double)e, 0x1.62e42fefa39efp-1, lm);
}

inline float det_log_f(float x) {
  u32 b;
  std::memcpy(&b, &x, 4);
  int e = (int)((b >> 23) & 0xff) - 127;
  b = (b & 0x007fffffu) | 0x3f800000u;
  float m;
  std::memcpy(&m, &b, 4);
  if (m > 0x1.6a09e6p+0f) {
    m = m * 0.5f;
    e += 1;
  }
  float f = m - 1.0f;
  float t = f / (2.0f + f);
  float t2 = t * t;
  float p = 0x1.745d18p-4f;               // 1/11
  p = std::fma(t2, p, 0x1.c71c72p-4f);    // 1/9
  p = std::fma(t2, p, 0x1.24924ap-3f);    // 1/7
  p = std::fma(t2, p, 0x1.99999ap-3f);    // 1/5
  p = std::fma(t2, p, 0x1.555556p-2f);    // 1/3
  float s = t2 * p;
  float tt = t + t;
  float lm = std::fma(tt, s, tt);
  return std::fma((float)e, 0x1.62e430p-1f, lm);
}

// U in (0,1) from a 32-bit word, and the acceptance threshold
// theta = log(U) * nbinv with nbinv = -1/beta (theta = +inf at beta = 0).
inline double uniform_d(u32 u) { return ((double)u + 0.5) * 0x1p-32; }
inline float uniform_f(u32 u) { return ((float)(u >> 9) + 0.5f) * 0x1p-23f; }
inline double theta_d(u32 u, double nbinv) { return det_log_d(uniform_d(u)) * nbinv; }
inline float theta_f(u32 u, float nbinv) { return det_log_f(uniform_f(u)) * nbinv; }
// Heat-bath threshold: the reference flips iff 1 / (1 + exp(beta dE)) > U
// (include/openjij/updater/single_spin_flip.hpp:365-371), i.e. iff
// dE < log(U / (1 - U)) / beta.  1 - U = uniform(~u) exactly, in both precisions.
inline double theta_hb_d(u32 u, double nbinv) {
  return (det_log_d(uniform_d(u)) - det_log_d(uniform_d(~u))) * nbinv;
}
inline float theta_hb_f(u32 u, float nbinv) {
  return (det_log_f(uniform_f(u)) - det_log_f(uniform_f(~u))) * nbinv;
}
inline double theta_rule_d(int rule, u32 u, double nbinv) {
  return rule ? theta_hb_d(u, nbinv) : theta_d(u, nbinv);
}
inline float theta_rule_f(int rule, u32 u, float nbinv) {
  return rule ? theta_hb_f(u, nbinv) : theta_f(u, nbinv);
}

// double-double accumulation (Knuth TwoSum); terms here are exact products.
struct DD {
  double hi = 0.0, lo = 0.0;
  void add(double a) {
    double s = hi + a;
    double bp = s - hi;
    double err = (hi - (s - bp)) + (a - bp);
    hi = s;
    lo += err;
  }
  double value() const { return hi + lo; }
};

inline bool is_integer_matrix(const double* v, i64 n) {
  for (i64 k = 0; k < n; ++k)
    if (v[k] != std::floor(v[k]) || std::fabs(v[k]) > 1e9) return false;
  return true;
}

}  // namespace

extern "C" {

// ===========================================================================
// Primitives
void orc_philox4x32_10(const u32* ctr, const u32* key, u32* out) { philox(ctr, key, out); }
u32 orc_philox_word(u64 seed, u64 t, u64 r, u64 m, u64 i) { return philox_word(seed, t, r, m, i); }
double orc_det_log_d(double x) { return det_log_d(x); }
float orc_det_log_f(float x) { return det_log_f(x); }
double orc_theta_d(u32 u, double nbinv) { return theta_d(u, nbinv); }
float orc_theta_f(u32 u, float nbinv) { return theta_f(u, nbinv); }
double orc_theta_hb_d(u32 u, double nbinv) { return theta_hb_d(u, nbinv); }
float orc_theta_hb_f(u32 u, float nbinv) { return theta_hb_f(u, nbinv); }

// Reference Xorshift raw stream (for pinning against the recurrence).
void orc_ref_xorshift(unsigned seed, i64 count, u32* out) {
  RefXorshift g(seed);
  for (i64 k = 0; k < count; ++k) out[k] = g();
}

// uniform_real_distribution<double>(0,1) over the reference engine.
void orc_ref_uniform(unsigned seed, i64 count, double* out) {
  RefXorshift g(seed);
  std::uniform_real_distribution<> urd(0, 1.0);
  for (i64 k = 0; k < count; ++k) out[k] = urd(g);
}

// ===========================================================================
// Schedules (include/openjij/utility/schedule_list.hpp:165-207): beta by
// repeated multiplication, s by accumulated addition.
void orc_make_classical_schedule(double beta_min, double beta_max, i64 L, double* beta) {
  double r = std::pow(beta_max / beta_min, 1.0 / static_cast<double>(L - 1));
  double b = beta_min;
  for (i64 k = 0; k < L; ++k) {
    beta[k] = b;
    b *= r;
  }
}

void orc_make_transverse_schedule(i64 L, double* s_out) {
  double ds = 1.0 / static_cast<double>(L - 1);
  double s = 0;
  for (i64 k = 0; k < L; ++k) {
    s_out[k] = s;
    s += ds;
  }
}

// Per-sweep Trotter parameters, computed with the host libm exactly as the
// engine's host code does: a = s*(beta/M), B = 1/2 log tanh(beta*gamma*(1-s)/M)
// (include/openjij/updater/single_spin_flip.hpp:147-148, :216-217).
i64 orc_expand_transverse(const double* beta, const double* s, const i64* mcs, i64 L,
                          double gamma, i64 M, double* a, double* B) {
  i64 t = 0;
  const double Md = (double)M;
  for (i64 l = 0; l < L; ++l) {
    const double av = s[l] * (beta[l] / Md);
    const double Bv = (1 / 2.) * std::log(std::tanh(beta[l] * gamma * (1.0 - s[l]) / Md));
    for (i64 k = 0; k < mcs[l]; ++k, ++t) {
      if (a) a[t] = av;
      if (B) B[t] = Bv;
    }
  }
  return t;
}

// ===========================================================================
// Reference restatement: initial spins, include/openjij/graph/graph.hpp:54-65
// with the engine seeded as in openjij/declare.hpp:46-52 (size_t -> unsigned).
void orc_ref_gen_spin(i64 n, u64 seed, i8* out) {
  RefXorshift g((unsigned)seed);
  std::uniform_int_distribution<> uid(0, 1);
  for (i64 k = 0; k < n; ++k) out[k] = (i8)(2 * uid(g) - 1);
}

// Classical sweep, dense (N+1)^2 symmetric J_ext with h in column n and
// corner 1.  include/openjij/updater/single_spin_flip.hpp:67-95 and
// include/openjij/system/classical_ising.hpp:87-90.
static void ref_sa_dense_one(i64 n, const double* J, i8* spins, unsigned seed,
                             const double* beta, const i64* mcs, i64 L) {
  const i64 n1 = n + 1;
  std::vector<double> s(n1), dE(n1);
  for (i64 i = 0; i < n; ++i) s[i] = spins[i];
  s[n] = 1;
  for (i64 i = 0; i < n1; ++i) {
    double acc = 0;
    for (i64 j = 0; j < n1; ++j) acc += J[i * n1 + j] * s[j];
    dE[i] = -2.0 * s[i] * acc;
  }
  RefXorshift rng(seed);
  std::uniform_real_distribution<> urd(0, 1.0);
  for (i64 l = 0; l < L; ++l) {
    const double b = beta[l];
    for (i64 step = 0; step < mcs[l]; ++step) {
      for (i64 i = 0; i < n; ++i) {
        if (dE[i] <= 0 || std::exp(-b * dE[i]) > urd(rng)) {
          const double c = 4 * s[i];
          const double* row = J + i * n1;
          for (i64 j = 0; j < n1; ++j) dE[j] += c * (row[j] * s[j]);
          dE[i] *= -1;
          s[i] *= -1;
        }
        s[n] = 1;
      }
    }
  }
  for (i64 i = 0; i < n; ++i) spins[i] = (i8)(s[i] * s[n]);
}

int orc_ref_sa_dense(i64 n, const double* J_ext, i8* spins, u64 seed, const double* beta,
                     const i64* mcs, i64 L) {
  if (n < 1 || L < 0) return 1;
  ref_sa_dense_one(n, J_ext, spins, (unsigned)seed, beta, mcs, L);
  return 0;
}

// Many reads, each with its own seed (OpenMP over reads, as the C++ sampler
// does at include/openjij/sampler/sa_sampler.hpp:285).  Used as the CPU baseline.
int orc_ref_sa_dense_batch(i64 n, const double* J_ext, i64 R, i8* spins, const u32* seeds,
                           const double* beta, const i64* mcs, i64 L, int nthreads) {
  if (n < 1) return 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (i64 r = 0; r < R; ++r) ref_sa_dense_one(n, J_ext, spins + r * n, seeds[r], beta, mcs, L);
  return 0;
}

// Classical sweep, CSR with (n+1) rows (symmetric, dummy row/column, corner).
static void ref_sa_csr_one(i64 n, const i64* rp, const i64* ci, const double* v, i8* spins,
                           unsigned seed, const double* beta, const i64* mcs, i64 L) {
  const i64 n1 = n + 1;
  std::vector<double> s(n1), dE(n1);
  for (i64 i = 0; i < n; ++i) s[i] = spins[i];
  s[n] = 1;
  for (i64 i = 0; i < n1; ++i) {
    double acc = 0;
    for (i64 e = rp[i]; e < rp[i + 1]; ++e) acc += v[e] * s[ci[e]];
    dE[i] = -2.0 * s[i] * acc;
  }
  RefXorshift rng(seed);
  std::uniform_real_distribution<> urd(0, 1.0);
  for (i64 l = 0; l < L; ++l) {
    const double b = beta[l];
    for (i64 step = 0; step < mcs[l]; ++step) {
      for (i64 i = 0; i < n; ++i) {
        if (dE[i] <= 0 || std::exp(-b * dE[i]) > urd(rng)) {
          const double c = 4 * s[i];
          for (i64 e = rp[i]; e < rp[i + 1]; ++e) dE[ci[e]] += c * (v[e] * s[ci[e]]);
          dE[i] *= -1;
          s[i] *= -1;
        }
        s[n] = 1;
      }
    }
  }
  for (i64 i = 0; i < n; ++i) spins[i] = (i8)(s[i] * s[n]);
}

int orc_ref_sa_csr(i64 n, const i64* rowptr, const i64* col, const double* val, i8* spins,
                   u64 seed, const double* beta, const i64* mcs, i64 L) {
  if (n < 1) return 1;
  ref_sa_csr_one(n, rowptr, col, val, spins, (unsigned)seed, beta, mcs, L);
  return 0;
}

int orc_ref_sa_csr_batch(i64 n, const i64* rowptr, const i64* col, const double* val, i64 R,
                         i8* spins, const u32* seeds, const double* beta, const i64* mcs,
                         i64 L, int nthreads) {
  if (n < 1) return 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (i64 r = 0; r < R; ++r)
    ref_sa_csr_one(n, rowptr, col, val, spins + r * n, seeds[r], beta, mcs, L);
  return 0;
}

// Trotter sweep, dense.  tspins is M x n (slice-major), spins +-1.
// include/openjij/updater/single_spin_flip.hpp:128-251.
int orc_ref_sqa_dense(i64 n, const double* J, i64 M, double gamma, i8* tspins, u64 seed,
                      const double* beta, const double* sarr, const i64* mcs, i64 L) {
  if (M < 2 || n < 1) return 1;
  const i64 n1 = n + 1;
  // column-major (n+1) x M, as the reference's TrotterMatrix
  std::vector<double> S(n1 * M), dE(n1 * M), rp(n1 * M);
  for (i64 t = 0; t < M; ++t) {
    for (i64 i = 0; i < n; ++i) S[t * n1 + i] = tspins[t * n + i];
    S[t * n1 + n] = 1;
  }
  for (i64 t = 0; t < M; ++t)
    for (i64 i = 0; i < n1; ++i) {
      double acc = 0;
      for (i64 j = 0; j < n1; ++j) acc += J[i * n1 + j] * S[t * n1 + j];
      dE[t * n1 + i] = -2 * S[t * n1 + i] * acc;
    }
  RefXorshift rng((unsigned)seed);
  std::uniform_real_distribution<> urd(0, 1.0);
  auto do_calc = [&](i64 i, i64 t, double b, double s, double B) {
    const i64 tp = (t + 1 + M) % M, tm = (t - 1 + M) % M;
    double dEtrot = -2 * S[t * n1 + i] * (S[tp * n1 + i] + S[tm * n1 + i]);
    double d = s * (b / M) * dE[t * n1 + i] + B * dEtrot;
    if (d < 0 || std::exp(-d) > rp[t * n1 + i]) {
      const double c = 4 * S[t * n1 + i];
      const double* row = J + i * n1;
      for (i64 j = 0; j < n1; ++j) dE[t * n1 + j] += c * (row[j] * S[t * n1 + j]);
      dE[t * n1 + i] *= -1;
      S[t * n1 + i] *= -1;
    }
  };
  for (i64 l = 0; l < L; ++l) {
    const double b = beta[l], s = sarr[l];
    double B = (1 / 2.) * std::log(std::tanh(b * gamma * (1.0 - s) / M));
    for (i64 step = 0; step < mcs[l]; ++step) {
      for (i64 t = 0; t < M; ++t)
        for (i64 i = 0; i < n; ++i) rp[t * n1 + i] = urd(rng);
      i64 upper = (M % 2 != 0) ? M - 1 : M;
      for (i64 t = 0; t < upper; t += 2)
        for (i64 i = 0; i < n; ++i) do_calc(i, t, b, s, B);
      for (i64 t = 1; t < M; t += 2)
        for (i64 i = 0; i < n; ++i) do_calc(i, t, b, s, B);
      if (M % 2 != 0)
        for (i64 i = 0; i < n; ++i) do_calc(i, M - 1, b, s, B);
    }
  }
  for (i64 t = 0; t < M; ++t)
    for (i64 i = 0; i < n; ++i) tspins[t * n + i] = (i8)S[t * n1 + i];
  return 0;
}

// get_solution for a Trotter system (include/openjij/result/get_solution.hpp:59-84):
// the first slice minimising s^T J_ext s over the full (n+1) layout.
int orc_ref_solution_transverse(i64 n, const double* J, i64 M, const i8* tspins, i8* out) {
  const i64 n1 = n + 1;
  std::vector<double> s(n1);
  double best = std::numeric_limits<double>::max();
  i64 tb = 0;
  for (i64 t = 0; t < M; ++t) {
    for (i64 i = 0; i < n; ++i) s[i] = tspins[t * n + i];
    s[n] = 1;
    double e = 0;
    for (i64 i = 0; i < n1; ++i) {
      double acc = 0;
      for (i64 j = 0; j < n1; ++j) acc += J[i * n1 + j] * s[j];
      e += s[i] * acc;
    }
    if (e < best) {
      best = e;
      tb = t;
    }
  }
  for (i64 i = 0; i < n; ++i) out[i] = tspins[tb * n + i];
  return (int)tb;
}

// ===========================================================================
// Exact energies: E = sum_{i<j} J_ij s_i s_j + sum_i h_i s_i (the reference's
// Dense::energy, include/openjij/graph/dense.hpp:176-192, without its rounding).
double orc_energy_exact_dense(i64 n, const double* J, const i8* s) {
  const i64 n1 = n + 1;
  DD acc;
  for (i64 i = 0; i < n; ++i) {
    for (i64 j = i + 1; j < n; ++j) {
      double v = J[i * n1 + j];
      if (v != 0) acc.add(s[i] * s[j] > 0 ? v : -v);
    }
    double h = J[i * n1 + n];
    if (h != 0) acc.add(s[i] > 0 ? h : -h);
  }
  return acc.value();
}

void orc_energy_exact_dense_batch(i64 n, const double* J, i64 R, const i8* s, double* out) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (i64 r = 0; r < R; ++r) out[r] = orc_energy_exact_dense(n, J, s + r * n);
}

// CSR (n+1 rows, both triangles).  Off-diagonal pairs are counted once (i<j).
double orc_energy_exact_csr(i64 n, const i64* rp, const i64* ci, const double* v, const i8* s) {
  DD acc;
  for (i64 i = 0; i < n; ++i)
    for (i64 e = rp[i]; e < rp[i + 1]; ++e) {
      i64 j = ci[e];
      double w = v[e];
      if (w == 0) continue;
      if (j == n) acc.add(s[i] > 0 ? w : -w);
      else if (j > i) acc.add(s[i] * s[j] > 0 ? w : -w);
    }
  return acc.value();
}

// ===========================================================================
// Brute force (n <= 24): Gray-code walk over all 2^n states with incremental
// fields; the best state's energy is then recomputed exactly.  Ties keep the
// first state visited.
int orc_brute_force_dense(i64 n, const double* J, double* emin, i8* best) {
  if (n < 1 || n > 24) return 1;
  const i64 n1 = n + 1;
  // symmetric couplings from the upper triangle (the lower one is ignored)
  std::vector<double> Js((size_t)(n * n), 0.0);
  for (i64 i = 0; i < n; ++i)
    for (i64 j = i + 1; j < n; ++j) Js[i * n + j] = Js[j * n + i] = J[i * n1 + j];
  std::vector<double> s(n, 1.0), F(n);
  for (i64 i = 0; i < n; ++i) {
    double acc = J[i * n1 + n];
    for (i64 j = 0; j < n; ++j) acc += Js[i * n + j] * s[j];
    F[i] = acc;
  }
  std::vector<i8> cur(n, 1);
  double E = orc_energy_exact_dense(n, J, cur.data());
  double bestE = E;
  std::vector<i8> bs(cur);
  const u64 total = 1ull << n;
  for (u64 k = 1; k < total; ++k) {
    int i = __builtin_ctzll(k);
    E += -2.0 * s[i] * F[i];
    const double c = -2.0 * s[i];
    for (i64 j = 0; j < n; ++j) F[j] += c * Js[i * n + j];
    s[i] = -s[i];
    cur[i] = (i8)s[i];
    if (E < bestE - 1e-9 * (1.0 + std::fabs(bestE))) {
      bestE = E;
      bs = cur;
    }
  }
  // refine: exact energy of the winner
  *emin = orc_energy_exact_dense(n, J, bs.data());
  std::memcpy(best, bs.data(), n);
  return 0;
}

// ===========================================================================
// GPU-rule replay.
//   mode 0: fp64 fields, fp64 couplings, theta in fp64
//   mode 1: fp32 fields, fp32 couplings (J rounded to float), theta in fp32
//   mode 2: int32 fields, integer couplings, theta in fp64 (so that for
//           integer couplings modes 0 and 2 give the same trajectory)
// Fields F_j = h_j + sum_k J_jk s_k are initialised lane-sequentially:
// F = h; for k = 0..n-1: F += s_k * J[k][.]; a flip of spin i adds
// -2 s_i(old) * J[i][.].  The spin of index i is accepted iff
// dE_i = -2 s_i F_i < theta(u(r,t,0,i)); rule 1 (heat bath) uses theta_hb.
// nbinv[t] = -1/beta_t for each expanded sweep t (fp32 modes use (float)nbinv).

}  // extern "C"
template <typename F>
static void replay_dense_T(int mode, i64 n, const double* J, i64 R, i8* spins, u64 seed,
                           u64 t0, const double* nbinv, i64 T, u32 word3, int rule,
                           const double* F_init = nullptr) {
  const i64 n1 = n + 1;
  std::vector<F> Jc((size_t)n * n), h(n);
  for (i64 i = 0; i < n; ++i) {
    for (i64 j = 0; j < n; ++j) Jc[i * n + j] = (F)J[i * n1 + j];
    h[i] = (F)J[i * n1 + n];
  }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (i64 r = 0; r < R; ++r) {
    i8* s = spins + r * n;
    std::vector<F> Fv(h);
    if (F_init) {
      // the fields the run starts from, given exactly in the field type
      for (i64 j = 0; j < n; ++j) Fv[j] = (F)F_init[r * n + j];
    } else {
      for (i64 k = 0; k < n; ++k) {
        const F* row = &Jc[k * n];
        if (s[k] > 0)
          for (i64 j = 0; j < n; ++j) Fv[j] = Fv[j] + row[j];
        else
          for (i64 j = 0; j < n; ++j) Fv[j] = Fv[j] - row[j];
      }
    }
    for (i64 t = 0; t < T; ++t) {
      const u64 tg = t0 + t;
      const double nb = nbinv[t];
      const float nbf = (float)nb;
      for (i64 i = 0; i < n; ++i) {
        const u32 u = philox_word(seed, tg, r, word3, i);
        const F dE = (F)(-2 * s[i]) * Fv[i];
        bool acc;
        if (mode == 1) acc = (float)dE < theta_rule_f(rule, u, nbf);
        else acc = (double)dE < theta_rule_d(rule, u, nb);
        if (acc) {
          const F c = (F)(-2 * s[i]);
          const F* row = &Jc[i * n];
          for (i64 j = 0; j < n; ++j) Fv[j] = Fv[j] + c * row[j];
          s[i] = (i8)-s[i];
        }
      }
    }
  }
}
extern "C" {

// word3: the fourth Philox counter word (0 for a single problem; the problem
// index inside a batch of problems).
// rule: 0 = Metropolis, 1 = heat bath.
int orc_replay_dense_wr(int mode, i64 n, const double* J_ext, i64 R, i8* spins, u64 seed,
                        u64 sweep_offset, const double* nbinv, i64 T, u32 word3, int rule) {
  if (n < 1 || R < 0 || T < 0 || rule < 0 || rule > 1) return 1;
  if (mode == 0)
    replay_dense_T<double>(0, n, J_ext, R, spins, seed, sweep_offset, nbinv, T, word3, rule);
  else if (mode == 1)
    replay_dense_T<float>(1, n, J_ext, R, spins, seed, sweep_offset, nbinv, T, word3, rule);
  else if (mode == 2) {
    if (!is_integer_matrix(J_ext, (n + 1) * (n + 1))) return 2;
    replay_dense_T<int32_t>(2, n, J_ext, R, spins, seed, sweep_offset, nbinv, T, word3, rule);
  } else return 1;
  return 0;
}

// As orc_replay_dense_wr, starting every replica r from the fields
// F_init[r][0..n-1] instead of accumulating them from the spins (the engine's
// matrix-core refresh points).  FP32 mode rounds them to float, which is exact
// for fields downloaded from an FP32 context.
int orc_replay_dense_from_fields(int mode, i64 n, const double* J_ext, i64 R, i8* spins,
                                 const double* F_init, u64 seed, u64 sweep_offset,
                                 const double* nbinv, i64 T, u32 word3, int rule) {
  if (n < 1 || R < 0 || T < 0 || rule < 0 || rule > 1 || !F_init) return 1;
  if (mode == 0)
    replay_dense_T<double>(0, n, J_ext, R, spins, seed, sweep_offset, nbinv, T, word3, rule, F_init);
  else if (mode == 1)
    replay_dense_T<float>(1, n, J_ext, R, spins, seed, sweep_offset, nbinv, T, word3, rule, F_init);
  else if (mode == 2) {
    if (!is_integer_matrix(J_ext, (n + 1) * (n + 1))) return 2;
    replay_dense_T<int32_t>(2, n, J_ext, R, spins, seed, sweep_offset, nbinv, T, word3, rule,
                            F_init);
  } else return 1;
  return 0;
}

int orc_replay_dense_w(int mode, i64 n, const double* J_ext, i64 R, i8* spins, u64 seed,
                       u64 sweep_offset, const double* nbinv, i64 T, u32 word3) {
  return orc_replay_dense_wr(mode, n, J_ext, R, spins, seed, sweep_offset, nbinv, T, word3, 0);
}

int orc_replay_dense(int mode, i64 n, const double* J_ext, i64 R, i8* spins, u64 seed,
                     u64 sweep_offset, const double* nbinv, i64 T) {
  return orc_replay_dense_w(mode, n, J_ext, R, spins, seed, sweep_offset, nbinv, T, 0);
}

// CSR replay: the field of spin i is recomputed at each attempt in stored
// order, F_i = h_i + sum_e val_e s_col(e) (diagonal and dummy-column entries
// excluded, h_i taken from the dummy column).  rowptr/col/val describe the
// (n+1)-row symmetric layout of the reference's CSRSparse.
}  // extern "C"
template <typename F>
static void replay_csr_T(int mode, i64 n, const i64* rp, const i64* ci, const double* v, i64 R,
                         i8* spins, u64 seed, u64 t0, const double* nbinv, i64 T, int rule) {
  std::vector<F> h(n, (F)0);
  std::vector<i64> rp2(n + 1, 0), c2;
  std::vector<F> v2;
  for (i64 i = 0; i < n; ++i) {
    rp2[i] = (i64)c2.size();
    for (i64 e = rp[i]; e < rp[i + 1]; ++e) {
      if (ci[e] == n) h[i] = (F)v[e];
      else if (ci[e] != i) {
        c2.push_back(ci[e]);
        v2.push_back((F)v[e]);
      }
    }
  }
  rp2[n] = (i64)c2.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (i64 r = 0; r < R; ++r) {
    i8* s = spins + r * n;
    for (i64 t = 0; t < T; ++t) {
      const u64 tg = t0 + t;
      const double nb = nbinv[t];
      const float nbf = (float)nb;
      for (i64 i = 0; i < n; ++i) {
        F acc = h[i];
        for (i64 e = rp2[i]; e < rp2[i + 1]; ++e)
          acc = s[c2[e]] > 0 ? acc + v2[e] : acc - v2[e];
        const u32 u = philox_word(seed, tg, r, 0, i);
        const F dE = (F)(-2 * s[i]) * acc;
        bool accp;
        if (mode == 1) accp = (float)dE < theta_rule_f(rule, u, nbf);
        else accp = (double)dE < theta_rule_d(rule, u, nb);
        if (accp) s[i] = (i8)-s[i];
      }
    }
  }
}
extern "C" {

int orc_replay_csr_r(int mode, i64 n, const i64* rowptr, const i64* col, const double* val,
                     i64 R, i8* spins, u64 seed, u64 sweep_offset, const double* nbinv, i64 T,
                     int rule) {
  if (n < 1 || R < 0 || T < 0 || rule < 0 || rule > 1) return 1;
  if (mode == 0)
    replay_csr_T<double>(0, n, rowptr, col, val, R, spins, seed, sweep_offset, nbinv, T, rule);
  else if (mode == 1)
    replay_csr_T<float>(1, n, rowptr, col, val, R, spins, seed, sweep_offset, nbinv, T, rule);
  else if (mode == 2) {
    if (!is_integer_matrix(val, rowptr[n + 1])) return 2;
    replay_csr_T<int32_t>(2, n, rowptr, col, val, R, spins, seed, sweep_offset, nbinv, T, rule);
  } else return 1;
  return 0;
}

int orc_replay_csr(int mode, i64 n, const i64* rowptr, const i64* col, const double* val, i64 R,
                   i8* spins, u64 seed, u64 sweep_offset, const double* nbinv, i64 T) {
  return orc_replay_csr_r(mode, n, rowptr, col, val, R, spins, seed, sweep_offset, nbinv, T, 0);
}

// Trotter replay, dense.  spins: R x M x n.  Per sweep t: a[t] = s*(beta/M),
// B[t] = 0.5*log(tanh(beta*gamma*(1-s)/M)) (host-computed, as the GPU gets
// them).  Slices: even phase, odd phase, then slice M-1 if M is odd.
//   dEcl = -2 s F,  dEtr = -2 s (s_up + s_dn),  x = B*dEtr,
//   dE = fma(a, dEcl, x), accept iff dE < -log(U).
}  // extern "C"
template <typename F>
static void replay_sqa_dense_T(int mode, i64 n, const double* J, i64 R, i64 M, i8* spins,
                               u64 seed, u64 t0, const double* a, const double* B, i64 T) {
  const i64 n1 = n + 1;
  std::vector<F> Jc((size_t)n * n), h(n);
  for (i64 i = 0; i < n; ++i) {
    for (i64 j = 0; j < n; ++j) Jc[i * n + j] = (F)J[i * n1 + j];
    h[i] = (F)J[i * n1 + n];
  }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (i64 r = 0; r < R; ++r) {
    i8* S = spins + r * M * n;
    std::vector<F> Fv((size_t)M * n);
    for (i64 m = 0; m < M; ++m) {
      F* Fm = &Fv[m * n];
      for (i64 j = 0; j < n; ++j) Fm[j] = h[j];
      for (i64 k = 0; k < n; ++k) {
        const F* row = &Jc[k * n];
        if (S[m * n + k] > 0)
          for (i64 j = 0; j < n; ++j) Fm[j] = Fm[j] + row[j];
        else
          for (i64 j = 0; j < n; ++j) Fm[j] = Fm[j] - row[j];
      }
    }
    for (i64 t = 0; t < T; ++t) {
      const u64 tg = t0 + t;
      const F at = (F)a[t], Bt = (F)B[t];
      auto slice = [&](i64 m) {
        i8* s = S + m * n;
        const i8* su = S + ((m + 1) % M) * n;
        const i8* sd = S + ((m - 1 + M) % M) * n;
        F* Fm = &Fv[m * n];
        for (i64 i = 0; i < n; ++i) {
          const u32 u = philox_word(seed, tg, r, m, i);
          const F dEcl = (F)(-2 * s[i]) * Fm[i];
          const F dEtr = (F)(-2 * s[i] * (su[i] + sd[i]));
          const F x = Bt * dEtr;
          const F dE = std::fma(at, dEcl, x);
          bool acc;
          if (mode == 0) acc = (double)dE < -det_log_d(uniform_d(u));
          else acc = (float)dE < -det_log_f(uniform_f(u));
          if (acc) {
            const F c = (F)(-2 * s[i]);
            const F* row = &Jc[i * n];
            for (i64 j = 0; j < n; ++j) Fm[j] = Fm[j] + c * row[j];
            s[i] = (i8)-s[i];
          }
        }
      };
      const i64 upper = (M % 2 != 0) ? M - 1 : M;
      for (i64 m = 0; m < upper; m += 2) slice(m);
      for (i64 m = 1; m < M; m += 2) slice(m);
      if (M % 2 != 0) slice(M - 1);
    }
  }
}
extern "C" {

int orc_replay_sqa_dense(int mode, i64 n, const double* J_ext, i64 R, i64 M, i8* spins, u64 seed,
                         u64 sweep_offset, const double* a, const double* B, i64 T) {
  if (n < 1 || M < 2 || R < 0 || T < 0) return 1;
  if (mode == 0) replay_sqa_dense_T<double>(0, n, J_ext, R, M, spins, seed, sweep_offset, a, B, T);
  else if (mode == 1) replay_sqa_dense_T<float>(1, n, J_ext, R, M, spins, seed, sweep_offset, a, B, T);
  else return 1;
  return 0;
}

// Trotter replay, CSR (fields recomputed per attempt, as in orc_replay_csr).
}  // extern "C"
template <typename F>
static void replay_sqa_csr_T(int mode, i64 n, const i64* rp, const i64* ci, const double* v,
                             i64 R, i64 M, i8* spins, u64 seed, u64 t0, const double* a,
                             const double* B, i64 T) {
  std::vector<F> h(n, (F)0);
  std::vector<i64> rp2(n + 1, 0), c2;
  std::vector<F> v2;
  for (i64 i = 0; i < n; ++i) {
    rp2[i] = (i64)c2.size();
    for (i64 e = rp[i]; e < rp[i + 1]; ++e) {
      if (ci[e] == n) h[i] = (F)v[e];
      else if (ci[e] != i) {
        c2.push_back(ci[e]);
        v2.push_back((F)v[e]);
      }
    }
  }
  rp2[n] = (i64)c2.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (i64 r = 0; r < R; ++r) {
    i8* S = spins + r * M * n;
    for (i64 t = 0; t < T; ++t) {
      const u64 tg = t0 + t;
      const F at = (F)a[t], Bt = (F)B[t];
      auto slice = [&](i64 m) {
        i8* s = S + m * n;
        const i8* su = S + ((m + 1) % M) * n;
        const i8* sd = S + ((m - 1 + M) % M) * n;
        for (i64 i = 0; i < n; ++i) {
          F acc = h[i];
          for (i64 e = rp2[i]; e < rp2[i + 1]; ++e)
            acc = s[c2[e]] > 0 ? acc + v2[e] : acc - v2[e];
          const u32 u = philox_word(seed, tg, r, m, i);
          const F dEcl = (F)(-2 * s[i]) * acc;
          const F dEtr = (F)(-2 * s[i] * (su[i] + sd[i]));
          const F x = Bt * dEtr;
          const F dE = std::fma(at, dEcl, x);
          bool accp;
          if (mode == 0) accp = (double)dE < -det_log_d(uniform_d(u));
          else accp = (float)dE < -det_log_f(uniform_f(u));
          if (accp) s[i] = (i8)-s[i];
        }
      };
      const i64 upper = (M % 2 != 0) ? M - 1 : M;
      for (i64 m = 0; m < upper; m += 2) slice(m);
      for (i64 m = 1; m < M; m += 2) slice(m);
      if (M % 2 != 0) slice(M - 1);
    }
  }
}
extern "C" {

int orc_replay_sqa_csr(int mode, i64 n, const i64* rowptr, const i64* col, const double* val,
                       i64 R, i64 M, i8* spins, u64 seed, u64 sweep_offset, const double* a,
                       const double* B, i64 T) {
  if (n < 1 || M < 2 || R < 0 || T < 0) return 1;
  if (mode == 0) replay_sqa_csr_T<double>(0, n, rowptr, col, val, R, M, spins, seed, sweep_offset, a, B, T);
  else if (mode == 1) replay_sqa_csr_T<float>(1, n, rowptr, col, val, R, M, spins, seed, sweep_offset, a, B, T);
  else return 1;
  return 0;
}

}  // extern "C"
