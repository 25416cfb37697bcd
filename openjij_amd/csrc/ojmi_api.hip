// ojmi_api.hip -- the C ABI (include/ojmi.h): contexts, validation, layout
// conversion, schedule expansion and kernel dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/ojmi.h"
#include "ojmi_kernels.h"
#include "ojmi_persist.h"
#include "ojmi_tuned.h"
#include "ojmi_wave.h"

using namespace ojmi;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define OJMI_HIP(call)                                                               \
  do {                                                                               \
    hipError_t _e = (call);                                                          \
    if (_e != hipSuccess)                                                            \
      return fail(_e == hipErrorOutOfMemory ? OJMI_ENOMEM : OJMI_EDEVICE,            \
                  std::string(#call) + ": " + hipGetErrorString(_e));                \
  } while (0)

template <typename T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? OJMI_ENOMEM : OJMI_EDEVICE,
                std::string("hipMalloc: ") + hipGetErrorString(e));
  return OJMI_OK;
}

// bfloat16 of a float, round to nearest even (finite inputs), and its value.
uint16_t bf16_rn(float x) {
  uint32_t b;
  std::memcpy(&b, &x, 4);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}
float bf16_value(uint16_t v) {
  const uint32_t b = (uint32_t)v << 16;
  float x;
  std::memcpy(&x, &b, 4);
  return x;
}

int pow2_slots(int64_t n) {
  int64_t need = (n + 63) / 64;
  int ns = 1;
  while (ns < need) ns <<= 1;
  return ns;
}

}  // namespace

struct ojmi_ctx {
  int kind = 0;  // 0 dense, 1 csr
  int64_t n = 0, R = 0, M = 1;
  double gamma = 0;
  int prec = OJMI_PREC_FP64;
  int family = OJMI_KERNEL_DENSE_WAVE;
  int ftype = FT_F64, jtype = JT_F64;
  int NS = 1;
  int64_t Npad = 64;
  int sqa_ppw = 1, sqa_w = 1;
  int csr_wpb = 1, csr_waves = 0;
  int rule = RULE_METROPOLIS;  // classical acceptance rule
  int refresh_every = 0;
  int group = 1;         // 1: one wave per replica; 4/8/16: replicas per workgroup
  int xcd_window = 2;    // shared kernel: XCD pacing window in blocks (0 = off)
  int lds_budget = 65536;
  int ring_rows_opt = 0, round_rows_opt = 0, ring_rows_used = 0, lds_used = 0;
  int colown = 1;  // shared family: 1 = column-owned kernel when the shape allows
  // dense classical FP32 / INT32 (int8 couplings): 1 = row updates on the
  // matrix cores (k_dense_mfma), when the shape allows
  // (2: per 16-sweep segment, by the acceptance rate the GPU measured in the
  // previous one, against the column-owned kernel)
  int mfma_sweep = 0;
  int mfma_accept = 85;  // per mille: the matrix-core sweep at or above this rate
  int mfma_form = 1;     // 1: four waves update the fields while four decide; 0: all do both
  int sqa_rows_lds = 1;  // Trotter kernel: first slots of every coupling row in LDS (0: none)
  int csr_pack8 = 1;     // INT32 CSR sweeps: int8 values packed four per word (0: int32 values)
  int sqa_csr_records = 1;  // CSR Trotter kernel: per-position records where built (0: word loads)
  void* dJpk = nullptr;  // packed MFMA operands of the couplings (k_dense_mfma)
  int32_t* dsel = nullptr;  // segment kernel choice and flip snapshot (mode 2)
  int lds_rows = 1;      // wave family: coupling rows staged in LDS when they fit 64 KiB
  int lds_rows_wpb = 0;  // replicas per workgroup of that form (0: by R)
  // wave family: threshold helper waves per replica (0: none; -1: 2 while the
  // replicas fit two workgroups per CU and n <= 512, else none)
  int wave_helpers = -1;
  int helpers_used = 0;
  int num_cus = 256;
  void* dJnat = nullptr;   // natural-layout couplings when dJ is packed
  uint32_t* dprog = nullptr;
  int64_t prog_cap = 0;

  // kernel-layout couplings
  void* dJ = nullptr;
  void* dh = nullptr;
  // FP64 natural padded couplings for energies (Npad x Npad) and h (dense)
  double* dJ64 = nullptr;
  double* dh64 = nullptr;
  // CSR level-order data; FP64 neighbour values and fields at level-order
  // positions for energies (the sweep's own arrays when those are FP64)
  int32_t *d_order = nullptr, *d_lvl = nullptr, *d_deg = nullptr, *d_ell = nullptr;
  void* d_ellval = nullptr;
  double* d_ellv64 = nullptr;
  uint32_t* d_ell8 = nullptr;  // INT32 CSR: values in int8, four per word ([Dp/4][n]), or null
  uint4 *d_rec = nullptr, *d_rec8 = nullptr;  // CSR workgroup form: position records, or null
  double* d_hpos64 = nullptr;
  int nlev = 0, D = 0;

  int8_t* dspins = nullptr;
  double* dE = nullptr;    // [C]
  double* ddrift = nullptr;
  void* dF = nullptr;      // [R][Npad] fields (FP32 or int32): MFMA init / refresh, sweep output
  bool fields_valid = false;  // dF holds the fields of the current spins (FP32 / FP64 dense)
  bool energies_valid = false;  // dE / ddrift hold the current spins' energies and drift
  // matrix-core operands: bf16 hi/lo couplings (FP32 path), spins zero-padded
  // to [C][Npad] as int8 and bf16, energy partials
  uint16_t* dJhi = nullptr;
  uint16_t* dJlo = nullptr;
  int8_t* dS8 = nullptr;
  uint16_t* dSbf = nullptr;
  double* dpart = nullptr;  // 2 x energy_part_count(C, Npad)
  std::vector<hipEvent_t> ev_ref;  // refresh begin/end pairs of the last run
  int64_t last_refreshes = 0;
  unsigned long long* dstats = nullptr;  // [2] accepted flips, coupling rows streamed
  int64_t last_flips = 0, last_rows = 0;

  double *dp0_d = nullptr, *dp1_d = nullptr;
  float *dp0_f = nullptr, *dp1_f = nullptr;
  int64_t Tcap = 0;

  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int device = 0;  // current device when the context was created

  int64_t last_T = 0, last_attempts = 0;
  int64_t last_mfma_segments = 0;  // segments of the last run on the matrix-core sweep
  double last_ms = 0, last_refresh_ms = 0;

  // batch of independent problems (kind 2): per-problem sizes and offsets
  int64_t K = 0;
  std::vector<int32_t> bn;
  int32_t* d_bn = nullptr;
  int64_t *d_bj_off = nullptr, *d_bh_off = nullptr, *d_bs_off = nullptr, *d_brb = nullptr;
  std::vector<std::pair<int, int32_t*>> bclass;  // (slot class NS, device problem list)
  std::vector<int> bclass_count;
  // persistent batch launch: per-XCD LPT queues of (problem, replica) items
  // 0: one launch of every problem, largest first (the hardware's in-order
  // workgroup dispatch schedules them); 1: one launch per slot class;
  // 2: one persistent launch over per-XCD software queues
  int batch_mode = 0;
  int32_t* d_fused = nullptr;  // every problem, largest first
  int batch_waves = 4;  // waves per workgroup of the persistent launch
  int batch_ns_max = 1;
  int32_t *d_q_items = nullptr, *d_q_begin = nullptr;
  uint32_t* d_q_head = nullptr;
  double* db_p0_d = nullptr;
  float* db_p0_f = nullptr;
  int64_t bTcap = 0;

  ~ojmi_ctx() {
    if (dJ == dJ64) dJ = nullptr;
    if (dh == dh64) dh = nullptr;
    if (d_ellv64 == d_ellval) d_ellv64 = nullptr;
    if (d_hpos64 == dh) d_hpos64 = nullptr;
    void* bufs[] = {dJ, dh, dJ64, dh64, d_order, d_lvl, d_deg, d_ell, d_ellval, dspins,
                    d_ellv64, d_hpos64, d_ell8, d_rec, d_rec8,
                    dE, ddrift, dF, dp0_d, dp1_d, dp0_f, dp1_f, dJnat, dprog,
                    d_bn, d_bj_off, d_bh_off, d_bs_off, d_brb, db_p0_d, db_p0_f,
                    dJhi, dJlo, dS8, dSbf, dpart, dstats, d_q_items, d_q_begin, d_q_head,
                    d_fused, dJpk, dsel};
    for (void* b : bufs)
      if (b) (void)hipFree(b);
    for (auto& cl : bclass)
      if (cl.second) (void)hipFree(cl.second);
    for (hipEvent_t e : ev_ref) (void)hipEventDestroy(e);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

// Read the upper triangle of the dense (n+1)^2 layout: Jup[i][j] for j>i
// mirrored, h from column n.
int check_dense_input(int64_t n, const double* J_ext) {
  const int64_t n1 = n + 1;
  for (int64_t i = 0; i < n; ++i) {
    if (J_ext[i * n1 + i] != 0.0)
      return fail(OJMI_ERUNTIME, "The diagonal elements of interaction matrix must be zero.");
    for (int64_t j = i; j < n1; ++j)
      if (!std::isfinite(J_ext[i * n1 + j]))
        return fail(OJMI_EINVAL, "interaction matrix holds a non-finite value");
  }
  return OJMI_OK;
}

bool all_integer(const std::vector<double>& v) {
  for (double x : v)
    if (x != std::floor(x) || std::fabs(x) > 1e9) return false;
  return true;
}

int resolve_precision(int32_t req, bool integer, double max_abs_field, int* out) {
  if (req == OJMI_PREC_AUTO) {
    *out = (integer && 2.0 * max_abs_field < 2147483647.0) ? OJMI_PREC_INT32 : OJMI_PREC_FP64;
    return OJMI_OK;
  }
  if (req == OJMI_PREC_INT32) {
    if (!integer) return fail(OJMI_EINVAL, "INT32 precision needs integer couplings");
    if (2.0 * max_abs_field >= 2147483647.0)
      return fail(OJMI_EINVAL, "INT32 precision: local fields may overflow int32");
    *out = req;
    return OJMI_OK;
  }
  if (req == OJMI_PREC_FP64 || req == OJMI_PREC_FP32) {
    *out = req;
    return OJMI_OK;
  }
  return fail(OJMI_EINVAL, "unknown precision");
}

// LDS a workgroup may use on the current device (at least 64 KiB).
int device_lds_budget() {
  int dev = 0, lds = 0, optin = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    lds = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess)
    optin = 0;
  return std::max(65536, std::max(lds, optin));
}

int common_init(ojmi_ctx* c) {
  OJMI_HIP(hipGetDevice(&c->device));
  c->lds_budget = device_lds_budget();
  if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, c->device) !=
          hipSuccess ||
      c->num_cus < 1)
    c->num_cus = 256;
  OJMI_HIP(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  OJMI_HIP(hipEventCreate(&c->e0));
  OJMI_HIP(hipEventCreate(&c->e1));
  int rc;
  if ((rc = dalloc(&c->dspins, (size_t)(c->R * c->M * c->n))) != OJMI_OK) return rc;
  std::vector<int8_t> ones((size_t)(c->R * c->M * c->n), 1);
  OJMI_HIP(hipMemcpy(c->dspins, ones.data(), ones.size(), hipMemcpyHostToDevice));
  if ((rc = dalloc(&c->dE, (size_t)(c->R * c->M))) != OJMI_OK) return rc;
  if ((rc = dalloc(&c->ddrift, (size_t)(c->R * c->M))) != OJMI_OK) return rc;
  return OJMI_OK;
}

// Upload the dense couplings in the kernel layout of (ftype, jtype).
int upload_dense(ojmi_ctx* c, const std::vector<double>& Jn /* n x n symmetric */,
                 const std::vector<double>& h) {
  const int64_t n = c->n, Np = c->Npad;
  const int NS = c->NS;
  int rc;
  // FP64 natural padded copy (energies, drift)
  {
    std::vector<double> J64((size_t)(Np * Np), 0.0), h64((size_t)Np, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      std::memcpy(&J64[i * Np], &Jn[i * n], n * sizeof(double));
      h64[i] = h[i];
    }
    if ((rc = dalloc(&c->dJ64, J64.size())) != OJMI_OK) return rc;
    if ((rc = dalloc(&c->dh64, h64.size())) != OJMI_OK) return rc;
    OJMI_HIP(hipMemcpy(c->dJ64, J64.data(), J64.size() * 8, hipMemcpyHostToDevice));
    OJMI_HIP(hipMemcpy(c->dh64, h64.data(), h64.size() * 8, hipMemcpyHostToDevice));
  }
  if (c->jtype == JT_F64) {
    c->dJ = c->dJ64;
    c->dh = c->dh64;
    return OJMI_OK;
  }
  if (c->jtype == JT_F32) {
    std::vector<float> Jf((size_t)(Np * Np), 0.0f), hf((size_t)Np, 0.0f);
    for (int64_t i = 0; i < n; ++i) {
      for (int64_t j = 0; j < n; ++j) Jf[i * Np + j] = (float)Jn[i * n + j];
      hf[i] = (float)h[i];
    }
    float *dJ, *dh;
    if ((rc = dalloc(&dJ, Jf.size())) != OJMI_OK) return rc;
    c->dJ = dJ;
    if ((rc = dalloc(&dh, hf.size())) != OJMI_OK) return rc;
    c->dh = dh;
    OJMI_HIP(hipMemcpy(dJ, Jf.data(), Jf.size() * 4, hipMemcpyHostToDevice));
    OJMI_HIP(hipMemcpy(dh, hf.data(), hf.size() * 4, hipMemcpyHostToDevice));
    if (c->M == 1) {
      // matrix-core operands of the field init / refresh: J = J_hi + J_lo in bf16
      std::vector<uint16_t> hi(Jf.size()), lo(Jf.size());
      for (size_t q = 0; q < Jf.size(); ++q) {
        hi[q] = bf16_rn(Jf[q]);
        lo[q] = bf16_rn(Jf[q] - bf16_value(hi[q]));
      }
      if ((rc = dalloc(&c->dJhi, hi.size())) != OJMI_OK) return rc;
      if ((rc = dalloc(&c->dJlo, lo.size())) != OJMI_OK) return rc;
      OJMI_HIP(hipMemcpy(c->dJhi, hi.data(), hi.size() * 2, hipMemcpyHostToDevice));
      OJMI_HIP(hipMemcpy(c->dJlo, lo.data(), lo.size() * 2, hipMemcpyHostToDevice));
    }
    return OJMI_OK;
  }
  // integer paths: h as int32
  std::vector<int32_t> hi((size_t)Np, 0);
  for (int64_t i = 0; i < n; ++i) hi[i] = (int32_t)h[i];
  int32_t* dh;
  if ((rc = dalloc(&dh, hi.size())) != OJMI_OK) return rc;
  c->dh = dh;
  OJMI_HIP(hipMemcpy(dh, hi.data(), hi.size() * 4, hipMemcpyHostToDevice));
  if (c->jtype == JT_I32) {
    std::vector<int32_t> Ji((size_t)(Np * Np), 0);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < n; ++j) Ji[i * Np + j] = (int32_t)Jn[i * n + j];
    int32_t* dJ;
    if ((rc = dalloc(&dJ, Ji.size())) != OJMI_OK) return rc;
    c->dJ = dJ;
    OJMI_HIP(hipMemcpy(dJ, Ji.data(), Ji.size() * 4, hipMemcpyHostToDevice));
    return OJMI_OK;
  }
  // int8 packed: element (m, l) of row i at byte ((m>>2)*64 + l)*4 + (m&3)
  const int64_t rb = ((NS + 3) / 4) * 256;
  std::vector<int8_t> J8((size_t)(Np * rb), 0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) {
      const int64_t m = j >> 6, l = j & 63;
      J8[i * rb + ((m >> 2) * 64 + l) * 4 + (m & 3)] = (int8_t)Jn[i * n + j];
    }
  int8_t* dJ;
  if ((rc = dalloc(&dJ, J8.size())) != OJMI_OK) return rc;
  c->dJ = dJ;
  OJMI_HIP(hipMemcpy(dJ, J8.data(), J8.size(), hipMemcpyHostToDevice));
  // natural row layout for the workgroup-shared kernel
  std::vector<int8_t> Jn8((size_t)(Np * Np), 0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) Jn8[i * Np + j] = (int8_t)Jn[i * n + j];
  int8_t* dn8;
  if ((rc = dalloc(&dn8, Jn8.size())) != OJMI_OK) return rc;
  c->dJnat = dn8;
  OJMI_HIP(hipMemcpy(dn8, Jn8.data(), Jn8.size(), hipMemcpyHostToDevice));
  return OJMI_OK;
}

int setup_dense_family(ojmi_ctx* c) {
  const int64_t n = c->n;
  c->NS = pow2_slots(n);
  c->Npad = 64LL * c->NS;
  if (c->M == 1) {
    if (c->NS > 64) {
      // fields in LDS, rows of 64 * ceil(n / 64)
      c->NS = (int)((n + 63) / 64);
      c->Npad = 64LL * c->NS;
      const int fb = (c->ftype == FT_F64) ? 8 : 4;
      if (lds_kernel_bytes((int)c->Npad, fb) > (size_t)device_lds_budget())
        return fail(OJMI_ENOTSUP, "dense GPU path: n too large for LDS-resident fields");
      c->family = OJMI_KERNEL_DENSE_LDS;
      return OJMI_OK;
    }
    c->family = OJMI_KERNEL_DENSE_WAVE;
  } else {
    const int64_t P = (c->M + 1) / 2;
    const int words = (c->ftype == FT_F64) ? 2 : 1;
    // 64 slots (n <= 4096) hold one slice's fields in 64 (FP32) or 128 (FP64)
    // registers per lane; the register budget below bounds the slices
    if (c->NS > 64) return fail(OJMI_ENOTSUP, "Trotter GPU path supports n <= 4096 spins");
    auto pick = [&](int wmax, int ppw_max, int budget, int* Wout, int* Pout) -> bool {
      const int W = (int)std::min<int64_t>(P, wmax);
      const int64_t need = (P + W - 1) / W;
      int PPW = 1;
      while (PPW < need) PPW <<= 1;
      if (PPW > ppw_max || 2 * PPW * c->NS * words > budget) return false;
      *Wout = W;
      *Pout = PPW;
      return true;
    };
    int W = 0, PPW = 0;
    // 16 waves per replica when the fields fit 128 VGPRs, else 4 waves with 512
    if (!(P > 4 && pick(16, 2, 48, &W, &PPW)) && !pick(4, 8, 320, &W, &PPW))
      return fail(OJMI_ENOTSUP, "Trotter GPU path: too many spins x slices per replica");
    c->sqa_ppw = PPW;
    c->sqa_w = W;
    c->family = OJMI_KERNEL_SQA_WAVE;
  }
  return OJMI_OK;
}

void pick_types(ojmi_ctx* c, double maxabsJ) {
  if (c->prec == OJMI_PREC_FP64) {
    c->ftype = FT_F64;
    c->jtype = JT_F64;
  } else if (c->prec == OJMI_PREC_FP32) {
    c->ftype = FT_F32;
    c->jtype = JT_F32;
  } else {
    c->ftype = FT_I32;
    c->jtype = (maxabsJ <= 127.0) ? JT_I8 : JT_I32;
  }
}

int ensure_sched(ojmi_ctx* c, int64_t T) {
  if (T <= c->Tcap) return OJMI_OK;
  void* b[] = {c->dp0_d, c->dp1_d, c->dp0_f, c->dp1_f};
  for (void* p : b)
    if (p) (void)hipFree(p);
  c->dp0_d = c->dp1_d = nullptr;
  c->dp0_f = c->dp1_f = nullptr;
  int rc;
  if ((rc = dalloc(&c->dp0_d, (size_t)T)) != OJMI_OK) return rc;
  if ((rc = dalloc(&c->dp1_d, (size_t)T)) != OJMI_OK) return rc;
  if ((rc = dalloc(&c->dp0_f, (size_t)T)) != OJMI_OK) return rc;
  if ((rc = dalloc(&c->dp1_f, (size_t)T)) != OJMI_OK) return rc;
  c->Tcap = T;
  return OJMI_OK;
}

// ---------------------------------------------------------------------------
// Batch of problems.  Problem k's couplings are stored in the wave kernel's
// layout for its own slot class NS_k (n_k rows of row_bytes), its fields'
// constant term as Npad_k entries, its spins as [R][n_k]; all concatenated.
template <typename JT>
static void pack_rows(const double* Je, int64_t n, int NS, std::vector<char>& out, size_t off) {
  const int64_t n1 = n + 1;
  const int64_t rb = (sizeof(JT) == 1) ? (int64_t)((NS + 3) / 4) * 256 : (int64_t)NS * 64 * sizeof(JT);
  for (int64_t i = 0; i < n; ++i) {
    char* row = out.data() + off + (size_t)(i * rb);
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double v = (j > i) ? Je[i * n1 + j] : Je[j * n1 + i];
      if (sizeof(JT) == 1) {
        const int64_t m = j >> 6, l = j & 63;
        row[((m >> 2) * 64 + l) * 4 + (m & 3)] = (char)(int8_t)v;
      } else {
        const JT x = (JT)v;
        std::memcpy(row + j * sizeof(JT), &x, sizeof(JT));
      }
    }
  }
}

static int64_t batch_row_bytes(int jtype, int NS) {
  if (jtype == JT_I8) return (int64_t)((NS + 3) / 4) * 256;
  return (int64_t)NS * 64 * (jtype == JT_F64 ? 8 : 4);
}

int begin_stats(ojmi_ctx* c);
int end_stats(ojmi_ctx* c);

static int run_batch(ojmi_ctx* c, const std::vector<double>& p0 /* [K][T] */, int64_t T,
                     uint64_t seed, uint64_t sweep_offset) {
  c->last_T = T;
  int64_t sum_n = 0;
  for (int32_t v : c->bn) sum_n += v;
  c->last_attempts = T * c->R * sum_n;
  c->last_ms = 0;
  c->last_refresh_ms = 0;
  c->last_refreshes = 0;
  c->last_flips = c->last_rows = 0;
  if (T == 0) return OJMI_OK;
  int rc;
  if (T > c->bTcap) {
    if (c->db_p0_d) (void)hipFree(c->db_p0_d);
    if (c->db_p0_f) (void)hipFree(c->db_p0_f);
    c->db_p0_d = nullptr;
    c->db_p0_f = nullptr;
    c->bTcap = 0;
    if ((rc = dalloc(&c->db_p0_d, (size_t)(c->K * T))) || (rc = dalloc(&c->db_p0_f, (size_t)(c->K * T))))
      return rc;
    c->bTcap = T;
  }
  std::vector<float> pf(p0.size());
  for (size_t q = 0; q < p0.size(); ++q) pf[q] = (float)p0[q];
  OJMI_HIP(hipMemcpyAsync(c->db_p0_d, p0.data(), p0.size() * 8, hipMemcpyHostToDevice, c->st));
  OJMI_HIP(hipMemcpyAsync(c->db_p0_f, pf.data(), pf.size() * 4, hipMemcpyHostToDevice, c->st));
  BatchArgs b{};
  b.J = c->dJ;
  b.h = c->dh;
  b.spins = c->dspins;
  b.j_off = c->d_bj_off;
  b.h_off = c->d_bh_off;
  b.s_off = c->d_bs_off;
  b.n = c->d_bn;
  b.R = (int)c->R;
  b.T = (int)T;
  b.seed = seed;
  b.t0 = sweep_offset;
  b.p0_d = c->db_p0_d;
  b.p0_f = c->db_p0_f;
  b.rule = c->rule;
  if ((rc = begin_stats(c))) return rc;
  b.stats = c->dstats;
  if (c->batch_mode == 2) OJMI_HIP(hipMemsetAsync(c->d_q_head, 0, 8 * sizeof(uint32_t), c->st));
  OJMI_HIP(hipEventRecord(c->e0, c->st));
  if (c->batch_mode == 0) {
    b.inst = c->d_fused;
    b.num_items = (int)c->K;
    hipError_t e = launch_batch_wave(c->ftype, c->jtype, b, c->batch_ns_max, c->st);
    if (e != hipSuccess)
      return fail(OJMI_EDEVICE, std::string("batch kernel launch: ") + hipGetErrorString(e));
  } else if (c->batch_mode == 2) {
    BatchQueues q{c->d_q_items, c->d_q_begin, c->d_q_head};
    const int blocks = batch_persist_blocks(c->ftype, c->jtype, c->batch_ns_max, c->batch_waves);
    if (blocks <= 0) return fail(OJMI_EDEVICE, "persistent batch kernel: no occupancy");
    hipError_t e = launch_batch_persist(c->ftype, c->jtype, b, q, c->batch_ns_max, c->batch_waves,
                                        blocks, c->st);
    if (e != hipSuccess)
      return fail(OJMI_EDEVICE, std::string("batch kernel launch: ") + hipGetErrorString(e));
  } else {
    // largest slot class first: the longest waves start first
    for (size_t q = c->bclass.size(); q-- > 0;) {
      b.inst = c->bclass[q].second;
      b.num_items = c->bclass_count[q];
      hipError_t e = launch_batch_wave(c->ftype, c->jtype, b, c->bclass[q].first, c->st);
      if (e != hipSuccess)
        return fail(OJMI_EDEVICE, std::string("batch kernel launch: ") + hipGetErrorString(e));
    }
  }
  OJMI_HIP(hipEventRecord(c->e1, c->st));
  OJMI_HIP(hipEventSynchronize(c->e1));
  float ms = 0;
  OJMI_HIP(hipEventElapsedTime(&ms, c->e0, c->e1));
  c->last_ms = ms;
  return end_stats(c);
}

static int batch_energies(ojmi_ctx* c) {
  BatchArgs b{};
  b.J = c->dJ;
  b.h = c->dh;
  b.spins = c->dspins;
  b.j_off = c->d_bj_off;
  b.h_off = c->d_bh_off;
  b.s_off = c->d_bs_off;
  b.n = c->d_bn;
  b.R = (int)c->R;
  hipError_t e = launch_batch_energy(c->ftype, c->jtype, b, c->d_brb, (int)c->K, c->dE, c->st);
  if (e != hipSuccess) return fail(OJMI_EDEVICE, std::string("batch energy: ") + hipGetErrorString(e));
  OJMI_HIP(hipStreamSynchronize(c->st));
  return OJMI_OK;
}

// Device counters of a run (accepted flips, coupling rows streamed), zeroed on
// the context's stream before its first launch.
int begin_stats(ojmi_ctx* c) {
  int rc;
  if (!c->dstats && (rc = dalloc(&c->dstats, 2))) return rc;
  OJMI_HIP(hipMemsetAsync(c->dstats, 0, 2 * sizeof(unsigned long long), c->st));
  return OJMI_OK;
}
// After the run's final synchronisation.
int end_stats(ojmi_ctx* c) {
  unsigned long long v[2] = {0, 0};
  OJMI_HIP(hipMemcpy(v, c->dstats, sizeof(v), hipMemcpyDeviceToHost));
  c->last_flips = (int64_t)v[0];
  c->last_rows = (int64_t)v[1];
  return OJMI_OK;
}

// [R][Npad] field buffer (4-byte elements: FP32 or int32 fields).
int ensure_fields(ojmi_ctx* c) {
  if (c->dF) return OJMI_OK;
  int rc;
  if (c->ftype == FT_F64) {
    double* f;
    if ((rc = dalloc(&f, (size_t)(c->R * c->Npad)))) return rc;
    c->dF = f;
  } else {
    float* f;
    if ((rc = dalloc(&f, (size_t)(c->R * c->Npad)))) return rc;
    c->dF = f;
  }
  return OJMI_OK;
}

// The classical dense families carry real-valued fields from one run to the
// next (F_out, then F_in while the spins are unchanged): a run split into
// pieces, or one sweep per call with a callback, follows the trajectory of a
// single run.  Integer fields are rebuilt exactly instead.
bool carries_fields(const ojmi_ctx* c) {
  return c->kind == 0 && c->M == 1 && (c->ftype == FT_F32 || c->ftype == FT_F64);
}

// The matrix-core sweep applies: classical dense FP32 (bf16 hi/lo operands
// present) or INT32 with int8 couplings, 4..64 slots, field buffer offsets
// within 2 GiB (32-bit buffer offsets).
bool mfma_sweep_shape(const ojmi_ctx* c) {
  return c->kind == 0 && c->M == 1 &&
         (c->family == OJMI_KERNEL_DENSE_WAVE || c->family == OJMI_KERNEL_DENSE_SHARED) &&
         mfma_sweep_supported(c->NS) &&
         ((c->ftype == FT_F32 && c->dJhi) || (c->ftype == FT_I32 && c->jtype == JT_I8 && c->dJnat)) &&
         c->R * c->Npad * 4 < ((int64_t)1 << 31);
}
// Mode 2 alternates with the column-owned kernel, which needs the
// workgroup-shared family in its column-owned form; with another launch
// shape chosen (group, colown) that shape runs instead.
bool colown_form(const ojmi_ctx* c) {
  return c->family == OJMI_KERNEL_DENSE_SHARED && c->colown &&
         colown_supported((int)c->NS, c->group) && c->rule == RULE_METROPOLIS;
}
bool uses_mfma_sweep(const ojmi_ctx* c) {
  return mfma_sweep_shape(c) && (c->mfma_sweep == 1 || (c->mfma_sweep == 2 && colown_form(c)));
}
bool hybrid_sweep(const ojmi_ctx* c) { return c->mfma_sweep == 2 && uses_mfma_sweep(c); }

// Packed operands of the matrix-core sweep, built once on the device.
int ensure_mfma_packed(ojmi_ctx* c) {
  if (c->dJpk) return OJMI_OK;
  char* p;
  int rc;
  if ((rc = dalloc(&p, mfma_packed_bytes(c->ftype, c->NS)))) return rc;
  c->dJpk = p;
  const void* src = c->ftype == FT_F32 ? (const void*)c->dJhi : (const void*)c->dJnat;
  OJMI_HIP(launch_pack_mfma(c->ftype, src, c->dJlo, c->NS, c->dJpk, c->st));
  return OJMI_OK;
}

// Zero-padded spins for the matrix-core kernels: int8 [R*M][Npad] (fields of
// INT32 contexts, energies of every context) and bf16 [R][Npad] (FP32 fields).
int ensure_mfma_spins(ojmi_ctx* c) {
  int rc;
  if (!c->dS8 && (rc = dalloc(&c->dS8, (size_t)(c->R * c->M * c->Npad)))) return rc;
  if (c->ftype == FT_F32 && c->M == 1 && !c->dSbf &&
      (rc = dalloc(&c->dSbf, (size_t)(c->R * c->Npad))))
    return rc;
  return OJMI_OK;
}

// XCD pacing counters of the workgroup kernels, one per workgroup.
int ensure_progress(ojmi_ctx* c) {
  const int64_t nwg = (c->R + 3) / 4;  // the most workgroups any group size launches
  if (c->prog_cap >= nwg) return OJMI_OK;
  if (c->dprog) (void)hipFree(c->dprog);
  c->dprog = nullptr;
  c->prog_cap = 0;
  int rc = dalloc(&c->dprog, (size_t)nwg);
  if (rc) return rc;
  c->prog_cap = nwg;
  return OJMI_OK;
}

// Fields of every replica from the current spins on the matrix cores, into
// dF: bf16 hi/lo with FP32 accumulation (FP32 contexts) or int8 with exact
// int32 accumulation (INT32 contexts with int8 couplings).
hipError_t mfma_fields(ojmi_ctx* c) {
  if (ensure_mfma_spins(c) != OJMI_OK) return hipErrorOutOfMemory;
  const int R = (int)c->R, n = (int)c->n, Np = (int)c->Npad;
  hipError_t e;
  if (c->ftype == FT_F32) {
    e = launch_pack_spins(c->dspins, n, R, Np, nullptr, c->dSbf, c->st);
    if (e != hipSuccess) return e;
    return launch_fields_bf16x2(c->dSbf, c->dJhi, c->dJlo, (const float*)c->dh, R, Np,
                                (float*)c->dF, c->st);
  }
  e = launch_pack_spins(c->dspins, n, R, Np, c->dS8, nullptr, c->st);
  if (e != hipSuccess) return e;
  return launch_fields_i8(c->dS8, (const int8_t*)c->dJnat, (const int32_t*)c->dh, R, Np,
                          (int32_t*)c->dF, c->st);
}

// One launch of the context's sweep kernel over the expanded sweeps
// [t_begin, t_begin + T): global sweep indices sweep_offset + t_begin + t.
// F_in: initial fields ([R][Npad], the field type) or null (the kernel builds
// them in sweep order).  FP32 contexts write their final fields to dF.
hipError_t launch_shared_family(ojmi_ctx* c, int64_t t_begin, int64_t T, uint64_t seed,
                                uint64_t sweep_offset, const void* F_in, void* F_out,
                                const int32_t* sel, int sel_id);

hipError_t launch_sweeps(ojmi_ctx* c, int64_t t_begin, int64_t T, uint64_t seed,
                         uint64_t sweep_offset, const void* F_in) {
  if (c->family == OJMI_KERNEL_CSR_LEVEL) {
    CsrArgs a{};
    a.order = c->d_order;
    a.lvl_ptr = c->d_lvl;
    a.nlev = c->nlev;
    a.deg = c->d_deg;
    a.ell_col = c->d_ell;
    a.val = c->d_ellval;
    a.D = c->D;
    a.h = c->dh;
    a.spins = c->dspins;
    a.n = (int)c->n;
    a.R = (int)c->R;
    a.M = 1;
    a.T = (int)T;
    a.seed = seed;
    a.t0 = sweep_offset + (uint64_t)t_begin;
    a.rep_offset = 0;
    a.p0_d = c->dp0_d + t_begin;
    a.p0_f = c->dp0_f + t_begin;
    a.rule = c->rule;
    a.stats = c->dstats;
    a.val8 = c->csr_pack8 ? c->d_ell8 : nullptr;
    a.pack8 = (c->csr_pack8 && c->d_rec8) ? 1 : 0;
    a.rec = a.pack8 ? c->d_rec8 : c->d_rec;
    return launch_csr_level(c->ftype, a, c->csr_wpb, c->csr_waves, c->st);
  }
  if (c->family == OJMI_KERNEL_SQA_CSR) {
    CsrArgs a{};
    a.order = c->d_order;
    a.lvl_ptr = c->d_lvl;
    a.nlev = c->nlev;
    a.deg = c->d_deg;
    a.ell_col = c->d_ell;
    a.val = c->d_ellval;
    a.D = c->D;
    a.h = c->dh;
    a.spins = c->dspins;
    a.n = (int)c->n;
    a.R = (int)c->R;
    a.M = (int)c->M;
    a.T = (int)T;
    a.seed = seed;
    a.t0 = sweep_offset + (uint64_t)t_begin;
    a.rep_offset = 0;
    a.p0_d = c->dp0_d + t_begin;
    a.p0_f = c->dp0_f + t_begin;
    a.p1_d = c->dp1_d + t_begin;
    a.p1_f = c->dp1_f + t_begin;
    a.rule = RULE_METROPOLIS;
    a.stats = c->dstats;
    a.pack8 = (c->sqa_csr_records && c->csr_pack8 && c->d_rec8) ? 1 : 0;
    a.rec = !c->sqa_csr_records ? nullptr : (a.pack8 ? c->d_rec8 : c->d_rec);
    // The slices' Philox words are staged in LDS (one Philox per 4 spins
    // instead of one per spin) where they fit beside the spins and leave as
    // many workgroups resident per CU as the replicas can use (at most 28
    // waves per CU at this kernel's register count).
    const size_t l0 = sqa_csr_lds_bytes((int)c->n, (int)c->M);
    const size_t l1 = sqa_csr_staged_lds_bytes((int)c->n, (int)c->M);
    const int64_t need = (c->R + c->num_cus - 1) / c->num_cus;
    auto resident = [&](size_t lds) -> int64_t {
      return std::min<int64_t>((int64_t)((size_t)c->lds_budget / std::max<size_t>(lds, 1)),
                               28 / sqa_csr_waves((int)c->M));
    };
    const bool stage = l1 <= (size_t)c->lds_budget && resident(l1) >= std::min(need, resident(l0));
    c->lds_used = (int)(stage ? sqa_csr_staged_lds_bytes((int)c->n, (int)c->M)
                              : sqa_csr_lds_bytes((int)c->n, (int)c->M));
    return launch_sqa_csr(c->ftype, a, stage, c->st);
  }
  void* F_out = carries_fields(c) ? c->dF : nullptr;
  DenseArgs a{};
  a.J = c->dJ;
  a.h = c->dh;
  a.spins = c->dspins;
  a.n = (int)c->n;
  a.R = (int)c->R;
  a.M = (int)c->M;
  a.T = (int)T;
  a.seed = seed;
  a.t0 = sweep_offset + (uint64_t)t_begin;
  a.rep_offset = 0;
  a.p0_d = c->dp0_d + t_begin;
  a.p0_f = c->dp0_f + t_begin;
  a.p1_d = c->dp1_d + t_begin;
  a.p1_f = c->dp1_f + t_begin;
  a.rule = c->rule;
  a.F_in = F_in;
  a.F_out = F_out;
  a.stats = c->dstats;
  if (c->family == OJMI_KERNEL_SQA_WAVE) {
    a.F_in = a.F_out = nullptr;
    // The first slots of every coupling row are read from LDS (copied once
    // per workgroup) and the others from L2: a flip's row then costs the
    // vector-memory return path, which binds this kernel, only its other
    // slots.  The workgroups of one CU share its LDS when the replicas
    // outnumber the CUs.
    const size_t spin_words = (size_t)c->M * (size_t)c->NS * 8;
    const size_t fb = c->ftype == FT_F64 ? 8 : 4;
    int kc = 0;
    if (c->sqa_rows_lds && c->NS <= 8 && c->sqa_ppw < 8) {  // (8 pairs: rows from L2)
      const int64_t per_cu = std::min<int64_t>(2, (c->R + c->num_cus - 1) / c->num_cus);
      kc = sqa_lds_slots(c->ftype, (int)c->NS, (int)c->n, (int)c->M,
                         (size_t)c->lds_budget / (size_t)std::max<int64_t>(1, per_cu));
    }
    a.cache_rows = kc;
    c->lds_used = (int)(spin_words + (size_t)kc * (size_t)c->n * 64 * fb);
    return launch_sqa_wave(c->ftype, a, c->NS, c->sqa_ppw, c->sqa_w, c->st);
  }
  if (uses_mfma_sweep(c)) {
    SharedArgs sa{};
    sa.J = c->ftype == FT_F32 ? c->dJ : c->dJnat;
    sa.h = c->dh;
    sa.spins = c->dspins;
    sa.n = (int)c->n;
    sa.R = (int)c->R;
    sa.T = (int)T;
    sa.seed = seed;
    sa.t0 = sweep_offset + (uint64_t)t_begin;
    sa.rule = c->rule;
    sa.rep_offset = 0;
    sa.p0_d = c->dp0_d + t_begin;
    sa.p0_f = c->dp0_f + t_begin;
    sa.F_in = F_in;
    sa.F_out = F_out;
    sa.stats = c->dstats;
    const int64_t nwg = (c->R + 7) / 8;
    sa.xcd_window = (nwg >= 16) ? c->xcd_window : 0;
    sa.xcd_spin = 4000;
    if (sa.xcd_window > 0) {
      if (c->prog_cap < nwg) return hipErrorInvalidValue;
      hipError_t e = hipMemsetAsync(c->dprog, 0, (size_t)nwg * 4, c->st);
      if (e != hipSuccess) return e;
      sa.progress = c->dprog;
    }
    c->lds_used = (int)mfma_sweep_lds_bytes(c->NS, 4, c->ftype == FT_F32 ? 4 : 1);
    c->ring_rows_used = 0;
    if (!hybrid_sweep(c)) return launch_dense_mfma(c->ftype, sa, c->dJpk, c->NS, c->mfma_form, c->st);
    // one segment: the kernel the selector chose runs, the other returns at once
    hipError_t e;
    if (t_begin == 0 &&
        (e = launch_sweep_select(c->dstats, c->dsel, 0, c->mfma_accept, 1, c->st)) != hipSuccess)
      return e;
    if ((e = launch_shared_family(c, t_begin, T, seed, sweep_offset, F_in, F_out, c->dsel, 0)) !=
        hipSuccess)
      return e;
    sa.sel = c->dsel;
    sa.sel_id = 1;
    if (sa.xcd_window > 0 &&
        (e = hipMemsetAsync(c->dprog, 0, (size_t)nwg * 4, c->st)) != hipSuccess)
      return e;
    if ((e = launch_dense_mfma(c->ftype, sa, c->dJpk, c->NS, c->mfma_form, c->st)) != hipSuccess) return e;
    const uint64_t attempts = (uint64_t)c->R * (uint64_t)c->n * (uint64_t)T;
    return launch_sweep_select(c->dstats, c->dsel, attempts, c->mfma_accept, 0, c->st);
  }
  if (c->family == OJMI_KERNEL_DENSE_LDS) {
    if (c->jtype == JT_I8) a.J = c->dJnat;
    c->lds_used = (int)lds_kernel_bytes((int)c->Npad, c->ftype == FT_F64 ? 8 : 4);
    return launch_dense_lds(c->ftype, c->jtype, a, (int)c->Npad, c->st);
  }
  if (c->family == OJMI_KERNEL_DENSE_WAVE) {
    const int jb = c->jtype == JT_I8 ? 1 : (c->jtype == JT_F64 ? 8 : 4);
    const size_t lds = wave_lds_bytes((int)c->n, c->NS, jb);
    const bool rows_in_lds = c->lds_rows != 0 && c->NS <= 4 && lds <= (size_t)65536 &&
                             lds <= (size_t)c->lds_budget;
    c->helpers_used = 0;
    const int helpers = c->wave_helpers >= 0
                            ? c->wave_helpers
                            : ((c->R <= 2 * (int64_t)c->num_cus && c->NS <= 8) ? 2 : 0);
    if (helpers > 0 && c->NS <= 8) {
      const int tb = c->ftype == FT_F32 ? 4 : 8;
      const bool rl = rows_in_lds &&
                      wave_pipe_lds_bytes((int)c->n, c->NS, jb, tb, true) <= (size_t)c->lds_budget;
      c->lds_used = (int)wave_pipe_lds_bytes((int)c->n, c->NS, jb, tb, rl);
      c->helpers_used = helpers;
      return launch_dense_wave_pipe(c->ftype, c->jtype, a, c->NS, helpers, rl, c->st);
    }
    if (rows_in_lds) {
      // fewer replicas per workgroup when there are few: more CUs, one LDS copy each
      const int wpb = c->lds_rows_wpb > 0 ? c->lds_rows_wpb
                                          : (int)std::min<int64_t>(4, std::max<int64_t>(1, (c->R + 255) / 256));
      c->lds_used = (int)lds;
      return launch_dense_wave_lds(c->ftype, c->jtype, a, c->NS, wpb, c->st);
    }
    c->lds_used = 0;
    return launch_dense_wave(c->ftype, c->jtype, a, c->NS, c->st);
  }

  return launch_shared_family(c, t_begin, T, seed, sweep_offset, F_in, F_out, nullptr, 0);
}

hipError_t launch_shared_family(ojmi_ctx* c, int64_t t_begin, int64_t T, uint64_t seed,
                                uint64_t sweep_offset, const void* F_in, void* F_out,
                                const int32_t* sel, int sel_id) {
  SharedArgs sa{};
  sa.sel = sel;
  sa.sel_id = sel_id;
  sa.J = c->dJnat ? c->dJnat : c->dJ;
  sa.h = c->dh;
  sa.spins = c->dspins;
  sa.n = (int)c->n;
  sa.R = (int)c->R;
  sa.T = (int)T;
  sa.seed = seed;
  sa.t0 = sweep_offset + (uint64_t)t_begin;
  sa.rule = c->rule;
  sa.rep_offset = 0;
  sa.p0_d = c->dp0_d + t_begin;
  sa.p0_f = c->dp0_f + t_begin;
  sa.F_in = F_in;
  sa.F_out = F_out;
  sa.stats = c->dstats;
  // heat-bath colown kernels exist for G = 4 only; a replica's trajectory
  // does not depend on G, so such runs take G = 4 (grid and pacing below too)
  const bool hb_colown = c->rule == RULE_HEAT_BATH && c->colown &&
                         colown_supported((int)c->NS, c->group) &&
                         colown_supported((int)c->NS, 4);
  const int G = hb_colown ? 4 : c->group;
  const int jb = (c->jtype == JT_I8) ? 1 : 4;
  const bool use_colown = c->colown && colown_supported((int)c->NS, G);
  const size_t fixed = shared_lds_bytes((int)c->NS, G, 0, jb);
  const size_t per_row = (size_t)jb * 64 * c->NS;
  int rb = (int)std::min<size_t>(64, ((size_t)c->lds_budget - fixed) / per_row);
  if (c->ring_rows_opt > 0) rb = std::min(rb, c->ring_rows_opt);
  // rounds of S rows, at least two buffers (one consumed, others in flight)
  const int S = std::max(1, std::min(c->round_rows_opt > 0 ? c->round_rows_opt : 2, rb / 2));
  rb = (rb / S) * S;
  if (rb < 2 * S) return hipErrorInvalidValue;
  sa.ring_rows = rb;
  sa.round_rows = S;
  c->ring_rows_used = rb;
  c->lds_used = (int)shared_lds_bytes((int)c->NS, G, rb, jb);
  const int64_t nwg = (c->R + G - 1) / G;
  sa.xcd_window = (nwg >= 16) ? c->xcd_window : 0;
  sa.xcd_spin = 4000;
  if (sa.xcd_window > 0) {
    if (c->prog_cap < nwg) return hipErrorInvalidValue;
    hipError_t e = hipMemsetAsync(c->dprog, 0, (size_t)nwg * 4, c->st);
    if (e != hipSuccess) return e;
    sa.progress = c->dprog;
  }
  if (use_colown) {
    c->ring_rows_used = 0;
    c->lds_used = (int)colown_lds_bytes((int)c->n, (int)c->NS, G, jb, 4);
    return launch_dense_colown(c->ftype, c->jtype, sa, (int)c->NS, G, c->st);
  }
  return launch_dense_shared(c->ftype, c->jtype, sa, (int)c->NS, G, c->st);
}

}  // namespace

extern "C" {

const char* ojmi_last_error(void) { return g_err.c_str(); }

int ojmi_device_count(int* count) {
  if (!count) return fail(OJMI_EINVAL, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = (e == hipSuccess) ? c : 0;
  return OJMI_OK;
}

int ojmi_create_dense(int64_t n, const double* J_ext, int64_t num_replicas, int64_t num_trotter,
                      double gamma, int32_t precision, ojmi_ctx** out) {
  if (!out) return fail(OJMI_EINVAL, "null output pointer");
  *out = nullptr;
  if (n < 1 || !J_ext) return fail(OJMI_EINVAL, "need n >= 1 and a coupling matrix");
  if (num_replicas < 1) return fail(OJMI_EINVAL, "num_replicas must be >= 1");
  if (num_trotter < 1) return fail(OJMI_EINVAL, "num_trotter must be >= 1");
  int rc = check_dense_input(n, J_ext);
  if (rc) return rc;
  const int64_t n1 = n + 1;
  std::vector<double> Jn((size_t)(n * n), 0.0), h((size_t)n);
  double maxabsJ = 0, maxfield = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t j = i + 1; j < n; ++j) {
      const double v = J_ext[i * n1 + j];
      Jn[i * n + j] = v;
      Jn[j * n + i] = v;
    }
    h[i] = J_ext[i * n1 + n];
  }
  for (int64_t i = 0; i < n; ++i) {
    double rs = std::fabs(h[i]);
    for (int64_t j = 0; j < n; ++j) {
      rs += std::fabs(Jn[i * n + j]);
      maxabsJ = std::max(maxabsJ, std::fabs(Jn[i * n + j]));
    }
    maxfield = std::max(maxfield, rs);
  }
  const bool integer = all_integer(Jn) && all_integer(h);
  int prec;
  if ((rc = resolve_precision(precision, integer, maxfield, &prec))) return rc;
  if (num_trotter > 1 && prec == OJMI_PREC_INT32) prec = OJMI_PREC_FP64;

  ojmi_ctx* c = new (std::nothrow) ojmi_ctx;
  if (!c) return fail(OJMI_ENOMEM, "host allocation failed");
  c->kind = 0;
  c->n = n;
  c->R = num_replicas;
  c->M = num_trotter;
  c->gamma = gamma;
  c->prec = prec;
  pick_types(c, maxabsJ);
  if ((rc = setup_dense_family(c)) || (rc = common_init(c)) || (rc = upload_dense(c, Jn, h))) {
    delete c;
    return rc;
  }
  if (c->M == 1 && (c->ftype == FT_F32 || c->ftype == FT_I32) &&
      c->family == OJMI_KERNEL_DENSE_WAVE) {
    // workgroup shape from the autotuner's table (openjij_amd/tune.py,
    // ojmi_tuned.h); replicas share coupling rows only when there are at
    // least two workgroups' worth of them
    for (const tuned::DenseShape& e : tuned::kDense) {
      if (e.ftype != c->ftype || e.slots != c->NS || c->R > e.r_max) continue;
      if (e.group > 1 && c->R >= 2 * e.group && c->R >= 8) {
        c->group = e.group;
        c->colown = e.colown;
        c->xcd_window = e.xcd_window;
        c->family = OJMI_KERNEL_DENSE_SHARED;
      }
      break;
    }
  }
  // Large FP32 systems with replicas enough to fill the device (8 per
  // workgroup): per 16-sweep segment the matrix-core sweep while the
  // acceptance rate is at least the autotuner's crossover for this size, the
  // column-owned kernel below (ojmi_tuned.h)
  if (c->ftype == FT_F32 && c->M == 1 && c->R >= 8 * (int64_t)c->num_cus && colown_form(c) &&
      mfma_sweep_shape(c)) {
    for (const tuned::MfmaChoice& e : tuned::kMfma) {
      if (e.slots != c->NS || e.accept_permille <= 0) continue;
      if ((rc = dalloc(&c->dsel, 4))) {
        delete c;
        return rc;
      }
      c->mfma_sweep = 2;
      c->mfma_accept = e.accept_permille;
      break;
    }
  }
  *out = c;
  return OJMI_OK;
}

int ojmi_create_csr(int64_t n, const int64_t* rowptr, const int64_t* col, const double* val,
                    int64_t num_replicas, int64_t num_trotter, double gamma, int32_t precision,
                    ojmi_ctx** out) {
  if (!out) return fail(OJMI_EINVAL, "null output pointer");
  *out = nullptr;
  if (n < 1 || !rowptr || !col || !val)
    return fail(OJMI_EINVAL, "need n >= 1 and CSR arrays");
  if (num_replicas < 1) return fail(OJMI_EINVAL, "num_replicas must be >= 1");
  if (num_trotter < 1) return fail(OJMI_EINVAL, "num_trotter must be >= 1");
  if (rowptr[0] != 0) return fail(OJMI_EINVAL, "rowptr[0] must be 0");
  for (int64_t i = 0; i <= n; ++i)
    if (rowptr[i + 1] < rowptr[i]) return fail(OJMI_EINVAL, "rowptr must be non-decreasing");
  // upper triangle -> symmetric adjacency without diagonal / dummy
  std::vector<std::vector<std::pair<int64_t, double>>> adj((size_t)n);
  std::vector<double> h((size_t)n, 0.0);
  for (int64_t i = 0; i <= n; ++i)
    for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      const int64_t j = col[e];
      const double v = val[e];
      if (j < 0 || j > n) return fail(OJMI_ERANGE, "column index out of range");
      if (!std::isfinite(v)) return fail(OJMI_EINVAL, "non-finite coupling");
      if (j < i || i == n) continue;
      if (j == i) {
        if (v != 0.0)
          return fail(OJMI_ERUNTIME,
                      "The diagonal elements of interaction matrix must be zero.");
        continue;
      }
      if (j == n) {
        h[i] += v;
        continue;
      }
      if (v == 0.0) continue;
      adj[i].push_back({j, v});
      adj[j].push_back({i, v});
    }
  double maxabsJ = 0, maxfield = 0;
  bool integer = all_integer(h);
  int64_t D = 0, nnz = 0;
  for (int64_t i = 0; i < n; ++i) {
    auto& a = adj[i];
    std::sort(a.begin(), a.end(),
              [](const std::pair<int64_t, double>& x, const std::pair<int64_t, double>& y) {
                return x.first < y.first;
              });
    // merge duplicates (an upper-triangle input may list a pair twice)
    std::vector<std::pair<int64_t, double>> m;
    for (auto& p : a) {
      if (!m.empty() && m.back().first == p.first) m.back().second += p.second;
      else m.push_back(p);
    }
    a.swap(m);
    double rs = std::fabs(h[i]);
    for (auto& p : a) {
      rs += std::fabs(p.second);
      maxabsJ = std::max(maxabsJ, std::fabs(p.second));
      if (p.second != std::floor(p.second) || std::fabs(p.second) > 1e9) integer = false;
    }
    maxfield = std::max(maxfield, rs);
    D = std::max<int64_t>(D, (int64_t)a.size());
    nnz += (int64_t)a.size();
  }
  int rc, prec;
  if ((rc = resolve_precision(precision, integer, maxfield, &prec))) return rc;
  if (num_trotter > 1 && prec == OJMI_PREC_INT32) prec = OJMI_PREC_FP64;

  ojmi_ctx* c = new (std::nothrow) ojmi_ctx;
  if (!c) return fail(OJMI_ENOMEM, "host allocation failed");
  c->kind = 1;
  c->n = n;
  c->R = num_replicas;
  c->M = num_trotter;
  c->gamma = gamma;
  c->prec = prec;
  pick_types(c, maxabsJ);
  if (c->jtype == JT_I8) c->jtype = JT_I32;

  // Graphs with a node of degree > 64 (the widest neighbour list of the CSR
  // kernels) run on the dense kernels, and so do Trotter systems whose slices'
  // spins do not fit the Trotter CSR kernel's LDS.
  const bool sqa_csr = num_trotter > 1 && D <= 64 &&
                       sqa_csr_lds_bytes((int)n, (int)num_trotter) <= (size_t)device_lds_budget();
  const bool dense_like = (num_trotter > 1 && !sqa_csr) || (D > 64);
  if (dense_like) {
    std::vector<double> Jn((size_t)(n * n), 0.0);
    for (int64_t i = 0; i < n; ++i)
      for (auto& p : adj[i]) Jn[i * n + p.first] = p.second;
    if (c->prec == OJMI_PREC_INT32) c->jtype = (maxabsJ <= 127.0) ? JT_I8 : JT_I32;
    if ((rc = setup_dense_family(c)) || (rc = common_init(c)) ||
        (rc = upload_dense(c, Jn, h))) {
      delete c;
      return rc;
    }
    *out = c;
    return OJMI_OK;
  }

  // Level order: level(i) = 1 + max level of smaller-index neighbours.
  std::vector<int32_t> level((size_t)n, 0);
  int32_t nlev = 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t lv = 0;
    for (auto& p : adj[i])
      if (p.first < i) lv = std::max(lv, level[p.first] + 1);
    level[i] = lv;
    nlev = std::max(nlev, lv + 1);
  }
  std::vector<int32_t> order((size_t)n);
  for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return level[a] < level[b]; });
  std::vector<int32_t> lvl_ptr((size_t)nlev + 1, 0);
  for (int64_t i = 0; i < n; ++i) lvl_ptr[level[i] + 1]++;
  for (int32_t l = 0; l < nlev; ++l) lvl_ptr[l + 1] += lvl_ptr[l];
  if (D == 0) D = 1;
  // neighbour lists padded to the kernel's fixed width (zero couplings)
  const int64_t Dp = csr_padded_width((int)D);
  std::vector<int32_t> deg((size_t)n), ell((size_t)(Dp * n), 0);
  std::vector<double> ellv((size_t)(Dp * n), 0.0), hpos((size_t)n);
  for (int64_t pos = 0; pos < n; ++pos) {
    const int64_t i = order[pos];
    deg[pos] = (int32_t)adj[i].size();
    hpos[pos] = h[i];
    for (size_t d = 0; d < adj[i].size(); ++d) {
      ell[d * n + pos] = (int32_t)adj[i][d].first;
      ellv[d * n + pos] = adj[i][d].second;
    }
  }
  c->family = sqa_csr ? OJMI_KERNEL_SQA_CSR : OJMI_KERNEL_CSR_LEVEL;
  c->nlev = nlev;
  c->D = (int)Dp;
  c->NS = pow2_slots(n);
  c->Npad = 64LL * c->NS;
  // per wave: Philox words (16 B per 4 spins) + int8 spins; <= 64 KiB per block
  const size_t per_wave = (size_t)((n + 3) / 4) * 20;
  c->csr_wpb = (int)std::max<size_t>(1, std::min<size_t>(4, 65536 / per_wave));
  if (per_wave > 65536 && !sqa_csr) {
    delete c;
    return fail(OJMI_ENOTSUP, "CSR GPU path: n too large for LDS-resident spins");
  }
  if ((rc = common_init(c))) {
    delete c;
    return rc;
  }
  auto up32 = [&](int32_t** d, const std::vector<int32_t>& v) -> int {
    int r = dalloc(d, v.size());
    if (r) return r;
    OJMI_HIP(hipMemcpy(*d, v.data(), v.size() * 4, hipMemcpyHostToDevice));
    return OJMI_OK;
  };
  if ((rc = up32(&c->d_order, order)) || (rc = up32(&c->d_lvl, lvl_ptr)) ||
      (rc = up32(&c->d_deg, deg)) || (rc = up32(&c->d_ell, ell))) {
    delete c;
    return rc;
  }
  // values and h at level-order position, in the field type
  size_t esz = (c->ftype == FT_F64) ? 8 : 4;
  std::vector<char> vb(ellv.size() * esz), hb(hpos.size() * esz);
  for (size_t k = 0; k < ellv.size(); ++k) {
    if (c->ftype == FT_F64) ((double*)vb.data())[k] = ellv[k];
    else if (c->ftype == FT_F32) ((float*)vb.data())[k] = (float)ellv[k];
    else ((int32_t*)vb.data())[k] = (int32_t)ellv[k];
  }
  for (size_t k = 0; k < hpos.size(); ++k) {
    if (c->ftype == FT_F64) ((double*)hb.data())[k] = hpos[k];
    else if (c->ftype == FT_F32) ((float*)hb.data())[k] = (float)hpos[k];
    else ((int32_t*)hb.data())[k] = (int32_t)hpos[k];
  }
  char *dv, *dhb;
  if ((rc = dalloc(&dv, vb.size())) || (rc = dalloc(&dhb, hb.size()))) {
    delete c;
    return rc;
  }
  c->d_ellval = dv;
  c->dh = dhb;
  if (hipMemcpy(dv, vb.data(), vb.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dhb, hb.data(), hb.size(), hipMemcpyHostToDevice) != hipSuccess) {
    delete c;
    return fail(OJMI_EDEVICE, "hipMemcpy of CSR values failed");
  }
  // position records of the CSR kernels (the classical kernels hold at most
  // 32768 spins in LDS, so 16-bit columns): per position the column pairs, the values,
  // the spin index and h, as 16-byte chunks [words/4][n] (the layout of
  // CsrRec in ojmi_csr.h); INT32 couplings within int8 get a second set with
  // the values as int8 quadruples
  auto build_rec = [&](bool v8, uint4** dst) -> int {
    const int64_t vwords = v8 ? Dp / 4 : Dp * (int64_t)esz / 4;
    const int64_t kvals = Dp / 2, kidx = kvals + vwords, kh = kidx + 1;
    const int64_t words = (kh + (v8 ? 1 : (int64_t)esz / 4) + 3) / 4 * 4;
    std::vector<uint32_t> w((size_t)(words * n), 0u);
    auto at = [&](int64_t pos, int64_t k) -> uint32_t& {
      return w[(size_t)(((k / 4) * n + pos) * 4 + k % 4)];
    };
    for (int64_t pos = 0; pos < n; ++pos) {
      for (int64_t d = 0; d < Dp; ++d) {
        at(pos, d / 2) |= (uint32_t)ell[d * n + pos] << (16 * (d % 2));
        const double v = ellv[d * n + pos];
        if (v8) {
          at(pos, kvals + d / 4) |= (uint32_t)(uint8_t)(int8_t)(int32_t)v << (8 * (d % 4));
        } else if (c->ftype == FT_F64) {
          uint64_t b;
          std::memcpy(&b, &v, 8);
          at(pos, kvals + 2 * d) = (uint32_t)b;
          at(pos, kvals + 2 * d + 1) = (uint32_t)(b >> 32);
        } else if (c->ftype == FT_F32) {
          const float f = (float)v;
          std::memcpy(&at(pos, kvals + d), &f, 4);
        } else {
          at(pos, kvals + d) = (uint32_t)(int32_t)v;
        }
      }
      at(pos, kidx) = (uint32_t)order[pos];
      const double hv = hpos[pos];
      if (v8 || c->ftype == FT_I32) {
        at(pos, kh) = (uint32_t)(int32_t)hv;
      } else if (c->ftype == FT_F64) {
        uint64_t b;
        std::memcpy(&b, &hv, 8);
        at(pos, kh) = (uint32_t)b;
        at(pos, kh + 1) = (uint32_t)(b >> 32);
      } else {
        const float f = (float)hv;
        std::memcpy(&at(pos, kh), &f, 4);
      }
    }
    int r = dalloc(dst, (size_t)(words / 4 * n));
    if (r) return r;
    if (hipMemcpy(*dst, w.data(), w.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
      return fail(OJMI_EDEVICE, "hipMemcpy of CSR position records failed");
    return OJMI_OK;
  };
  // Integer couplings within int8 (fitsv) and, for real field types, integer
  // fields whose sums with them stay exact in the field type (|h| <= 2^20:
  // below 2^24 with 32 neighbours of |J| <= 127) take the int8 forms
  bool fitsv = true;
  for (double v : ellv) fitsv = fitsv && v >= -127.0 && v <= 127.0 && v == std::floor(v);
  bool fits8 = fitsv;
  if (c->ftype != FT_I32)
    for (double hv : hpos) fits8 = fits8 && std::fabs(hv) <= 1048576.0 && hv == std::floor(hv);
  // (Trotter systems: the records of neighbour lists of <= 16, k_sqa_csr's prefetching form)
  if ((sqa_csr ? Dp <= 16 : Dp <= 64) && n < 65536) {
    if ((rc = build_rec(false, &c->d_rec)) || (fits8 && (rc = build_rec(true, &c->d_rec8)))) {
      delete c;
      return rc;
    }
  }
  // INT32 sweeps of couplings within int8: the values also packed four per
  // word (word (q, pos) holds entries 4q .. 4q+3 of position pos, padding 0),
  // for the dot4 field form of the CSR kernels
  if (!sqa_csr && c->ftype == FT_I32) {
    if (fitsv) {
      std::vector<uint32_t> p8((size_t)(Dp / 4 * n), 0u);
      for (int64_t d = 0; d < Dp; ++d)
        for (int64_t pos = 0; pos < n; ++pos) {
          const uint32_t b = (uint32_t)(uint8_t)(int8_t)(int32_t)ellv[d * n + pos];
          p8[(d / 4) * n + pos] |= b << (8 * (d % 4));
        }
      if ((rc = dalloc(&c->d_ell8, p8.size()))) {
        delete c;
        return rc;
      }
      if (hipMemcpy(c->d_ell8, p8.data(), p8.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        delete c;
        return fail(OJMI_EDEVICE, "hipMemcpy of packed CSR values failed");
      }
    }
  }
  // energies run over the same ELL arrays with FP64 values (k_csr_energy);
  // FP32 and int32 contexts keep an FP64 copy of the values (Dp x n)
  if (c->ftype == FT_F64) {
    c->d_ellv64 = reinterpret_cast<double*>(dv);
    c->d_hpos64 = reinterpret_cast<double*>(dhb);
  } else {
    if ((rc = dalloc(&c->d_ellv64, ellv.size())) || (rc = dalloc(&c->d_hpos64, hpos.size()))) {
      delete c;
      return rc;
    }
    if (hipMemcpy(c->d_ellv64, ellv.data(), ellv.size() * 8, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(c->d_hpos64, hpos.data(), hpos.size() * 8, hipMemcpyHostToDevice) !=
            hipSuccess) {
      delete c;
      return fail(OJMI_EDEVICE, "hipMemcpy of FP64 CSR values failed");
    }
  }
  *out = c;
  return OJMI_OK;
}

int ojmi_create_batch(int64_t num_problems, const int64_t* n, const double* J_ext,
                      int64_t num_replicas, int32_t precision, ojmi_ctx** out) {
  if (!out) return fail(OJMI_EINVAL, "null output pointer");
  *out = nullptr;
  if (num_problems < 1 || !n || !J_ext) return fail(OJMI_EINVAL, "need >= 1 problem and arrays");
  if (num_problems > INT32_MAX / 2) return fail(OJMI_EINVAL, "too many problems");
  if (num_replicas < 1) return fail(OJMI_EINVAL, "num_replicas must be >= 1");
  const int64_t K = num_problems;
  std::vector<int64_t> ext_off((size_t)K + 1, 0);
  int64_t sum_n = 0;
  for (int64_t k = 0; k < K; ++k) {
    if (n[k] < 1) return fail(OJMI_EINVAL, "every problem needs n >= 1");
    if (n[k] > 4096) return fail(OJMI_ENOTSUP, "batch GPU path supports n <= 4096 per problem");
    ext_off[k + 1] = ext_off[k] + (n[k] + 1) * (n[k] + 1);
    sum_n += n[k];
  }
  if ((double)sum_n * (double)num_replicas > 4.0e9)
    return fail(OJMI_ERANGE, "batch spins exceed 4e9 entries");
  bool integer = true;
  double maxabsJ = 0, maxfield = 0;
  int rc;
  for (int64_t k = 0; k < K; ++k) {
    const double* Je = J_ext + ext_off[k];
    const int64_t nk = n[k], n1 = nk + 1;
    if ((rc = check_dense_input(nk, Je))) return fail(rc, "problem " + std::to_string(k) + ": " + g_err);
    std::vector<double> rs((size_t)nk, 0.0);
    for (int64_t i = 0; i < nk; ++i) {
      const double hv = Je[i * n1 + nk];
      if (hv != std::floor(hv) || std::fabs(hv) > 1e9) integer = false;
      rs[i] += std::fabs(hv);
      for (int64_t j = i + 1; j < nk; ++j) {
        const double v = Je[i * n1 + j];
        if (v == 0.0) continue;
        if (v != std::floor(v) || std::fabs(v) > 1e9) integer = false;
        maxabsJ = std::max(maxabsJ, std::fabs(v));
        rs[i] += std::fabs(v);
        rs[j] += std::fabs(v);
      }
    }
    for (double x : rs) maxfield = std::max(maxfield, x);
  }
  int prec;
  if ((rc = resolve_precision(precision, integer, maxfield, &prec))) return rc;

  ojmi_ctx* c = new (std::nothrow) ojmi_ctx;
  if (!c) return fail(OJMI_ENOMEM, "host allocation failed");
  c->kind = 2;
  c->K = K;
  c->n = sum_n;
  c->R = num_replicas;
  c->M = 1;
  c->prec = prec;
  c->family = OJMI_KERNEL_BATCH_WAVE;
  pick_types(c, maxabsJ);
  if ((rc = common_init(c))) {
    delete c;
    return rc;
  }
  (void)hipFree(c->dE);
  c->dE = nullptr;
  if ((rc = dalloc(&c->dE, (size_t)(K * num_replicas)))) {
    delete c;
    return rc;
  }
  // layout: offsets, then one host image of couplings and fields
  const size_t fsz = (c->ftype == FT_F64) ? 8 : 4;
  std::vector<int64_t> j_off((size_t)K), h_off((size_t)K), s_off((size_t)K), rbv((size_t)K);
  // rows of ceil(n_k / 64) slots (NSk); kernels of the power-of-two slot
  // class pow2_slots(n_k) run them with that slot count as a bound
  std::vector<int> NSk((size_t)K);
  int64_t jbytes = 0, hcount = 0, scount = 0;
  c->bn.resize((size_t)K);
  for (int64_t k = 0; k < K; ++k) {
    NSk[k] = (int)((n[k] + 63) / 64);
    rbv[k] = batch_row_bytes(c->jtype, NSk[k]);
    j_off[k] = jbytes;
    jbytes += ((n[k] * rbv[k] + 255) / 256) * 256;
    h_off[k] = hcount;
    hcount += 64LL * NSk[k];
    s_off[k] = scount;
    scount += num_replicas * n[k];
    c->bn[k] = (int32_t)n[k];
  }
  std::vector<char> Jimg((size_t)jbytes, 0), himg((size_t)hcount * fsz, 0);
  for (int64_t k = 0; k < K; ++k) {
    const double* Je = J_ext + ext_off[k];
    const int64_t nk = n[k], n1 = nk + 1;
    if (c->jtype == JT_I8) pack_rows<int8_t>(Je, nk, NSk[k], Jimg, (size_t)j_off[k]);
    else if (c->jtype == JT_I32) pack_rows<int32_t>(Je, nk, NSk[k], Jimg, (size_t)j_off[k]);
    else if (c->jtype == JT_F32) pack_rows<float>(Je, nk, NSk[k], Jimg, (size_t)j_off[k]);
    else pack_rows<double>(Je, nk, NSk[k], Jimg, (size_t)j_off[k]);
    for (int64_t i = 0; i < nk; ++i) {
      const double hv = Je[i * n1 + nk];
      char* dst = himg.data() + (size_t)(h_off[k] + i) * fsz;
      if (c->ftype == FT_F64) std::memcpy(dst, &hv, 8);
      else if (c->ftype == FT_F32) { const float x = (float)hv; std::memcpy(dst, &x, 4); }
      else { const int32_t x = (int32_t)hv; std::memcpy(dst, &x, 4); }
    }
  }
  auto up = [&](auto** d, const auto& v) -> int {
    int r = dalloc(d, v.size());
    if (r) return r;
    OJMI_HIP(hipMemcpy(*d, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
    return OJMI_OK;
  };
  char* dJc = nullptr;
  char* dhc = nullptr;
  rc = up(&dJc, Jimg);
  c->dJ = dJc;  // owned by the context from here on
  if (!rc) {
    rc = up(&dhc, himg);
    c->dh = dhc;
  }
  if (!rc) rc = up(&c->d_bj_off, j_off);
  if (!rc) rc = up(&c->d_bh_off, h_off);
  if (!rc) rc = up(&c->d_bs_off, s_off);
  if (!rc) rc = up(&c->d_brb, rbv);
  if (!rc) rc = up(&c->d_bn, c->bn);
  if (rc) {
    delete c;
    return rc;
  }
  // one problem list per slot class, largest problems first within a class
  for (int NS = 1; NS <= 64; NS <<= 1) {
    std::vector<int32_t> lst;
    for (int64_t k = 0; k < K; ++k)
      if (pow2_slots(n[k]) == NS) lst.push_back((int32_t)k);
    if (lst.empty()) continue;
    std::stable_sort(lst.begin(), lst.end(), [&](int32_t a, int32_t b) { return n[a] > n[b]; });
    int32_t* d = nullptr;
    if ((rc = up(&d, lst))) {
      delete c;
      return rc;
    }
    c->bclass.push_back({NS, d});
    c->bclass_count.push_back((int)lst.size());
    c->batch_ns_max = std::max(c->batch_ns_max, NS);
  }
  // longest first: cost of a replica ~ n_k attempts per sweep x 64 NS_k
  // columns per row update
  std::vector<int32_t> byc((size_t)K);
  for (int64_t k = 0; k < K; ++k) byc[k] = (int32_t)k;
  auto cost = [&](int32_t k) { return (double)n[k] * 64.0 * (double)NSk[k]; };
  std::stable_sort(byc.begin(), byc.end(), [&](int32_t a, int32_t b) { return cost(a) > cost(b); });
  if ((rc = up(&c->d_fused, byc))) {
    delete c;
    return rc;
  }
  c->family = OJMI_KERNEL_BATCH_FUSED;
  c->batch_mode = 0;
  // software queues: whole problems to 8 per-XCD queues in LPT order,
  // replicas consecutive
  if ((double)K * (double)num_replicas <= (double)INT32_MAX) {
    std::vector<std::vector<int32_t>> qs(8);
    double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t k : byc) {
      int best = 0;
      for (int q = 1; q < 8; ++q)
        if (load[q] < load[best]) best = q;
      load[best] += cost(k);
      qs[best].push_back(k);
    }
    std::vector<int32_t> items, begin(9, 0);
    items.reserve((size_t)(K * num_replicas));
    for (int q = 0; q < 8; ++q) {
      begin[q] = (int32_t)items.size();
      for (int32_t k : qs[q])
        for (int64_t r = 0; r < num_replicas; ++r) items.push_back((int32_t)(k * num_replicas + r));
    }
    begin[8] = (int32_t)items.size();
    if ((rc = up(&c->d_q_items, items)) || (rc = up(&c->d_q_begin, begin)) ||
        (rc = dalloc(&c->d_q_head, 8))) {
      delete c;
      return rc;
    }
  }
  // default launch from the autotuner's table
  if (tuned::kBatchMode == 1 || (tuned::kBatchMode == 2 && c->d_q_items)) {
    c->batch_mode = tuned::kBatchMode;
    c->family = tuned::kBatchMode == 1 ? OJMI_KERNEL_BATCH_WAVE : OJMI_KERNEL_BATCH_PERSIST;
  }
  *out = c;
  return OJMI_OK;
}

int ojmi_run_batch(ojmi_ctx* c, const double* beta, const int64_t* mcs, int64_t L, uint64_t seed,
                   uint64_t sweep_offset) {
  if (!c) return fail(OJMI_EINVAL, "null context");
  if (c->kind != 2) return fail(OJMI_EINVAL, "ojmi_run_batch needs a batch context");
  if (L < 0 || (L > 0 && (!beta || !mcs))) return fail(OJMI_EINVAL, "bad schedule arrays");
  int64_t T = 0;
  for (int64_t l = 0; l < L; ++l) {
    if (mcs[l] < 0) return fail(OJMI_EINVAL, "one_mc_step must be >= 0");
    T += mcs[l];
  }
  for (int64_t q = 0; q < c->K * L; ++q)
    if (!(beta[q] >= 0.0)) return fail(OJMI_EINVAL, "schedule beta range is '0 <= beta'.");
  if (T > INT32_MAX || sweep_offset + (uint64_t)T > 0xffffffffull)
    return fail(OJMI_EINVAL, "global sweep index exceeds 32 bits");
  std::vector<double> p0((size_t)(c->K * T));
  for (int64_t k = 0; k < c->K; ++k) {
    int64_t t = 0;
    for (int64_t l = 0; l < L; ++l) {
      const double a0 = -1.0 / beta[k * L + l];
      for (int64_t m = 0; m < mcs[l]; ++m, ++t) p0[(size_t)(k * T + t)] = a0;
    }
  }
  return run_batch(c, p0, T, seed, sweep_offset);
}

int ojmi_batch_beta_stats(ojmi_ctx* c, const int8_t* spins, double* out) {
  if (!c || !spins || !out) return fail(OJMI_EINVAL, "null argument");
  if (c->kind != 2) return fail(OJMI_EINVAL, "ojmi_batch_beta_stats needs a batch context");
  const int64_t K = c->K;
  std::vector<int64_t> off((size_t)K);
  int64_t tot = 0;
  int nmax = 1;
  for (int64_t k = 0; k < K; ++k) {
    off[k] = tot;
    tot += 2 * (int64_t)c->bn[k];
    nmax = std::max(nmax, (int)c->bn[k]);
  }
  for (int64_t q = 0; q < tot; ++q)
    if (spins[q] != 1 && spins[q] != -1) return fail(OJMI_EINVAL, "spins must be +1 or -1");
  if (beta_stats_lds_bytes(nmax) > (size_t)c->lds_budget)
    return fail(OJMI_ENOTSUP, "beta statistics: problem too large for LDS");
  int8_t* dS = nullptr;
  int64_t* doff = nullptr;
  double* dout = nullptr;
  int rc;
  if ((rc = dalloc(&dS, (size_t)tot)) || (rc = dalloc(&doff, (size_t)K)) ||
      (rc = dalloc(&dout, (size_t)(5 * K)))) {
    if (dS) (void)hipFree(dS);
    if (doff) (void)hipFree(doff);
    return rc;
  }
  BatchArgs b{};
  b.J = c->dJ;
  b.h = c->dh;
  b.j_off = c->d_bj_off;
  b.h_off = c->d_bh_off;
  b.n = c->d_bn;
  hipError_t e = hipMemcpy(dS, spins, (size_t)tot, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(doff, off.data(), (size_t)K * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_batch_beta_stats(c->ftype, c->jtype, b, c->d_brb, (int)K, nmax, dS, doff, dout,
                                c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)(5 * K) * 8, hipMemcpyDeviceToHost);
  (void)hipFree(dS);
  (void)hipFree(doff);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(OJMI_EDEVICE, std::string("beta statistics: ") + hipGetErrorString(e));
  return OJMI_OK;
}

int ojmi_set_option(ojmi_ctx* c, const char* key, int64_t value) {
  if (!c || !key) return fail(OJMI_EINVAL, "null argument");
  const std::string k(key);
  if (k == "refresh_every") {
    if (value < 0 || value > INT32_MAX) return fail(OJMI_EINVAL, "refresh_every must be >= 0");
    // FP32 dense classical contexts: fields from the matrix cores (bf16 hi/lo
    // couplings, FP32 accumulation) at the start of each run and every
    // `value` sweeps of it.  0: the sweep kernel builds them in sweep order and
    // carries them through the run (the FP32 replay's arithmetic).
    if (value > 0 && !(c->kind == 0 && c->M == 1 && c->ftype == FT_F32 && c->dJhi))
      return fail(OJMI_ENOTSUP, "field refresh applies to classical dense FP32 contexts");
    c->refresh_every = (int)value;
    return OJMI_OK;
  }
  if (k == "group") {
    if (!(value == 1 || value == 4 || value == 8 || value == 16))
      return fail(OJMI_EINVAL, "group must be 1, 4, 8 or 16");
    if (value > 1 && (c->M != 1 || c->ftype == FT_F64))
      return fail(OJMI_EINVAL, "group > 1 needs a classical FP32 or INT32 system");
    if (c->family != OJMI_KERNEL_DENSE_WAVE && c->family != OJMI_KERNEL_DENSE_SHARED)
      return fail(OJMI_EINVAL, "group applies to dense classical systems");
    if (value == 16 && c->NS > 32)
      return fail(OJMI_EINVAL, "group 16 supports n <= 2048 (register budget)");
    c->group = (int)value;
    c->family = value > 1 ? OJMI_KERNEL_DENSE_SHARED : OJMI_KERNEL_DENSE_WAVE;
    return OJMI_OK;
  }
  if (k == "ring_rows") {
    if (value < 0 || value > 64) return fail(OJMI_EINVAL, "ring_rows must be in 0..64");
    c->ring_rows_opt = (int)value;
    return OJMI_OK;
  }
  if (k == "colown") {
    if (value != 0 && value != 1) return fail(OJMI_EINVAL, "colown must be 0 or 1");
    c->colown = (int)value;
    return OJMI_OK;
  }
  if (k == "mfma_sweep") {
    // 0: off; 1: the matrix-core sweep; 2: per 16-sweep segment, the
    // matrix-core sweep while the acceptance rate is at least mfma_accept per
    // mille, else the column-owned kernel
    if (value < 0 || value > 2) return fail(OJMI_EINVAL, "mfma_sweep must be 0, 1 or 2");
    if (value >= 1 && !mfma_sweep_shape(c))
      return fail(OJMI_ENOTSUP,
                  "mfma_sweep applies to classical dense FP32 or int8-coupling INT32 contexts "
                  "of 513..4096 spins");
    if (value == 2 && !colown_form(c))
      return fail(OJMI_ENOTSUP,
                  "mfma_sweep 2 alternates with the column-owned kernel (Metropolis rule)");
    if (value == 2 && !c->dsel) {
      int rc = dalloc(&c->dsel, 4);
      if (rc) return rc;
    }
    c->mfma_sweep = (int)value;
    return OJMI_OK;
  }
  if (k == "mfma_form") {
    if (value != 0 && value != 1) return fail(OJMI_EINVAL, "mfma_form must be 0 or 1");
    c->mfma_form = (int)value;
    return OJMI_OK;
  }
  if (k == "sqa_csr_records") {
    if (value != 0 && value != 1) return fail(OJMI_EINVAL, "sqa_csr_records must be 0 or 1");
    c->sqa_csr_records = (int)value;
    return OJMI_OK;
  }
  if (k == "csr_pack8") {
    if (value != 0 && value != 1) return fail(OJMI_EINVAL, "csr_pack8 must be 0 or 1");
    c->csr_pack8 = (int)value;
    return OJMI_OK;
  }
  if (k == "sqa_rows_lds") {
    if (value != 0 && value != 1) return fail(OJMI_EINVAL, "sqa_rows_lds must be 0 or 1");
    c->sqa_rows_lds = (int)value;
    return OJMI_OK;
  }
  if (k == "mfma_accept") {
    if (value < 0 || value > 1000) return fail(OJMI_EINVAL, "mfma_accept must be 0..1000");
    c->mfma_accept = (int)value;
    return OJMI_OK;
  }
  if (k == "lds_rows") {
    if (value != 0 && value != 1) return fail(OJMI_EINVAL, "lds_rows must be 0 or 1");
    c->lds_rows = value ? 1 : 0;
    return OJMI_OK;
  }
  if (k == "wave_helpers") {
    if (value < -1 || value > 3) return fail(OJMI_EINVAL, "wave_helpers must be in -1..3");
    if (value > 0 && (c->family != OJMI_KERNEL_DENSE_WAVE || c->NS > 8))
      return fail(OJMI_EINVAL, "wave_helpers applies to one-wave dense contexts of n <= 512");
    c->wave_helpers = (int)value;
    return OJMI_OK;
  }
  if (k == "lds_rows_wpb") {
    if (value < 0 || value > 4) return fail(OJMI_EINVAL, "lds_rows_wpb must be 0..4");
    c->lds_rows_wpb = (int)value;
    return OJMI_OK;
  }
  if (k == "round_rows") {
    if (value < 0 || value > 32) return fail(OJMI_EINVAL, "round_rows must be in 0..32");
    c->round_rows_opt = (int)value;
    return OJMI_OK;
  }
  if (k == "update_method") {
    // 0 = METROPOLIS, 1 = HEAT_BATH, the first two values of
    // algorithm::UpdateMethod (include/openjij/algorithm/algorithm.hpp:63-76),
    // whose heat-bath rule exists only for classical sweeps.
    if (value != RULE_METROPOLIS && value != RULE_HEAT_BATH)
      return fail(OJMI_EINVAL, "update_method must be 0 (METROPOLIS) or 1 (HEAT_BATH)");
    if (c->M != 1 && value != RULE_METROPOLIS)
      return fail(OJMI_EINVAL, "HEAT_BATH applies to classical systems (num_trotter = 1)");
    c->rule = (int)value;
    return OJMI_OK;
  }
  if (k == "batch_mode") {
    if (c->kind != 2) return fail(OJMI_EINVAL, "batch_mode applies to batch contexts");
    if (value < 0 || value > 2) return fail(OJMI_EINVAL, "batch_mode must be 0, 1 or 2");
    if (value == 2 && !c->d_q_items)
      return fail(OJMI_ENOTSUP, "persistent batch launch: too many (problem, replica) items");
    c->batch_mode = (int)value;
    c->family = value == 0 ? OJMI_KERNEL_BATCH_FUSED
                           : (value == 1 ? OJMI_KERNEL_BATCH_WAVE : OJMI_KERNEL_BATCH_PERSIST);
    return OJMI_OK;
  }
  if (k == "batch_waves") {
    if (c->kind != 2) return fail(OJMI_EINVAL, "batch_waves applies to batch contexts");
    if (value < 1 || value > 4) return fail(OJMI_EINVAL, "batch_waves must be 1..4");
    c->batch_waves = (int)value;
    return OJMI_OK;
  }
  if (k == "csr_waves") {
    if (!(value == 0 || value == 1 || value == 2 || value == 3 || value == 4 || value == 8))
      return fail(OJMI_EINVAL, "csr_waves must be 0, 1, 2, 3, 4 or 8");
    c->csr_waves = (int)value;
    return OJMI_OK;
  }
  if (k == "xcd_window") {
    if (value < 0 || value > 1000000) return fail(OJMI_EINVAL, "xcd_window must be >= 0");
    c->xcd_window = (int)value;
    return OJMI_OK;
  }
  return fail(OJMI_EINVAL, "unknown option: " + k);
}

int ojmi_set_spins(ojmi_ctx* c, const int8_t* spins) {
  if (!c || !spins) return fail(OJMI_EINVAL, "null argument");
  const size_t cnt = (size_t)(c->R * c->M * c->n);
  for (size_t k = 0; k < cnt; ++k)
    if (spins[k] != 1 && spins[k] != -1)
      return fail(OJMI_EINVAL, "spins must be +1 or -1");
  OJMI_HIP(hipMemcpy(c->dspins, spins, cnt, hipMemcpyHostToDevice));
  c->fields_valid = false;
  c->energies_valid = false;
  return OJMI_OK;
}

int ojmi_get_spins(ojmi_ctx* c, int8_t* spins) {
  if (!c || !spins) return fail(OJMI_EINVAL, "null argument");
  OJMI_HIP(hipStreamSynchronize(c->st));
  OJMI_HIP(hipMemcpy(spins, c->dspins, (size_t)(c->R * c->M * c->n), hipMemcpyDeviceToHost));
  return OJMI_OK;
}

int ojmi_set_gamma(ojmi_ctx* c, double gamma) {
  if (!c) return fail(OJMI_EINVAL, "null context");
  c->gamma = gamma;
  return OJMI_OK;
}

int ojmi_run(ojmi_ctx* c, const double* beta, const double* s, const int64_t* mcs, int64_t L,
             uint64_t seed, uint64_t sweep_offset) {
  if (!c) return fail(OJMI_EINVAL, "null context");
  if (L < 0 || (L > 0 && (!beta || !mcs))) return fail(OJMI_EINVAL, "bad schedule arrays");
  if (c->kind == 2) {  // one schedule for every problem of a batch
    std::vector<double> rows((size_t)(c->K * L));
    for (int64_t k = 0; k < c->K; ++k)
      for (int64_t l = 0; l < L; ++l) rows[(size_t)(k * L + l)] = beta[l];
    return ojmi_run_batch(c, rows.data(), mcs, L, seed, sweep_offset);
  }
  const bool transverse = c->M > 1;
  if (transverse && L > 0 && !s) return fail(OJMI_EINVAL, "transverse system needs s[]");
  int64_t T = 0;
  for (int64_t l = 0; l < L; ++l) {
    if (mcs[l] < 0) return fail(OJMI_EINVAL, "one_mc_step must be >= 0");
    if (!(beta[l] >= 0.0)) return fail(OJMI_EINVAL, "schedule beta range is '0 <= beta'.");
    T += mcs[l];
  }
  if (sweep_offset + (uint64_t)T > 0xffffffffull)
    return fail(OJMI_EINVAL, "global sweep index exceeds 32 bits");
  c->last_T = T;
  c->last_attempts = T * c->R * c->M * c->n;
  c->last_ms = 0;
  c->last_refresh_ms = 0;
  c->last_refreshes = 0;
  c->last_flips = c->last_rows = 0;
  c->last_mfma_segments = 0;
  if (T == 0) return OJMI_OK;
  int rc;
  if ((rc = ensure_sched(c, T))) return rc;
  std::vector<double> p0((size_t)T), p1((size_t)T, 0.0);
  std::vector<float> p0f((size_t)T), p1f((size_t)T);
  {
    int64_t t = 0;
    const double Md = (double)c->M;
    for (int64_t l = 0; l < L; ++l) {
      double a0, a1 = 0.0;
      if (!transverse) {
        a0 = -1.0 / beta[l];
      } else {
        // include/openjij/updater/single_spin_flip.hpp:147-148,216-217
        a0 = s[l] * (beta[l] / Md);
        a1 = (1 / 2.) * std::log(std::tanh(beta[l] * c->gamma * (1.0 - s[l]) / Md));
      }
      for (int64_t k = 0; k < mcs[l]; ++k, ++t) {
        p0[t] = a0;
        p1[t] = a1;
      }
    }
    for (int64_t k = 0; k < T; ++k) {
      p0f[k] = (float)p0[k];
      p1f[k] = (float)p1[k];
    }
  }
  OJMI_HIP(hipMemcpyAsync(c->dp0_d, p0.data(), T * 8, hipMemcpyHostToDevice, c->st));
  OJMI_HIP(hipMemcpyAsync(c->dp1_d, p1.data(), T * 8, hipMemcpyHostToDevice, c->st));
  OJMI_HIP(hipMemcpyAsync(c->dp0_f, p0f.data(), T * 4, hipMemcpyHostToDevice, c->st));
  OJMI_HIP(hipMemcpyAsync(c->dp1_f, p1f.data(), T * 4, hipMemcpyHostToDevice, c->st));
  if (c->family != OJMI_KERNEL_CSR_LEVEL && c->family != OJMI_KERNEL_SQA_WAVE &&
      c->family != OJMI_KERNEL_SQA_CSR && (rc = ensure_fields(c)))
    return rc;
  const bool msweep = uses_mfma_sweep(c);
  if ((c->family == OJMI_KERNEL_DENSE_SHARED || msweep) && (rc = ensure_progress(c))) return rc;
  if (msweep && (rc = ensure_mfma_packed(c))) return rc;
  c->energies_valid = false;
  if ((rc = begin_stats(c))) return rc;

  OJMI_HIP(hipEventRecord(c->e0, c->st));
  hipError_t e = hipSuccess;
  if (c->family == OJMI_KERNEL_CSR_LEVEL || c->family == OJMI_KERNEL_SQA_WAVE ||
      c->family == OJMI_KERNEL_SQA_CSR) {
    e = launch_sweeps(c, 0, T, seed, sweep_offset, nullptr);
  } else {
    // Fields: from the matrix cores (exact int8 for INT32 couplings; bf16
    // hi/lo for FP32 when a refresh period is set, refreshed every
    // refresh_every sweeps, or when the matrix-core sweep starts without
    // carried fields), else built by the sweep kernel itself in sweep order
    // (the FP32 replay's order).
    const bool mfma_i8 = c->ftype == FT_I32 && c->dJnat && c->M == 1;
    const bool refresh_bf = c->ftype == FT_F32 && c->dJhi && c->refresh_every > 0;
    const int64_t chunk = refresh_bf ? c->refresh_every : (hybrid_sweep(c) ? 16 : T);
    for (int64_t t = 0; t < T && e == hipSuccess; t += chunk) {
      const int64_t Tc = std::min<int64_t>(chunk, T - t);
      // carried fields: from the previous run, or the previous segment
      const void* F_in =
          ((t == 0 && c->fields_valid) || t > 0) && carries_fields(c) ? c->dF : nullptr;
      const bool mfma_bf = refresh_bf || (msweep && c->ftype == FT_F32 && !F_in);
      if (mfma_i8 || mfma_bf) {
        if (mfma_bf) {
          if ((size_t)(2 * c->last_refreshes + 2) > c->ev_ref.size()) {
            hipEvent_t ev[2];
            OJMI_HIP(hipEventCreate(&ev[0]));
            OJMI_HIP(hipEventCreate(&ev[1]));
            c->ev_ref.push_back(ev[0]);
            c->ev_ref.push_back(ev[1]);
          }
          OJMI_HIP(hipEventRecord(c->ev_ref[2 * c->last_refreshes], c->st));
        }
        e = mfma_fields(c);
        if (mfma_bf) OJMI_HIP(hipEventRecord(c->ev_ref[2 * c->last_refreshes + 1], c->st));
        if (mfma_bf) ++c->last_refreshes;
        if (e != hipSuccess) break;
        F_in = c->dF;
      }
      e = launch_sweeps(c, t, Tc, seed, sweep_offset, F_in);
      if (msweep && !hybrid_sweep(c)) ++c->last_mfma_segments;
    }
    c->fields_valid = carries_fields(c);
  }
  if (e != hipSuccess) return fail(OJMI_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  OJMI_HIP(hipEventRecord(c->e1, c->st));
  OJMI_HIP(hipEventSynchronize(c->e1));
  float ms = 0;
  OJMI_HIP(hipEventElapsedTime(&ms, c->e0, c->e1));
  c->last_ms = ms;
  for (int64_t q = 0; q < c->last_refreshes; ++q) {
    float rm = 0;
    OJMI_HIP(hipEventElapsedTime(&rm, c->ev_ref[2 * q], c->ev_ref[2 * q + 1]));
    c->last_refresh_ms += rm;
  }
  if (hybrid_sweep(c)) {
    int32_t sel[2] = {0, 0};
    OJMI_HIP(hipMemcpy(sel, c->dsel, sizeof(sel), hipMemcpyDeviceToHost));
    c->last_mfma_segments = sel[1];
  }
  return end_stats(c);
}

int ojmi_refresh_fields(ojmi_ctx* c) {
  if (!c) return fail(OJMI_EINVAL, "null context");
  const bool ok = c->kind == 0 && c->M == 1 &&
                  ((c->ftype == FT_F32 && c->dJhi) || (c->ftype == FT_I32 && c->dJnat));
  if (!ok)
    return fail(OJMI_ENOTSUP,
                "matrix-core fields need a classical dense FP32 or int8-coupling INT32 context");
  int rc;
  if ((rc = ensure_fields(c))) return rc;
  hipError_t e = mfma_fields(c);
  if (e != hipSuccess) return fail(OJMI_EDEVICE, std::string("field kernel: ") + hipGetErrorString(e));
  OJMI_HIP(hipStreamSynchronize(c->st));
  c->fields_valid = (c->ftype == FT_F32);
  c->energies_valid = false;
  return OJMI_OK;
}

int ojmi_get_fields(ojmi_ctx* c, double* out) {
  if (!c || !out) return fail(OJMI_EINVAL, "null argument");
  if (!c->dF || c->M != 1 || c->kind != 0)
    return fail(OJMI_EINVAL, "no fields held: run or refresh a classical dense context first");
  const int64_t n = c->n, R = c->R, Np = c->Npad;
  OJMI_HIP(hipStreamSynchronize(c->st));
  if (c->ftype == FT_F32) {
    std::vector<float> f((size_t)(R * Np));
    OJMI_HIP(hipMemcpy(f.data(), c->dF, f.size() * 4, hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < R; ++r)
      for (int64_t j = 0; j < n; ++j) out[r * n + j] = (double)f[(size_t)(r * Np + j)];
  } else if (c->ftype == FT_F64) {
    std::vector<double> f((size_t)(R * Np));
    OJMI_HIP(hipMemcpy(f.data(), c->dF, f.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < R; ++r)
      for (int64_t j = 0; j < n; ++j) out[r * n + j] = f[(size_t)(r * Np + j)];
  } else {
    std::vector<int32_t> f((size_t)(R * Np));
    OJMI_HIP(hipMemcpy(f.data(), c->dF, f.size() * 4, hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < R; ++r)
      for (int64_t j = 0; j < n; ++j) out[r * n + j] = (double)f[(size_t)(r * Np + j)];
  }
  return OJMI_OK;
}

static int compute_energies(ojmi_ctx* c) {
  if (c->energies_valid) return OJMI_OK;
  const int64_t C = c->R * c->M;
  if (c->d_ellv64) {  // sparse system on the CSR kernel
    hipError_t e = launch_csr_energy(c->dspins, (int)c->n, (int)C, c->d_order, c->d_deg,
                                     c->d_ell, c->d_ellv64, c->d_hpos64, c->dE, c->st);
    if (e != hipSuccess)
      return fail(OJMI_EDEVICE, std::string("CSR energy kernel: ") + hipGetErrorString(e));
    OJMI_HIP(hipStreamSynchronize(c->st));
    c->energies_valid = true;
    return OJMI_OK;
  }
  if (!c->dJ64) return fail(OJMI_ENOTSUP, "energies need the FP64 coupling copy");
  int rc;
  if ((rc = ensure_mfma_spins(c))) return rc;
  if (!c->dpart && (rc = dalloc(&c->dpart, 2 * energy_part_count((int)C, (int)c->Npad))))
    return rc;
  const float* F32 = (c->fields_valid && c->dF && c->M == 1 && c->ftype == FT_F32)
                         ? (const float*)c->dF
                         : nullptr;
  hipError_t e = launch_pack_spins(c->dspins, (int)c->n, (int)C, (int)c->Npad, c->dS8, nullptr,
                                   c->st);
  if (e == hipSuccess)
    e = launch_energy_f64(c->dS8, c->dJ64, c->dh64, (int)C, (int)c->Npad, F32, c->dpart,
                          c->dpart + energy_part_count((int)C, (int)c->Npad), c->dE,
                          F32 ? c->ddrift : nullptr, c->st);
  if (e != hipSuccess) return fail(OJMI_EDEVICE, std::string("energy kernel: ") + hipGetErrorString(e));
  OJMI_HIP(hipStreamSynchronize(c->st));
  c->energies_valid = true;
  return OJMI_OK;
}

int ojmi_energies(ojmi_ctx* c, double* out) {
  if (!c || !out) return fail(OJMI_EINVAL, "null argument");
  if (c->kind == 2) {
    int rc = batch_energies(c);
    if (rc) return rc;
    OJMI_HIP(hipMemcpy(out, c->dE, (size_t)(c->K * c->R) * 8, hipMemcpyDeviceToHost));
    return OJMI_OK;
  }
  int rc = compute_energies(c);
  if (rc) return rc;
  OJMI_HIP(hipMemcpy(out, c->dE, (size_t)(c->R * c->M) * 8, hipMemcpyDeviceToHost));
  return OJMI_OK;
}

int ojmi_energies_static(ojmi_ctx* c, double* out) {
  if (!c || !out) return fail(OJMI_EINVAL, "null argument");
  if (c->kind == 2) return fail(OJMI_ENOTSUP, "ojmi_energies_static: batch contexts use ojmi_energies");
  const int64_t C = c->R * c->M, Np = c->Npad;
  if (c->d_ellv64) {  // sparse system: the CSR energy kernel needs only E[C]
    if ((size_t)C * 8 > STATIC_SCRATCH_BYTES)
      return fail(OJMI_ERANGE, "ojmi_energies_static: " + std::to_string(C) +
                                   " configurations exceed the static buffer");
    void* base = nullptr;
    hipError_t e = static_scratch(c->device, &base);
    if (e != hipSuccess)
      return fail(OJMI_EDEVICE, std::string("static scratch lookup: ") + hipGetErrorString(e));
    double* E = static_cast<double*>(base);
    e = launch_csr_energy(c->dspins, (int)c->n, (int)C, c->d_order, c->d_deg, c->d_ell,
                          c->d_ellv64, c->d_hpos64, E, c->st);
    if (e != hipSuccess)
      return fail(OJMI_EDEVICE, std::string("CSR energy kernel: ") + hipGetErrorString(e));
    OJMI_HIP(hipMemcpyAsync(out, E, (size_t)C * 8, hipMemcpyDeviceToHost, c->st));
    OJMI_HIP(hipStreamSynchronize(c->st));
    return OJMI_OK;
  }
  if (!c->dJ64) return fail(OJMI_ENOTSUP, "energies need the FP64 coupling copy");
  const size_t parts = energy_part_count((int)C, (int)Np);
  const size_t s8 = (((size_t)(C * Np) + 255) / 256) * 256;
  const size_t need = s8 + 2 * parts * 8 + (size_t)C * 8;
  if (need > STATIC_SCRATCH_BYTES)
    return fail(OJMI_ERANGE, "ojmi_energies_static: " + std::to_string(C) + " configurations x " +
                                 std::to_string(Np) + " padded spins need " +
                                 std::to_string(need) + " bytes of scratch; the static buffer holds " +
                                 std::to_string(STATIC_SCRATCH_BYTES));
  void* base = nullptr;
  hipError_t e = static_scratch(c->device, &base);
  if (e != hipSuccess)
    return fail(OJMI_EDEVICE, std::string("static scratch lookup: ") + hipGetErrorString(e));
  int8_t* S8 = static_cast<int8_t*>(base);
  double* part_e = reinterpret_cast<double*>(static_cast<char*>(base) + s8);
  double* part_d = part_e + parts;
  double* E = part_d + parts;
  e = launch_pack_spins(c->dspins, (int)c->n, (int)C, (int)Np, S8, nullptr, c->st);
  if (e == hipSuccess)
    e = launch_energy_f64(S8, c->dJ64, c->dh64, (int)C, (int)Np, nullptr, part_e, part_d, E,
                          nullptr, c->st);
  if (e != hipSuccess) return fail(OJMI_EDEVICE, std::string("energy kernel: ") + hipGetErrorString(e));
  OJMI_HIP(hipMemcpyAsync(out, E, (size_t)C * 8, hipMemcpyDeviceToHost, c->st));
  OJMI_HIP(hipStreamSynchronize(c->st));
  return OJMI_OK;
}

int ojmi_field_drift(ojmi_ctx* c, double* out) {
  if (!c || !out) return fail(OJMI_EINVAL, "null argument");
  *out = 0.0;
  if (!(c->fields_valid && c->dF && c->M == 1 && c->kind == 0 && c->ftype == FT_F32))
    return OJMI_OK;
  int rc = compute_energies(c);
  if (rc) return rc;
  std::vector<double> d((size_t)c->R);
  OJMI_HIP(hipMemcpy(d.data(), c->ddrift, d.size() * 8, hipMemcpyDeviceToHost));
  double m = 0;
  for (double x : d) m = std::max(m, x);
  *out = m;
  return OJMI_OK;
}

int ojmi_solution(ojmi_ctx* c, int8_t* out) {
  if (!c || !out) return fail(OJMI_EINVAL, "null argument");
  const int64_t n = c->n, M = c->M, R = c->R;
  std::vector<int8_t> sp((size_t)(R * M * n));
  OJMI_HIP(hipStreamSynchronize(c->st));
  OJMI_HIP(hipMemcpy(sp.data(), c->dspins, sp.size(), hipMemcpyDeviceToHost));
  if (M == 1) {
    std::memcpy(out, sp.data(), sp.size());
    return OJMI_OK;
  }
  std::vector<double> E((size_t)(R * M));
  int rc = ojmi_energies(c, E.data());
  if (rc) return rc;
  for (int64_t r = 0; r < R; ++r) {
    int64_t tb = 0;
    for (int64_t t = 1; t < M; ++t)
      if (E[r * M + t] < E[r * M + tb]) tb = t;
    std::memcpy(out + r * n, &sp[(size_t)((r * M + tb) * n)], n);
  }
  return OJMI_OK;
}

int ojmi_get_info(ojmi_ctx* c, ojmi_info* info) {
  if (!c || !info) return fail(OJMI_EINVAL, "null argument");
  info->num_spins = c->n;
  info->num_replicas = c->R;
  info->num_trotter = c->M;
  info->precision = c->prec;
  info->kernel = c->family;
  if (c->family == OJMI_KERNEL_DENSE_SHARED && c->colown && colown_supported((int)c->NS, c->group))
    info->kernel = OJMI_KERNEL_DENSE_COLOWN;
  if (uses_mfma_sweep(c)) info->kernel = OJMI_KERNEL_DENSE_MFMA;
  info->refresh_every = c->refresh_every;
  info->coupling_bytes = (c->jtype == JT_F64) ? 8 : (c->jtype == JT_I8 ? 1 : 4);
  info->last_sweeps = c->last_T;
  info->last_attempts = c->last_attempts;
  info->last_kernel_ms = c->last_ms;
  info->last_refresh_ms = c->last_refresh_ms;
  info->group = uses_mfma_sweep(c) ? 8 : c->group;
  info->ring_rows = c->ring_rows_used;
  info->lds_bytes = c->lds_used;
  info->lds_budget = c->lds_budget;
  info->num_problems = (c->kind == 2) ? c->K : 1;
  info->waves_per_replica =
      (c->family == OJMI_KERNEL_CSR_LEVEL)
          ? csr_waves_per_replica(c->R, c->D, c->csr_waves, (int)c->n, c->nlev, c->ftype)
          : 1 + c->helpers_used;
  info->mfma_segments = (int32_t)c->last_mfma_segments;
  info->last_refreshes = c->last_refreshes;
  info->last_flips = c->last_flips;
  info->last_rows = c->last_rows;
  return OJMI_OK;
}

int ojmi_destroy(ojmi_ctx* c) {
  if (!c) return OJMI_OK;
  (void)hipStreamSynchronize(c->st);
  delete c;
  return OJMI_OK;
}

}  // extern "C"
