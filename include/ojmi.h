/*
 * ojmi.h -- C ABI of the MI355X (gfx950) single-spin-flip annealing engine.
 *
 * Drop-in boundary for OpenJij's single-spin-flip hot path:
 *   Algorithm<SingleSpinFlip>::run over ClassicalIsing / TransverseIsing on
 *   Dense / CSRSparse graphs, driven by a schedule list.
 * Plain C: pointers and sizes only, no exceptions cross it.  Host buffers are
 * borrowed for the duration of a call (copied in or out).  One context owns
 * its device buffers and one HIP stream; distinct contexts may be used from
 * distinct threads concurrently, one context must not.
 *
 * Problem layout (the reference's, include/openjij/graph/dense.hpp:117-140 and
 * include/openjij/utility/eigen.hpp:66-81): an (n+1) x (n+1) matrix J_ext with
 * J_ext[i][j] = J_ij (i != j < n), J_ext[i][n] = h_i, zero diagonal for i < n,
 * J_ext[n][n] = 1.  Only the upper triangle (j >= i) is read, as the
 * reference symmetrises with selfadjointView<Upper> (dense.hpp:139,
 * csr_sparse.hpp:96-101).  Energy E(s) = sum_{i<j} J_ij s_i s_j + sum_i h_i s_i
 * (dense.hpp:176-192 without the dummy-corner constant).
 *
 * Random stream: Philox4x32-10, key = seed, counter (i>>2, t, r, m), word i&3,
 * for spin i, global sweep t, replica r, Trotter slice m.  A spin flips iff
 * dE < log(U) * (-1/beta) (classical) or dE < -log(U) (Trotter, dE already
 * dimensionless), which is the reference's Metropolis rule
 * (include/openjij/updater/single_spin_flip.hpp:80-82, :240) in threshold form.
 *
 * Status codes map to the exceptions the reference raises through pybind11:
 * OJMI_EINVAL -> ValueError, OJMI_ERANGE -> IndexError, others -> RuntimeError.
 */
#ifndef OJMI_H
#define OJMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OJMI_OK 0
#define OJMI_EINVAL 1   /* bad argument (std::invalid_argument)            */
#define OJMI_ERANGE 2   /* size / index mismatch (std::out_of_range)       */
#define OJMI_EDEVICE 3  /* HIP error or no device                          */
#define OJMI_ENOMEM 4   /* device allocation failed                        */
#define OJMI_ENOTSUP 5  /* shape not supported by the GPU kernels          */
#define OJMI_ERUNTIME 6 /* invalid model data (std::runtime_error)         */

/* arithmetic of the sweep */
#define OJMI_PREC_AUTO 0  /* INT32 for integer couplings, else FP64          */
#define OJMI_PREC_FP64 1  /* plain: FP64 couplings and fields               */
#define OJMI_PREC_FP32 2  /* mixed: FP32 couplings/fields, FP64 energies    */
#define OJMI_PREC_INT32 3 /* exact: int8/int32 couplings, int32 fields      */

/* kernel family that runs a context (reported by ojmi_get_info) */
#define OJMI_KERNEL_DENSE_WAVE 1   /* one wave64 per replica, fields in VGPRs   */
#define OJMI_KERNEL_SQA_WAVE 2     /* one wave64 per Trotter slice              */
#define OJMI_KERNEL_CSR_LEVEL 3    /* 1-8 wave64s per replica, level-parallel  */
#define OJMI_KERNEL_DENSE_SHARED 4 /* replicas of a workgroup share J rows     */
#define OJMI_KERNEL_DENSE_COLOWN 5 /* as 4, each wave owns a slice of the fields */
#define OJMI_KERNEL_BATCH_WAVE 6   /* batch of problems, one wave per replica   */
#define OJMI_KERNEL_DENSE_LDS 7    /* N > 4096: one workgroup per replica, fields in LDS */
#define OJMI_KERNEL_BATCH_PERSIST 8 /* batch of problems, one persistent launch, LPT queues */
#define OJMI_KERNEL_BATCH_FUSED 9   /* batch of problems, one launch, largest problem first */
#define OJMI_KERNEL_DENSE_MFMA 10   /* 8 replicas per workgroup, row updates on the matrix cores */
#define OJMI_KERNEL_SQA_CSR 11      /* Trotter CSR: one workgroup per replica, a wave per slice */

typedef struct ojmi_ctx ojmi_ctx;

typedef struct ojmi_info {
  int64_t num_spins;      /* n                                               */
  int64_t num_replicas;   /* R                                               */
  int64_t num_trotter;    /* M (1 for classical)                             */
  int32_t precision;      /* resolved OJMI_PREC_*                            */
  int32_t kernel;         /* OJMI_KERNEL_*                                   */
  int32_t refresh_every;  /* sweeps between matrix-core field refreshes (0: none) */
  int32_t coupling_bytes; /* stored bytes per coupling (1, 4 or 8)           */
  int64_t last_sweeps;    /* expanded sweeps of the last ojmi_run            */
  int64_t last_attempts;  /* flip attempts of the last ojmi_run              */
  double last_kernel_ms;  /* device time of the last ojmi_run (HIP events)   */
  double last_refresh_ms; /* part of it spent in matrix-core field refreshes */
  int32_t group;          /* replicas per workgroup (1 = one wave each)      */
  int32_t ring_rows;      /* coupling rows per LDS staging round (shared)    */
  int32_t lds_bytes;      /* LDS per workgroup of the last launch            */
  int32_t lds_budget;     /* device LDS limit per workgroup                  */
  int64_t num_problems;   /* problems of a batch context (1 otherwise)       */
  int32_t waves_per_replica; /* CSR: waves sharing one replica's sweep;     */
                             /* dense one-wave form: 1 + threshold helpers */
  int32_t mfma_segments;  /* sweep segments of the last run on the matrix-core sweep */
  int64_t last_refreshes; /* matrix-core field computations of the last run  */
  int64_t last_flips;     /* accepted flips of the last run (counted on the GPU) */
  int64_t last_rows;      /* coupling rows the last run streamed: each row read
                             once per workgroup per block it was flipped in,
                             (per replica for one-wave-per-replica kernels)  */
} ojmi_info;

/* Thread-local message of the last failing call on this thread. */
const char* ojmi_last_error(void);

/* Number of visible HIP devices (0 when none). */
int ojmi_device_count(int* count);

/*
 * Dense context.  Replaces the construction of
 *   ClassicalIsing<Dense>   (include/openjij/system/classical_ising.hpp:60-69)
 *   TransverseIsing<Dense>  (include/openjij/system/transverse_ising.hpp:70-119)
 * for num_replicas independent copies.  num_trotter = 1 selects the classical
 * system; num_trotter >= 2 the Suzuki-Trotter one with transverse field gamma
 * (num_trotter = 0 or < 0 is rejected; the reference rejects M < 2 for the
 * transverse system, transverse_ising.hpp:78-81).
 * J_ext: (n+1)*(n+1) doubles, row-major.  Spins start at +1.
 */
int ojmi_create_dense(int64_t n, const double* J_ext, int64_t num_replicas, int64_t num_trotter,
                      double gamma, int32_t precision, ojmi_ctx** out);

/*
 * CSR context.  Replaces ClassicalIsing<CSRSparse> / TransverseIsing<CSRSparse>
 * (include/openjij/system/classical_ising.hpp:204-212,
 *  include/openjij/system/transverse_ising.hpp:441-459) built from
 * graph::CSRSparse (include/openjij/graph/csr_sparse.hpp:96-101).
 * rowptr has n+2 entries (n+1 rows), col/val have rowptr[n+1] entries.
 */
int ojmi_create_csr(int64_t n, const int64_t* rowptr, const int64_t* col, const double* val,
                    int64_t num_replicas, int64_t num_trotter, double gamma, int32_t precision,
                    ojmi_ctx** out);

/*
 * Batch context: num_problems independent classical dense problems of
 * different sizes n[k] (1..4096), each annealed as num_replicas replicas in
 * one launch whose workgroups hold the problems largest first, one wave per
 * (problem, replica) (option "batch_mode": 1 runs one launch per slot class,
 * 2 one persistent launch over per-XCD work queues; results are identical).  No single reference function corresponds: it
 * replaces running SASampler.sample / Algorithm<SingleSpinFlip>::run once per
 * problem and read (openjij/sampler/sampler.py:115-140).  J_ext holds the
 * problems' (n[k]+1)^2 matrices back to back, each in the ojmi_create_dense
 * layout (upper triangle read).  Spins and solutions are [k][replica][n[k]]
 * back to back; ojmi_energies returns [k][replica].  Problem k's random
 * stream uses k as the fourth counter word (the Trotter slice word of the
 * single-problem contexts), replica r as the third.
 */
int ojmi_create_batch(int64_t num_problems, const int64_t* n, const double* J_ext,
                      int64_t num_replicas, int32_t precision, ojmi_ctx** out);

/*
 * Anneal every replica of every problem of a batch context, problem k
 * through beta[k*L + l] for l = 0..L-1 with one_mc_step[l] sweeps each (the
 * sweep counts are shared, the temperatures are per problem).  ojmi_run on a
 * batch context applies the same beta[L] to every problem.
 */
int ojmi_run_batch(ojmi_ctx* ctx, const double* beta, const int64_t* one_mc_step, int64_t L,
                   uint64_t seed, uint64_t sweep_offset);

/*
 * Statistics of the automatic beta bounds for every problem of a batch
 * context, computed on the GPU from the batch's couplings (the dE analysis of
 * geometric_ising_beta_schedule, openjij/sampler/sa_sampler.py:882-911,937-938).
 * spins: [K][2][n_k] int8 +-1, the two random spin vectors of each problem
 * (the reference's random_spin[:-1, c] for c = 0, 1).  out: [K][5] doubles
 * {dE_max, dE_min, lin_max, mean_abs, int_min}: with dE_quad = -2 s (J s) and
 * threshold 1e-8, dE_max / dE_min are the median / minimum of the entries
 * above the threshold (1.1 / 1 if none); over the entries with |dE_quad|
 * above it, lin_max = max |h s| and mean_abs = mean |dE_quad| (-1 if none),
 * so rate_dE = lin_max / (mean_abs + 1e-8) (0.5 if none); int_min is the
 * smallest |J_ij| or |h_i| above the threshold (+inf if none; the reference
 * also counts the dummy corner 1).  The bounds follow by the reference's
 * scalar formulas (sa_sampler.py:914-968).
 */
int ojmi_batch_beta_stats(ojmi_ctx* ctx, const int8_t* spins, double* out);

/*
 * Tuning knobs (no reference counterpart).  Keys:
 *   "refresh_every": classical dense FP32 contexts: K > 0 computes the local
 *                    fields on the matrix cores (ojmi_refresh_fields) at the
 *                    start of each run and again every K sweeps of it, so that
 *                    FP32 rounding never accumulates over more than K sweeps;
 *                    0 (default) builds them in the sweep kernel and carries
 *                    them through the run (the FP32 replay's arithmetic).
 *                    Refresh points count from the start of each ojmi_run.
 *   "group":         replicas per workgroup sharing coupling rows (1, 4, 8, 16;
 *                    classical FP32 / INT32 dense systems)
 *   "xcd_window":    shared kernel: pace workgroups of one XCD to within this
 *                    many 64-spin blocks of each other (0 = off)
 *   "ring_rows":     shared kernel: coupling rows per LDS staging round (0 = auto)
 *   "round_rows":    shared kernel: rows per staging round (0 = auto)
 *   "colown":        1 (default) = run shared contexts with the column-owned
 *                    kernel where the shape allows it, 0 = LDS row staging
 *   "batch_mode":    batch contexts: 0 (default) one launch, problems largest
 *                    first; 1 one launch per slot class; 2 one persistent
 *                    launch pulling (problem, replica) items from per-XCD
 *                    longest-first queues; results are identical
 *   "batch_waves":   batch contexts, batch_mode 2: waves per workgroup (1..4)
 *   "lds_rows":      one-wave-per-replica dense contexts of <= 256 slots' worth
 *                    of rows (NS <= 4, rows <= 64 KiB): 1 (default) copies the
 *                    coupling rows to LDS once per workgroup and reads flipped
 *                    rows from there, 0 reads them from L2
 *   "lds_rows_wpb":  replicas per workgroup of that form (0 = by R: 1 up to
 *                    256 replicas, then up to 4)
 *   "sqa_csr_records": sparse Trotter contexts of < 65536 spins and degree
 *                    <= 16 (<= 8 for FP64 values that are not int8
 *                    integers): 1 (default) loads each 64-spin pass's neighbour
 *                    lists as per-position 16-byte records; 0 loads one word
 *                    per column and value; results are identical
 *   "csr_pack8":     sparse contexts whose couplings are integers in
 *                    -127..127 (FP32 / FP64 fields: and whose h are integers
 *                    of magnitude <= 2^20): 1 (default) reads the couplings
 *                    packed four per word as int8 and forms fields with 4-way
 *                    integer dot products (exact); 0 reads them in the field
 *                    type; results are identical
 *   "sqa_rows_lds":  dense Trotter contexts of <= 512 spins: 1 (default)
 *                    keeps the first 1, 2 or 4 slots (64 columns each) of
 *                    every coupling row in LDS, as many as fit beside the
 *                    slices' spins (2 of 4 at 256 FP32 spins), and reads only
 *                    the other slots of a flipped row from L2; 0 reads whole
 *                    rows from L2; results are identical
 *   "wave_helpers":  one-wave-per-replica dense contexts of n <= 512: helper
 *                    waves (1..3) per replica that compute its thresholds one
 *                    phase of 4 sweeps ahead, 0 = none, -1 (default) = 2 while
 *                    the replicas number at most twice the CUs, else none;
 *                    info.waves_per_replica reports 1 + helpers of the last run
 *   "csr_waves":     CSR contexts of degree <= 32 (16 with FP64 fields): waves
 *                    sharing one replica's sweep (1, 2, 3, 4, 8; 0 = chosen by
 *                    the replica count)
 *   "update_method": acceptance rule of classical contexts (dense, CSR,
 *                    batch): 0 = METROPOLIS (default), 1 = HEAT_BATH, the
 *                    reference's algorithm::UpdateMethod values
 *                    (include/openjij/algorithm/algorithm.hpp:63-76,
 *                    include/openjij/updater/single_spin_flip.hpp:352-376).
 *                    Heat bath flips iff dE < log(U / (1 - U)) * (-1/beta),
 *                    the threshold form of 1 / (1 + exp(beta dE)) > U.
 */
int ojmi_set_option(ojmi_ctx* ctx, const char* key, int64_t value);

/*
 * Spins: int8 +-1, [num_replicas][num_trotter][n].  Replaces reset_spins /
 * the rw `spin` and `trotter_spins` attributes
 * (include/openjij/system/classical_ising.hpp:76-80, openjij/declare.hpp:426,615).
 */
int ojmi_set_spins(ojmi_ctx* ctx, const int8_t* spins);
int ojmi_get_spins(ojmi_ctx* ctx, int8_t* spins);

/* Transverse field Gamma (the rw `gamma` attribute, openjij/declare.hpp:619). */
int ojmi_set_gamma(ojmi_ctx* ctx, double gamma);

/*
 * Anneal every replica through the schedule.  Replaces
 *   Algorithm<SingleSpinFlip>::run (include/openjij/algorithm/algorithm.hpp:26-56)
 * as bound by Algorithm_SingleSpinFlip_run (openjij/declare.hpp:687-793):
 * for each entry l, one_mc_step[l] sweeps at beta[l] (and s[l] for the
 * transverse system; pass s = NULL for the classical one).  Sweeps are
 * numbered sweep_offset, sweep_offset+1, ... in the random stream, so a run
 * split in two calls with consecutive offsets equals one call.
 * beta must be >= 0 (the reference sampler rejects beta < 0,
 * openjij/sampler/sa_sampler.py:122-125).
 */
int ojmi_run(ojmi_ctx* ctx, const double* beta, const double* s, const int64_t* one_mc_step,
             int64_t num_entries, uint64_t seed, uint64_t sweep_offset);

/*
 * get_solution (include/openjij/result/get_solution.hpp:39-84): [R][n] int8.
 * Classical: the spins.  Transverse: the first Trotter slice of minimum
 * energy, per replica.
 */
int ojmi_solution(ojmi_ctx* ctx, int8_t* out);

/* FP64 energies E(s) of every replica and slice, [R][M], evaluated on the GPU
 * (H = S J in FP64 on the matrix cores, then E = sum_j s_j (H_j / 2 + h_j) in
 * a fixed order): Dense::energy, include/openjij/graph/dense.hpp:176-192. */
int ojmi_energies(ojmi_ctx* ctx, double* out);

/*
 * ojmi_energies without allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same result as
 * ojmi_energies, bit for bit (same kernels); the padded spins and per-tile
 * partials it needs live in a static device array of 16 MiB per GPU
 * (allocated with the library's code object, held until the process ends),
 * so (R * M) * Npad + 16 * (R * M) * (Npad / 64) + 8 * (R * M) bytes must fit
 * in 16 MiB (Npad = 64 * 2^ceil(log2(ceil(n / 64))) up to 4096 spins, else
 * 64 * ceil(n / 64)); larger contexts return OJMI_ERANGE.  Every call on a
 * GPU uses that GPU's one array: calls that overlap in time on one GPU (from
 * several threads or contexts) race and must be serialised by the caller.
 * Batch contexts return OJMI_ENOTSUP.
 */
int ojmi_energies_static(ojmi_ctx* ctx, double* out);

/*
 * Local fields F = h + J s of every replica from its current spins, on the
 * matrix cores, kept in the context (the fields the next sweep starts from
 * when refresh_every > 0).  Replaces ClassicalIsing::reset_dE,
 * include/openjij/system/classical_ising.hpp:87-90 (which stores
 * dE = -2 s F), for all replicas at once.  FP32 contexts: J split into bf16
 * J_hi + J_lo (|J - J_hi - J_lo| <= 2^-18 |J|), +-1 spins exact in bf16,
 * FP32 accumulation.  INT32 contexts with |J| <= 127: int8 operands, exact
 * int32 accumulation.  Other contexts return OJMI_ENOTSUP.
 */
int ojmi_refresh_fields(ojmi_ctx* ctx);

/*
 * The context's local fields, [R][n] as doubles (exact conversions of the
 * FP32 or int32 values): after ojmi_refresh_fields, or after a run of an FP32
 * context (the fields the sweep carried to its end).  The counterpart of
 * reading ClassicalIsing::dE (classical_ising.hpp:111), as F = -dE / (2 s).
 */
int ojmi_get_fields(ojmi_ctx* ctx, double* out);

/*
 * FP32 path: max over replicas and spins of |F32 - F64| between the fields the
 * sweep carried and fields recomputed in FP64 from the final spins (0 for the
 * FP64 and INT32 paths, whose fields are exact or not kept).
 */
int ojmi_field_drift(ojmi_ctx* ctx, double* max_abs_drift);

int ojmi_get_info(ojmi_ctx* ctx, ojmi_info* info);

int ojmi_destroy(ojmi_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* OJMI_H */
