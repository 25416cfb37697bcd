// ojmi_kernels.h -- argument blocks and launchers shared between the C ABI
// (ojmi_api.hip) and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ojmi {

enum FieldType { FT_F64 = 0, FT_F32 = 1, FT_I32 = 2 };
enum StoreType { JT_F64 = 0, JT_F32 = 1, JT_I8 = 2, JT_I32 = 3 };

// Per-sweep parameters (one entry per expanded sweep t):
//   classical: p0 = -1/beta
//   transverse: p0 = s*beta/M, p1 = 1/2 log tanh(beta*Gamma*(1-s)/M)
struct DenseArgs {
  const void* J;       // device coupling rows (layout in ojmi_wave.h)
  const void* h;       // Npad fields' constant term (h_i), field type
  int8_t* spins;       // [R][M][n] +-1
  int n, R, M, T;
  uint64_t seed, t0;   // Philox key, first global sweep index
  int rep_offset;      // replica index offset of replica 0 in the stream
  const double* p0_d;
  const float* p0_f;
  const double* p1_d;
  const float* p1_f;
  const void* F_in;    // optional [R][Npad] initial fields (field type)
  void* F_out;         // optional [R][Npad] final fields
  int rule;            // classical acceptance rule: RULE_METROPOLIS / RULE_HEAT_BATH
  unsigned long long* stats = nullptr;  // optional [2]: += accepted flips, coupling rows streamed
  int cache_rows = 0;  // Trotter kernel: slots of every coupling row held in LDS
};

// Batch of independent dense classical problems (k_batch.hip, k_batch_wave):
// per-problem offsets into concatenated buffers; `inst` lists the problems of
// one launch (one slot class), each annealed as R replicas.
struct BatchArgs {
  const void* J;         // couplings, problem k at byte offset j_off[k]
  const void* h;         // fields' constant term, problem k at element h_off[k]
  int8_t* spins;         // problem k: [R][n_k] at s_off[k]
  const int64_t* j_off;
  const int64_t* h_off;
  const int64_t* s_off;
  const int32_t* n;      // [K] spins per problem
  const int32_t* inst;   // [num_items] problems of this launch
  int num_items, R, T;
  uint64_t seed, t0;
  const double* p0_d;    // [K][T] -1/beta per sweep
  const float* p0_f;
  int rule;
  unsigned long long* stats = nullptr;  // optional [2]: += accepted flips, coupling rows streamed
};

// CSR in level order (k_csr.hip): spins sorted by level, neighbours in ELL
// form [D][n] indexed by level-order position.
struct CsrArgs {
  const int32_t* order;    // [n] spin index at level-order position
  const int32_t* lvl_ptr;  // [nlev+1]
  int nlev;
  const int32_t* deg;      // [n] degree at position
  const int32_t* ell_col;  // [D][n]
  const void* val;         // [D][n], field type
  int D;
  const void* h;           // [n] at position, field type
  int8_t* spins;           // [R][n]
  int n, R, M, T;
  uint64_t seed, t0;
  int rep_offset;
  const double* p0_d;
  const float* p0_f;
  int rule;
  unsigned long long* stats = nullptr;  // optional [2]: += accepted flips, coupling rows streamed
  const double* p1_d = nullptr;  // Trotter: B per sweep (p0 = s beta / M)
  const float* p1_f = nullptr;
  // INT32 fields whose couplings fit int8: the values packed four per word,
  // word (q, pos) at q * n + pos holding entries 4q .. 4q+3 (else nullptr)
  const uint32_t* val8 = nullptr;
  // position records of the workgroup form (n < 65536): 16-byte chunk k of
  // position pos at k * n + pos; words: 16-bit column pairs, the values (int8
  // quadruples when val8 is set, else the field type), the spin index, h
  const uint4* rec = nullptr;
  int pack8 = 0;  // rec holds int8 value quadruples and integer h (integer couplings within int8)
};

// Workgroup-shared dense sweep (k_shared.hip): J in natural row layout
// (element j of row i at i*Npad + j), G replicas per workgroup.
struct SharedArgs {
  const void* J;
  const void* h;
  int8_t* spins;
  int n, R, T;
  uint64_t seed, t0;
  int rep_offset;
  const double* p0_d;
  const float* p0_f;
  const void* F_in;
  void* F_out;
  int ring_rows;        // coupling rows held in the LDS ring (multiple of round_rows)
  int round_rows;       // rows per staging round (ring_rows / round_rows >= 2 buffers)
  int xcd_window;       // 0: no pacing; else max lead (blocks) over XCD peers
  int xcd_spin;         // bound on pacing polls per block
  uint32_t* progress;   // [workgroups] blocks completed
  int rule;
  unsigned long long* stats = nullptr;  // optional [2]: += accepted flips, coupling rows streamed
  // optional kernel selector (column-owned and matrix-core sweeps): the kernel
  // returns at once unless *sel == sel_id
  const int32_t* sel = nullptr;
  int sel_id = 0;
};

size_t shared_lds_bytes(int NS, int G, int ring_rows, int fbytes);
// True when every workgroup of a grid of `blocks` can be resident on the device
// at once (occupancy of `kern` at this block size and LDS times the CU count).
bool all_resident(const void* kern, int blocks, int threads, size_t lds);
hipError_t launch_dense_shared(int ftype, int jtype, const SharedArgs& a, int NS, int G,
                               hipStream_t st);

size_t colown_lds_bytes(int n, int NS, int G, int jbytes, int fbytes);
bool colown_supported(int NS, int G);
hipError_t launch_dense_colown(int ftype, int jtype, const SharedArgs& a, int NS, int G,
                               hipStream_t st);

// Dense sweep with the row updates on the matrix cores (k_mfma_sweep.hip):
// 8 replicas per workgroup, FP32 fields with bf16 hi/lo couplings or INT32
// fields with int8 couplings; a.J natural couplings (diagonal tiles), Jpk the
// packed operands written by launch_pack_mfma, a.F_in required.
size_t mfma_sweep_lds_bytes(int NS, int fbytes, int jbytes);
bool mfma_sweep_supported(int NS);
size_t mfma_packed_bytes(int ftype, int NS);
// FP32: Jhi, Jlo natural bf16 couplings; INT32: Jhi natural int8, Jlo unused.
hipError_t launch_pack_mfma(int ftype, const void* Jhi, const void* Jlo, int NS, void* out,
                            hipStream_t st);
// form 0: every wave decides and updates; form 1: four waves update the
// fields while four decide (the same results)
hipError_t launch_dense_mfma(int ftype, const SharedArgs& a, const void* Jpk, int NS, int form,
                             hipStream_t st);
// Kernel choice of the next sweep segment from the accepted flips of the last
// one: sel = 1 (matrix-core sweep) when flips * 1000 >= permille * attempts,
// else 0; init: start a run (sel = 1, flip snapshot taken).  sel[0] is the
// choice, sel[1] counts the segments run on the matrix-core sweep, sel[2..3]
// the snapshot.
hipError_t launch_sweep_select(const unsigned long long* stats, int32_t* sel, uint64_t attempts,
                               int permille, int init, hipStream_t st);
hipError_t launch_dense_wave(int ftype, int jtype, const DenseArgs& a, int NS, hipStream_t st);
// The wave kernel with the coupling rows staged in LDS (NS <= 4): bytes of
// LDS per workgroup, and the launch with wpb waves (replicas) per workgroup.
size_t wave_lds_bytes(int n, int NS, int jbytes);
hipError_t launch_dense_wave_lds(int ftype, int jtype, const DenseArgs& a, int NS, int wpb,
                                 hipStream_t st);
// The wave anneal of one replica per workgroup with H = 1..3 helper waves
// making its thresholds one phase ahead (NS <= 8); rows from LDS or L2.
size_t wave_pipe_lds_bytes(int n, int NS, int jbytes, int tbytes, bool rows_lds);
hipError_t launch_dense_wave_pipe(int ftype, int jtype, const DenseArgs& a, int NS, int H,
                                  bool rows_lds, hipStream_t st);
hipError_t launch_batch_wave(int ftype, int jtype, const BatchArgs& b, int NS, hipStream_t st);
// Classical dense systems too large for register-resident fields (k_lds.hip):
// natural coupling rows of Npad = 64 * ceil(n / 64), fields in LDS.
size_t lds_kernel_bytes(int Npad, int fbytes);
hipError_t launch_dense_lds(int ftype, int jtype, const DenseArgs& a, int Npad, hipStream_t st);
hipError_t launch_batch_energy(int ftype, int jtype, const BatchArgs& b, const int64_t* rb, int K,
                               double* E, hipStream_t st);
hipError_t launch_sqa_wave(int ftype, const DenseArgs& a, int NS, int PPW, int W, hipStream_t st);
// Trotter kernel: a.cache_rows = slots (of 64 columns) of every coupling row
// held in LDS, as chosen by sqa_lds_slots for an LDS budget per workgroup.
int sqa_lds_slots(int ftype, int NS, int n, int M, size_t budget);
// ELL width of the CSR kernels for maximum degree D (8, 16, 32 or 64).
int csr_padded_width(int D);
// Waves per replica the CSR launcher uses (1 = one wave per replica, several
// replicas per workgroup); `requested` 0 chooses by R; D, n, nlev and the
// field type decide whether the workgroup form applies at all.
int csr_waves_per_replica(int64_t R, int D, int requested, int n, int nlev, int ftype);
// waves_per_block: replicas per workgroup of the wave-per-replica form;
// waves: 0 (choose), 1, 2, 4 or 8 waves per replica.
hipError_t launch_csr_level(int ftype, const CsrArgs& a, int waves_per_block, int waves,
                            hipStream_t st);

// Trotter sweeps of CSR systems (k_sqa_csr.hip): one workgroup per replica,
// spins of the M slices in LDS (n * M bytes), a.M >= 2, a.D the ELL width.
// stage: each wave also holds its slice's Philox words in LDS
// (sqa_csr_staged_lds_bytes in total).
size_t sqa_csr_lds_bytes(int n, int M);
size_t sqa_csr_staged_lds_bytes(int n, int M);
int sqa_csr_waves(int M);  // waves per workgroup (one slice of a phase each)
hipError_t launch_sqa_csr(int ftype, const CsrArgs& a, bool stage, hipStream_t st);

// Automatic-schedule statistics of every problem of a batch (k_beta_stats.hip):
// out[k] = {dE_max, dE_min, lin_max, mean_abs, int_min} from the spin pair
// S2[s2_off[k] .. + 2 n_k) (two rows of n_k +-1 values).
size_t beta_stats_lds_bytes(int nmax);
hipError_t launch_batch_beta_stats(int ftype, int jtype, const BatchArgs& b, const int64_t* rb,
                                   int K, int nmax, const int8_t* S2, const int64_t* s2_off,
                                   double* out, hipStream_t st);

// Matrix-core contractions (k_spin_gemm.hip).  S8 / Sbf: spins zero-padded to
// [C][Npad] (int8 / bf16 bits), written by launch_pack_spins; J matrices are
// the natural zero-padded Npad x Npad couplings (Npad a multiple of 64).
hipError_t launch_pack_spins(const int8_t* spins, int n, int C, int Kp, int8_t* S8,
                             uint16_t* Sbf, hipStream_t st);
// F[c][j] = h[j] + sum_i S[c][i] (Jhi + Jlo)[i][j], FP32 accumulation (bf16 MFMA)
hipError_t launch_fields_bf16x2(const uint16_t* Sbf, const uint16_t* Jhi, const uint16_t* Jlo,
                                const float* h, int C, int Npad, float* F, hipStream_t st);
// F[c][j] = h[j] + sum_i S[c][i] J8[i][j], exact int32 (int8 MFMA)
hipError_t launch_fields_i8(const int8_t* S8, const int8_t* J8, const int32_t* h, int C, int Npad,
                            int32_t* F, hipStream_t st);
// FP64 MFMA: E[c] = sum_j s_cj (H_cj / 2 + h_j) with H = S J64, and, when F32 is
// given, drift[c] = max_j |F32[c][j] - (H_cj + h_j)|.  part_e / part_d hold
// energy_part_count(C, Npad) doubles each.
size_t energy_part_count(int C, int Npad);
// Energies of C configurations of a sparse system from the level-order ELL
// arrays (FP64 values and fields at level-order positions), E[C].
hipError_t launch_csr_energy(const int8_t* spins, int n, int C, const int32_t* order,
                             const int32_t* deg, const int32_t* ell, const double* val,
                             const double* hpos, double* E, hipStream_t st);
// Device address (on `device`) of the static scratch array of the
// allocation-free energies, STATIC_SCRATCH_BYTES long; looked up once per
// device with that device current.
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)16 << 20;
constexpr int MAX_DEVICES = 64;
hipError_t static_scratch(int device, void** ptr);
hipError_t launch_energy_f64(const int8_t* S8, const double* J64, const double* h64, int C,
                             int Npad, const float* F32, double* part_e, double* part_d,
                             double* E, double* drift, hipStream_t st);

}  // namespace ojmi
