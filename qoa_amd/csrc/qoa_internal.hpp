// qoa_internal.hpp — shared declarations of the gfx950 backend (not public).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/qoa_hip.h"

namespace qoa {

// ---- errors ---------------------------------------------------------------
int set_error(int code, const char* fmt, ...);
#define QOA_HIP_CHECK(expr)                                                      \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess)                                                        \
      return ::qoa::set_error(QOA_ERR_HIP, "%s failed: %s (%s:%d)", #expr,       \
                              hipGetErrorString(_e), __FILE__, __LINE__);        \
  } while (0)

// ---- kernel timing (per-kind totals via HIP events) ------------------------
enum KernelKind : int {
  KK_DENSE = 0,   // 2x2 pair sweeps (H, RX, RY, Y, U2)
  KK_DIAG = 1,    // diagonal sweeps (Z, S, T, PHASE, RZ, CZ, CPHASE)
  KK_PERM = 2,    // permutations (X, CNOT, SWAP, CSWAP, TOFFOLI)
  KK_FUSED = 3,   // LDS-tiled multi-gate passes
  KK_REDUCE = 4,  // norm / probability reductions and collapses
  KK_OTHER = 5,
  KK_BLOCK = 6,   // dense 5-qubit blocks (complex64, MFMA)
  KK_COUNT = 7
};
const char* kind_name(int k);

struct TimedLaunch {
  hipEvent_t start, stop;
  int kind;
  double bytes;
};

// Philox4x32-10 host stream for QuantumState-level draws.
struct HostRng {
  uint32_t key[2] = {0, 0};
  uint64_t ctr = 0;
  uint32_t buf[4] = {0, 0, 0, 0};
  int have = 0;
  void seed(uint64_t s);
  uint64_t next_u64();
  double next_f64();
  uint32_t below3();
};

}  // namespace qoa

// ---- the state handle -----------------------------------------------------
struct qoa_state {
  uint32_t magic;
  uint32_t n;          // qubits
  uint64_t len;        // 2^n amplitudes
  uint32_t precision;  // QOA_C128 / QOA_C64
  int device;
  void* amps;          // device buffer (double2 or float2)
  hipStream_t stream;
  double* red;         // reduction scratch: partials + result slots
  void* fused_cache;   // last fused plan and its device copy (fused.hip)
  void* block_cache;   // last dense-block plan and its device tables (blocks.hip)
  // QuantumState-level fields (src/runtime/quantum_state.rs:37-53)
  int noise_kind;
  double noise_p;
  uint32_t flags;
  qoa::HostRng rng;
  // timing
  bool timing;
  std::vector<qoa::TimedLaunch> pending;
  std::vector<hipEvent_t> event_pool;
  double kind_ms[qoa::KK_COUNT];
  double kind_bytes[qoa::KK_COUNT];
  uint64_t kind_launches[qoa::KK_COUNT];
};

namespace qoa {

constexpr uint32_t kMagic = 0x514F4148u;  // "QOAH"
constexpr int kRedBlocks = 1024;          // fixed reduction grid (deterministic order)
// Layout of qoa_state::red (doubles): [0, 2*kRedBlocks) partials, then slots.
constexpr int kSlotBase = 2 * kRedBlocks;
constexpr int kRedDoubles = kSlotBase + 16;

bool valid(const qoa_state* s);

// 32-bit index arithmetic unless the state needs more (or it is forced off).
extern bool g_force_index64;
// Batches replay their launch sequence as a captured HIP graph (option
// "batch_graphs", default on).
extern bool g_batch_graphs;
inline bool use32(const qoa_state* s) { return s->n <= 31 && !g_force_index64; }

// launch bracketing for timing
void timing_begin(qoa_state* s, int kind, double bytes);
void timing_end(qoa_state* s);
int timing_collect(qoa_state* s);

// ---- kernel launchers (kernels_*.hip); all enqueue on s->stream ------------
// Coefficients are complex numbers as (re, im) pairs.
struct Cx { double re, im; };

int launch_hadamard(qoa_state* s, Cx nf, uint32_t t, uint64_t ctrl_mask);
int launch_matrix2(qoa_state* s, const Cx m[4], uint32_t t, uint64_t ctrl_mask);
int launch_pauli_x(qoa_state* s, uint32_t t, uint64_t ctrl_mask);
int launch_pauli_y(qoa_state* s, uint32_t t, uint64_t ctrl_mask);
// Diagonal: every amplitude with (i & mask) == mask gets negated (neg=true)
// or multiplied by f.
int launch_diag_mask(qoa_state* s, uint64_t mask, bool neg, Cx f);
int launch_rz(qoa_state* s, uint64_t bit, Cx m00, Cx m11);
int launch_swap2(qoa_state* s, uint32_t q1, uint32_t q2, uint64_t ctrl_mask);
int launch_matrix4(qoa_state* s, uint32_t q1, uint32_t q2, const double* m32);
int launch_pauli_codes(qoa_state* s, const uint8_t* codes, uint32_t n_codes);
int launch_set_basis0(qoa_state* s);

// reductions: results land in device slots of s->red
// sums |a|^2 over indices with (i & selbit) == selval (selbit == 0: all) into
// slot `slot`; with two_way, the complementary sum goes into slot+1.
int launch_norm_sum(qoa_state* s, uint64_t selbit, bool two_way, int slot);
int launch_normalize_apply(qoa_state* s, int slot);
// NaN (bit 0) / Inf (bit 1) flags of the state into result slot `slot`
// (read as an unsigned int).
int launch_nonfinite(qoa_state* s, int slot);
int launch_collapse(qoa_state* s, uint64_t bit, int outcome, double factor);
int launch_reset_apply(qoa_state* s, uint64_t bit, int slot);
int launch_probabilities(qoa_state* s, double* dev_out);
int launch_convert(qoa_state* dst, const qoa_state* src);
int launch_upload_c64(qoa_state* s, const double* dev_src);
// sample.hip: `count` samples from each of `ntraj` n-qubit states `stride`
// amplitudes apart; results XOR the state's X frame (frames may be NULL).
// cs: sample_scratch_doubles(n, ntraj) doubles of device scratch (per
// trajectory the 2^n / min(2^n, 4096) chunk running sums and two slots).
inline uint64_t sample_scratch_doubles(uint32_t n, uint32_t ntraj) {
  const uint64_t len = 1ULL << n;
  return (uint64_t)ntraj * (len / (len < 4096 ? len : 4096) + 2);
}
int launch_sample(hipStream_t st, const double2* states, uint64_t stride, uint32_t n,
                  uint32_t ntraj, uint64_t seed, uint32_t count, const uint64_t* frames,
                  uint32_t frame_stride, double* cs, uint64_t* out);
int launch_download_c64(qoa_state* s, double* dev_dst);

// dense 5-qubit blocks (blocks.hip)
void block_cache_free(qoa_state* s);

// fused passes (fused.hip / planner.cpp)
struct PlanOptions {
  uint32_t tile_qubits;
  uint32_t low_qubits;
  uint32_t flags;
};
int plan_order(uint32_t n, const qoa_gate* g, uint32_t ng, const PlanOptions& opt,
               std::vector<uint32_t>* order, uint32_t* n_passes);
int run_fused(qoa_state* s, const qoa_gate* g, uint32_t ng, const double* mats,
              const PlanOptions& opt);
void fused_cache_free(qoa_state* s);

// one gate with layer-1 semantics (api.cpp)
int apply_gate(qoa_state* s, const qoa_gate* g, const double* mats, uint32_t flags);

}  // namespace qoa
