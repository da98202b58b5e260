/*
 * qoa_hip.h — C ABI of the MI355X (gfx950) state-vector backend for QOA's
 * gate-application path.
 *
 * Every entry point is plain C: opaque handles, plain pointers and sizes.
 * A state lives in HBM for its whole life; the host only touches it through
 * qoa_state_upload / qoa_state_download / qoa_probabilities.
 *
 * Conventions
 *   - Return value: 0 (QOA_OK) or a negative QOA_ERR_* code. Never aborts.
 *     qoa_last_error() returns a human-readable message for the calling thread.
 *   - Amplitudes are interleaved (re, im) doubles; basis index i has qubit q
 *     in bit q (LSB = qubit 0), as in src/vectorization.rs:920-929.
 *   - Each handle owns one HIP stream. Calls on a handle are stream-ordered;
 *     calls that return host data synchronise that stream. Distinct handles
 *     may be driven from distinct threads.
 *   - Gate coefficients (cos/sin, 1/sqrt 2) are computed on the host with the
 *     C library, like the reference (src/vectorization.rs:529-535, 650-651).
 *
 * Layer 1 (qoa_apply_*) mirrors the free functions of src/vectorization.rs
 * one-for-one, including their parameter meaning (qubit index vs bit mask)
 * and their out-of-range behaviour. Layer 2 (qoa_qs_*) mirrors the
 * QuantumState wrappers of src/runtime/quantum_state.rs (bounds check,
 * noise, renormalisation after every gate, measurement).
 */
#ifndef QOA_HIP_H
#define QOA_HIP_H

#if defined(__HIPCC_RTC__)  /* kernels compiled at run time see the types only */
#include <cstdint>
#else
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define QOA_OK 0
#define QOA_ERR_HANDLE (-1)     /* null or destroyed handle                  */
#define QOA_ERR_QUBIT (-2)      /* qubit index out of range (layer 2)        */
#define QOA_ERR_OOM (-3)        /* device allocation failed                  */
#define QOA_ERR_HIP (-4)        /* a HIP runtime call failed                 */
#define QOA_ERR_ARG (-5)        /* malformed argument                        */
#define QOA_ERR_ADMISSION (-6)  /* request does not fit the memory budget    */
#define QOA_ERR_UNSUPPORTED (-7)/* operation not available for this handle  */

/* ---- precisions ------------------------------------------------------ */
#define QOA_C128 0u  /* complex<double>, 16 B per amplitude (default)  */
#define QOA_C64 1u   /* complex<float>, 8 B per amplitude (fused FP32)  */

/* ---- gate op codes ---------------------------------------------------
 * Where the reference bytecode has an opcode for the gate, the same byte is
 * used (src/instructions.rs:631-830); ops without a bytecode use 0xE0+.    */
enum qoa_op {
  QOA_OP_H = 0x05,        /* q0                          (0x05 APPLYHADAMARD) */
  QOA_OP_Z = 0x06,        /* q0                          (0x06 APPLYPHASEFLIP) */
  QOA_OP_X = 0x07,        /* q0                          (0x07 APPLYBITFLIP) */
  QOA_OP_PHASESHIFT = 0x08, /* q0, theta                 (0x08 PHASESHIFT)   */
  QOA_OP_RESET = 0x0A,    /* q0                          (0x0A RESET)        */
  QOA_OP_SWAP = 0x0B,     /* q0, q1                      (0x0B SWAP)         */
  QOA_OP_CSWAP = 0x0C,    /* q0 = control, q1, q2        (0x0C CONTROLLEDSWAP) */
  QOA_OP_T = 0x0D,        /* q0                          (0x0D APPLYTGATE)   */
  QOA_OP_S = 0x0E,        /* q0                          (0x0E APPLYSGATE)   */
  QOA_OP_RZ = 0x0F,       /* q0, theta                   (0x0F RZ)           */
  QOA_OP_CNOT = 0x17,     /* q0 = control, q1 = target   (0x17 CNOT)         */
  QOA_OP_RESET_ALL = 0x1C,/*                             (0x1C RESETALL)     */
  QOA_OP_CZ = 0x1E,       /* q0 = control, q1 = target   (0x1E CZ)           */
  QOA_OP_RX = 0x22,       /* q0, theta                   (0x22 RX)           */
  QOA_OP_RY = 0x23,       /* q0, theta                   (0x23 RY)           */
  QOA_OP_PHASE = 0x24,    /* q0, theta  == PHASESHIFT    (0x24 PHASE)        */
  QOA_OP_MEASURE = 0x32,  /* q0                          (0x32 QMEAS)        */
  QOA_OP_CPHASE = 0x35,   /* q0 = control, q1 = target, theta (0x35 CPHASE)  */
  QOA_OP_Y = 0xE1,        /* q0  (reached in the reference only via noise)   */
  QOA_OP_U2 = 0xE2,       /* q0; aux = index of a 2x2 matrix (8 doubles)     */
  QOA_OP_U4 = 0xE3,       /* q0, q1; aux = index of a 4x4 matrix (32 doubles)*/
  QOA_OP_TOFFOLI = 0xE4,  /* q0 = c1, q1 = c2, q2 = target                   */
  QOA_OP_CY = 0xE5        /* q0 = control, q1 = target                       */
};

/* One gate of a circuit (32 bytes). */
typedef struct qoa_gate {
  uint32_t op;     /* enum qoa_op                         */
  uint32_t q0, q1, q2;
  double theta;    /* angle for parametrised gates        */
  uint64_t aux;    /* matrix index for U2/U4, else 0      */
} qoa_gate;

/* Flags for circuit execution. */
#define QOA_FUSE_LDS_ONLY 2u /* fused passes keep the whole tile in LDS and
                                apply every gate there (no register-resident
                                segments); for comparison and testing.      */
#define QOA_FUSE_FMA 4u      /* register-resident passes compute complex
                                products with fused multiply-adds: fewer
                                FP64 instructions, results within a few ulp
                                of the exact (num-complex) arithmetic
                                instead of bit-identical to it.              */
#define QOA_FUSE_MERGE_DIAG 16u /* register-resident passes merge each run of
                                diagonal ops (Z/S/T/PHASE, CZ/CPHASE, RZ)
                                into one op: per thread the factors that
                                apply are multiplied together first, then
                                into the amplitudes. Far fewer instructions
                                for phase ladders; results within a few ulp
                                (tested <= 1e-12) instead of bit-identical.
                                Implies QOA_FUSE_FMA. Batches take it when
                                none of their jobs is noisy; the server
                                ignores it.                                 */
#define QOA_FUSE_SPECIALIZE 64u /* register-resident passes run as kernels
                                compiled for this plan (hipRTC, on first use;
                                the code is cached per process by its
                                source): every op's words are constants, so
                                the per-op decoding of the op-loop kernel
                                folds away. Results are bit-identical to the
                                op-loop kernel with the same other flags.
                                Compiling takes seconds per circuit
                                structure; see qoa_specialize_info.        */
#define QOA_FUSE_FACTOR 128u /* register-resident passes take a scalar factor
                                out of uncontrolled H, RX, RY (m00 = m11)
                                and RZ gates and apply the pass's product of
                                those factors once (in one of its ops): H
                                becomes sums and differences, RX/RY one
                                fused multiply-add per component, RZ a phase
                                on half the amplitudes. Results within a few
                                ulp (tested <= 1e-12). Implies QOA_FUSE_FMA
                                for single states; batches (qoa_batch_create)
                                and the server (qoa_server_create) take it
                                for their jobs, noisy ones included (the
                                scalars commute with the Pauli errors).     */
#define QOA_FUSE_RELABEL 256u /* register-resident passes may store their
                                tile with its qubits at other positions of
                                the tile: the state's qubit order changes
                                from pass to pass (the order is put back
                                before the run returns), so each pass can
                                bring the qubits the next pass needs to the
                                low (contiguous) positions instead of always
                                keeping qubits 0..2 in every tile. Fewer
                                passes; data movement only (results as
                                without it). Single-state fused runs only.  */
#define QOA_CZ_REF_MASKS 1u  /* CZ/CPHASE use the integer c|t as the bit mask,
                                as QuantumState::apply_cz does
                                (src/runtime/quantum_state.rs:338, 356-361).
                                Default: textbook masks (1<<c)|(1<<t).       */

/* ---- library --------------------------------------------------------- */
int qoa_version(void);                      /* 10000*major + 100*minor + patch */
const char* qoa_last_error(void);           /* thread-local, never NULL        */
int qoa_device_info(char* name, size_t name_len, uint64_t* hbm_bytes, int* num_cus);
/* Library options. "index64" (0/1): use 64-bit index arithmetic in every
   kernel even for states of at most 31 qubits (the path states of 32+ qubits
   take), so that it can be tested on small states. "batch_graphs" (0/1,
   default 1): qoa_batch_run captures a batch's launch sequence once as a
   HIP graph and replays it; 0 launches every kernel from the host. */
int qoa_set_option(const char* key, int64_t value);

/* ---- state lifecycle: QuantumState::new (quantum_state.rs:56-90) ------- */
typedef struct qoa_state qoa_state;

/* Allocates 2^max(n,1) amplitudes in HBM initialised to |0...0>. */
int qoa_state_create(uint32_t n_qubits, uint32_t precision, qoa_state** out);
int qoa_state_destroy(qoa_state* s);
int qoa_state_num_qubits(const qoa_state* s, uint32_t* n_out);
/* Host <-> device copies of the whole vector (interleaved re,im doubles;
   for QOA_C64 handles the device converts to/from float). */
int qoa_state_upload(qoa_state* s, const double* amps, uint64_t n_amps);
int qoa_state_download(qoa_state* s, double* amps, uint64_t n_amps);
int qoa_state_synchronize(qoa_state* s);
/* Copies the amplitudes of src into dst (same n); c128 -> c64 and c64 ->
   c128 conversions are allowed. Starts after src's work and returns when the
   copy is complete, so work may follow at once on either handle. */
int qoa_state_copy(qoa_state* dst, const qoa_state* src);

/* ---- layer 1: src/vectorization.rs free functions ---------------------- */
/* apply_hadamard_vectorized(amps, norm_factor, mask_bit)   vectorization.rs:16   */
int qoa_apply_hadamard(qoa_state* s, double nf_re, double nf_im, uint32_t mask_bit);
/* apply_x_vectorized(amps, mask_bit)                       vectorization.rs:69   */
int qoa_apply_x(qoa_state* s, uint32_t mask_bit);
/* apply_y_vectorized(amps, mask_bit)                       vectorization.rs:118  */
int qoa_apply_y(qoa_state* s, uint32_t mask_bit);
/* apply_z_vectorized(amps, mask_bit)                       vectorization.rs:173  */
int qoa_apply_z(qoa_state* s, uint32_t mask_bit);
/* apply_t_vectorized(amps, mask_bit)                       vectorization.rs:217  */
int qoa_apply_t(qoa_state* s, uint32_t mask_bit);
/* apply_s_vectorized(amps, mask_bit)                       vectorization.rs:262  */
int qoa_apply_s(qoa_state* s, uint32_t mask_bit);
/* apply_phaseshift_vectorized(amps, mask_bit, angle)       vectorization.rs:307  */
int qoa_apply_phaseshift(qoa_state* s, uint32_t mask_bit, double angle);
/* apply_reset_vectorized(amps, mask_bit)                   vectorization.rs:352  */
int qoa_apply_reset(qoa_state* s, uint32_t mask_bit);
/* apply_swap_vectorized(amps, q1, q2)   (qubit indices)    vectorization.rs:421  */
int qoa_apply_swap(qoa_state* s, uint32_t q1, uint32_t q2);
/* apply_controlled_swap_vectorized(amps, c, t1, t2)        vectorization.rs:469  */
int qoa_apply_controlled_swap(qoa_state* s, uint32_t control, uint32_t target1,
                              uint32_t target2);
/* apply_rx_vectorized(amps, mask_bit, angle)               vectorization.rs:528  */
int qoa_apply_rx(qoa_state* s, uint32_t mask_bit, double angle);
/* apply_ry_vectorized(amps, mask_bit, angle)               vectorization.rs:588  */
int qoa_apply_ry(qoa_state* s, uint32_t mask_bit, double angle);
/* apply_rz_vectorized(amps, mask_bit, angle)               vectorization.rs:648  */
int qoa_apply_rz(qoa_state* s, uint32_t mask_bit, double angle);
/* apply_cnot_vectorized(amps, control, target) (indices)   vectorization.rs:699  */
int qoa_apply_cnot(qoa_state* s, uint32_t control, uint32_t target);
/* apply_cz_vectorized(amps, control_mask, target_mask) (MASKS) vectorization.rs:745 */
int qoa_apply_cz(qoa_state* s, uint64_t control_mask, uint64_t target_mask);
/* apply_controlled_phase_rotation_vectorized(amps, cmask, tmask, angle)
                                                           vectorization.rs:788  */
int qoa_apply_controlled_phase_rotation(qoa_state* s, uint64_t control_mask,
                                        uint64_t target_mask, double angle);
/* apply_reset_all_vectorized(amps)                         vectorization.rs:857  */
int qoa_apply_reset_all(qoa_state* s);

/* General 2x2 on qubit q: m = {m00re,m00im,m01re,m01im,m10re,m10im,m11re,m11im},
   computed as m00*a0 + m01*a1, m10*a0 + m11*a1 (apply_custom_unitary,
   quantum_state.rs:1010-1062, without its noise/normalize step). */
int qoa_apply_matrix2(qoa_state* s, uint32_t q, const double* m);
/* General 4x4 on (q1,q2) with basis order (bit_high<<1)|bit_low, row-major
   32 doubles (apply_two_qubit_unitary, quantum_state.rs:1064-1127). */
int qoa_apply_matrix4(qoa_state* s, uint32_t q1, uint32_t q2, const double* m);
/* Toffoli (quantum_state.rs:960-1008) and controlled-Y (:1163-1212) sweeps. */
int qoa_apply_toffoli(qoa_state* s, uint32_t c1, uint32_t c2, uint32_t t);
int qoa_apply_controlled_y(qoa_state* s, uint32_t c, uint32_t t);

/* ---- reductions and collapse (quantum_state.rs:128-133, 399-458, 888-907) */
/* validate_state (quantum_state.rs:135-167) on the device: whether any
   amplitude has a NaN / infinite component, and sum |a_i|^2; the caller
   applies the reference's order of checks and its 1e-9 norm tolerance. */
int qoa_validate(qoa_state* s, int* has_nan, int* has_inf, double* norm_sqr);
/* sum_i |a_i|^2 (deterministic fixed-shape tree on the device). */
int qoa_norm_sqr(qoa_state* s, double* out);
/* normalize(): s = sum |a|^2; if s > 1e-12 divide every amplitude by sqrt(s),
   else reset to |0...0>. Optional outputs (may be NULL). */
int qoa_normalize(qoa_state* s, double* norm_sqr_out, int* was_reset);
/* measure(q) with the uniform variate r supplied by the caller: p0/p1 sums,
   outcome = (r < p0) ? 0 : 1, collapse with factor 1/sqrt(p) or 0 when
   p <= 1e-12. Does NOT renormalise (the caller's wrapper does). */
int qoa_measure(qoa_state* s, uint32_t q, double r, int* outcome, double* p0,
                double* p1);
/* |a_i|^2 for every i into out[0 .. n_amps). */
int qoa_probabilities(qoa_state* s, double* out, uint64_t n_amps);
/* Draws `count` Z-basis outcomes from |a_i|^2 (get_probabilities,
   quantum_state.rs:128-133) by inverse CDF on the device. Uniform s is
   Philox4x32-10(counter {0xFFFFFFFF, s, 0, 0x5A3E1D2C}, key seed); the CDF is
   summed in a fixed order (chunks of 4096, runs of 16), so results are
   reproducible and independent of timing. The state is not collapsed. */
int qoa_sample(qoa_state* s, uint64_t seed, uint32_t count, uint64_t* out);
/* qoa_sample without any device allocation. No counterpart in the
   reference: for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc). Its
   scratch is a fixed per-device array of about 8.5 MiB, held until the
   process ends, so it accepts states of at most 32 qubits and at most 65536
   samples per call (QOA_ERR_ARG beyond). Same results as qoa_sample. Calls
   that overlap on one GPU (from several threads, or several handles on one
   device) share that scratch and race; calls in sequence do not. */
int qoa_sample_noalloc(qoa_state* s, uint64_t seed, uint32_t count, uint64_t* out);
/* One depolarizing event set: for every qubit q with code[q] in {1,2,3}
   apply X, Y or Z on q (0 = none), in qubit order, as one fused sweep
   (apply_depolarizing_noise, quantum_state.rs:872-885). */
int qoa_apply_pauli_codes(qoa_state* s, const uint8_t* codes, uint32_t n_codes);

/* ---- circuits ---------------------------------------------------------- */
/* Runs gates in order with one sweep kernel per gate (layer-1 semantics: no
   renormalisation, no noise). MEASURE/RESET are not allowed here. */
int qoa_run_gates(qoa_state* s, const qoa_gate* gates, uint32_t n_gates,
                  const double* matrices, uint32_t flags);

/* One noisy trajectory as a plain per-gate loop: each gate's sweep, then
   for each operand qubit an X, Y or Z sweep with probability p1 (one-qubit
   gate) or p2 (two- and three-qubit gates), drawn by Philox4x32-10 with
   counter (gate index, qubit, traj, 0x7A3C5E1F) and key seed, the draws the
   batched and serving kernels make in-kernel (the serving workload's
   operand channel; the reference's own channel is qoa_qs_apply_noise,
   quantum_state.rs:872-885). order (NULL: 0..n_gates-1) is the execution
   order; the draw counter is always a gate's index in `gates`. Operands must
   be distinct and in range; MEASURE/RESET are rejected. */
int qoa_run_trajectory(qoa_state* s, const qoa_gate* gates, uint32_t n_gates,
                       const double* matrices, uint32_t flags, uint64_t seed, uint32_t traj,
                       double p1, double p2, const uint32_t* order);

/* Fused execution: the gate list is planned into passes; each pass loads
   2^tile_qubits-amplitude tiles (the low `low_qubits` qubits plus qubits
   chosen per pass) into LDS once, applies every gate of the pass there and
   writes the tile back. tile_qubits = 0 picks the default (12; 11 for
   states of 13..17 qubits). Results equal the per-gate run of the same
   gates in planned order (bit for bit). */
typedef struct qoa_fuse_opts {
  uint32_t tile_qubits;  /* k: 2^k amplitudes per tile (0 = default)     */
  uint32_t low_qubits;   /* m: qubits 0..m-1 always in the tile (0 = default) */
  uint32_t flags;        /* QOA_CZ_REF_MASKS                            */
  uint32_t reserved;
} qoa_fuse_opts;
int qoa_run_fused(qoa_state* s, const qoa_gate* gates, uint32_t n_gates,
                  const double* matrices, const qoa_fuse_opts* opts);
/* Planner only (host): writes the executed order of gate indices into
   order[0..n_gates) and the pass boundary count into *n_passes. */
int qoa_plan_fused(uint32_t n_qubits, const qoa_gate* gates, uint32_t n_gates,
                   const qoa_fuse_opts* opts, uint32_t* order, uint32_t* n_passes);

/* Introspection of a fused plan (host only): copies the op program (records
   of qoa_tile_op_size() bytes) and, per pass, the tile qubit positions
   (16 bytes), and per pass three words: op begin, op end, kernel mode.
   Returns the counts; pass NULL buffers to query sizes; buffers too small
   for the plan give QOA_ERR_ARG. */
int qoa_tile_op_size(void);
int qoa_plan_program(uint32_t n_qubits, const qoa_gate* gates, uint32_t n_gates,
                     const double* matrices, const qoa_fuse_opts* opts, void* ops_out,
                     uint32_t max_ops, uint8_t* pass_spos_out, uint32_t* pass_range_out,
                     uint32_t max_passes, uint32_t* n_ops, uint32_t* n_passes, uint32_t* tile_k);

/* The same plan's store positions (host only): per pass the physical
   position local bit j is stored to (16 bytes; the load positions of
   qoa_plan_program unless the plan relabels qubits, QOA_FUSE_RELABEL). */
int qoa_plan_store_positions(uint32_t n_qubits, const qoa_gate* gates, uint32_t n_gates,
                             const double* matrices, const qoa_fuse_opts* opts,
                             uint8_t* pass_spos_st_out, uint32_t max_passes, uint32_t* n_passes);

/* The same plan's coefficient pool (host only): MAT4 matrices (32 doubles)
   and merged diagonal terms (6 doubles each: rs, tl as two 32-bit words,
   then gmask, gval, re, im, pad). NULL pool_out queries the size. */
int qoa_plan_pool(uint32_t n_qubits, const qoa_gate* gates, uint32_t n_gates,
                  const double* matrices, const qoa_fuse_opts* opts, double* pool_out,
                  uint32_t max_doubles, uint32_t* n_doubles);

/* Specialised passes (QOA_FUSE_SPECIALIZE; no reference counterpart: the
   reference has no fused passes). qoa_specialize_compile plans the circuit
   and compiles its per-pass kernels for the architecture `arch` (e.g.
   "gfx950") into the process-wide cache without loading them: host only,
   no device needed (a build check, or a warm cache before the first run).
   qoa_specialize_info reports the cache: programs held, programs compiled
   so far and the wall time spent compiling. */
int qoa_specialize_compile(uint32_t n_qubits, const qoa_gate* gates, uint32_t n_gates,
                           const double* matrices, const qoa_fuse_opts* opts, const char* arch);
int qoa_specialize_info(uint32_t* programs, uint32_t* compiles, double* compile_seconds);

/* Dense 5-qubit blocks on complex64 states (the block-fusion primitive the
   boundary lists as apply_fused_block; no reference counterpart: the
   reference applies every gate as its own FP64 sweep, vectorization.rs).
   Gates are packed greedily into blocks acting on at most block_qubits
   (<= 5) qubits; each block's 32x32 unitary is the FP64 product of its
   gates rounded to FP32 and is applied with FP32 matrix cores. After the
   last block the state is renormalised with an FP64 sum unless
   QOA_BLOCK_NO_RENORM is set. MEASURE/RESET are rejected; the state must be
   QOA_C64 with at least 5 qubits. */
#define QOA_BLOCK_NO_RENORM 8u
typedef struct qoa_block_opts {
  uint32_t block_qubits;  /* 1..5 (0 = 5)                                */
  uint32_t flags;         /* QOA_CZ_REF_MASKS, QOA_BLOCK_NO_RENORM         */
} qoa_block_opts;
int qoa_run_blocks(qoa_state* s, const qoa_gate* gates, uint32_t n_gates,
                   const double* matrices, const qoa_block_opts* opts, uint32_t* n_blocks);
/* Fuser only (host): per block the 5 qubits (ascending; local bit b of the
   block index is qubit qubits[5*blk+b]) and the 32x32 FP64 unitary,
   row-major, interleaved (re, im): 2048 doubles per block. NULL outputs
   query the count. */
int qoa_plan_blocks(uint32_t n_qubits, const qoa_gate* gates, uint32_t n_gates,
                    const double* matrices, const qoa_block_opts* opts, uint32_t* qubits_out,
                    double* unitaries_out, uint32_t max_blocks, uint32_t* n_blocks);
/* Compares two states of equal size and any precision on the device:
   max |a_i - b_i| and fidelity |<a|b>|^2 / (<a|a><b|b>), FP64 sums. */
int qoa_state_compare(qoa_state* a, qoa_state* b, double* max_abs_diff, double* fidelity);

/* Kernel timing: when enabled, every kernel launched on the handle's stream
   is bracketed by HIP events; qoa_timing_read returns per-kind totals. */
int qoa_timing_enable(qoa_state* s, int enable);
/* kind names: "sweep_dense", "sweep_diag", "sweep_perm", "fused_pass",
   "reduce", "other", "fused_block". Returns total ms and launch count for
   that kind. */
int qoa_timing_read(qoa_state* s, const char* kind, double* total_ms,
                    uint64_t* launches, double* bytes);

/* ---- layer 2: QuantumState wrappers (src/runtime/quantum_state.rs) ------ */
#define QOA_NOISE_NONE 0   /* noise_config: None           */
#define QOA_NOISE_IDEAL 1  /* Some(NoiseConfig::Ideal)     */
#define QOA_NOISE_RANDOM 2 /* Some(NoiseConfig::Random)    */
#define QOA_NOISE_FIXED 3  /* Some(NoiseConfig::Fixed(p))  */
int qoa_qs_set_noise(qoa_state* s, int kind, double p);
/* seed_rng (quantum_state.rs:956-958): the handle's Philox4x32-10 stream. */
int qoa_qs_seed_rng(qoa_state* s, uint64_t seed);
int qoa_qs_set_flags(qoa_state* s, uint32_t flags);  /* QOA_CZ_REF_MASKS */
/* Applies one gate as QuantumState::apply_* does: bounds check (error ->
   QOA_ERR_QUBIT, state untouched), the sweep, apply_noise, normalize. */
int qoa_qs_apply(qoa_state* s, const qoa_gate* g, const double* matrices);
/* QuantumState::measure (quantum_state.rs:399-458): draws r from the
   handle's stream, collapses, then apply_noise + normalize. */
int qoa_qs_measure(qoa_state* s, uint32_t q, int* outcome);
/* measure_all (quantum_state.rs:909-918): measures qubits 0..n-1 in order on
   a copy; the handle's state is left unchanged. out has n entries. */
int qoa_qs_measure_all(qoa_state* s, uint8_t* out, uint32_t n_out);
/* apply_noise + normalize as run after every gate (quantum_state.rs:839-907). */
int qoa_qs_apply_noise(qoa_state* s);

/* ---- batches: many independent circuits, one launch per pass round ------- */
typedef struct qoa_batch qoa_batch;
typedef struct qoa_job_desc {
  uint32_t n_qubits;     /* 4..30                                         */
  uint32_t n_gates;      /* gates of this job in the shared gate array     */
  uint64_t gate_offset;  /* first gate index in the shared gate array      */
  uint64_t noise_seed;   /* Philox key for the job's noise trajectories    */
  uint32_t trajectories; /* >= 1; each is an independent state             */
  uint32_t precision;    /* QOA_C128                                       */
  double p1, p2;         /* Pauli error prob after 1q / 2q gates (0 = none)*/
} qoa_job_desc;
/* Plans every job and allocates every instance's state in one device buffer
   (QOA_ERR_OOM when they do not fit). Gates use layer-1 semantics
   (opts->flags may hold QOA_CZ_REF_MASKS, QOA_BATCH_AUTOTUNE, QOA_FUSE_FACTOR,
   QOA_FUSE_MERGE_DIAG).
   QOA_C128 jobs (the default): fused FP64 register passes, tile =
   min(opts->tile_qubits or 12, n), trajectory noise allowed.
   QOA_C64 jobs (all jobs of a batch share one precision): the FP32-fused
   form, dense 5-qubit blocks (as qoa_run_blocks, no renormalisation) on FP32
   matrix cores, round r applying block r of every job; noise-free jobs of
   5..30 qubits. */
#define QOA_BATCH_AUTOTUNE 32u /* qoa_fuse_opts.flags for qoa_batch_create:
                                  pick each qubit count's tile (10, 11 or 12
                                  qubits) by timing one of its jobs with each;
                                  choices are cached for the process.        */
int qoa_batch_create(const qoa_job_desc* jobs, uint32_t n_jobs, const qoa_gate* gates,
                     uint64_t n_gates_total, const double* matrices,
                     const qoa_fuse_opts* opts, qoa_batch** out);
/* Runs every instance from |0...0>: one launch per round (round r = pass r
   of every job that has it). Asynchronous; see qoa_batch_synchronize. */
int qoa_batch_run(qoa_batch* b);
/* Waits for the last run; *run_ms = its device time (HIP events). */
int qoa_batch_synchronize(qoa_batch* b, double* run_ms);
int qoa_batch_info(qoa_batch* b, uint32_t* n_rounds, uint64_t* total_amps, double* alg_bytes);
int qoa_batch_download(qoa_batch* b, uint32_t job, uint32_t traj, double* amps,
                       uint64_t n_amps);
/* Executed gate order of a job (for bit-exact checks against the oracle;
   QOA_C128 batches). */
int qoa_batch_order(qoa_batch* b, uint32_t job, uint32_t* order, uint32_t n_order);
/* A job's tile (QOA_C128: qubits per pass tile, after autotuning) or its
   number of blocks (QOA_C64). */
int qoa_batch_job_plan(qoa_batch* b, uint32_t job, uint32_t* tile_qubits, uint32_t* n_steps);
int qoa_batch_destroy(qoa_batch* b);

/* ---- serving: continuous batching of arriving jobs ------------------------
   An engine thread plans arrived jobs, admits them in arrival order while
   their footprint (trajectories x 2^n x 16 B plus scratch) fits the arena,
   and launches one step at a time: one kernel that runs the next pass of
   every admitted job. Each trajectory ends with one Z-basis sample
   (Philox key noise_seed ^ 0x53414D504C45, fixed-order CDF as qoa_sample).
   A job's result does not depend on the other jobs in flight. */
typedef struct qoa_server qoa_server;
typedef struct qoa_server_opts {
  uint64_t budget_bytes;    /* HBM arena for job states (0: 80% of free memory) */
  uint32_t tile_qubits;     /* 8..12 (0: 12)                                   */
  uint32_t planner_threads; /* host planning threads (0: 4)                    */
  uint32_t max_items;       /* work items per step (0: 65536)                  */
  uint32_t max_ops;         /* op slots for all resident jobs (0: 2^21)         */
  uint32_t max_jobs;        /* resident jobs (0: 4096)                          */
  uint32_t flags;           /* QOA_CZ_REF_MASKS, QOA_FUSE_FACTOR                */
} qoa_server_opts;
typedef struct qoa_job_result {
  uint64_t job_id;
  int32_t status;           /* QOA_OK, QOA_ERR_ADMISSION, ...                  */
  uint32_t n_qubits, trajectories, n_passes, n_samples;
  double t_arrival, t_admit, t_done;  /* server clock, seconds               */
} qoa_job_result;
typedef struct qoa_server_stats {
  uint64_t submitted, completed, rejected, steps;
  uint64_t arena_bytes, arena_peak;
  int32_t fatal;
} qoa_server_stats;
int qoa_server_create(const qoa_server_opts* opts, qoa_server** out);
/* Copies the job's gates (job->gate_offset is ignored); the job becomes
   eligible at arrival_s on the server clock (qoa_server_now). A job that can
   never fit the arena is rejected with QOA_ERR_ADMISSION. */
int qoa_server_submit(qoa_server* s, const qoa_job_desc* job, const qoa_gate* gates,
                      const double* matrices, uint32_t n_matrices, double arrival_s,
                      uint64_t* job_id);
double qoa_server_now(qoa_server* s);
/* Blocks until the job finished (or timeout); copies up to max_samples
   samples (one per trajectory, logical basis indices). Returns the status. */
int qoa_server_wait(qoa_server* s, uint64_t job_id, double timeout_s, qoa_job_result* res,
                    uint64_t* samples, uint32_t max_samples);
int qoa_server_drain(qoa_server* s, double timeout_s);
int qoa_server_get_stats(qoa_server* s, qoa_server_stats* st);
int qoa_server_destroy(qoa_server* s);

#ifdef __cplusplus
}
#endif
#endif /* QOA_HIP_H */
