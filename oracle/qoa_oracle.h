/*
 * qoa_oracle.h — CPU restatement of the reference gate path (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This library is the correctness oracle for the HIP backend. Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it; the
 * product (libqoa_hip.so) never links or calls it.
 *
 * Every function restates one reference function with the same arithmetic
 * (num-complex 0.4.6 operator order, no FMA contraction) so that a single
 * gate on the GPU and here agree bit for bit. Citations are to the
 * reference tree (planetryan/qoa).
 *
 * State layout: `a` holds `len` complex128 amplitudes as interleaved doubles
 * (re at a[2i], im at a[2i+1]); qubit q is bit q of the basis index.
 */
#ifndef QOA_ORACLE_H
#define QOA_ORACLE_H
#include <stdint.h>
#include "../include/qoa_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- layer 1: src/vectorization.rs x86_64 module ---- */
void orc_hadamard(double* a, uint64_t len, double nf_re, double nf_im, uint64_t mask_bit);
void orc_x(double* a, uint64_t len, uint64_t mask_bit);
void orc_y(double* a, uint64_t len, uint64_t mask_bit);
void orc_z(double* a, uint64_t len, uint64_t mask_bit);
void orc_phase_factor(double* a, uint64_t len, uint64_t mask_bit, double pf_re, double pf_im);
void orc_t(double* a, uint64_t len, uint64_t mask_bit);
void orc_s(double* a, uint64_t len, uint64_t mask_bit);
void orc_phaseshift(double* a, uint64_t len, uint64_t mask_bit, double angle);
void orc_reset(double* a, uint64_t len, uint64_t mask_bit);
void orc_swap(double* a, uint64_t len, uint64_t q1, uint64_t q2);
void orc_controlled_swap(double* a, uint64_t len, uint64_t c, uint64_t t1, uint64_t t2);
void orc_matrix2(double* a, uint64_t len, uint64_t mask_bit, const double* m);
void orc_rx(double* a, uint64_t len, uint64_t mask_bit, double angle);
void orc_ry(double* a, uint64_t len, uint64_t mask_bit, double angle);
void orc_rz(double* a, uint64_t len, uint64_t mask_bit, double angle);
void orc_cnot(double* a, uint64_t len, uint64_t control, uint64_t target);
void orc_cz(double* a, uint64_t len, uint64_t cmask, uint64_t tmask);
void orc_cphase(double* a, uint64_t len, uint64_t cmask, uint64_t tmask, double angle);
void orc_reset_all(double* a, uint64_t len);

/* ---- extended gate surface: src/runtime/quantum_state.rs ---- */
void orc_custom_unitary(double* a, uint64_t len, uint64_t q, const double* u);
void orc_two_qubit_unitary(double* a, uint64_t len, uint64_t q1, uint64_t q2, const double* u);
void orc_toffoli(double* a, uint64_t len, uint64_t c1, uint64_t c2, uint64_t t);
void orc_controlled_y(double* a, uint64_t len, uint64_t c, uint64_t t);

/* ---- wrappers: src/runtime/quantum_state.rs ---- */
double orc_norm_sqr(const double* a, uint64_t len);
int orc_normalize(double* a, uint64_t len);
int orc_measure(double* a, uint64_t len, uint64_t q, double r, double* p0, double* p1);
void orc_probabilities(const double* a, uint64_t len, double* out);
void orc_pauli_codes(double* a, uint64_t len, const uint8_t* codes, uint32_t n_codes);

/* ---- circuits ---- */
/* Applies gates in order with layer-1 semantics (flags: QOA_CZ_REF_MASKS).
   Returns 0, or -1 for an op that cannot appear in a unitary circuit. */
int orc_run_gates(double* a, uint64_t len, const qoa_gate* g, uint32_t n,
                  const double* matrices, uint32_t flags);
/* Same gates applied in the order given by `order` (a planner's output). */
int orc_run_gates_ordered(double* a, uint64_t len, const qoa_gate* g, const uint32_t* order,
                          uint32_t n, const double* matrices, uint32_t flags);
/* Threads for the gate loops (CPU baseline "strong"; default 1). */
void orc_set_threads(int threads);
/* CPU baseline "faithful": one-thread sweep + `threads`-thread normalize
   after every gate (the QuantumState execution model). */
int orc_run_gates_faithful(double* a, uint64_t len, const qoa_gate* g, uint32_t n,
                           const double* matrices, uint32_t flags, int threads);

/* ---- counter-based RNG (Salmon et al., SC'11) and circuit PRNG ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t orc_splitmix64_next(uint64_t* state);

/* ---- wrapper-level runs (QuantumState semantics) ---- */
/* Philox stream as used by the backend's QuantumState: key = seed, counter
   = draw index; u64 draws, f64 uniform = (u >> 11) * 2^-53.               */
typedef struct orc_rng { uint32_t key[2]; uint64_t ctr; uint32_t buf[4]; int have; } orc_rng;
void orc_rng_seed(orc_rng* r, uint64_t seed);
uint64_t orc_rng_u64(orc_rng* r);
double orc_rng_f64(orc_rng* r);
uint32_t orc_rng_below3(orc_rng* r);
/* One gate with QuantumState semantics: bounds check, sweep, apply_noise,
   normalize. noise_kind/p as QOA_NOISE_*. Returns 0 or -2 (out of range). */
int orc_qs_apply(double* a, uint64_t len, uint32_t n, const qoa_gate* g, const double* matrices,
                 uint32_t flags, int noise_kind, double noise_p, orc_rng* rng);
int orc_qs_measure(double* a, uint64_t len, uint32_t n, uint32_t q, int noise_kind,
                   double noise_p, orc_rng* rng, int* outcome);
void orc_qs_apply_noise(double* a, uint64_t len, uint32_t n, int noise_kind, double noise_p,
                        orc_rng* rng);

/* ---- trajectory noise (per-gate operand Pauli channel) ---- */
/* Decision for (seed, trajectory, gate_idx, qubit): 0 none, 1 X, 2 Y, 3 Z. */
uint32_t orc_traj_pauli(uint64_t seed, uint32_t traj, uint32_t gate_idx, uint32_t qubit, double p);
int orc_run_trajectory(double* a, uint64_t len, const qoa_gate* g, uint32_t n,
                       const double* matrices, uint32_t flags, uint64_t seed, uint32_t traj,
                       double p1, double p2);
int orc_run_trajectory_ordered(double* a, uint64_t len, const qoa_gate* g, const uint32_t* order,
                               uint32_t n, const double* matrices, uint32_t flags, uint64_t seed,
                               uint32_t traj, double p1, double p2);

/* ---- Z-basis sampling with the backend's fixed summation order ---- */
double orc_sample_uniform(uint64_t seed, uint32_t traj, uint32_t s);
uint64_t orc_sample(const double* a, uint64_t len, double u);

#ifdef __cplusplus
}
#endif
#endif
