/*
 * qoa_oracle.c — CPU restatement of planetryan/qoa's gate path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker. Never part of the product.
 *
 * Build with -ffp-contract=off and without fast-math: the reference is Rust
 * with no fast-math flags (Cargo.toml:94-103, .cargo/config.toml:28-35), so
 * every complex product below is the unfused num-complex 0.4.6 expression
 *   (a*b).re = a.re*b.re - a.im*b.im,  (a*b).im = a.re*b.im + a.im*b.re.
 *
 * Loops run over every basis index in increasing order with the same
 * predicates as the reference x86_64 module (src/vectorization.rs:904-1261).
 */
#define _GNU_SOURCE
#include "qoa_oracle.h"

#include <math.h>
#include <string.h>

#define RE(a, i) (a)[2 * (i)]
#define IM(a, i) (a)[2 * (i) + 1]

static inline uint64_t bit_of(uint64_t q) { return q < 64 ? (1ULL << q) : 0ULL; }

/* Gate loops may be split over threads for the all-core CPU baseline
   (orc_set_threads; 1 by default). Every iteration writes only its own
   amplitude or pair, so results do not depend on the split; sums stay
   sequential. */
static int orc_threads = 1;
void orc_set_threads(int t) { orc_threads = t < 1 ? 1 : t; }
#define ORC_PAR_FOR \
  _Pragma("omp parallel for schedule(static) if (orc_threads > 1) num_threads(orc_threads)")

/* cos x and sin x from one glibc sincos call. The reference computes them
   side by side (Complex64::from_polar, the rx/ry set-up at
   src/vectorization.rs:529-535), which LLVM lowers to sincos on linux-gnu;
   sincos and separate cos/sin can differ in the last bit. */
static inline void cis(double x, double* c, double* s) { sincos(x, s, c); }

/* c = x * y (num-complex Mul). */
static inline void cmul(double xr, double xi, double yr, double yi, double* cr, double* ci) {
  *cr = xr * yr - xi * yi;
  *ci = xr * yi + xi * yr;
}

/* apply_hadamard_simd, src/vectorization.rs:915-939:
   a' = nf*(a+b), b' = nf*(a-b) for each pair (i, i|bit). */
void orc_hadamard(double* a, uint64_t len, double nf_re, double nf_im, uint64_t mask_bit) {
  uint64_t bit = bit_of(mask_bit);
  if (!bit) return;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & bit) == 0) {
      uint64_t j = i | bit;
      if (j < len) {
        double ar = RE(a, i), ai = IM(a, i), br = RE(a, j), bi = IM(a, j);
        double sr = ar + br, si = ai + bi, dr = ar - br, di = ai - bi;
        cmul(nf_re, nf_im, sr, si, &RE(a, i), &IM(a, i));
        cmul(nf_re, nf_im, dr, di, &RE(a, j), &IM(a, j));
      }
    }
  }
}

/* apply_x_simd, src/vectorization.rs:942-961: swap (i, i|bit). */
void orc_x(double* a, uint64_t len, uint64_t mask_bit) {
  uint64_t bit = bit_of(mask_bit);
  if (!bit) return;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & bit) == 0) {
      uint64_t j = i | bit;
      if (j < len) {
        double tr = RE(a, i), ti = IM(a, i);
        RE(a, i) = RE(a, j); IM(a, i) = IM(a, j);
        RE(a, j) = tr; IM(a, j) = ti;
      }
    }
  }
}

/* apply_y_simd, src/vectorization.rs:964-985: a' = b*(-i), b' = a*(+i). */
void orc_y(double* a, uint64_t len, uint64_t mask_bit) {
  uint64_t bit = bit_of(mask_bit);
  if (!bit) return;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & bit) == 0) {
      uint64_t j = i | bit;
      if (j < len) {
        double ar = RE(a, i), ai = IM(a, i), br = RE(a, j), bi = IM(a, j);
        cmul(br, bi, 0.0, -1.0, &RE(a, i), &IM(a, i));
        cmul(ar, ai, 0.0, 1.0, &RE(a, j), &IM(a, j));
      }
    }
  }
}

/* apply_z_simd, src/vectorization.rs:988-999: a = a.scale(-1) where bit set. */
void orc_z(double* a, uint64_t len, uint64_t mask_bit) {
  uint64_t bit = bit_of(mask_bit);
  if (!bit) return;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & bit) != 0) {
      RE(a, i) = RE(a, i) * -1.0;
      IM(a, i) = IM(a, i) * -1.0;
    }
  }
}

/* apply_phaseshift_simd, src/vectorization.rs:1002-1017: a *= phase where bit set. */
void orc_phase_factor(double* a, uint64_t len, uint64_t mask_bit, double pf_re, double pf_im) {
  uint64_t bit = bit_of(mask_bit);
  if (!bit) return;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & bit) != 0) {
      double xr = RE(a, i), xi = IM(a, i);
      cmul(xr, xi, pf_re, pf_im, &RE(a, i), &IM(a, i));
    }
  }
}

/* apply_t_vectorized, src/vectorization.rs:217-218: from_polar(1, pi/4). */
void orc_t(double* a, uint64_t len, uint64_t mask_bit) {
  double c, s;
  cis(M_PI / 4.0, &c, &s);
  orc_phase_factor(a, len, mask_bit, 1.0 * c, 1.0 * s);
}

/* apply_s_vectorized, src/vectorization.rs:262-263: from_polar(1, pi/2) =
   (6.123e-17, 1.0), not exactly i. */
void orc_s(double* a, uint64_t len, uint64_t mask_bit) {
  double c, s;
  cis(M_PI / 2.0, &c, &s);
  orc_phase_factor(a, len, mask_bit, 1.0 * c, 1.0 * s);
}

/* apply_phaseshift_vectorized, src/vectorization.rs:307-308. */
void orc_phaseshift(double* a, uint64_t len, uint64_t mask_bit, double angle) {
  double c, s;
  cis(angle, &c, &s);
  orc_phase_factor(a, len, mask_bit, 1.0 * c, 1.0 * s);
}

/* apply_reset_simd, src/vectorization.rs:1020-1056. */
void orc_reset(double* a, uint64_t len, uint64_t mask_bit) {
  uint64_t bit = bit_of(mask_bit);
  double ns = 0.0;
  for (uint64_t i = 0; i < len; ++i)
    if ((i & bit) == 0) ns += RE(a, i) * RE(a, i) + IM(a, i) * IM(a, i);
  if (ns > 1e-12) {
    double scale = 1.0 / sqrt(ns);
    for (uint64_t i = 0; i < len; ++i) {
      if ((i & bit) == 0) {
        RE(a, i) *= scale;
        IM(a, i) *= scale;
      } else {
        RE(a, i) = 0.0;
        IM(a, i) = 0.0;
      }
    }
  } else if (len > 0) {
    orc_reset_all(a, len);
  }
}

/* apply_swap_simd, src/vectorization.rs:1176-1197 (masks built at :421-423). */
void orc_swap(double* a, uint64_t len, uint64_t q1, uint64_t q2) {
  uint64_t m1 = bit_of(q1), m2 = bit_of(q2), comb = m1 | m2;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & comb) == 0) {
      uint64_t x = i | m2, y = i | m1;
      if (x < len && y < len) {
        double tr = RE(a, x), ti = IM(a, x);
        RE(a, x) = RE(a, y); IM(a, x) = IM(a, y);
        RE(a, y) = tr; IM(a, y) = ti;
      }
    }
  }
}

/* apply_controlled_swap_simd, src/vectorization.rs:1200-1232. */
void orc_controlled_swap(double* a, uint64_t len, uint64_t c, uint64_t t1, uint64_t t2) {
  uint64_t cm = bit_of(c), m1 = bit_of(t1), m2 = bit_of(t2), comb = cm | m1 | m2;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & comb) == 0) {
      uint64_t x = i | cm | m2, y = i | cm | m1;
      if (x < len && y < len) {
        double tr = RE(a, x), ti = IM(a, x);
        RE(a, x) = RE(a, y); IM(a, x) = IM(a, y);
        RE(a, y) = tr; IM(a, y) = ti;
      }
    }
  }
}

/* apply_rx_simd, src/vectorization.rs:1059-1085:
   a0' = m00*a0 + m01*a1, a1' = m10*a0 + m11*a1.
   m = {m00re, m00im, m01re, m01im, m10re, m10im, m11re, m11im}. */
void orc_matrix2(double* a, uint64_t len, uint64_t mask_bit, const double* m) {
  uint64_t bit = bit_of(mask_bit);
  if (!bit) return;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & bit) == 0) {
      uint64_t j = i | bit;
      if (j < len) {
        double xr = RE(a, i), xi = IM(a, i), yr = RE(a, j), yi = IM(a, j);
        double p0r, p0i, p1r, p1i;
        cmul(m[0], m[1], xr, xi, &p0r, &p0i);
        cmul(m[2], m[3], yr, yi, &p1r, &p1i);
        RE(a, i) = p0r + p1r; IM(a, i) = p0i + p1i;
        cmul(m[4], m[5], xr, xi, &p0r, &p0i);
        cmul(m[6], m[7], yr, yi, &p1r, &p1i);
        RE(a, j) = p0r + p1r; IM(a, j) = p0i + p1i;
      }
    }
  }
}

/* apply_rx_vectorized, src/vectorization.rs:528-535. */
void orc_rx(double* a, uint64_t len, uint64_t mask_bit, double angle) {
  double c, s;
  cis(angle / 2.0, &c, &s);
  double m[8] = {c, 0.0, 0.0, -s, 0.0, -s, c, 0.0};
  orc_matrix2(a, len, mask_bit, m);
}

/* apply_ry_vectorized, src/vectorization.rs:588-595 (apply_ry_simd calls the
   rx loop, :1088-1100). */
void orc_ry(double* a, uint64_t len, uint64_t mask_bit, double angle) {
  double c, s;
  cis(angle / 2.0, &c, &s);
  double m[8] = {c, 0.0, -s, 0.0, s, 0.0, c, 0.0};
  orc_matrix2(a, len, mask_bit, m);
}

/* apply_rz_vectorized/apply_rz_simd, src/vectorization.rs:648-651, 1103-1122:
   a *= m00 = from_polar(1,-t/2) if bit clear, else a *= m11 = from_polar(1,t/2).
   With mask_bit >= n every index takes m00 (a global phase). */
void orc_rz(double* a, uint64_t len, uint64_t mask_bit, double angle) {
  double c0, s0, c1, s1;
  cis(-angle / 2.0, &c0, &s0);
  cis(angle / 2.0, &c1, &s1);
  double m00r = 1.0 * c0, m00i = 1.0 * s0, m11r = 1.0 * c1, m11i = 1.0 * s1;
  uint64_t bit = bit_of(mask_bit);
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    double xr = RE(a, i), xi = IM(a, i);
    if ((i & bit) == 0)
      cmul(xr, xi, m00r, m00i, &RE(a, i), &IM(a, i));
    else
      cmul(xr, xi, m11r, m11i, &RE(a, i), &IM(a, i));
  }
}

/* apply_cnot_vectorized/apply_cnot_simd, src/vectorization.rs:699-701,
   1141-1160: for i with control set and target clear, swap(i, i|tmask). */
void orc_cnot(double* a, uint64_t len, uint64_t control, uint64_t target) {
  uint64_t cm = bit_of(control), tm = bit_of(target);
  if (!cm || !tm) return;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & cm) == cm && (i & tm) == 0) {
      uint64_t j = i | tm;
      if (j < len) {
        double tr = RE(a, i), ti = IM(a, i);
        RE(a, i) = RE(a, j); IM(a, i) = IM(a, j);
        RE(a, j) = tr; IM(a, j) = ti;
      }
    }
  }
}

/* apply_cz_simd, src/vectorization.rs:1163-1173 (takes MASKS). */
void orc_cz(double* a, uint64_t len, uint64_t cmask, uint64_t tmask) {
  uint64_t m = cmask | tmask;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & m) == m) {
      RE(a, i) = RE(a, i) * -1.0;
      IM(a, i) = IM(a, i) * -1.0;
    }
  }
}

/* apply_controlled_phase_rotation_vectorized/_simd, src/vectorization.rs:
   788-794, 1235-1250 (MASKS; phase = from_polar(1, angle)). */
void orc_cphase(double* a, uint64_t len, uint64_t cmask, uint64_t tmask, double angle) {
  uint64_t m = cmask | tmask;
  double c, s;
  cis(angle, &c, &s);
  double pr = 1.0 * c, pi = 1.0 * s;
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & m) == m) {
      double xr = RE(a, i), xi = IM(a, i);
      cmul(xr, xi, pr, pi, &RE(a, i), &IM(a, i));
    }
  }
}

/* apply_reset_all_simd, src/vectorization.rs:1253-1260. */
void orc_reset_all(double* a, uint64_t len) {
  if (len == 0) return;
  memset(a, 0, sizeof(double) * 2 * len);
  RE(a, 0) = 1.0;
}

/* apply_custom_unitary, src/runtime/quantum_state.rs:1010-1062 (reads a copy
   of the old state; the per-pair formulas equal the rx loop). */
void orc_custom_unitary(double* a, uint64_t len, uint64_t q, const double* u) {
  orc_matrix2(a, len, q, u);
}

/* apply_two_qubit_unitary, src/runtime/quantum_state.rs:1064-1127:
   new(s) = u[4s]*a00 + u[4s+1]*a01 + u[4s+2]*a10 + u[4s+3]*a11 with
   s = (bit_high<<1)|bit_low. u is row-major, 16 complex = 32 doubles. */
void orc_two_qubit_unitary(double* a, uint64_t len, uint64_t q1, uint64_t q2, const double* u) {
  uint64_t lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
  uint64_t ml = bit_of(lo), mh = bit_of(hi);
  ORC_PAR_FOR
  for (uint64_t base = 0; base < len; ++base) {
    if (base & (ml | mh)) continue;
    uint64_t idx[4] = {base, base | ml, base | mh, base | ml | mh};
    double vr[4], vi[4];
    for (int k = 0; k < 4; ++k) { vr[k] = RE(a, idx[k]); vi[k] = IM(a, idx[k]); }
    for (int s = 0; s < 4; ++s) {
      double accr, acci, pr, pi;
      cmul(u[8 * s + 0], u[8 * s + 1], vr[0], vi[0], &accr, &acci);
      for (int k = 1; k < 4; ++k) {
        cmul(u[8 * s + 2 * k], u[8 * s + 2 * k + 1], vr[k], vi[k], &pr, &pi);
        accr = accr + pr;
        acci = acci + pi;
      }
      RE(a, idx[s]) = accr;
      IM(a, idx[s]) = acci;
    }
  }
}

/* apply_toffoli, src/runtime/quantum_state.rs:960-1008: where both controls
   are set, exchange the target-0 and target-1 amplitudes. */
void orc_toffoli(double* a, uint64_t len, uint64_t c1, uint64_t c2, uint64_t t) {
  uint64_t cm = bit_of(c1) | bit_of(c2), tm = bit_of(t);
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & cm) == cm && (i & tm) == 0) {
      uint64_t j = i | tm;
      double tr = RE(a, i), ti = IM(a, i);
      RE(a, i) = RE(a, j); IM(a, i) = IM(a, j);
      RE(a, j) = tr; IM(a, j) = ti;
    }
  }
}

/* apply_controlled_y, src/runtime/quantum_state.rs:1163-1212:
   new[t=0] = (-i)*old[t=1], new[t=1] = (i)*old[t=0] where control set. */
void orc_controlled_y(double* a, uint64_t len, uint64_t c, uint64_t t) {
  uint64_t cm = bit_of(c), tm = bit_of(t);
  ORC_PAR_FOR
  for (uint64_t i = 0; i < len; ++i) {
    if ((i & cm) == cm && (i & tm) == 0) {
      uint64_t j = i | tm;
      double ar = RE(a, i), ai = IM(a, i), br = RE(a, j), bi = IM(a, j);
      cmul(0.0, -1.0, br, bi, &RE(a, i), &IM(a, i));
      cmul(0.0, 1.0, ar, ai, &RE(a, j), &IM(a, j));
    }
  }
}

/* ---- wrappers ---------------------------------------------------------- */

/* sum |a|^2 in index order (the reference's rayon sum at
   src/runtime/quantum_state.rs:894 has no fixed order; results differ from
   any other order only at the ulp level). */
double orc_norm_sqr(const double* a, uint64_t len) {
  double s = 0.0;
  for (uint64_t i = 0; i < len; ++i) s += RE(a, i) * RE(a, i) + IM(a, i) * IM(a, i);
  return s;
}

/* normalize, src/runtime/quantum_state.rs:888-907. Returns 1 if the state was
   reset to |0...0>. */
int orc_normalize(double* a, uint64_t len) {
  if (len == 0) return 0;
  double s = orc_norm_sqr(a, len);
  if (s > 1e-12) {
    double nrm = sqrt(s);
    for (uint64_t i = 0; i < len; ++i) {
      RE(a, i) /= nrm;
      IM(a, i) /= nrm;
    }
    return 0;
  }
  orc_reset_all(a, len);
  return 1;
}

/* measure, src/runtime/quantum_state.rs:399-452 with the uniform r injected
   (the reference draws it at :424). Collapse only; the wrapper renormalises. */
int orc_measure(double* a, uint64_t len, uint64_t q, double r, double* p0_out, double* p1_out) {
  uint64_t bit = bit_of(q);
  double p0 = 0.0, p1 = 0.0;
  for (uint64_t i = 0; i < len; ++i) {
    double pr = RE(a, i) * RE(a, i) + IM(a, i) * IM(a, i);
    if ((i & bit) == 0) p0 += pr; else p1 += pr;
  }
  int outcome = (r < p0) ? 0 : 1;
  double f;
  if (outcome == 0) f = p0 > 1e-12 ? 1.0 / sqrt(p0) : 0.0;
  else f = p1 > 1e-12 ? 1.0 / sqrt(p1) : 0.0;
  for (uint64_t i = 0; i < len; ++i) {
    int set = (i & bit) != 0;
    if ((outcome == 0 && set) || (outcome == 1 && !set)) {
      RE(a, i) = 0.0;
      IM(a, i) = 0.0;
    } else {
      RE(a, i) *= f;
      IM(a, i) *= f;
    }
  }
  if (p0_out) *p0_out = p0;
  if (p1_out) *p1_out = p1;
  return outcome;
}

/* get_probabilities, src/runtime/quantum_state.rs:128-133. */
void orc_probabilities(const double* a, uint64_t len, double* out) {
  for (uint64_t i = 0; i < len; ++i) out[i] = RE(a, i) * RE(a, i) + IM(a, i) * IM(a, i);
}

/* The Pauli part of apply_depolarizing_noise, src/runtime/quantum_state.rs:
   872-885: in qubit order, code 1/2/3 applies the X/Y/Z sweep on that qubit. */
void orc_pauli_codes(double* a, uint64_t len, const uint8_t* codes, uint32_t n_codes) {
  for (uint32_t q = 0; q < n_codes; ++q) {
    if (codes[q] == 1) orc_x(a, len, q);
    else if (codes[q] == 2) orc_y(a, len, q);
    else if (codes[q] == 3) orc_z(a, len, q);
  }
}

/* ---- circuits ---------------------------------------------------------- */

static int apply_gate(double* a, uint64_t len, const qoa_gate* g, const double* mats, uint32_t flags) {
  uint64_t cm, tm;
  switch (g->op) {
    case QOA_OP_H: orc_hadamard(a, len, 1.0 / M_SQRT2, 0.0, g->q0); return 0;
    case QOA_OP_X: orc_x(a, len, g->q0); return 0;
    case QOA_OP_Y: orc_y(a, len, g->q0); return 0;
    case QOA_OP_Z: orc_z(a, len, g->q0); return 0;
    case QOA_OP_S: orc_s(a, len, g->q0); return 0;
    case QOA_OP_T: orc_t(a, len, g->q0); return 0;
    case QOA_OP_PHASESHIFT:
    case QOA_OP_PHASE: orc_phaseshift(a, len, g->q0, g->theta); return 0;
    case QOA_OP_RX: orc_rx(a, len, g->q0, g->theta); return 0;
    case QOA_OP_RY: orc_ry(a, len, g->q0, g->theta); return 0;
    case QOA_OP_RZ: orc_rz(a, len, g->q0, g->theta); return 0;
    case QOA_OP_CNOT: orc_cnot(a, len, g->q0, g->q1); return 0;
    case QOA_OP_CZ:
    case QOA_OP_CPHASE:
      if (flags & QOA_CZ_REF_MASKS) { cm = g->q0; tm = g->q1; }
      else { cm = bit_of(g->q0); tm = bit_of(g->q1); }
      if (g->op == QOA_OP_CZ) orc_cz(a, len, cm, tm);
      else orc_cphase(a, len, cm, tm, g->theta);
      return 0;
    case QOA_OP_SWAP: orc_swap(a, len, g->q0, g->q1); return 0;
    case QOA_OP_CSWAP: orc_controlled_swap(a, len, g->q0, g->q1, g->q2); return 0;
    case QOA_OP_U2: orc_matrix2(a, len, g->q0, mats + 8 * g->aux); return 0;
    case QOA_OP_U4: orc_two_qubit_unitary(a, len, g->q0, g->q1, mats + 32 * g->aux); return 0;
    case QOA_OP_TOFFOLI: orc_toffoli(a, len, g->q0, g->q1, g->q2); return 0;
    case QOA_OP_CY: orc_controlled_y(a, len, g->q0, g->q1); return 0;
    case QOA_OP_RESET_ALL: orc_reset_all(a, len); return 0;
    default: return -1;
  }
}

int orc_run_gates(double* a, uint64_t len, const qoa_gate* g, uint32_t n, const double* mats,
                  uint32_t flags) {
  for (uint32_t k = 0; k < n; ++k)
    if (apply_gate(a, len, &g[k], mats, flags)) return -1;
  return 0;
}

int orc_run_gates_ordered(double* a, uint64_t len, const qoa_gate* g, const uint32_t* order,
                          uint32_t n, const double* mats, uint32_t flags) {
  for (uint32_t k = 0; k < n; ++k)
    if (apply_gate(a, len, &g[order[k]], mats, flags)) return -1;
  return 0;
}

/* The reference's QuantumState execution model as a timed baseline: each
   gate is a one-thread sweep (src/vectorization.rs:926-938) followed by
   normalize (src/runtime/quantum_state.rs:868, 888-907), whose sum and
   division run on `threads` threads as rayon's par_iter does (:894-900).
   The parallel sum's order differs from the sequential one at the ulp
   level only. */
int orc_run_gates_faithful(double* a, uint64_t len, const qoa_gate* g, uint32_t n,
                           const double* mats, uint32_t flags, int threads) {
  const int saved = orc_threads;
  orc_threads = 1;
  int rc = 0;
  for (uint32_t k = 0; k < n && rc == 0; ++k) {
    if (apply_gate(a, len, &g[k], mats, flags)) {
      rc = -1;
      break;
    }
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(threads)
    for (uint64_t i = 0; i < len; ++i) s += RE(a, i) * RE(a, i) + IM(a, i) * IM(a, i);
    if (s > 1e-12) {
      const double nrm = sqrt(s);
#pragma omp parallel for schedule(static) num_threads(threads)
      for (uint64_t i = 0; i < len; ++i) {
        RE(a, i) /= nrm;
        IM(a, i) /= nrm;
      }
    } else {
      orc_reset_all(a, len);
    }
  }
  orc_threads = saved;
  return rc;
}

/* ---- RNG ----------------------------------------------------------------- */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
   easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with multipliers
   0xD2511F53/0xCD9E8D57 and Weyl key increments 0x9E3779B9/0xBB67AE85. */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* SplitMix64 (Steele, Lea, Flood, OOPSLA'14; constants as in Vigna's
   splitmix64.c). */
uint64_t orc_splitmix64_next(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_rng_seed(orc_rng* r, uint64_t seed) {
  r->key[0] = (uint32_t)seed;
  r->key[1] = (uint32_t)(seed >> 32);
  r->ctr = 0;
  r->have = 0;
}

/* Draw k of the stream = u64 half (k & 1) of Philox block (k >> 1) with
   counter {lo32(k>>1), hi32(k>>1), 0, 0}. */
uint64_t orc_rng_u64(orc_rng* r) {
  if (r->have == 0) {
    uint64_t blk = r->ctr >> 1;
    uint32_t c[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), 0u, 0u};
    orc_philox4x32_10(c, r->key, r->buf);
    r->have = 2;
  }
  int half = (int)(r->ctr & 1);
  r->ctr++;
  r->have--;
  return ((uint64_t)r->buf[2 * half + 1] << 32) | r->buf[2 * half];
}

double orc_rng_f64(orc_rng* r) { return (double)(orc_rng_u64(r) >> 11) * 0x1.0p-53; }

uint32_t orc_rng_below3(orc_rng* r) { return (uint32_t)(((orc_rng_u64(r) >> 32) * 3ULL) >> 32); }

/* ---- QuantumState semantics --------------------------------------------- */

/* apply_noise + apply_depolarizing_noise + normalize,
   src/runtime/quantum_state.rs:839-907. */
void orc_qs_apply_noise(double* a, uint64_t len, uint32_t n, int kind, double p, orc_rng* rng) {
  double level = -1.0;
  if (kind == QOA_NOISE_RANDOM) level = 0.1 + (1.0 - 0.1) * orc_rng_f64(rng);
  else if (kind == QOA_NOISE_FIXED && p > 0.0) level = p;
  if (level >= 0.0) {
    for (uint32_t q = 0; q < n; ++q) {
      double r = orc_rng_f64(rng);
      if (r < level) {
        uint32_t e = orc_rng_below3(rng);
        if (e == 0) orc_x(a, len, q);
        else if (e == 1) orc_y(a, len, q);
        else orc_z(a, len, q);
      }
    }
  }
  orc_normalize(a, len);
}

static int gate_in_range(uint32_t n, const qoa_gate* g) {
  switch (g->op) {
    case QOA_OP_RESET_ALL: return 1;
    case QOA_OP_CNOT: case QOA_OP_CZ: case QOA_OP_CPHASE: case QOA_OP_SWAP:
    case QOA_OP_U4: case QOA_OP_CY:
      return g->q0 < n && g->q1 < n && g->q0 != g->q1;
    case QOA_OP_CSWAP: case QOA_OP_TOFFOLI:
      return g->q0 < n && g->q1 < n && g->q2 < n && g->q0 != g->q1 && g->q0 != g->q2 &&
             g->q1 != g->q2;
    default: return g->q0 < n;
  }
}

int orc_qs_apply(double* a, uint64_t len, uint32_t n, const qoa_gate* g, const double* mats,
                 uint32_t flags, int noise_kind, double noise_p, orc_rng* rng) {
  if (!gate_in_range(n, g)) return -2;
  if (g->op == QOA_OP_RESET) orc_reset(a, len, g->q0);
  else if (apply_gate(a, len, g, mats, flags)) return -1;
  orc_qs_apply_noise(a, len, n, noise_kind, noise_p, rng);
  return 0;
}

int orc_qs_measure(double* a, uint64_t len, uint32_t n, uint32_t q, int noise_kind,
                   double noise_p, orc_rng* rng, int* outcome) {
  if (q >= n) return -2;
  double r = orc_rng_f64(rng);
  *outcome = orc_measure(a, len, q, r, 0, 0);
  orc_qs_apply_noise(a, len, n, noise_kind, noise_p, rng);
  return 0;
}

/* ---- trajectory noise ----------------------------------------------------- */

uint32_t orc_traj_pauli(uint64_t seed, uint32_t traj, uint32_t gate_idx, uint32_t qubit, double p) {
  uint32_t c[4] = {gate_idx, qubit, traj, 0x7A3C5E1Fu};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32_10(c, k, o);
  double u = (double)((((uint64_t)o[1] << 32) | o[0]) >> 11) * 0x1.0p-53;
  if (!(u < p)) return 0;
  return 1u + (uint32_t)(((uint64_t)o[2] * 3ULL) >> 32);
}

static int gate_arity(uint32_t op) {
  switch (op) {
    case QOA_OP_CNOT: case QOA_OP_CZ: case QOA_OP_CPHASE: case QOA_OP_SWAP:
    case QOA_OP_U4: case QOA_OP_CY: return 2;
    case QOA_OP_CSWAP: case QOA_OP_TOFFOLI: return 3;
    default: return 1;
  }
}

/* One noisy trajectory: after gate k, each operand qubit q of the gate gets
   X/Y/Z with probability p1 (one-qubit gate) or p2 (multi-qubit gate). */
int orc_run_trajectory(double* a, uint64_t len, const qoa_gate* g, uint32_t n,
                       const double* mats, uint32_t flags, uint64_t seed, uint32_t traj,
                       double p1, double p2) {
  for (uint32_t k = 0; k < n; ++k) {
    if (apply_gate(a, len, &g[k], mats, flags)) return -1;
    int ar = gate_arity(g[k].op);
    double p = ar == 1 ? p1 : p2;
    if (p <= 0.0) continue;
    uint32_t qs[3] = {g[k].q0, g[k].q1, g[k].q2};
    for (int j = 0; j < ar; ++j) {
      uint32_t e = orc_traj_pauli(seed, traj, k, qs[j], p);
      if (e == 1) orc_x(a, len, qs[j]);
      else if (e == 2) orc_y(a, len, qs[j]);
      else if (e == 3) orc_z(a, len, qs[j]);
    }
  }
  return 0;
}

/* orc_run_trajectory with the gates taken in the order `order` (a planner's
   executed order); the noise counter is the gate's original index. */
int orc_run_trajectory_ordered(double* a, uint64_t len, const qoa_gate* g, const uint32_t* order,
                               uint32_t n, const double* mats, uint32_t flags, uint64_t seed,
                               uint32_t traj, double p1, double p2) {
  for (uint32_t k = 0; k < n; ++k) {
    const qoa_gate* gk = &g[order[k]];
    if (apply_gate(a, len, gk, mats, flags)) return -1;
    int ar = gate_arity(gk->op);
    double p = ar == 1 ? p1 : p2;
    if (p <= 0.0) continue;
    uint32_t qs[3] = {gk->q0, gk->q1, gk->q2};
    for (int j = 0; j < ar; ++j) {
      uint32_t e = orc_traj_pauli(seed, traj, order[k], qs[j], p);
      if (e == 1) orc_x(a, len, qs[j]);
      else if (e == 2) orc_y(a, len, qs[j]);
      else if (e == 3) orc_z(a, len, qs[j]);
    }
  }
  return 0;
}

/* ---- sampling ---------------------------------------------------------------
   The measurement distribution |a_i|^2 (get_probabilities,
   src/runtime/quantum_state.rs:128-133) sampled by inverse CDF with sums in a
   fixed order: chunks of C = min(len, 4096), runs of C / min(256, C)
   amplitudes, every sum left to right. Returns the first index whose running
   sum exceeds u * total (or the last non-zero one if rounding leaves none). */
double orc_sample_uniform(uint64_t seed, uint32_t traj, uint32_t s) {
  uint32_t c[4] = {0xFFFFFFFFu, s, traj, 0x5A3E1D2Cu};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32_10(c, k, o);
  return (double)((((uint64_t)o[1] << 32) | o[0]) >> 11) * 0x1.0p-53;
}

static double run_sum(const double* a, uint64_t start, uint64_t cnt) {
  double s = 0.0;
  for (uint64_t j = 0; j < cnt; ++j)
    s += RE(a, start + j) * RE(a, start + j) + IM(a, start + j) * IM(a, start + j);
  return s;
}

uint64_t orc_sample(const double* a, uint64_t len, double u) {
  const uint64_t C = len < 4096 ? len : 4096, nch = len / C;
  const uint64_t runs = C < 256 ? C : 256, per = C / runs;
  double total = 0.0;
  for (uint64_t c = 0; c < nch; ++c) {
    double cs = 0.0;
    for (uint64_t r = 0; r < runs; ++r) cs += run_sum(a, c * C + r * per, per);
    total += cs;
  }
  const double target = u * total;
  double acc = 0.0, acc_last = 0.0;
  int64_t chosen = -1, lastnz = -1;
  for (uint64_t c = 0; c < nch; ++c) {
    double cs = 0.0;
    for (uint64_t r = 0; r < runs; ++r) cs += run_sum(a, c * C + r * per, per);
    if (cs > 0.0) { lastnz = (int64_t)c; acc_last = acc; }
    if (acc + cs > target) { chosen = (int64_t)c; break; }
    acc += cs;
  }
  if (chosen < 0) { chosen = lastnz < 0 ? 0 : lastnz; acc = lastnz < 0 ? 0.0 : acc_last; }
  const double acc_chunk = acc;
  int64_t rsel = -1, rlast = -1;
  acc_last = acc;
  for (uint64_t r = 0; r < runs; ++r) {
    const double ps = run_sum(a, (uint64_t)chosen * C + r * per, per);
    if (ps > 0.0) { rlast = (int64_t)r; acc_last = acc; }
    if (acc + ps > target) { rsel = (int64_t)r; break; }
    acc += ps;
  }
  if (rsel < 0) { rsel = rlast < 0 ? 0 : rlast; acc = rlast < 0 ? acc_chunk : acc_last; }
  const uint64_t base = (uint64_t)chosen * C + (uint64_t)rsel * per;
  int64_t pick = -1, jlast = -1;
  for (uint64_t j = 0; j < per; ++j) {
    const double p = RE(a, base + j) * RE(a, base + j) + IM(a, base + j) * IM(a, base + j);
    if (p > 0.0) jlast = (int64_t)j;
    if (acc + p > target) { pick = (int64_t)j; break; }
    acc += p;
  }
  if (pick < 0) pick = jlast < 0 ? 0 : jlast;
  return base + (uint64_t)pick;
}
