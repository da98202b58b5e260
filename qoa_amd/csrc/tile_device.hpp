// tile_device.hpp — device code shared by the fused and batched pass kernels.
//
// Arithmetic follows the per-gate sweeps exactly (num-complex products, no
// FMA: every includer is built with -ffp-contract=off), so a gate applied in
// a tile and by its own sweep give the same bits. The register-resident pass
// also has an FMA instantiation (F = true) whose complex products use fused
// multiply-adds; it is selected per call and is not bit-identical.
#pragma once
#include "planner.hpp"
#if !defined(__HIPCC_RTC__)
#include "qoa_internal.hpp"
#endif
#include "traj_noise.hpp"

#include <type_traits>

#pragma clang fp contract(off)

namespace qoa {
namespace tile {

constexpr int kThreads = 256;

// Tile geometry: k and, four to a word (entry j is byte j % 4 of word j / 4),
// the physical position local bit j is loaded from (spw) and stored to (stw),
// and the same positions in ascending order (srt; the tile base inserts a
// zero bit at each). Without relabelling, spw = stw = srt. Positions are read
// through the *_at helpers, whose word selects keep them in scalar registers:
// a byte array indexed by a loop variable becomes one vector memory load,
// and a wait for it, per access (at the start of every tile, under the full
// HBM load).
struct Geo {
  int k;
  uint32_t spw[4];
  uint32_t stw[4];
  uint32_t srt[4];
};

__device__ __forceinline__ uint32_t byte_at(const uint32_t (&w4)[4], uint32_t j) {
  const uint32_t w = j < 4 ? w4[0] : j < 8 ? w4[1] : j < 12 ? w4[2] : w4[3];
  return (w >> (8 * (j & 3u))) & 0xFFu;
}
__device__ __forceinline__ uint32_t spos_at(const Geo& g, uint32_t j) { return byte_at(g.spw, j); }
__device__ __forceinline__ uint32_t stpos_at(const Geo& g, uint32_t j) { return byte_at(g.stw, j); }
__device__ __forceinline__ uint32_t srt_at(const Geo& g, uint32_t j) { return byte_at(g.srt, j); }

// The geometry of plans that never relabel (batches, serving): one set of
// ascending positions for loads, stores and the tile base, in five words
// (the batched kernels read them per work item and are short of registers).
struct GeoSame {
  int k;
  uint32_t spw[4];
};
__device__ __forceinline__ uint32_t spos_at(const GeoSame& g, uint32_t j) { return byte_at(g.spw, j); }
__device__ __forceinline__ uint32_t stpos_at(const GeoSame& g, uint32_t j) { return byte_at(g.spw, j); }
__device__ __forceinline__ uint32_t srt_at(const GeoSame& g, uint32_t j) { return byte_at(g.spw, j); }

// The geometry of batch work items, read where it lies in the item (global
// memory, uniform: scalar loads): load positions, store positions (equal
// unless the item's plan relabels qubits) and the positions ascending (the
// tile base). A copy in a local struct would sit in scratch memory, since the
// noise code indexes it with a loop variable.
struct GeoItem {
  int k;
  const uint32_t* spw;
  const uint32_t* stw;
  const uint32_t* srt;
};
__device__ __forceinline__ uint32_t word_byte(const uint32_t* w, uint32_t j) {
  return (w[j >> 2] >> (8 * (j & 3u))) & 0xFFu;
}
__device__ __forceinline__ uint32_t spos_at(const GeoItem& g, uint32_t j) { return word_byte(g.spw, j); }
__device__ __forceinline__ uint32_t stpos_at(const GeoItem& g, uint32_t j) { return word_byte(g.stw, j); }
__device__ __forceinline__ uint32_t srt_at(const GeoItem& g, uint32_t j) { return word_byte(g.srt, j); }


// The tile base of tile index t: t's bits spread over the positions outside
// the tile (a zero bit inserted at each tile position, ascending).
template <typename I, typename G>
__device__ __forceinline__ I tile_base(I t, const G& g) {
  for (int j = 0; j < g.k; ++j) {
    const int b = (int)srt_at(g, (uint32_t)j);
    t = ((t >> b) << (b + 1)) | (t & ((I(1) << b) - I(1)));
  }
  return t;
}

__device__ __forceinline__ double2 cmul(double2 x, double2 y) {
  return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
// c ? a : b per component (a conditional on the vector type itself would
// select addresses and keep the operands in memory)
__device__ __forceinline__ double2 sel2(bool c, double2 a, double2 b) {
  return make_double2(c ? a.x : b.x, c ? a.y : b.y);
}

// insert zero bits at the op's sorted local positions
__device__ __forceinline__ uint32_t dep(uint32_t x, const TileOp& op) {
  for (int j = 0; j < op.nins; ++j) {
    const int b = op.ins[j];
    x = ((x >> b) << (b + 1)) | (x & ((1u << b) - 1u));
  }
  return x;
}

template <typename I, typename G>
__device__ __forceinline__ I scatter(uint32_t l, const G& g) {
  I off = 0;
  for (int j = 0; j < g.k; ++j) off |= (I)((l >> j) & 1u) << spos_at(g, j);
  return off;
}

// Tile loads and stores of the register passes: every amplitude is read and
// written once per pass, so they carry the nontemporal hint (measured 2.7 %
// faster on the 28-qubit layered circuit than cached accesses).
typedef double qoa_dv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_stream(const double2* p) {
  const qoa_dv2 t = __builtin_nontemporal_load(reinterpret_cast<const qoa_dv2*>(p));
  return make_double2(t.x, t.y);
}
__device__ __forceinline__ void st_stream(double2* p, double2 v) {
  const qoa_dv2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<qoa_dv2*>(p));
}

// ---- complex products, exact (F = false) or with FMA (F = true) -------------
template <bool F>
__device__ __forceinline__ double2 mulc(double2 x, double2 y) {
  if (F) return make_double2(__builtin_fma(x.x, y.x, -(x.y * y.y)), __builtin_fma(x.x, y.y, x.y * y.x));
  return cmul(x, y);
}

// a*x + b*y
template <bool F>
__device__ __forceinline__ double2 mul2c(double2 a, double2 x, double2 b, double2 y) {
  if (F) {
    double re = __builtin_fma(a.x, x.x, -(a.y * x.y));
    re = __builtin_fma(b.x, y.x, re);
    re = __builtin_fma(-b.y, y.y, re);
    double im = __builtin_fma(a.x, x.y, a.y * x.x);
    im = __builtin_fma(b.x, y.y, im);
    im = __builtin_fma(b.y, y.x, im);
    return make_double2(re, im);
  }
  return cadd(cmul(a, x), cmul(b, y));
}

// acc + a*x
template <bool F>
__device__ __forceinline__ double2 macc(double2 acc, double2 a, double2 x) {
  if (F) {
    return make_double2(__builtin_fma(-a.y, x.y, __builtin_fma(a.x, x.x, acc.x)),
                        __builtin_fma(a.y, x.x, __builtin_fma(a.x, x.y, acc.y)));
  }
  return cadd(acc, cmul(a, x));
}

// ---- register-resident passes -------------------------------------------------
// Thread t of a 2^(k-4)-thread block holds the 16 amplitudes whose local
// indices are thr_l(t) | regl(r), r = 0..15: slot j of the register index r
// stands for local bit rpos[j] and the thread index fills the other local
// bits in ascending order. Ops whose mixed bits are register bits (and every
// diagonal op) run on registers; a TK_LAYOUT op changes the register bits
// through one LDS round trip. Controls come split by the planner: register
// slots (rsel, a test that is uniform over the block per register) and
// thread bits (tl, one per-thread test per op).
constexpr int kR = kRegAmps;

// A register permutation: logical register r (its slot bits) lives in
// physical register (pi >> 4r) & 15. The op-loop kernels always pass the
// identity; specialised kernels pass the permutation their X, CNOT and SWAP
// ops on register slots have accumulated instead of moving data (the
// primitives below take it as an argument that is a literal at every call).
constexpr uint64_t kPermId = 0xFEDCBA9876543210ULL;
__host__ __device__ constexpr int pm(uint64_t pi, int r) { return (int)((pi >> (4 * r)) & 15u); }

struct RegLayout {
  uint32_t rb[kRegBits];  // 1 << rpos[j] (uniform)
  uint32_t thr_l;         // this thread's local bits, register bits clear
};

__device__ __forceinline__ uint32_t regl(const RegLayout& L, int r) {
  uint32_t o = 0;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j)
    if ((r >> j) & 1) o |= L.rb[j];
  return o;
}

// set_layout for a layout whose register bit positions are also given in
// ascending order (a byte each in `sorted`, as exchange ops carry them): the
// thread's local bits are tid with a zero inserted at each of them.
__device__ __forceinline__ void set_layout_sorted(RegLayout& L, uint32_t tid, const uint8_t* rpos,
                                                  uint32_t sorted) {
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) L.rb[j] = 1u << rpos[j];
  uint32_t l = tid;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) {
    const uint32_t p = (sorted >> (8 * j)) & 0xFFu;
    l = ((l >> p) << (p + 1)) | (l & ((1u << p) - 1u));
  }
  L.thr_l = l;
}

__device__ __forceinline__ void set_layout(RegLayout& L, uint32_t tid, int k, const uint8_t* rpos) {
  uint32_t rmask = 0;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) {
    L.rb[j] = 1u << rpos[j];
    rmask |= L.rb[j];
  }
  uint32_t l = 0, src = 0;
  for (int b = 0; b < k; ++b) {
    if (rmask & (1u << b)) continue;
    l |= ((tid >> src) & 1u) << b;
    ++src;
  }
  L.thr_l = l;
}

constexpr int kHReal = 16;  // TK_H with a real factor: a' = (nf*Re, nf*Im)
constexpr int kMReal = 17;  // TK_M of shape kM2Real
constexpr int kMRx = 18;    // TK_M of shape kM2Rx
constexpr int kHRaw = 19;   // TK_H with nf = 1: sums and differences only
constexpr int kMRealU = 20; // TK_M of shape kM2RealU (m00 = m11 = 1, real)
constexpr int kMRxU = 21;   // TK_M of shape kM2RxU (m00 = m11 = 1, m01, m10 imaginary)

// Coefficients of a 2x2 op: c[0..3] = m00, m01, m10, m11 (TK_H: c[0] = nf).
struct Coef4 {
  double2 c[4];
};

// ---- in-place register arithmetic ------------------------------------------
// Each primitive writes its results into the registers that held its inputs
// ("+v" operands), so every amplitude keeps one register for the whole op
// loop and the loop's branches merge without register copies (the compiler
// otherwise places results in fresh registers and copies them back at every
// merge). The instruction sequences are the num-complex expressions of the
// per-gate sweeps, every product and sum rounded on its own; the *_fma forms
// are the FMA arithmetic of mulc<true> / mul2c<true>.

// The factor operands of these primitives are VGPRs (S = false: values that
// may differ per lane) or SGPR pairs (S = true: wave-uniform values such as
// an op's coefficients, read straight from the op words; every VOP3
// instruction here reads at most one of them).

// Inline asm statements are given several amplitudes at a time: the compiler
// cannot see into an asm statement, so it separates two consecutive ones that
// touch a common register (here the temporaries, which it allocates to the
// same registers every time) by a wait state (s_nop 0). One statement per
// two pairs, or per four factors, leaves one such wait where there were two
// or four; the instructions and their order per amplitude are unchanged.

// ---- complex products by a factor (c, d) ------------------------------------
// (a, b) *= (c, d): (a c - b d, a d + b c)
#define QOA_T_CMUL(A, B)                      \
  "v_mul_f64 %[t1], %[" A "], %[c]\n\t"       \
  "v_mul_f64 %[t2], %[" B "], %[d]\n\t"       \
  "v_mul_f64 %[t3], %[" A "], %[d]\n\t"       \
  "v_mul_f64 %[t4], %[" B "], %[c]\n\t"       \
  "v_add_f64 %[" A "], %[t1], -%[t2]\n\t"     \
  "v_add_f64 %[" B "], %[t3], %[t4]\n\t"
// (a, b) *= (c, d) as fma(a, c, -(b d)), fma(a, d, b c)
#define QOA_T_CMUL_FMA(A, B)                        \
  "v_mul_f64 %[t1], %[" B "], %[d]\n\t"             \
  "v_mul_f64 %[t2], %[" B "], %[c]\n\t"             \
  "v_fma_f64 %[" B "], %[" A "], %[d], %[t2]\n\t"   \
  "v_fma_f64 %[" A "], %[" A "], %[c], -%[t1]\n\t"

#define QOA_O1 [a0] "+v"(x0.x), [b0] "+v"(x0.y)
#define QOA_O2 QOA_O1, [a1] "+v"(x1.x), [b1] "+v"(x1.y)
#define QOA_O4 QOA_O2, [a2] "+v"(x2.x), [b2] "+v"(x2.y), [a3] "+v"(x3.x), [b3] "+v"(x3.y)

// ip_mulc{1,2,4}<F, S>(x0.., f): every x_i *= f (F: FMA form; S: f uniform)
#define QOA_DEF_MULC(NAME, PARAMS, TEXT_FMA, TEXT, OPS)                                   \
  template <bool F, bool S = false>                                                     \
  __device__ __forceinline__ void NAME PARAMS {                                         \
    if constexpr (F) {                                                                  \
      double t1, t2;                                                                    \
      if constexpr (S)                                                                  \
        asm(TEXT_FMA : OPS, [t1] "=&v"(t1), [t2] "=&v"(t2) : [c] "s"(f.x), [d] "s"(f.y)); \
      else                                                                              \
        asm(TEXT_FMA : OPS, [t1] "=&v"(t1), [t2] "=&v"(t2) : [c] "v"(f.x), [d] "v"(f.y)); \
    } else {                                                                            \
      double t1, t2, t3, t4;                                                            \
      if constexpr (S)                                                                  \
        asm(TEXT : OPS, [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3), [t4] "=&v"(t4)  \
            : [c] "s"(f.x), [d] "s"(f.y));                                              \
      else                                                                              \
        asm(TEXT : OPS, [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3), [t4] "=&v"(t4)  \
            : [c] "v"(f.x), [d] "v"(f.y));                                              \
    }                                                                                   \
  }
QOA_DEF_MULC(ip_mulc1, (double2& x0, double2 f), QOA_T_CMUL_FMA("a0", "b0"),
             QOA_T_CMUL("a0", "b0"), QOA_O1)
QOA_DEF_MULC(ip_mulc2, (double2& x0, double2& x1, double2 f),
             QOA_T_CMUL_FMA("a0", "b0") QOA_T_CMUL_FMA("a1", "b1"),
             QOA_T_CMUL("a0", "b0") QOA_T_CMUL("a1", "b1"), QOA_O2)
QOA_DEF_MULC(ip_mulc4, (double2& x0, double2& x1, double2& x2, double2& x3, double2 f),
             QOA_T_CMUL_FMA("a0", "b0") QOA_T_CMUL_FMA("a1", "b1")
                 QOA_T_CMUL_FMA("a2", "b2") QOA_T_CMUL_FMA("a3", "b3"),
             QOA_T_CMUL("a0", "b0") QOA_T_CMUL("a1", "b1") QOA_T_CMUL("a2", "b2")
                 QOA_T_CMUL("a3", "b3"),
             QOA_O4)
#undef QOA_DEF_MULC

template <bool F, bool S = false>
__device__ __forceinline__ void ip_mulc(double2& x, double2 f) {
  ip_mulc1<F, S>(x, f);
}

// (a, b) *= (c, d), exact form, factor in VGPRs
__device__ __forceinline__ void ip_cmul(double& a, double& b, double c, double d) {
  double2 x0 = make_double2(a, b);
  ip_mulc1<false, false>(x0, make_double2(c, d));
  a = x0.x;
  b = x0.y;
}

// (a, b) *= s (s = +-1: exact)
__device__ __forceinline__ void ip_scale(double2& x, double s) {
  asm("v_mul_f64 %[a], %[a], %[s]\n\t"
      "v_mul_f64 %[b], %[b], %[s]"
      : [a] "+v"(x.x), [b] "+v"(x.y)
      : [s] "v"(s));
}

// (a, b) *= -1 (exact)
__device__ __forceinline__ void ip_neg(double2& x) {
  asm("v_mul_f64 %[a], %[a], -1.0\n\t"
      "v_mul_f64 %[b], %[b], -1.0"
      : [a] "+v"(x.x), [b] "+v"(x.y));
}

// ---- 2x2 ops on pairs (x, y) = ((a, b), (c, d)) -------------------------------
#define QOA_P1 [a0] "+v"(x0.x), [b0] "+v"(x0.y), [c0] "+v"(y0.x), [d0] "+v"(y0.y)
#define QOA_P2 QOA_P1, [a1] "+v"(x1.x), [b1] "+v"(x1.y), [c1] "+v"(y1.x), [d1] "+v"(y1.y)

// x' = nf (x + y), y' = nf (x - y), nf real and uniform
#define QOA_T_HREAL(A, B, C, D)                  \
  "v_add_f64 %[t1], %[" A "], %[" C "]\n\t"      \
  "v_add_f64 %[t2], %[" B "], %[" D "]\n\t"      \
  "v_add_f64 %[t3], %[" A "], -%[" C "]\n\t"     \
  "v_add_f64 %[t4], %[" B "], -%[" D "]\n\t"     \
  "v_mul_f64 %[" A "], %[n], %[t1]\n\t"          \
  "v_mul_f64 %[" B "], %[n], %[t2]\n\t"          \
  "v_mul_f64 %[" C "], %[n], %[t3]\n\t"          \
  "v_mul_f64 %[" D "], %[n], %[t4]\n\t"
#define QOA_HREAL_TAIL                                                                   \
  [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3), [t4] "=&v"(t4) : [n] "s"(nf)
__device__ __forceinline__ void ip_h_real(double2& x0, double2& y0, double nf) {
  double t1, t2, t3, t4;
  asm(QOA_T_HREAL("a0", "b0", "c0", "d0") : QOA_P1, QOA_HREAL_TAIL);
}
__device__ __forceinline__ void ip_h_real2(double2& x0, double2& y0, double2& x1, double2& y1,
                                           double nf) {
  double t1, t2, t3, t4;
  asm(QOA_T_HREAL("a0", "b0", "c0", "d0") QOA_T_HREAL("a1", "b1", "c1", "d1")
      : QOA_P2, QOA_HREAL_TAIL);
}
#undef QOA_HREAL_TAIL
#undef QOA_T_HREAL

// Factored 2x2 forms: four independent instructions per pair, the new x in
// fresh registers (no copy back), y in place. Each equals the corresponding
// plain form with those coefficients exactly (products by 1 are exact).
// x' = x + y, y' = x - y (ip_h_real with nf = 1)
__device__ __forceinline__ void ip_h_raw(double2& x, double2& y) {
  double na, nb;
  asm("v_add_f64 %[na], %[a], %[c]\n\t"
      "v_add_f64 %[nb], %[b], %[d]\n\t"
      "v_add_f64 %[c], %[a], -%[c]\n\t"
      "v_add_f64 %[d], %[b], -%[d]"
      : [na] "=&v"(na), [nb] "=&v"(nb), [c] "+v"(y.x), [d] "+v"(y.y)
      : [a] "v"(x.x), [b] "v"(x.y));
  x = make_double2(na, nb);
}
// x' = x + q y, y' = r x + y, q, r real
__device__ __forceinline__ void ip_mat2_real_u(double2& x, double2& y, double q, double r) {
  double na, nb;
  asm("v_fma_f64 %[na], %[q], %[c], %[a]\n\t"
      "v_fma_f64 %[nb], %[q], %[d], %[b]\n\t"
      "v_fma_f64 %[c], %[r], %[a], %[c]\n\t"
      "v_fma_f64 %[d], %[r], %[b], %[d]"
      : [na] "=&v"(na), [nb] "=&v"(nb), [c] "+v"(y.x), [d] "+v"(y.y)
      : [a] "v"(x.x), [b] "v"(x.y), [q] "s"(q), [r] "s"(r));
  x = make_double2(na, nb);
}
// x' = x + (i q) y, y' = (i r) x + y, q, r real
__device__ __forceinline__ void ip_mat2_rx_u(double2& x, double2& y, double q, double r) {
  double na, nb;
  asm("v_fma_f64 %[na], -%[q], %[d], %[a]\n\t"
      "v_fma_f64 %[nb], %[q], %[c], %[b]\n\t"
      "v_fma_f64 %[c], -%[r], %[b], %[c]\n\t"
      "v_fma_f64 %[d], %[r], %[a], %[d]"
      : [na] "=&v"(na), [nb] "=&v"(nb), [c] "+v"(y.x), [d] "+v"(y.y)
      : [a] "v"(x.x), [b] "v"(x.y), [q] "s"(q), [r] "s"(r));
  x = make_double2(na, nb);
}

// x' = nf (x + y), y' = nf (x - y), nf complex and uniform: sums first, then
// in-place products nf * s (the product's operand order is the per-gate sweep's)
template <bool F>
__device__ __forceinline__ void ip_h_cplx(double2& x, double2& y, double2 nf) {
  double2 s, d;
  asm("v_add_f64 %[sr], %[a], %[c]\n\t"
      "v_add_f64 %[si], %[b], %[d]\n\t"
      "v_add_f64 %[dr], %[a], -%[c]\n\t"
      "v_add_f64 %[di], %[b], -%[d]"
      : [sr] "=&v"(s.x), [si] "=&v"(s.y), [dr] "=&v"(d.x), [di] "=&v"(d.y)
      : [a] "v"(x.x), [b] "v"(x.y), [c] "v"(y.x), [d] "v"(y.y));
  // nf * s: (nf.x s.x - nf.y s.y, nf.x s.y + nf.y s.x), written into x and y
  double t1, t2;
  if (F) {
    asm("v_mul_f64 %[t1], %[ni], %[si]\n\t"
        "v_mul_f64 %[t2], %[ni], %[sr]\n\t"
        "v_fma_f64 %[xa], %[nr], %[sr], -%[t1]\n\t"
        "v_fma_f64 %[xb], %[nr], %[si], %[t2]\n\t"
        "v_mul_f64 %[t1], %[ni], %[di]\n\t"
        "v_mul_f64 %[t2], %[ni], %[dr]\n\t"
        "v_fma_f64 %[yc], %[nr], %[dr], -%[t1]\n\t"
        "v_fma_f64 %[yd], %[nr], %[di], %[t2]"
        : [xa] "+v"(x.x), [xb] "+v"(x.y), [yc] "+v"(y.x), [yd] "+v"(y.y), [t1] "=&v"(t1),
          [t2] "=&v"(t2)
        : [nr] "s"(nf.x), [ni] "s"(nf.y), [sr] "v"(s.x), [si] "v"(s.y), [dr] "v"(d.x),
          [di] "v"(d.y));
  } else {
    asm("v_mul_f64 %[t1], %[nr], %[sr]\n\t"
        "v_mul_f64 %[t2], %[ni], %[si]\n\t"
        "v_add_f64 %[xa], %[t1], -%[t2]\n\t"
        "v_mul_f64 %[t1], %[nr], %[si]\n\t"
        "v_mul_f64 %[t2], %[ni], %[sr]\n\t"
        "v_add_f64 %[xb], %[t1], %[t2]\n\t"
        "v_mul_f64 %[t1], %[nr], %[dr]\n\t"
        "v_mul_f64 %[t2], %[ni], %[di]\n\t"
        "v_add_f64 %[yc], %[t1], -%[t2]\n\t"
        "v_mul_f64 %[t1], %[nr], %[di]\n\t"
        "v_mul_f64 %[t2], %[ni], %[dr]\n\t"
        "v_add_f64 %[yd], %[t1], %[t2]"
        : [xa] "+v"(x.x), [xb] "+v"(x.y), [yc] "+v"(y.x), [yd] "+v"(y.y), [t1] "=&v"(t1),
          [t2] "=&v"(t2)
        : [nr] "s"(nf.x), [ni] "s"(nf.y), [sr] "v"(s.x), [si] "v"(s.y), [dr] "v"(d.x),
          [di] "v"(d.y));
  }
}

// x' = m00 x + m01 y, y' = m10 x + m11 y (cadd of two num-complex products;
// uniform coefficients). One complex product m * z into (pr, pi), num-complex
// order: (m.x z.x - m.y z.y, m.x z.y + m.y z.x).
#define QOA_CMUL_TO(PR, PI, MR, MI, ZR, ZI)            \
  "v_mul_f64 %[t1], %[" MR "], %[" ZR "]\n\t"            \
  "v_mul_f64 %[t2], %[" MI "], %[" ZI "]\n\t"            \
  "v_mul_f64 %[t3], %[" MR "], %[" ZI "]\n\t"            \
  "v_mul_f64 %[t4], %[" MI "], %[" ZR "]\n\t"            \
  "v_add_f64 %[" PR "], %[t1], -%[t2]\n\t"               \
  "v_add_f64 %[" PI "], %[t3], %[t4]\n\t"
#define QOA_T_MAT2(A, B, C, D)                           \
  QOA_CMUL_TO("pr", "pi", "m0r", "m0i", A, B)            \
  QOA_CMUL_TO("qr", "qi", "m1r", "m1i", C, D)            \
  QOA_CMUL_TO("rr", "ri", "m2r", "m2i", A, B)            \
  QOA_CMUL_TO("sr", "si", "m3r", "m3i", C, D)            \
  "v_add_f64 %[" A "], %[pr], %[qr]\n\t"                 \
  "v_add_f64 %[" B "], %[pi], %[qi]\n\t"                 \
  "v_add_f64 %[" C "], %[rr], %[sr]\n\t"                 \
  "v_add_f64 %[" D "], %[ri], %[si]\n\t"
#define QOA_MAT2_COEFS                                                                   \
  [m0r] "s"(k.c[0].x), [m0i] "s"(k.c[0].y), [m1r] "s"(k.c[1].x), [m1i] "s"(k.c[1].y),     \
      [m2r] "s"(k.c[2].x), [m2i] "s"(k.c[2].y), [m3r] "s"(k.c[3].x), [m3i] "s"(k.c[3].y)
#define QOA_MAT2_TEMPS                                                                   \
  [pr] "=&v"(pr), [pi] "=&v"(pi), [qr] "=&v"(qr), [qi] "=&v"(qi), [rr] "=&v"(rr),          \
      [ri] "=&v"(ri), [sr] "=&v"(sr), [si] "=&v"(si), [t1] "=&v"(t1), [t2] "=&v"(t2),     \
      [t3] "=&v"(t3), [t4] "=&v"(t4)
__device__ __forceinline__ void ip_mat2(double2& x0, double2& y0, const Coef4& k) {
  double pr, pi, qr, qi, rr, ri, sr, si, t1, t2, t3, t4;
  asm(QOA_T_MAT2("a0", "b0", "c0", "d0") : QOA_P1, QOA_MAT2_TEMPS : QOA_MAT2_COEFS);
}
__device__ __forceinline__ void ip_mat2_2(double2& x0, double2& y0, double2& x1, double2& y1,
                                          const Coef4& k) {
  double pr, pi, qr, qi, rr, ri, sr, si, t1, t2, t3, t4;
  asm(QOA_T_MAT2("a0", "b0", "c0", "d0") QOA_T_MAT2("a1", "b1", "c1", "d1")
      : QOA_P2, QOA_MAT2_TEMPS : QOA_MAT2_COEFS);
}
#undef QOA_MAT2_TEMPS
#undef QOA_T_MAT2
#undef QOA_CMUL_TO

// FMA form of ip_mat2 (mul2c<true>): re = fma(-m.y', z'.y, fma(m.x', z'.x,
// fma(m.x, z.x, -(m.y z.y)))) and likewise im.
#define QOA_T_MAT2_FMA(A, B, C, D)                             \
  "v_mul_f64 %[t], %[m0i], %[" B "]\n\t"                       \
  "v_fma_f64 %[o0r], %[m0r], %[" A "], -%[t]\n\t"              \
  "v_fma_f64 %[o0r], %[m1r], %[" C "], %[o0r]\n\t"             \
  "v_fma_f64 %[o0r], -%[m1i], %[" D "], %[o0r]\n\t"            \
  "v_mul_f64 %[t], %[m0i], %[" A "]\n\t"                       \
  "v_fma_f64 %[o0i], %[m0r], %[" B "], %[t]\n\t"               \
  "v_fma_f64 %[o0i], %[m1r], %[" D "], %[o0i]\n\t"             \
  "v_fma_f64 %[o0i], %[m1i], %[" C "], %[o0i]\n\t"             \
  "v_mul_f64 %[t], %[m2i], %[" B "]\n\t"                       \
  "v_fma_f64 %[o1r], %[m2r], %[" A "], -%[t]\n\t"              \
  "v_fma_f64 %[o1r], %[m3r], %[" C "], %[o1r]\n\t"             \
  "v_fma_f64 %[o1r], -%[m3i], %[" D "], %[o1r]\n\t"            \
  "v_mul_f64 %[t], %[m2i], %[" A "]\n\t"                       \
  "v_fma_f64 %[t], %[m2r], %[" B "], %[t]\n\t"                 \
  "v_fma_f64 %[t], %[m3r], %[" D "], %[t]\n\t"                 \
  "v_fma_f64 %[" D "], %[m3i], %[" C "], %[t]\n\t"             \
  "v_mov_b64 %[" A "], %[o0r]\n\t"                             \
  "v_mov_b64 %[" B "], %[o0i]\n\t"                             \
  "v_mov_b64 %[" C "], %[o1r]\n\t"
#define QOA_MAT2_FMA_TEMPS [o0r] "=&v"(o0r), [o0i] "=&v"(o0i), [o1r] "=&v"(o1r), [t] "=&v"(t)
__device__ __forceinline__ void ip_mat2_fma(double2& x0, double2& y0, const Coef4& k) {
  double o0r, o0i, o1r, t;
  asm(QOA_T_MAT2_FMA("a0", "b0", "c0", "d0") : QOA_P1, QOA_MAT2_FMA_TEMPS : QOA_MAT2_COEFS);
}
__device__ __forceinline__ void ip_mat2_fma2(double2& x0, double2& y0, double2& x1, double2& y1,
                                             const Coef4& k) {
  double o0r, o0i, o1r, t;
  asm(QOA_T_MAT2_FMA("a0", "b0", "c0", "d0") QOA_T_MAT2_FMA("a1", "b1", "c1", "d1")
      : QOA_P2, QOA_MAT2_FMA_TEMPS : QOA_MAT2_COEFS);
}
#undef QOA_MAT2_FMA_TEMPS
#undef QOA_T_MAT2_FMA
#undef QOA_MAT2_COEFS

// TK_M with every imaginary part 0 (m_ij real, uniform):
// x' = (m00 a + m01 c, m00 b + m01 d), y' = (m10 a + m11 c, m10 b + m11 d)
// for x = (a, b), y = (c, d).
#define QOA_T_MAT2_REAL(A, B, C, D)              \
  "v_mul_f64 %[t1], %[p], %[" A "]\n\t"          \
  "v_mul_f64 %[t2], %[q], %[" C "]\n\t"          \
  "v_mul_f64 %[t3], %[r], %[" A "]\n\t"          \
  "v_mul_f64 %[t4], %[s], %[" C "]\n\t"          \
  "v_add_f64 %[" A "], %[t1], %[t2]\n\t"         \
  "v_add_f64 %[" C "], %[t3], %[t4]\n\t"         \
  "v_mul_f64 %[t1], %[p], %[" B "]\n\t"          \
  "v_mul_f64 %[t2], %[q], %[" D "]\n\t"          \
  "v_mul_f64 %[t3], %[r], %[" B "]\n\t"          \
  "v_mul_f64 %[t4], %[s], %[" D "]\n\t"          \
  "v_add_f64 %[" B "], %[t1], %[t2]\n\t"         \
  "v_add_f64 %[" D "], %[t3], %[t4]\n\t"
// FMA form: x' = (fma(m01, c, m00 a), fma(m01, d, m00 b)), likewise y'
#define QOA_T_MAT2_REAL_FMA(A, B, C, D)                  \
  "v_mul_f64 %[t1], %[p], %[" A "]\n\t"                  \
  "v_mul_f64 %[t3], %[r], %[" A "]\n\t"                  \
  "v_fma_f64 %[" A "], %[q], %[" C "], %[t1]\n\t"        \
  "v_fma_f64 %[" C "], %[s], %[" C "], %[t3]\n\t"        \
  "v_mul_f64 %[t1], %[p], %[" B "]\n\t"                  \
  "v_mul_f64 %[t3], %[r], %[" B "]\n\t"                  \
  "v_fma_f64 %[" B "], %[q], %[" D "], %[t1]\n\t"        \
  "v_fma_f64 %[" D "], %[s], %[" D "], %[t3]\n\t"
// TK_M with m00 = (p, 0), m01 = (0, q), m10 = (0, r), m11 = (s, 0):
// x' = (p a - q d, p b + q c), y' = (s c - r b, r a + s d).
#define QOA_T_MAT2_RX(A, B, C, D)                \
  "v_mul_f64 %[t1], %[p], %[" A "]\n\t"          \
  "v_mul_f64 %[t3], %[p], %[" B "]\n\t"          \
  "v_mul_f64 %[t7], %[r], %[" A "]\n\t"          \
  "v_mul_f64 %[t6], %[r], %[" B "]\n\t"          \
  "v_mul_f64 %[t2], %[q], %[" D "]\n\t"          \
  "v_add_f64 %[" A "], %[t1], -%[t2]\n\t"        \
  "v_mul_f64 %[t2], %[q], %[" C "]\n\t"          \
  "v_add_f64 %[" B "], %[t3], %[t2]\n\t"         \
  "v_mul_f64 %[t1], %[s], %[" C "]\n\t"          \
  "v_add_f64 %[" C "], %[t1], -%[t6]\n\t"        \
  "v_mul_f64 %[t1], %[s], %[" D "]\n\t"          \
  "v_add_f64 %[" D "], %[t7], %[t1]\n\t"
// FMA form: x' = (fma(-q, d, p a), fma(q, c, p b)), y' = (fma(s, c, -(r b)), fma(s, d, r a))
#define QOA_T_MAT2_RX_FMA(A, B, C, D)                    \
  "v_mul_f64 %[t1], %[p], %[" A "]\n\t"                  \
  "v_mul_f64 %[t3], %[p], %[" B "]\n\t"                  \
  "v_mul_f64 %[t7], %[r], %[" A "]\n\t"                  \
  "v_mul_f64 %[t6], %[r], %[" B "]\n\t"                  \
  "v_fma_f64 %[" A "], -%[q], %[" D "], %[t1]\n\t"       \
  "v_fma_f64 %[" B "], %[q], %[" C "], %[t3]\n\t"        \
  "v_fma_f64 %[" C "], %[s], %[" C "], -%[t6]\n\t"       \
  "v_fma_f64 %[" D "], %[s], %[" D "], %[t7]\n\t"
#define QOA_PQRS : [p] "s"(p), [q] "s"(q), [r] "s"(r), [s] "s"(s)
#define QOA_T4 [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3), [t4] "=&v"(t4)
#define QOA_T13 [t1] "=&v"(t1), [t3] "=&v"(t3)
#define QOA_T5 [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3), [t6] "=&v"(t6), [t7] "=&v"(t7)
#define QOA_T4X [t1] "=&v"(t1), [t3] "=&v"(t3), [t6] "=&v"(t6), [t7] "=&v"(t7)
#define QOA_UNPAREN(...) __VA_ARGS__
#define QOA_DEF_PAIR(NAME, NAME2, TEXT, TEMPS_DECL, TEMPS)                                   \
  __device__ __forceinline__ void NAME(double2& x0, double2& y0, double p, double q, double r, \
                                       double s) {                                           \
    QOA_UNPAREN TEMPS_DECL;                                                                  \
    asm(TEXT("a0", "b0", "c0", "d0") : QOA_P1, TEMPS QOA_PQRS);                              \
  }                                                                                          \
  __device__ __forceinline__ void NAME2(double2& x0, double2& y0, double2& x1, double2& y1,   \
                                        double p, double q, double r, double s) {            \
    QOA_UNPAREN TEMPS_DECL;                                                                  \
    asm(TEXT("a0", "b0", "c0", "d0") TEXT("a1", "b1", "c1", "d1") : QOA_P2, TEMPS QOA_PQRS); \
  }
QOA_DEF_PAIR(ip_mat2_real, ip_mat2_real2, QOA_T_MAT2_REAL, (double t1, t2, t3, t4), QOA_T4)
QOA_DEF_PAIR(ip_mat2_real_fma, ip_mat2_real_fma2, QOA_T_MAT2_REAL_FMA, (double t1, t3), QOA_T13)
QOA_DEF_PAIR(ip_mat2_rx, ip_mat2_rx2, QOA_T_MAT2_RX, (double t1, t2, t3, t6, t7), QOA_T5)
QOA_DEF_PAIR(ip_mat2_rx_fma, ip_mat2_rx_fma2, QOA_T_MAT2_RX_FMA, (double t1, t3, t6, t7), QOA_T4X)
#undef QOA_DEF_PAIR
#undef QOA_UNPAREN
#undef QOA_T4X
#undef QOA_T5
#undef QOA_T13
#undef QOA_T4
#undef QOA_PQRS
#undef QOA_T_MAT2_RX_FMA
#undef QOA_T_MAT2_RX
#undef QOA_T_MAT2_REAL_FMA
#undef QOA_T_MAT2_REAL

// Y: x' = y (0, -1), y' = x (0, 1), as the per-gate sweep's products
__device__ __forceinline__ void ip_y(double2& x, double2& y) {
  double t1, t2, t3, t4, u1, u2, u3, u4;
  asm("v_mul_f64 %[t1], %[c], 0\n\t"
      "v_mul_f64 %[t2], %[d], -1.0\n\t"
      "v_mul_f64 %[t3], %[c], -1.0\n\t"
      "v_mul_f64 %[t4], %[d], 0\n\t"
      "v_mul_f64 %[u1], %[a], 0\n\t"
      "v_mul_f64 %[u2], %[b], 1.0\n\t"
      "v_mul_f64 %[u3], %[a], 1.0\n\t"
      "v_mul_f64 %[u4], %[b], 0\n\t"
      "v_add_f64 %[a], %[t1], -%[t2]\n\t"
      "v_add_f64 %[b], %[t3], %[t4]\n\t"
      "v_add_f64 %[c], %[u1], -%[u2]\n\t"
      "v_add_f64 %[d], %[u3], %[u4]"
      : [a] "+v"(x.x), [b] "+v"(x.y), [c] "+v"(y.x), [d] "+v"(y.y), [t1] "=&v"(t1),
        [t2] "=&v"(t2), [t3] "=&v"(t3), [t4] "=&v"(t4), [u1] "=&v"(u1), [u2] "=&v"(u2),
        [u3] "=&v"(u3), [u4] "=&v"(u4));
}

// exchange two doubles' registers
__device__ __forceinline__ void ip_swap(double& x, double& y) {
  uint64_t ux = __builtin_bit_cast(uint64_t, x), uy = __builtin_bit_cast(uint64_t, y);
  uint32_t xl = (uint32_t)ux, xh = (uint32_t)(ux >> 32);
  uint32_t yl = (uint32_t)uy, yh = (uint32_t)(uy >> 32);
  asm("v_swap_b32 %[xl], %[yl]\n\t"
      "v_swap_b32 %[xh], %[yh]"
      : [xl] "+v"(xl), [xh] "+v"(xh), [yl] "+v"(yl), [yh] "+v"(yh));
  x = __builtin_bit_cast(double, ((uint64_t)xh << 32) | xl);
  y = __builtin_bit_cast(double, ((uint64_t)yh << 32) | yl);
}

__device__ __forceinline__ void ip_swap2(double2& x, double2& y) {
  ip_swap(x.x, y.x);
  ip_swap(x.y, y.y);
}

// x <-> y in the lanes of `mask` only: 64-bit moves through one temporary
// with exec narrowed to those lanes (restored in the same statement), so no
// register beyond x, y and the temporary is written and every value stays in
// an aligned register pair.
__device__ __forceinline__ void ip_cswap2(double2& x, double2& y, uint64_t mask) {
  double t;
  uint64_t save;
  asm volatile(
      "s_and_saveexec_b64 %[s], %[m]\n\t"
      "v_mov_b64 %[t], %[a]\n\t"
      "v_mov_b64 %[a], %[c]\n\t"
      "v_mov_b64 %[c], %[t]\n\t"
      "v_mov_b64 %[t], %[b]\n\t"
      "v_mov_b64 %[b], %[d]\n\t"
      "v_mov_b64 %[d], %[t]\n\t"
      "s_mov_b64 exec, %[s]"
      : [a] "+v"(x.x), [b] "+v"(x.y), [c] "+v"(y.x), [d] "+v"(y.y), [t] "=&v"(t), [s] "=&s"(save)
      : [m] "s"(mask)
      : "scc");
}

// x = lanes of `mask` ? n : x (exec-narrowed 64-bit moves)
__device__ __forceinline__ void ip_select(double& x, double n, uint64_t mask) {
  uint64_t save;
  asm volatile(
      "s_and_saveexec_b64 %[s], %[m]\n\t"
      "v_mov_b64 %[x], %[n]\n\t"
      "s_mov_b64 exec, %[s]"
      : [x] "+v"(x), [s] "=&s"(save)
      : [n] "v"(n), [m] "s"(mask)
      : "scc");
}

__device__ __forceinline__ void ip_select2(double2& x, double2 n, uint64_t mask) {
  uint64_t save;
  asm volatile(
      "s_and_saveexec_b64 %[s], %[m]\n\t"
      "v_mov_b64 %[a], %[na]\n\t"
      "v_mov_b64 %[b], %[nb]\n\t"
      "s_mov_b64 exec, %[s]"
      : [a] "+v"(x.x), [b] "+v"(x.y), [s] "=&s"(save)
      : [na] "v"(n.x), [nb] "v"(n.y), [m] "s"(mask)
      : "scc");
}

template <int KIND, bool F>
__device__ __forceinline__ void pair_math(double2 x, double2 y, const Coef4& k, double2& o0,
                                          double2& o1) {
  if (KIND == kHReal) {
    // equals cmul((nf, 0), .) up to the sign of zero terms
    const double nf = k.c[0].x;
    o0 = make_double2(nf * (x.x + y.x), nf * (x.y + y.y));
    o1 = make_double2(nf * (x.x - y.x), nf * (x.y - y.y));
  } else if (KIND == kHRaw) {
    o0 = make_double2(x.x + y.x, x.y + y.y);
    o1 = make_double2(x.x - y.x, x.y - y.y);
  } else if (KIND == kMRealU) {
    const double q = k.c[1].x, r = k.c[2].x;
    o0 = make_double2(__builtin_fma(q, y.x, x.x), __builtin_fma(q, y.y, x.y));
    o1 = make_double2(__builtin_fma(r, x.x, y.x), __builtin_fma(r, x.y, y.y));
  } else if (KIND == kMRxU) {
    const double q = k.c[1].y, r = k.c[2].y;
    o0 = make_double2(__builtin_fma(-q, y.y, x.x), __builtin_fma(q, y.x, x.y));
    o1 = make_double2(__builtin_fma(-r, x.y, y.x), __builtin_fma(r, x.x, y.y));
  } else if (KIND == TK_H) {
    o0 = mulc<F>(k.c[0], cadd(x, y));
    o1 = mulc<F>(k.c[0], csub(x, y));
  } else if (KIND == TK_M || KIND == kMReal || KIND == kMRx) {
    o0 = mul2c<F>(k.c[0], x, k.c[1], y);
    o1 = mul2c<F>(k.c[2], x, k.c[3], y);
  } else if (KIND == TK_X) {
    o0 = y;
    o1 = x;
  } else {  // TK_Y
    o0 = cmul(y, make_double2(0.0, -1.0));
    o1 = cmul(x, make_double2(0.0, 1.0));
  }
}

// pair_math in place
template <int KIND, bool F>
__device__ __forceinline__ void pair_inplace(double2& x, double2& y, const Coef4& k) {
  if (KIND == kHReal) ip_h_real(x, y, k.c[0].x);
  else if (KIND == kHRaw) ip_h_raw(x, y);
  else if (KIND == kMRealU) ip_mat2_real_u(x, y, k.c[1].x, k.c[2].x);
  else if (KIND == kMRxU) ip_mat2_rx_u(x, y, k.c[1].y, k.c[2].y);
  else if (KIND == TK_H) ip_h_cplx<F>(x, y, k.c[0]);
  else if (KIND == TK_M) {
    if (F) ip_mat2_fma(x, y, k);
    else ip_mat2(x, y, k);
  } else if (KIND == kMReal) {
    if (F) ip_mat2_real_fma(x, y, k.c[0].x, k.c[1].x, k.c[2].x, k.c[3].x);
    else ip_mat2_real(x, y, k.c[0].x, k.c[1].x, k.c[2].x, k.c[3].x);
  } else if (KIND == kMRx) {
    if (F) ip_mat2_rx_fma(x, y, k.c[0].x, k.c[1].y, k.c[2].y, k.c[3].x);
    else ip_mat2_rx(x, y, k.c[0].x, k.c[1].y, k.c[2].y, k.c[3].x);
  } else if (KIND == TK_X) ip_swap2(x, y);
  else ip_y(x, y);
}

// pair_inplace on two pairs in one asm statement where the kind has one
template <int KIND, bool F>
__device__ __forceinline__ void pair_inplace2(double2& x0, double2& y0, double2& x1, double2& y1,
                                              const Coef4& k) {
  if constexpr (KIND == kHReal) {
    ip_h_real2(x0, y0, x1, y1, k.c[0].x);
  } else if constexpr (KIND == TK_M) {
    if (F) ip_mat2_fma2(x0, y0, x1, y1, k);
    else ip_mat2_2(x0, y0, x1, y1, k);
  } else if constexpr (KIND == kMReal) {
    if (F) ip_mat2_real_fma2(x0, y0, x1, y1, k.c[0].x, k.c[1].x, k.c[2].x, k.c[3].x);
    else ip_mat2_real2(x0, y0, x1, y1, k.c[0].x, k.c[1].x, k.c[2].x, k.c[3].x);
  } else if constexpr (KIND == kMRx) {
    if (F) ip_mat2_rx_fma2(x0, y0, x1, y1, k.c[0].x, k.c[1].y, k.c[2].y, k.c[3].x);
    else ip_mat2_rx2(x0, y0, x1, y1, k.c[0].x, k.c[1].y, k.c[2].y, k.c[3].x);
  } else {
    pair_inplace<KIND, F>(x0, y0, k);
    pair_inplace<KIND, F>(x1, y1, k);
  }
}

// Register indices r with (r & M) == V, ascending: how many, and the i-th.
__host__ __device__ constexpr int n_regs(int M, int V) {
  int n = 0;
  for (int r = 0; r < kR; ++r) n += (r & M) == V ? 1 : 0;
  return n;
}
__host__ __device__ constexpr int nth_reg(int M, int V, int i) {
  for (int r = 0; r < kR; ++r)
    if ((r & M) == V) {
      if (i == 0) return r;
      --i;
    }
  return 0;
}

// fn(std::integral_constant<int, I>) for I = B .. E-1, in order
template <int B, int E, class Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  if constexpr (B < E) {
    fn(std::integral_constant<int, B>{});
    static_for<B + 1, E>(fn);
  }
}

// 2x2 op on register slot J for the registers whose slots RS are all 1,
// two pairs per asm statement, pairs in ascending register order.
template <int J, int RS, int KIND, bool F>
__device__ __forceinline__ void reg_pair(double2 (&v)[kR], const Coef4& k, uint64_t pi = kPermId) {
  constexpr int M = (1 << J) | RS, V = RS & ~(1 << J), n = n_regs(M, V);
  static_for<0, n / 2>([&](auto i) {
    constexpr int r0 = nth_reg(M, V, 2 * decltype(i)::value);
    constexpr int r1 = nth_reg(M, V, 2 * decltype(i)::value + 1);
    pair_inplace2<KIND, F>(v[pm(pi, r0)], v[pm(pi, r0 | (1 << J))], v[pm(pi, r1)],
                           v[pm(pi, r1 | (1 << J))], k);
  });
  if constexpr (n & 1) {
    constexpr int r = nth_reg(M, V, n - 1);
    pair_inplace<KIND, F>(v[pm(pi, r)], v[pm(pi, r | (1 << J))], k);
  }
}

// Calls fn(std::integral_constant<int, RS>) for RS = rsel (0..15): one switch
// (a binary search of scalar compares) instead of a chain of tests.
template <class Fn>
__device__ __forceinline__ void rs_switch(uint32_t rsel, Fn&& fn) {
  switch (rsel) {
    case 0: fn(std::integral_constant<int, 0>{}); break;
    case 1: fn(std::integral_constant<int, 1>{}); break;
    case 2: fn(std::integral_constant<int, 2>{}); break;
    case 3: fn(std::integral_constant<int, 3>{}); break;
    case 4: fn(std::integral_constant<int, 4>{}); break;
    case 5: fn(std::integral_constant<int, 5>{}); break;
    case 6: fn(std::integral_constant<int, 6>{}); break;
    case 7: fn(std::integral_constant<int, 7>{}); break;
    case 8: fn(std::integral_constant<int, 8>{}); break;
    case 9: fn(std::integral_constant<int, 9>{}); break;
    case 10: fn(std::integral_constant<int, 10>{}); break;
    case 11: fn(std::integral_constant<int, 11>{}); break;
    case 12: fn(std::integral_constant<int, 12>{}); break;
    case 13: fn(std::integral_constant<int, 13>{}); break;
    case 14: fn(std::integral_constant<int, 14>{}); break;
    default: fn(std::integral_constant<int, 15>{}); break;
  }
}

// Register controls rsel picked at compile time (a switch of uniform tests).
template <int J, int KIND, bool F>
__device__ __forceinline__ void reg_pair_rs(double2 (&v)[kR], const Coef4& k, uint32_t rsel) {
  rs_switch(rsel, [&](auto rs) { reg_pair<J, decltype(rs)::value, KIND, F>(v, k); });
}

// Thread-bit controls: every lane computes, lanes whose control bits are
// not all set keep their values (tok == false).
template <int J, int KIND, bool F>
__device__ __forceinline__ void reg_pair_sel(double2 (&v)[kR], const Coef4& k, uint32_t rsel,
                                             bool tok, uint64_t pi = kPermId) {
  const uint64_t mask = __builtin_amdgcn_ballot_w64(tok);
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & (1 << J)) continue;
    if (((uint32_t)r & rsel) != rsel) continue;  // uniform
    const int r1 = r | (1 << J);
    if constexpr (KIND == TK_X) {  // a conditional swap in the lanes of tok
      ip_cswap2(v[pm(pi, r)], v[pm(pi, r1)], mask);
    } else {
      double2 o0, o1;
      pair_math<KIND, F>(v[pm(pi, r)], v[pm(pi, r1)], k, o0, o1);
      ip_select2(v[pm(pi, r)], o0, mask);
      ip_select2(v[pm(pi, r1)], o1, mask);
    }
  }
}

// Uncontrolled ops take the straight path; X (CNOT, Toffoli) with register
// controls has its control sets compiled in; other controlled ops test the
// register controls per pair.
template <int KIND, bool F>
__device__ __forceinline__ void reg_pair_any(double2 (&v)[kR], const Coef4& k, uint32_t j,
                                             uint32_t rsel, uint32_t tl, uint32_t thr_l) {
  if (tl == 0 && rsel == 0) {
    switch (j) {
      case 0: reg_pair<0, 0, KIND, F>(v, k); break;
      case 1: reg_pair<1, 0, KIND, F>(v, k); break;
      case 2: reg_pair<2, 0, KIND, F>(v, k); break;
      default: reg_pair<3, 0, KIND, F>(v, k); break;
    }
  } else if (KIND == TK_X && tl == 0) {
    switch (j) {
      case 0: reg_pair_rs<0, KIND, F>(v, k, rsel); break;
      case 1: reg_pair_rs<1, KIND, F>(v, k, rsel); break;
      case 2: reg_pair_rs<2, KIND, F>(v, k, rsel); break;
      default: reg_pair_rs<3, KIND, F>(v, k, rsel); break;
    }
  } else {  // shaped 2x2 matrices take the general code here (same values)
    constexpr int SK = (KIND == kMReal || KIND == kMRx) ? TK_M : KIND;
    const bool tok = (thr_l & tl) == tl;
    switch (j) {
      case 0: reg_pair_sel<0, SK, F>(v, k, rsel, tok); break;
      case 1: reg_pair_sel<1, SK, F>(v, k, rsel, tok); break;
      case 2: reg_pair_sel<2, SK, F>(v, k, rsel, tok); break;
      default: reg_pair_sel<3, SK, F>(v, k, rsel, tok); break;
    }
  }
}

// Factor f on the registers r with (r & M) == V, four (or two, or one) per
// asm statement, in ascending register order.
template <int M, int V, bool F, bool S>
__device__ __forceinline__ void reg_mulc_set(double2 (&v)[kR], double2 f, uint64_t pi = kPermId) {
  constexpr int n = n_regs(M, V);
  static_for<0, n / 4>([&](auto i) {
    constexpr int b = 4 * decltype(i)::value;
    ip_mulc4<F, S>(v[pm(pi, nth_reg(M, V, b))], v[pm(pi, nth_reg(M, V, b + 1))],
                   v[pm(pi, nth_reg(M, V, b + 2))], v[pm(pi, nth_reg(M, V, b + 3))], f);
  });
  if constexpr (n % 4 >= 2) {
    constexpr int b = n - n % 4;
    ip_mulc2<F, S>(v[pm(pi, nth_reg(M, V, b))], v[pm(pi, nth_reg(M, V, b + 1))], f);
  }
  if constexpr (n % 2) ip_mulc1<F, S>(v[pm(pi, nth_reg(M, V, n - 1))], f);
}

// Diagonal factor f on the registers whose slots RS are all 1.
template <int RS, bool F, bool S>
__device__ __forceinline__ void reg_phase(double2 (&v)[kR], double2 f, uint64_t pi = kPermId) {
  reg_mulc_set<RS, RS, F, S>(v, f, pi);
}

template <bool F, bool S>
__device__ __forceinline__ void reg_phase_rs(double2 (&v)[kR], double2 f, uint32_t rsel) {
  rs_switch(rsel, [&](auto rs) { reg_phase<decltype(rs)::value, F, S>(v, f); });
}

// RZ with its target in register slot RT: the factor of register r is
// m11 or m00 by bit RT of r (uniform).
template <int RT, bool F>
__device__ __forceinline__ void reg_rz(double2 (&v)[kR], double2 m00, double2 m11,
                                       uint64_t pi = kPermId) {
  reg_mulc_set<1 << RT, 0, F, true>(v, m00, pi);
  reg_mulc_set<1 << RT, 1 << RT, F, true>(v, m11, pi);
}

template <int RS>
__device__ __forceinline__ void reg_scale(double2 (&v)[kR], double s, uint64_t pi = kPermId) {
#pragma unroll
  for (int r = 0; r < kR; ++r)
    if ((r & RS) == RS) ip_scale(v[pm(pi, r)], s);
}

template <int RS>
__device__ __forceinline__ void reg_neg(double2 (&v)[kR], uint64_t pi = kPermId) {
#pragma unroll
  for (int r = 0; r < kR; ++r)
    if ((r & RS) == RS) ip_neg(v[pm(pi, r)]);
}

__device__ __forceinline__ void reg_neg_rs(double2 (&v)[kR], uint32_t rsel) {
  rs_switch(rsel, [&](auto rs) { reg_neg<decltype(rs)::value>(v); });
}

__device__ __forceinline__ void reg_scale_rs(double2 (&v)[kR], double sc, uint32_t rsel) {
  rs_switch(rsel, [&](auto rs) { reg_scale<decltype(rs)::value>(v, sc); });
}

template <int A, int B>
__device__ __forceinline__ void reg_swap(double2 (&v)[kR], uint32_t rsel, uint32_t tl, bool tok,
                                         uint64_t pi = kPermId) {
  const uint64_t mask = __builtin_amdgcn_ballot_w64(tok);
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & ((1 << A) | (1 << B))) continue;
    if (((uint32_t)r & rsel) != rsel) continue;
    if (tl == 0) {  // uniform
      ip_swap2(v[pm(pi, r | (1 << A))], v[pm(pi, r | (1 << B))]);
    } else {
      ip_cswap2(v[pm(pi, r | (1 << A))], v[pm(pi, r | (1 << B))], mask);
    }
  }
}

// A, B are the register slots of the lower and the higher local bit (basis
// order of the 4x4 matrix: index bit 0 = slot A).
template <int A, int B, bool F>
__device__ __forceinline__ void reg_mat4(double2 (&v)[kR], const double* u, uint64_t pi = kPermId) {
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & ((1 << A) | (1 << B))) continue;
    const int i0 = pm(pi, r), i1 = pm(pi, r | (1 << A)), i2 = pm(pi, r | (1 << B));
    const int i3 = pm(pi, r | (1 << A) | (1 << B));
    const double2 w0 = v[i0], w1 = v[i1], w2 = v[i2];
    const double2 w3 = v[i3];
    double2 o[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      double2 acc = mulc<F>(make_double2(u[8 * s], u[8 * s + 1]), w0);
      acc = macc<F>(acc, make_double2(u[8 * s + 2], u[8 * s + 3]), w1);
      acc = macc<F>(acc, make_double2(u[8 * s + 4], u[8 * s + 5]), w2);
      acc = macc<F>(acc, make_double2(u[8 * s + 6], u[8 * s + 7]), w3);
      o[s] = acc;
    }
    v[i0] = o[0];
    v[i1] = o[1];
    v[i2] = o[2];
    v[i3] = o[3];
  }
}

// ---- exchanges through LDS --------------------------------------------------
// One buffer of 2^(k-Q) amplitudes, indexed through swz (bits 4..7 of an
// index XORed into bits 0..3, which spreads the strided layouts over the
// banks without padding). Q = reg_xchg_split(k, batched) register slots (A, then B),
// whose bits are the same before and after the exchange, split it into 2^Q
// rounds: round c moves the amplitudes whose bits at those slots read c, and
// a buffer index is the local index without those bits. A round waits at a
// barrier (earlier reads of the buffer are done), writes, waits at a second
// barrier and reads back. (At k = 12 that is 4 barriers, as many as four
// double-buffered rounds of 1024 amplitudes took, while one slot fewer has
// to stay in place.) Every thread of the block calls these (act == false:
// barriers only).
struct Xchg {
  double2* lds;
};

// local index l without bit p
__device__ __forceinline__ uint32_t squeeze1(uint32_t l, uint32_t p) {
  return ((l >> (p + 1)) << p) | (l & ((1u << p) - 1u));
}

// l without the bits of the Q split slots (pa, then pb; Q <= 2)
template <int Q>
__device__ __forceinline__ uint32_t squeeze(uint32_t l, uint32_t pa, uint32_t pb) {
  if (Q == 0) return l;
  if (Q == 1) return squeeze1(l, pa);
  const uint32_t hi = pa > pb ? pa : pb, lo = pa > pb ? pb : pa;
  return squeeze1(squeeze1(l, hi), lo);
}

// 16-byte slots of a buffer: an index with bits 4..7 XORed into bits 0..3
// (a bijection of [0, 2^m) onto itself).
__device__ __forceinline__ uint32_t swz(uint32_t x) { return x ^ ((x >> 4) & 15u); }

// swz and squeeze are linear over XOR, and disjoint bit sets OR as they XOR,
// so a buffer index is a per-thread part XOR a per-register part (uniform,
// the XOR of per-slot terms; the split slots contribute nothing). Both parts
// are kept as byte offsets (index * 16), so an access address is one XOR.
template <int Q, int A, int B>
__device__ __forceinline__ void slot_offsets(const RegLayout& L, uint32_t pa, uint32_t pb,
                                             uint32_t (&o)[kRegBits]) {
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) {
    const bool split = (Q >= 1 && j == A) || (Q == 2 && j == B);
    o[j] = split ? 0u : swz(squeeze<Q>(L.rb[j], pa, pb)) << 4;
  }
}

template <int Q, int A, int B>
__device__ __forceinline__ uint32_t reg_offset(const uint32_t (&o)[kRegBits], int r) {
  uint32_t x = 0;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j)
    if (((r >> j) & 1) && !((Q >= 1 && j == A) || (Q == 2 && j == B))) x ^= o[j];
  return x;
}

template <int Q, int A, int B>
__device__ __forceinline__ constexpr int round_of(int r) {
  return Q == 0 ? 0 : Q == 1 ? ((r >> A) & 1) : (((r >> A) & 1) | (((r >> B) & 1) << 1));
}

// Register r of every thread moves from local index tw | regl(L, r) to
// tr | regl(N, r): tw / tr are the thread parts (register bits clear) on the
// writing and the reading side; ow / orr the slot offsets (slot_offsets) of
// the two layouts, pa / pb the positions of the slots A / B that key the
// rounds. Logical register r is physical register pm(pi, r) on both sides.
template <int Q, int A, int B>
__device__ __forceinline__ void xchg_off(double2 (&v)[kR], uint32_t pa, uint32_t pb,
                                         const uint32_t (&ow)[kRegBits],
                                         const uint32_t (&orr)[kRegBits], uint32_t tw,
                                         uint32_t tr, Xchg& X, bool act, uint64_t pi = kPermId) {
  const uint32_t bw = swz(squeeze<Q>(tw, pa, pb)) << 4, br = swz(squeeze<Q>(tr, pa, pb)) << 4;
  char* buf = reinterpret_cast<char*>(X.lds);
#pragma unroll
  for (int c = 0; c < (1 << Q); ++c) {
    __syncthreads();
    if (act) {
#pragma unroll
      for (int r = 0; r < kR; ++r)
        if (round_of<Q, A, B>(r) == c)
          *reinterpret_cast<double2*>(buf + (bw ^ reg_offset<Q, A, B>(ow, r))) = v[pm(pi, r)];
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int r = 0; r < kR; ++r)
        if (round_of<Q, A, B>(r) == c)
          v[pm(pi, r)] =
              *reinterpret_cast<const double2*>(buf + (br ^ reg_offset<Q, A, B>(orr, r)));
    }
  }
}

template <int Q, int A, int B>
__device__ __forceinline__ void xchg(double2 (&v)[kR], const RegLayout& L, const RegLayout& N,
                                     uint32_t tw, uint32_t tr, Xchg& X, bool act) {
  const uint32_t pa = __builtin_ctz(L.rb[A]), pb = __builtin_ctz(L.rb[B]);
  uint32_t ow[kRegBits], orr[kRegBits];
  slot_offsets<Q, A, B>(L, pa, pb, ow);
  slot_offsets<Q, A, B>(N, pa, pb, orr);
  xchg_off<Q, A, B>(v, pa, pb, ow, orr, tw, tr, X, act);
}

// xchg with the slot offsets the planner put in an exchange op (words w:
// kXchgOld, kXchgNew of its coefficients).
template <int Q, int A, int B>
__device__ __forceinline__ void xchg_op(double2 (&v)[kR], const RegLayout& L, const RegLayout& N,
                                        const uint32_t* w, Xchg& X, bool act,
                                        uint64_t pi = kPermId) {
  uint32_t ow[kRegBits], orr[kRegBits];
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) {
    ow[j] = w[kXchgOld + j];
    orr[j] = w[kXchgNew + j];
  }
  xchg_off<Q, A, B>(v, __builtin_ctz(L.rb[A]), __builtin_ctz(L.rb[B]), ow, orr, L.thr_l,
                    N.thr_l, X, act, pi);
}

// Exchange keyed by the first q of slots a < b (unchanged by it); 2^q rounds,
// q = reg_xchg_split(k, batched).
__device__ __forceinline__ void xchg_any(double2 (&v)[kR], const RegLayout& L, const RegLayout& N,
                                         uint32_t tw, uint32_t tr, uint32_t a, uint32_t b,
                                         uint32_t q, Xchg& X, bool act) {
  if (q == 0) {
    xchg<0, 0, 1>(v, L, N, tw, tr, X, act);
  } else if (q == 1) {
    switch (a) {  // B is not used for one split slot
      case 0: xchg<1, 0, 1>(v, L, N, tw, tr, X, act); break;
      case 1: xchg<1, 1, 2>(v, L, N, tw, tr, X, act); break;
      case 2: xchg<1, 2, 3>(v, L, N, tw, tr, X, act); break;
      default: xchg<1, 3, 0>(v, L, N, tw, tr, X, act); break;
    }
  } else {
    switch (a * 4 + b) {
      case 1: xchg<2, 0, 1>(v, L, N, tw, tr, X, act); break;
      case 2: xchg<2, 0, 2>(v, L, N, tw, tr, X, act); break;
      case 3: xchg<2, 0, 3>(v, L, N, tw, tr, X, act); break;
      case 6: xchg<2, 1, 2>(v, L, N, tw, tr, X, act); break;
      case 7: xchg<2, 1, 3>(v, L, N, tw, tr, X, act); break;
      default: xchg<2, 2, 3>(v, L, N, tw, tr, X, act); break;
    }
  }
}

// Whole-tile layout change L -> N (N's thread bits in ascending order) of an
// exchange op: a < b = slots holding the same bit in both, w = the op's
// coefficient words (the planner's slot offsets).
__device__ __forceinline__ void relayout(double2 (&v)[kR], const RegLayout& L, const RegLayout& N,
                                         uint32_t a, uint32_t b, uint32_t q, const uint32_t* w,
                                         Xchg& X, bool act) {
  if (q == 0) {
    xchg_op<0, 0, 1>(v, L, N, w, X, act);
  } else if (q == 1) {
    switch (a) {  // B is not used for one split slot
      case 0: xchg_op<1, 0, 1>(v, L, N, w, X, act); break;
      case 1: xchg_op<1, 1, 2>(v, L, N, w, X, act); break;
      case 2: xchg_op<1, 2, 3>(v, L, N, w, X, act); break;
      default: xchg_op<1, 3, 0>(v, L, N, w, X, act); break;
    }
  } else {
    switch (a * 4 + b) {
      case 1: xchg_op<2, 0, 1>(v, L, N, w, X, act); break;
      case 2: xchg_op<2, 0, 2>(v, L, N, w, X, act); break;
      case 3: xchg_op<2, 0, 3>(v, L, N, w, X, act); break;
      case 6: xchg_op<2, 1, 2>(v, L, N, w, X, act); break;
      case 7: xchg_op<2, 1, 3>(v, L, N, w, X, act); break;
      default: xchg_op<2, 2, 3>(v, L, N, w, X, act); break;
    }
  }
}

// ---- layout swaps ------------------------------------------------------------
// A layout op is a list of swaps (see the planner's TK_LAYOUT packing). Swap
// (slot j, thread index bit m): the local bits of register slot j and of
// thread bit m trade places. Per register pair x = v[r] (bit j of r clear),
// y = v[r | 2^j], a thread whose bit m is 0 keeps x and receives its
// partner's (thread ^ 2^m) x into y; a thread whose bit m is 1 keeps y and
// receives its partner's y into x. Half the amplitudes move.

// 32-bit halves of a double in place
struct U2 {
  uint32_t lo, hi;
};
__device__ __forceinline__ U2 split(double d) {
  const uint64_t u = __builtin_bit_cast(uint64_t, d);
  return {(uint32_t)u, (uint32_t)(u >> 32)};
}
__device__ __forceinline__ double join(U2 h) {
  return __builtin_bit_cast(double, ((uint64_t)h.hi << 32) | h.lo);
}

// Thread bits m < 4 within a wave: DPP-sourced selects. tm = lanes whose
// bit m is 1. y' = tm ? y : x[l + 2^m]; x' = tm ? y_old[l - 2^m] : x. Lanes
// whose DPP source lies outside their row are not written, and those are
// exactly lanes that keep their value.
#define QOA_DPP_SWAP(UP, DOWN)                                                              \
  asm("s_mov_b64 vcc, %[tm]\n\t"                                                            \
      "v_mov_b32 %[t0], %[y0]\n\t"                                                          \
      "v_mov_b32 %[t1], %[y1]\n\t"                                                          \
      "v_mov_b32 %[t2], %[y2]\n\t"                                                          \
      "v_mov_b32 %[t3], %[y3]\n\t"                                                          \
      "s_nop 1\n\t"                                                                         \
      "v_cndmask_b32_dpp %[y0], %[x0], %[y0], vcc " UP " row_mask:0xf bank_mask:0xf\n\t"    \
      "v_cndmask_b32_dpp %[y1], %[x1], %[y1], vcc " UP " row_mask:0xf bank_mask:0xf\n\t"    \
      "v_cndmask_b32_dpp %[y2], %[x2], %[y2], vcc " UP " row_mask:0xf bank_mask:0xf\n\t"    \
      "v_cndmask_b32_dpp %[y3], %[x3], %[y3], vcc " UP " row_mask:0xf bank_mask:0xf\n\t"    \
      "s_not_b64 vcc, vcc\n\t"                                                              \
      "s_nop 1\n\t"                                                                         \
      "v_cndmask_b32_dpp %[x0], %[t0], %[x0], vcc " DOWN " row_mask:0xf bank_mask:0xf\n\t"  \
      "v_cndmask_b32_dpp %[x1], %[t1], %[x1], vcc " DOWN " row_mask:0xf bank_mask:0xf\n\t"  \
      "v_cndmask_b32_dpp %[x2], %[t2], %[x2], vcc " DOWN " row_mask:0xf bank_mask:0xf\n\t"  \
      "v_cndmask_b32_dpp %[x3], %[t3], %[x3], vcc " DOWN " row_mask:0xf bank_mask:0xf"      \
      : [x0] "+v"(a0.lo), [x1] "+v"(a0.hi), [x2] "+v"(a1.lo), [x3] "+v"(a1.hi),             \
        [y0] "+v"(b0.lo), [y1] "+v"(b0.hi), [y2] "+v"(b1.lo), [y3] "+v"(b1.hi),             \
        [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3)                      \
      : [tm] "s"(tm)                                                                        \
      : "vcc")

template <int M>
__device__ __forceinline__ void lane_swap_pair(double2& x, double2& y) {
  U2 a0 = split(x.x), a1 = split(x.y), b0 = split(y.x), b1 = split(y.y);
  if constexpr (M >= 4) {  // rows 16 / halves 32 apart: x's upper part <-> y's lower part
    asm("s_nop 1\n\t"
        "v_permlane%c[w]_swap_b32 %[x0], %[y0]\n\t"
        "v_permlane%c[w]_swap_b32 %[x1], %[y1]\n\t"
        "v_permlane%c[w]_swap_b32 %[x2], %[y2]\n\t"
        "v_permlane%c[w]_swap_b32 %[x3], %[y3]"
        : [x0] "+v"(a0.lo), [x1] "+v"(a0.hi), [x2] "+v"(a1.lo), [x3] "+v"(a1.hi),
          [y0] "+v"(b0.lo), [y1] "+v"(b0.hi), [y2] "+v"(b1.lo), [y3] "+v"(b1.hi)
        : [w] "i"(M == 4 ? 16 : 32));
  } else {
    constexpr uint64_t kMask[4] = {0xAAAAAAAAAAAAAAAAull, 0xCCCCCCCCCCCCCCCCull,
                                   0xF0F0F0F0F0F0F0F0ull, 0xFF00FF00FF00FF00ull};
    const uint64_t tm = kMask[M];
    uint32_t t0, t1, t2, t3;
    if constexpr (M == 0) QOA_DPP_SWAP("quad_perm:[1,0,3,2]", "quad_perm:[1,0,3,2]");
    else if constexpr (M == 1) QOA_DPP_SWAP("quad_perm:[2,3,0,1]", "quad_perm:[2,3,0,1]");
    else if constexpr (M == 2) QOA_DPP_SWAP("row_shl:4", "row_shr:4");
    else QOA_DPP_SWAP("row_shl:8", "row_shr:8");
  }
  x = make_double2(join(a0), join(a1));
  y = make_double2(join(b0), join(b1));
}
#undef QOA_DPP_SWAP

template <int J, int M>
__device__ __forceinline__ void lane_swap(double2 (&v)[kR]) {
#pragma unroll
  for (int r = 0; r < kR; ++r)
    if (!(r & (1 << J))) lane_swap_pair<M>(v[r], v[r | (1 << J)]);
}

template <int J>
__device__ __forceinline__ void lane_swap_m(double2 (&v)[kR], uint32_t m) {
  switch (m) {
    case 0: lane_swap<J, 0>(v); break;
    case 1: lane_swap<J, 1>(v); break;
    case 2: lane_swap<J, 2>(v); break;
    case 3: lane_swap<J, 3>(v); break;
    case 4: lane_swap<J, 4>(v); break;
    default: lane_swap<J, 5>(v); break;
  }
}

// Thread bits m >= 6 (between waves; bit m is the same in a whole wave):
// through LDS, element pair * T + thread (T = threads of the tile), with
// barriers before (earlier reads of the area), between and after. Every
// wave passes the same three barriers; the branches between them are
// wave-uniform.
template <int J>
__device__ __forceinline__ void wave_swap(double2 (&v)[kR], uint32_t m, uint32_t T, double2* lds,
                                          bool act) {
  const uint32_t tid = threadIdx.x;
  const bool hi = __builtin_amdgcn_readfirstlane((tid >> m) & 1u) != 0u;
  const uint32_t src = tid ^ (1u << m);
  __syncthreads();
  if (act) {
    if (hi) {  // keeps y = v[r | 2^J], sends x
#pragma unroll
      for (int r = 0, p = 0; r < kR; ++r)
        if (!(r & (1 << J))) lds[(uint32_t)(p++) * T + tid] = v[r];
    } else {  // keeps x, sends y
#pragma unroll
      for (int r = 0, p = 0; r < kR; ++r)
        if (!(r & (1 << J))) lds[(uint32_t)(p++) * T + tid] = v[r | (1 << J)];
    }
  }
  __syncthreads();
  if (act) {
    if (hi) {
#pragma unroll
      for (int r = 0, p = 0; r < kR; ++r)
        if (!(r & (1 << J))) v[r] = lds[(uint32_t)(p++) * T + src];
    } else {
#pragma unroll
      for (int r = 0, p = 0; r < kR; ++r)
        if (!(r & (1 << J))) v[r | (1 << J)] = lds[(uint32_t)(p++) * T + src];
    }
  }
  __syncthreads();
}

// Register slots i <-> j: a renaming of the registers.
template <int I, int J>
__device__ __forceinline__ void slot_swap(double2 (&v)[kR]) {
#pragma unroll
  for (int r = 0; r < kR; ++r)
    if ((r & (1 << I)) && !(r & (1 << J))) ip_swap2(v[r], v[r ^ (1 << I) ^ (1 << J)]);
}

// Applies the swap list of a layout op; tbp packs the thread map (4 bits per
// thread index bit: its local bit).
__device__ __forceinline__ void apply_swaps(double2 (&v)[kR], RegLayout& L, uint64_t& tbp,
                                            uint32_t sw, int k, double2* lds, bool act) {
  const uint32_t n = sw >> 28;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t e = (sw >> (7 * i)) & 0x7Fu;
    const uint32_t j = e & 3u, opd = (e >> 2) & 15u;
    if (e & 64u) {  // slot j <-> slot opd
      const uint32_t lo = j < opd ? j : opd, hi = j < opd ? opd : j;
      switch (lo * 4 + hi) {
        case 1: slot_swap<1, 0>(v); break;
        case 2: slot_swap<2, 0>(v); break;
        case 3: slot_swap<3, 0>(v); break;
        case 6: slot_swap<2, 1>(v); break;
        case 7: slot_swap<3, 1>(v); break;
        default: slot_swap<3, 2>(v); break;
      }
      const uint32_t t = L.rb[j];
      L.rb[j] = L.rb[opd];
      L.rb[opd] = t;
      continue;
    }
    // slot j <-> thread bit m runs through slot 3: slots j and 3 trade their
    // registers first (slot j then holds slot 3's bit), then slot 3 and
    // thread bit m trade places
    const uint32_t m = opd;
    switch (j) {
      case 0: slot_swap<3, 0>(v); break;
      case 1: slot_swap<3, 1>(v); break;
      case 2: slot_swap<3, 2>(v); break;
      default: break;
    }
    const uint32_t t3 = L.rb[3];
    L.rb[3] = L.rb[j];
    L.rb[j] = t3;
    const uint32_t pb = __builtin_ctz(L.rb[3]);                // leaves the registers
    const uint32_t qb = (uint32_t)(tbp >> (4 * m)) & 15u;      // enters them
    if (m < 6) lane_swap_m<3>(v, m);
    else wave_swap<3>(v, m, 1u << (k - kRegBits), lds, act);
    L.thr_l = (L.thr_l & ~(1u << qb)) | (((tid >> m) & 1u) << pb);
    L.rb[3] = 1u << qb;
    tbp = (tbp & ~(15ull << (4 * m))) | ((uint64_t)pb << (4 * m));
  }
}

template <int J, bool Y>
__device__ __forceinline__ void reg_flip(double2 (&v)[kR]);

// X flips of the local bits in xl: register bits by exchanging registers,
// thread bits through LDS (thread t reads the amplitudes of thr_l ^ x).
// Trajectory noise runs in the batched kernels only: their exchange split.
__device__ __forceinline__ void flip_bits(double2 (&v)[kR], const RegLayout& L, uint32_t xl,
                                          int k, Xchg& X, bool act) {
  uint32_t rmask = 0;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) rmask |= L.rb[j];
  if (xl & L.rb[0]) reg_flip<0, false>(v);
  if (xl & L.rb[1]) reg_flip<1, false>(v);
  if (xl & L.rb[2]) reg_flip<2, false>(v);
  if (xl & L.rb[3]) reg_flip<3, false>(v);
  const uint32_t xt = xl & ~rmask;
  if (xt) xchg_any(v, L, L, L.thr_l, L.thr_l ^ xt, 2, 3, reg_xchg_split((uint32_t)k, true), X, act);
}

// Trajectory noise state of one tile pass (on == 0: no noise).
struct NoiseArgs {
  // the pass's trajectory draws: word o holds 2 bits per operand of op o
  // (0 = no error, 1..3 = X, Y, Z), drawn by traj_pauli before the passes
  const uint32_t* codes;
  uint32_t on;
  uint64_t frame;  // X frame entering the pass: logical index = physical ^ frame
};

// traj_pauli (traj_noise.hpp) draws the trajectory errors.

template <int J, bool Y>
__device__ __forceinline__ void reg_flip(double2 (&v)[kR]) {
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & (1 << J)) continue;
    if (Y) ip_y(v[r], v[r | (1 << J)]);
    else ip_swap2(v[r], v[r | (1 << J)]);
  }
}

// Pauli `code` (1 X, 2 Y, 3 Z) on local bit b of a register-resident tile,
// with the arithmetic of apply_{x,y,z}_simd (vectorization.rs:942-999).
__device__ __forceinline__ void reg_pauli(double2 (&v)[kR], uint32_t code, uint32_t b,
                                          const RegLayout& L, int k, Xchg& X, bool act) {
  if (code == 3) {
#pragma unroll
    for (int r = 0; r < kR; ++r) ip_scale(v[r], (((L.thr_l | regl(L, r)) >> b) & 1u) ? -1.0 : 1.0);
    return;
  }
  int slot = -1;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j)
    if (L.rb[j] == (1u << b)) slot = j;
  if (slot >= 0) {
    const bool y = code == 2;
    switch (slot) {
      case 0: if (y) reg_flip<0, true>(v); else reg_flip<0, false>(v); break;
      case 1: if (y) reg_flip<1, true>(v); else reg_flip<1, false>(v); break;
      case 2: if (y) reg_flip<2, true>(v); else reg_flip<2, false>(v); break;
      default: if (y) reg_flip<3, true>(v); else reg_flip<3, false>(v); break;
    }
    return;
  }
  // b is a thread bit: exchange partners through LDS
  flip_bits(v, L, 1u << b, k, X, act);
  if (code == 2) {
    const double2 f = make_double2(0.0, ((L.thr_l >> b) & 1u) ? 1.0 : -1.0);
#pragma unroll
    for (int r = 0; r < kR; ++r) ip_cmul(v[r].x, v[r].y, f.x, f.y);
  }
}

// A TileOp as its 28 32-bit words, held in scalar registers: uniform loads of
// whole dwords (a byte read would become a vector load and a wait), issued
// one op ahead so the op loop never waits on them.
// w0 = kind | t << 8 | t2 << 16 | nins << 24, w1 = ins[0..3],
// w2 = rsel | rt << 8 | outside << 16 | nq << 24, w3 = tl,
// w4..w5 = gmask, w6..w7 = gval, w8 = aux, w9 = lmask, w10 = gi,
// w11 = psel | qphys[0] << 8 | qphys[1] << 16 | qphys[2] << 24,
// w12..w27 = c[0..7] (the offsets below fix this layout).
constexpr int kOpWords = 28;
static_assert(sizeof(TileOp) == 4 * kOpWords && __builtin_offsetof(TileOp, ins) == 4 &&
                  __builtin_offsetof(TileOp, rsel) == 8 && __builtin_offsetof(TileOp, rt) == 9 &&
                  __builtin_offsetof(TileOp, nq) == 11 && __builtin_offsetof(TileOp, tl) == 12 &&
                  __builtin_offsetof(TileOp, gmask) == 16 && __builtin_offsetof(TileOp, gval) == 24 &&
                  __builtin_offsetof(TileOp, aux) == 32 && __builtin_offsetof(TileOp, gi) == 40 &&
                  __builtin_offsetof(TileOp, psel) == 44 && __builtin_offsetof(TileOp, qphys) == 45 &&
                  __builtin_offsetof(TileOp, c) == 48,
              "TileOp word layout");

struct OpRegs {
  uint32_t w[kOpWords];
  __device__ __forceinline__ uint32_t kind() const { return w[0] & 0xFFu; }
  __device__ __forceinline__ uint32_t aux() const { return w[8]; }
  __device__ __forceinline__ uint64_t gmask() const { return w[4] | (uint64_t)w[5] << 32; }
  __device__ __forceinline__ uint64_t gval() const { return w[6] | (uint64_t)w[7] << 32; }
  __device__ __forceinline__ double c(int i) const {
    return __builtin_bit_cast(double, w[12 + 2 * i] | (uint64_t)w[13 + 2 * i] << 32);
  }
  __device__ __forceinline__ double2 c2(int i) const { return make_double2(c(2 * i), c(2 * i + 1)); }
  __device__ __forceinline__ uint32_t gi() const { return w[10]; }
};

__device__ __forceinline__ OpRegs load_op(const TileOp* __restrict__ p) {
  const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
  OpRegs r;
#pragma unroll
  for (int i = 0; i < kOpWords; ++i) r.w[i] = q[i];
  return r;
}

// The words an op's dispatch reads (kind, slots, controls, tile condition:
// words 0..7) and the rest (the 4x4 pool offset or swap list, noise fields,
// coefficients), which only the work after dispatch reads.
constexpr int kCoefFirst = 8, kCoefEnd = kOpWords;
constexpr int kCoefWord = 12;  // c[0] as a 32-bit word of the op
__device__ __forceinline__ bool head_word(int i) { return i < kCoefFirst; }
__device__ __forceinline__ void load_head(OpRegs& r, const TileOp* __restrict__ p) {
  const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
  for (int i = 0; i < kOpWords; ++i)
    if (head_word(i)) r.w[i] = q[i];
}
__device__ __forceinline__ void load_coefs(OpRegs& r, const TileOp* __restrict__ p) {
  const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
  for (int i = kCoefFirst; i < kCoefEnd; ++i) r.w[i] = q[i];
}

// Applies one non-layout op whose tile condition holds. tok: this thread's
// control bits (tl) are all set; lanes without it keep their amplitudes
// (for diagonal ops through a factor of exactly 1). It is computed only on
// the paths with thread controls.
template <bool F>
__device__ __forceinline__ void reg_op(double2 (&v)[kR], const OpRegs& op, const RegLayout& L,
                                       const double* __restrict__ pool, uint64_t tkey) {
  const uint32_t kind = op.kind();
  const uint32_t rsel = op.w[2] & 0xFFu;
  const uint32_t tl = op.w[3];
  auto tok = [&]() { return (L.thr_l & tl) == tl; };
  const uint32_t s0 = op.w[1] & 0xFFu, s1 = (op.w[1] >> 8) & 0xFFu;
  switch (kind) {
    case TK_H: case TK_M: case TK_X: case TK_Y: {
      Coef4 k;
#pragma unroll
      for (int i = 0; i < 4; ++i) k.c[i] = op.c2(i);
      if (kind == TK_H) {
        if (op.c(1) == 0.0) {
          if (op.c(0) == 1.0) reg_pair_any<kHRaw, F>(v, k, s0, rsel, tl, L.thr_l);
          else reg_pair_any<kHReal, F>(v, k, s0, rsel, tl, L.thr_l);
        } else {
          reg_pair_any<TK_H, F>(v, k, s0, rsel, tl, L.thr_l);
        }
      } else if (kind == TK_M) {
        const uint32_t shape = (op.w[0] >> 16) & 0xFFu;  // t2
        if (shape == kM2Real) reg_pair_any<kMReal, F>(v, k, s0, rsel, tl, L.thr_l);
        else if (shape == kM2Rx) reg_pair_any<kMRx, F>(v, k, s0, rsel, tl, L.thr_l);
        else if (shape == kM2RealU) reg_pair_any<kMRealU, F>(v, k, s0, rsel, tl, L.thr_l);
        else if (shape == kM2RxU) reg_pair_any<kMRxU, F>(v, k, s0, rsel, tl, L.thr_l);
        else reg_pair_any<TK_M, F>(v, k, s0, rsel, tl, L.thr_l);
      } else if (kind == TK_X) {
        reg_pair_any<TK_X, F>(v, k, s0, rsel, tl, L.thr_l);
      } else {
        reg_pair_any<TK_Y, F>(v, k, s0, rsel, tl, L.thr_l);
      }
    } break;
    case TK_DNEG:
      if (tl == 0) reg_neg_rs(v, rsel);
      else reg_scale_rs(v, tok() ? -1.0 : 1.0, rsel);
      break;
    case TK_DPHASE:
      if (tl == 0) {  // no thread controls: one factor for the whole wave
        reg_phase_rs<F, true>(v, op.c2(0), rsel);
      } else {
        reg_phase_rs<F, false>(v, sel2(tok(), op.c2(0), make_double2(1.0, 0.0)), rsel);
      }
      break;
    case TK_RZ: {
      const double2 m00 = op.c2(0);
      const double2 m11 = op.c2(1);
      switch ((op.w[2] >> 8) & 0xFFu) {  // register slot of the target, 0xFF: a thread bit
        case 0: reg_rz<0, F>(v, m00, m11); break;
        case 1: reg_rz<1, F>(v, m00, m11); break;
        case 2: reg_rz<2, F>(v, m00, m11); break;
        case 3: reg_rz<3, F>(v, m00, m11); break;
        default: {
          const uint32_t t = (op.w[0] >> 8) & 0xFFu;
          const double2 f = sel2(((L.thr_l >> t) & 1u) != 0, m11, m00);
#pragma unroll
          for (int r = 0; r < kR; ++r) ip_mulc<F>(v[r], f);
        } break;
      }
    } break;
    case TK_SWAP: {
      const uint32_t a0 = s0 < s1 ? s0 : s1, b0 = s0 < s1 ? s1 : s0;
      switch (a0 * 4 + b0) {
        case 1: reg_swap<0, 1>(v, rsel, tl, tok()); break;
        case 2: reg_swap<0, 2>(v, rsel, tl, tok()); break;
        case 3: reg_swap<0, 3>(v, rsel, tl, tok()); break;
        case 6: reg_swap<1, 2>(v, rsel, tl, tok()); break;
        case 7: reg_swap<1, 3>(v, rsel, tl, tok()); break;
        default: reg_swap<2, 3>(v, rsel, tl, tok()); break;
      }
    } break;
    case TK_MAT4:  // the planner puts the lower local bit in slot 2, the higher in slot 3
      reg_mat4<2, 3, F>(v, pool + op.aux());
      break;
    case TK_DIAGN: {  // merged diagonal run: DiagTerm records, grouped by register slots
      // (tolerance mode: only the FMA instantiation carries this code, so the
      // exact kernels keep their register allocation). A term's words are
      // loaded while the term before it is applied (every term list ends
      // with a spare term), one scalar load wait per term.
      if constexpr (!F) break;
      const uint32_t* tw = reinterpret_cast<const uint32_t*>(pool + op.aux());
      const uint32_t nt = op.w[9];
      constexpr uint32_t W = kDiagTermDoubles * 2;  // words per term
      constexpr int kTW = 10;                       // words read: rs, tl, gmask, gval, c
      uint32_t cur[kTW], nxt[kTW];
#pragma unroll
      for (int j = 0; j < kTW; ++j) cur[j] = tw[j];
      double2 acc = make_double2(1.0, 0.0);
      for (uint32_t i = 0; i < nt; ++i) {
        const uint32_t* wn = tw + W * (i + 1);
#pragma unroll
        for (int j = 0; j < kTW; ++j) nxt[j] = wn[j];
        // a compiler memory barrier keeps the loads here (they would sink to
        // the end of the iteration, next to their first use)
        asm volatile("" ::: "memory");
        const uint64_t gm = cur[2] | (uint64_t)cur[3] << 32, gv = cur[4] | (uint64_t)cur[5] << 32;
        if ((tkey & gm) == gv) {  // uniform
          const double cr = __builtin_bit_cast(double, cur[6] | (uint64_t)cur[7] << 32);
          const double ci = __builtin_bit_cast(double, cur[8] | (uint64_t)cur[9] << 32);
          const uint32_t ttl = cur[1];
          if (ttl == 0) {
            ip_mulc<F, true>(acc, make_double2(cr, ci));
          } else {
            const bool on = (L.thr_l & ttl) == ttl;
            ip_mulc<F>(acc, make_double2(on ? cr : 1.0, on ? ci : 0.0));
          }
        }
        if (cur[0] & kDiagGroupEnd) {  // the group of slots cur[0] & 15 ends
          reg_phase_rs<F, false>(v, acc, cur[0] & 15u);
          acc = make_double2(1.0, 0.0);
        }
#pragma unroll
        for (int j = 0; j < kTW; ++j) cur[j] = nxt[j];
      }
    } break;
    default:
      break;
  }
}

// One register-resident pass over the tile with base `base` (bits of S clear).
// Threads with act == false (beyond 2^(k-4)) only take part in barriers.
// With trajectory noise, errors on qubits of the tile are applied in place;
// an X or Y on a qubit outside the tile flips that bit of the X frame
// (its phase, a tile constant, is applied at once), and the frame bits that
// fall inside the tile are applied when the tile is loaded. Returns the frame
// leaving the pass. NOISE = false compiles the noise handling out (nz.on
// must then be 0). Waves without active threads skip the arithmetic. SWAPS:
// the kernel takes layout ops that are swap lists (plans made with
// kPlanSwaps); without it every layout op is a whole-tile exchange.
template <typename I, bool F, bool NOISE, bool SCALAR_IDLE, bool SWAPS, typename G>
__device__ __forceinline__ uint64_t reg_tile(double2* __restrict__ a, I base, const G& geo,
                                             const TileOp* __restrict__ ops, int nops,
                                             const double* __restrict__ pool, double2* lds,
                                             bool act, bool init_zero, const NoiseArgs& nz) {
  const int k = geo.k;
  const uint32_t tid = threadIdx.x;
  uint8_t rpos[kRegBits];
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) rpos[j] = (uint8_t)(k - kRegBits + j);
  RegLayout L;
  set_layout(L, tid, k, rpos);  // thr_l = tid: thread bits are local bits 0..k-5
  uint64_t tbp = 0;             // thread index bit m stands for local bit (tbp >> 4m) & 15
  for (int m = 0; m < k - kRegBits; ++m) tbp |= (uint64_t)m << (4 * m);
  // HBM addresses: base | thread part | register part (uniform per r)
  I gthr = 0;
  for (int j = 0; j < k - kRegBits; ++j) gthr |= (I)((tid >> j) & 1u) << spos_at(geo, j);
  I gb[kRegBits];
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) gb[j] = (I)1 << spos_at(geo, k - kRegBits + j);
  auto gaddr = [&](int r) {
    I o = base | gthr;
#pragma unroll
    for (int j = 0; j < kRegBits; ++j)
      if ((r >> j) & 1) o |= gb[j];
    return o;
  };
  double2 v[kR];
  if (init_zero) {  // the state starts as |0...0>: nothing to read
#pragma unroll
    for (int r = 0; r < kR; ++r) v[r] = make_double2(gaddr(r) == 0 ? 1.0 : 0.0, 0.0);
  } else if (act) {
#pragma unroll
    for (int r = 0; r < kR; ++r) v[r] = ld_stream(a + gaddr(r));
  }
  // the tile's loads complete here, so the op loop carries no waits for them
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt, lgkmcnt not waited for
  Xchg X{lds};
  uint64_t frame = (NOISE && nz.on) ? nz.frame : 0ULL;
  // Threads are numbered wave by wave, so a wave is idle iff its first lane
  // is. SCALAR_IDLE: test that once per wave (a scalar branch; batched
  // launches, whose blocks can hold idle waves); else per lane (the fused
  // kernel, whose blocks have no idle waves: this form measured 8 % faster
  // there, through the code the compiler lays out around the op switch).
  const bool idle_l = (tid & ~63u) >= (1u << (k - kRegBits));
  const bool wave_act =
      SCALAR_IDLE ? __builtin_amdgcn_readfirstlane(idle_l ? 0u : 1u) != 0u : !idle_l;
  if (frame) {
    uint64_t smask = 0;
    uint32_t xl = 0;
    for (int j = 0; j < k; ++j) {
      smask |= 1ULL << spos_at(geo, j);
      if ((frame >> spos_at(geo, j)) & 1ULL) xl |= 1u << j;
    }
    if (xl) {  // frame bits inside the tile: apply those X flips now
      flip_bits(v, L, xl, k, X, act);
      frame &= ~smask;
    }
  }

  // The next op's dispatch words are loaded while the current op runs, and
  // the current op's other words at the top of its iteration (they are read
  // after dispatch): only the dispatch words pass from one iteration to the
  // next.
  // The op after the last one is read too (its words are not used): every
  // op buffer ends with one spare op.
  OpRegs op;
  if (nops > 0) {
    load_head(op, ops);
    // Consume the first op's words here: the wait for them then sits before
    // the loop.
    asm volatile("" ::"s"(op.w[0]));
  }
  const TileOp* const ops_end = ops + nops;
  for (const TileOp* p = ops; p != ops_end; ++p) {
    OpRegs next;
    load_head(next, p + 1);
    load_coefs(op, p);
    // the op's draws (read after its work; items without noise have no table)
    const uint32_t drawn = (NOISE && nz.on) ? nz.codes[p - ops] : 0u;
    if (op.kind() == TK_LAYOUT) {  // the new slot bits in ins
#pragma unroll
      for (int j = 0; j < kRegBits; ++j) rpos[j] = (uint8_t)(op.w[1] >> (8 * j));
      const uint32_t sw = op.aux();
      if (SWAPS && (sw >> 31)) {  // swap list
        apply_swaps(v, L, tbp, sw & 0x7FFFFFFFu, k, lds, act);
      } else {  // whole-tile exchange; two unchanged slots in t, t2
        RegLayout N;
        set_layout_sorted(N, tid, rpos, op.w[kCoefWord + kXchgSorted]);
        // batched kernels test idle waves per wave (SCALAR_IDLE): their split
        relayout(v, L, N, (op.w[0] >> 8) & 0xFFu, (op.w[0] >> 16) & 0xFFu,
                 reg_xchg_split((uint32_t)k, SCALAR_IDLE), op.w + kCoefWord, X, act);
        L = N;
        if (SWAPS) {
          uint32_t rmask = 0;
#pragma unroll
          for (int j = 0; j < kRegBits; ++j) rmask |= 1u << rpos[j];
          tbp = 0;
          for (int b = 0, m = 0; b < k; ++b)
            if (!(rmask & (1u << b))) tbp |= (uint64_t)b << (4 * m++);
        }
      }
    } else if (wave_act && (((uint64_t)base ^ frame) & op.gmask()) == op.gval()) {  // uniform
      reg_op<F>(v, op, L, pool, (uint64_t)base ^ frame);
    }
    const uint32_t nq = op.w[2] >> 24;
    if (NOISE && drawn) {  // trajectory errors after the gate, operand order
      for (uint32_t j = 0; j < nq; ++j) {
        const uint32_t q =
            j == 0 ? (op.w[11] >> 8) & 0xFFu : j == 1 ? (op.w[11] >> 16) & 0xFFu : op.w[11] >> 24;
        const uint32_t code = (drawn >> (2 * j)) & 3u;
        if (!code) continue;
        int lb = -1;
        for (int s = 0; s < k; ++s)
          if (spos_at(geo, s) == q) lb = s;
        if (lb >= 0) {
          reg_pauli(v, code, (uint32_t)lb, L, k, X, act);
        } else if (code == 3) {  // Z on an outside qubit: a tile-wide sign
          if ((((uint64_t)base ^ frame) >> q) & 1ULL) {
#pragma unroll
            for (int r = 0; r < kR; ++r) ip_scale(v[r], -1.0);
          }
        } else {
          frame ^= 1ULL << q;
          if (code == 2) {  // Y: phase -i / +i by the new logical bit
            const double2 f =
                make_double2(0.0, ((((uint64_t)base ^ frame) >> q) & 1ULL) ? 1.0 : -1.0);
#pragma unroll
            for (int r = 0; r < kR; ++r) ip_cmul(v[r].x, v[r].y, f.x, f.y);
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kOpWords; ++i)
      if (head_word(i)) op.w[i] = next.w[i];
  }
  if (act) {  // store from the final layout
    gthr = 0;
    for (int j = 0; j < k; ++j) gthr |= (I)((L.thr_l >> j) & 1u) << stpos_at(geo, j);
#pragma unroll
    for (int j = 0; j < kRegBits; ++j) gb[j] = (I)1 << stpos_at(geo, rpos[j]);
#pragma unroll
    for (int r = 0; r < kR; ++r) st_stream(a + gaddr(r), v[r]);
  }
  return frame;
}

}  // namespace tile
}  // namespace qoa
