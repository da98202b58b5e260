// planner.hpp — gate list -> LDS-tile passes (host side of the fused path).
#pragma once
#include <cstdint>
#if !defined(__HIPCC_RTC__)  // kernels compiled at run time see the op records only
#include <vector>
#endif

#include "../../include/qoa_hip.h"

namespace qoa {

// Tile-local op kinds (device op program).
enum TileKind : uint8_t {
  TK_H = 0,       // a' = nf*(a+b), b' = nf*(a-b)                      c[0..1] = nf
  TK_M = 1,       // a' = m00 a + m01 b, b' = m10 a + m11 b            c[0..7]; t2 = shape
  TK_X = 2,       // swap
  TK_Y = 3,       // a' = b*(-i), b' = a*(+i)
  TK_DNEG = 4,    // v = v.scale(-1) where (l & lmask) == lmask
  TK_DPHASE = 5,  // v *= c[0..1]   where (l & lmask) == lmask
  TK_RZ = 6,      // v *= bit t ? c[2..3] : c[0..1]
  TK_SWAP = 7,    // exchange (l|m2, l|m1) for l with both bits clear (and lmask set)
  TK_MAT4 = 8,    // 4x4 on local (t, t2) = (q1, q2); coefficients at pool[aux]
  TK_LAYOUT = 9,  // register passes: ins[0..3] = local bits of register slots 0..3;
                  // slots t < t2 keep their bit where the exchange needs it;
                  // a whole-tile exchange carries kXchgWords words in c
  TK_DIAGN = 10   // register passes (QOA_FUSE_MERGE_DIAG): a run of diagonal ops as
                  // lmask terms at pool[aux] (DiagTerm), grouped by register slots
};

// One term of a TK_DIAGN op: the factor c applies where the tile condition
// ((tile base ^ X frame) & gmask) == gval holds, the thread's local bits hold
// tl, and the register index holds the slots rs. Terms come grouped by rs
// (ascending); a group's factors are multiplied together (in term order)
// per thread, then into the registers whose index holds rs. 48 bytes = 6
// pool doubles. The last term of a group carries kDiagGroupEnd in rs; a
// TK_DIAGN op's terms are followed by one spare zero term.
constexpr uint32_t kDiagGroupEnd = 1u << 31;
struct DiagTerm {
  uint32_t rs, tl;
  uint64_t gmask, gval;
  double c[2];
  uint32_t pad[2];
};
static_assert(sizeof(DiagTerm) == 48, "DiagTerm layout");
constexpr uint32_t kDiagTermDoubles = sizeof(DiagTerm) / sizeof(double);

// A whole-tile exchange op's words in c (as 32-bit words): the new register
// bit positions in ascending order (a byte each), then the exchange buffer's
// byte offset of every register slot's bit under the old layout and under the
// new one (0 for the slots that key the rounds).
constexpr int kXchgSorted = 0, kXchgOld = 1, kXchgNew = 5, kXchgWords = 9;

// TK_M shapes (t2): general, real, or real diagonal with imaginary
// off-diagonal; the last two with m00 = m11 = 1 as kM2RealU / kM2RxU (the
// factored forms of RY / RX, QOA_FUSE_FACTOR), whose products by 1 are
// skipped.
constexpr uint8_t kM2General = 0, kM2Real = 1, kM2Rx = 2, kM2RealU = 3, kM2RxU = 4;

// Register-resident passes: each thread holds kRegAmps amplitudes of the tile
// whose local indices differ in kRegBits "register bits"; the other bits
// come from the thread index. In register mode ins[0] (and ins[1]) of a 2x2,
// SWAP or MAT4 op hold the register slot of its target bit(s); a MAT4 op's
// lower local bit is in slot 2 and its higher one in slot 3.
constexpr int kRegBits = 4;
constexpr int kRegAmps = 1 << kRegBits;
// Register passes exchange amplitudes through one LDS buffer of 2^(k-q)
// elements, q = reg_xchg_split(k, batched) (the exchange then takes 2^q
// rounds, and an exchange changes at most kRegBits - q register slots): the
// buffer holds at most 2048 elements (32 KiB). At k = 12 the fused kernel
// takes quarter-tile rounds (16 KiB): it holds 96 VGPRs, five waves per
// SIMD, and five 256-thread blocks fit a CU's LDS with 16 KiB buffers but
// not with 32 KiB ones (measured: four blocks, 3.6 % slower on the 28-qubit
// layered circuit despite fewer exchanges). The batched kernels hold more
// VGPRs (at most four waves per SIMD), so their occupancy does not depend on
// the buffer and they keep half-tile rounds, which change one more slot.
constexpr uint32_t reg_xchg_split(uint32_t k, bool batched) {
  return k <= 11 ? 0u : (k == 12 && batched) ? 1u : 2u;
}
constexpr uint32_t reg_xchg_elems(uint32_t k, bool batched) {
  return 1u << (k - reg_xchg_split(k, batched));
}
// PlanOptions.flags bit (internal): register mode for every tile of at least
// kRegBits qubits (the batched kernel runs small whole-state tiles this way);
// the default tile stays at 12 qubits for states of 13..17 qubits too.
constexpr uint32_t kPlanForceReg = 1u << 31;
// PlanOptions.flags bit (internal): attach trajectory-noise operands to ops.
constexpr uint32_t kPlanNoise = 1u << 30;
// PlanOptions.flags bit (internal): tiles of up to 10 qubits change their
// register layout by swaps (the batched kernel); else by whole-tile exchanges.
constexpr uint32_t kPlanSwaps = 1u << 29;

// One op of a pass program. `ins` holds the sorted local bit positions into
// which a zero is deposited to enumerate the visited indices (nins == 0xFF:
// too many bits, scan every index with the predicate instead). The first
// eight words are what the register kernels' op dispatch reads (they are
// prefetched for the next op); the rest is read after dispatch.
struct alignas(16) TileOp {
  uint8_t kind, t, t2, nins;
  uint8_t ins[4];
  // Register mode only (filled per layout by the segment scheduler): the
  // control bits of lmask split into register slots (rsel: slot j set means
  // local bit ins-of-layout[j] must be 1) and thread bits (tl = lmask & ~R);
  // rt is the register slot of a TK_RZ target, 0xFF when it is a thread bit.
  uint8_t rsel, rt;
  uint8_t outside;  // noise: the op reads or hits a qubit outside the tile
  uint8_t nq;       // noise: operand qubits (below)
  uint32_t tl;
  uint64_t gmask, gval;  // op applies to a tile iff (tile_base & gmask) == gval
  uint32_t aux;
  uint32_t lmask;
  // Trajectory noise after this op (evaluated in every tile, whether or not
  // the op's condition held there): for each of the nq operand qubits
  // qphys[j] a Philox draw keyed by (seed, trajectory, gi, qphys[j]) picks
  // X/Y/Z with probability p1 (psel == 1) or p2 (psel == 2).
  uint32_t gi;
  uint8_t psel;
  uint8_t qphys[3];
  double c[8];
};
static_assert(sizeof(TileOp) == 112, "TileOp layout");

#if !defined(__HIPCC_RTC__)
struct PassPlan {
  uint32_t k;                 // tile qubits
  uint32_t reg_mode;          // 1: register-resident kernel, 0: LDS-resident kernel
  uint8_t spos[16];           // physical position local bit j is loaded from
  uint8_t spos_st[16];        // and stored to (spos unless the pass relabels qubits)
  uint32_t op_begin, op_end;  // range in the flat op array
};

// A pass's geometry words (the kernels' Geo): load positions, store
// positions, and the positions in ascending order, four bytes to a word.
inline void pass_geo_words(const PassPlan& pp, uint32_t spw[4], uint32_t stw[4],
                           uint32_t srt[4]) {
  uint8_t s[16] = {};
  for (uint32_t j = 0; j < pp.k && j < 16; ++j) s[j] = pp.spos[j];
  for (uint32_t i = 1; i < pp.k && i < 16; ++i)  // insertion sort of k bytes
    for (uint32_t j = i; j > 0 && s[j - 1] > s[j]; --j) {
      const uint8_t t = s[j];
      s[j] = s[j - 1];
      s[j - 1] = t;
    }
  for (int w = 0; w < 4; ++w) {
    spw[w] = stw[w] = srt[w] = 0;
    for (int b = 0; b < 4; ++b) {
      spw[w] |= (uint32_t)pp.spos[4 * w + b] << (8 * b);
      stw[w] |= (uint32_t)pp.spos_st[4 * w + b] << (8 * b);
      srt[w] |= (uint32_t)s[4 * w + b] << (8 * b);
    }
  }
}

struct FusedPlan {
  uint32_t n = 0, k = 0;
  std::vector<PassPlan> passes;
  std::vector<TileOp> ops;
  std::vector<double> pool;     // MAT4 coefficients
  std::vector<uint32_t> order;  // executed order of gate indices
};

struct PlanOptions;
int build_plan(uint32_t n, const qoa_gate* g, uint32_t ng, const double* mats,
               const PlanOptions& opt, FusedPlan* plan);
#endif

}  // namespace qoa
