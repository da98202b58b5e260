// planner.cpp — splits a gate list into LDS-tile passes.
//
// A pass owns a set S of k qubits (always the low m qubits, for coalesced
// HBM runs of 2^m amplitudes, plus k-m qubits chosen per pass). Each
// workgroup loads the 2^k amplitudes that share one assignment of the other
// n-k bits, applies the pass's gates in LDS and writes them back, so every
// pass costs one read and one write of the state.
//
// A gate can run in a pass when every qubit it mixes (targets of non-diagonal
// actions) is in S and it commutes with every earlier gate left for a later
// pass (they share no qubit that either of them mixes). Qubits a gate only reads (controls, diagonal phases) may
// lie outside S: their bit is constant inside a tile, so the gate becomes a
// per-tile condition (CNOT with an outside control is "X or nothing"; a phase
// on an outside qubit is a tile-wide scalar). Gates keep their original
// relative order inside a pass; gates moved across passes act on disjoint
// qubits, so the planned order is a valid reordering of the circuit.
#include "planner.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "gates.hpp"
#include "qoa_internal.hpp"

namespace qoa {
namespace {

// Complex64::from_polar(1.0, th), cos/sin from one sincos call (as api.cpp)
inline void polar(double th, double* out) {
  double c, s;
  ::sincos(th, &s, &c);
  out[0] = 1.0 * c;
  out[1] = 1.0 * s;
}

inline uint64_t bit(int q) { return 1ULL << q; }

// Turns one gate into at most one fused op with the layer-1 semantics of the
// per-gate entry points (out-of-range qubits, degenerate controls).
int canonical(uint32_t n, const qoa_gate& g, const double* mats, uint32_t flags,
              std::vector<FOp>* out, uint32_t gi) {
  FOp f;
  f.gi = gi;
  auto in = [n](uint32_t q) { return q < n; };
  const uint32_t q0 = g.q0, q1 = g.q1, q2 = g.q2;
  switch (g.op) {
    case QOA_OP_H:
      if (!in(q0)) return 0;
      f.kind = FK_H; f.t = q0; f.c[0] = 1.0 / M_SQRT2; f.c[1] = 0.0;
      break;
    case QOA_OP_X:
      if (!in(q0)) return 0;
      f.kind = FK_X; f.t = q0;
      break;
    case QOA_OP_Y:
      if (!in(q0)) return 0;
      f.kind = FK_Y; f.t = q0;
      break;
    case QOA_OP_Z:
      if (!in(q0)) return 0;
      f.kind = FK_DNEG; f.pmask = bit(q0);
      break;
    case QOA_OP_S: case QOA_OP_T: case QOA_OP_PHASESHIFT: case QOA_OP_PHASE:
      if (!in(q0)) return 0;
      f.kind = FK_DPHASE; f.pmask = bit(q0);
      polar(g.op == QOA_OP_S ? M_PI / 2.0 : g.op == QOA_OP_T ? M_PI / 4.0 : g.theta, f.c);
      break;
    case QOA_OP_RX: case QOA_OP_RY: {
      if (!in(q0)) return 0;
      double c, s;
      ::sincos(g.theta / 2.0, &s, &c);
      f.kind = FK_M; f.t = q0;
      if (g.op == QOA_OP_RX) {
        const double m[8] = {c, 0.0, 0.0, -s, 0.0, -s, c, 0.0};
        std::copy(m, m + 8, f.c);
      } else {
        const double m[8] = {c, 0.0, -s, 0.0, s, 0.0, c, 0.0};
        std::copy(m, m + 8, f.c);
      }
      break;
    }
    case QOA_OP_RZ:
      polar(-g.theta / 2.0, f.c);
      polar(g.theta / 2.0, f.c + 2);
      if (in(q0)) {
        f.kind = FK_RZ; f.rzq = q0;
      } else {  // every index has the bit clear: global factor m00
        f.kind = FK_DPHASE; f.pmask = 0;
      }
      break;
    case QOA_OP_U2:
      if (!mats) return set_error(QOA_ERR_ARG, "U2 gate without a matrix table");
      if (!in(q0)) return 0;
      f.kind = FK_M; f.t = q0;
      std::copy(mats + 8 * g.aux, mats + 8 * g.aux + 8, f.c);
      break;
    case QOA_OP_CNOT:
      if (!in(q0) || !in(q1) || q0 == q1) return 0;
      f.kind = FK_X; f.t = q1; f.pmask = bit(q0);
      break;
    case QOA_OP_CY:
      if (!in(q0) || !in(q1)) return set_error(QOA_ERR_QUBIT, "controlled-y qubit out of bounds");
      if (q0 == q1) return set_error(QOA_ERR_ARG, "control and target qubits cannot be the same");
      f.kind = FK_Y; f.t = q1; f.pmask = bit(q0);
      break;
    case QOA_OP_TOFFOLI:
      if (!in(q0) || !in(q1) || !in(q2)) return set_error(QOA_ERR_QUBIT, "toffoli qubit out of bounds");
      if (q0 == q1 || q0 == q2 || q1 == q2)
        return set_error(QOA_ERR_ARG, "control and target qubits must be distinct for toffoli gate");
      f.kind = FK_X; f.t = q2; f.pmask = bit(q0) | bit(q1);
      break;
    case QOA_OP_CZ: case QOA_OP_CPHASE: {
      uint64_t m;
      if (flags & QOA_CZ_REF_MASKS) {
        m = (uint64_t)q0 | (uint64_t)q1;
      } else {
        if (q0 >= 64 || q1 >= 64) return set_error(QOA_ERR_ARG, "qubit index >= 64");
        m = bit(q0) | bit(q1);
      }
      if (n < 64 && (m >> n)) return 0;  // (i & m) == m never holds
      f.pmask = m;
      if (g.op == QOA_OP_CZ) f.kind = FK_DNEG;
      else { f.kind = FK_DPHASE; polar(g.theta, f.c); }
      break;
    }
    case QOA_OP_SWAP:
      if (!in(q0) || !in(q1) || q0 == q1) return 0;
      f.kind = FK_SWAP; f.t = q0; f.t2 = q1;
      break;
    case QOA_OP_CSWAP:
      if (!in(q0) || !in(q1) || !in(q2) || q1 == q2) return 0;
      if (q0 == q1) { f.kind = FK_X; f.t = q2; f.pmask = bit(q0); }
      else if (q0 == q2) { f.kind = FK_X; f.t = q1; f.pmask = bit(q0); }
      else { f.kind = FK_SWAP; f.t = q1; f.t2 = q2; f.pmask = bit(q0); }
      break;
    case QOA_OP_U4:
      if (!mats) return set_error(QOA_ERR_ARG, "U4 gate without a matrix table");
      if (!in(q0) || !in(q1)) return set_error(QOA_ERR_QUBIT, "two-qubit unitary qubit out of bounds");
      if (q0 == q1) return set_error(QOA_ERR_ARG, "qubits cannot be the same for two-qubit unitary");
      f.kind = FK_MAT4; f.t = q0; f.t2 = q1; f.m4 = mats + 32 * g.aux;
      break;
    default:
      return set_error(QOA_ERR_ARG, "op 0x%02x cannot run in a fused circuit", g.op);
  }
  f.active = (f.t >= 0 ? bit(f.t) : 0) | (f.t2 >= 0 ? bit(f.t2) : 0);
  f.all = f.active | f.pmask | (f.rzq >= 0 ? bit(f.rzq) : 0);
  f.mixb = f.active;
  // noise operands in gate order: one-qubit gates use p1, the others p2
  switch (g.op) {
    case QOA_OP_CNOT: case QOA_OP_CZ: case QOA_OP_CPHASE: case QOA_OP_SWAP:
    case QOA_OP_U4: case QOA_OP_CY: f.nq = 2; break;
    case QOA_OP_CSWAP: case QOA_OP_TOFFOLI: f.nq = 3; break;
    default: f.nq = 1; break;
  }
  f.psel = f.nq == 1 ? 1 : 2;
  f.qphys[0] = (uint8_t)q0;
  f.qphys[1] = (uint8_t)q1;
  f.qphys[2] = (uint8_t)q2;
  out->push_back(f);
  return 0;
}

// Gates left for later: an action may run ahead of them when its mixed
// qubits avoid every qubit they touch and its diagonal qubits avoid the
// qubits they mix (then it commutes with each of them).
template <typename M>
struct BlockerT {
  M any = 0, mix = 0;
  bool free(M all, M mixb) const { return (mixb & any) == 0 && ((all & ~mixb) & mix) == 0; }
  void add(M all, M mixb) {
    any |= all;
    mix |= mixb;
  }
};
using Blocker = BlockerT<uint64_t>;

// Number of ops of `rem` that a pass over S would run.
int simulate(const std::vector<FOp>& ops, const std::vector<uint32_t>& rem, uint64_t S,
             uint64_t allq) {
  Blocker b;
  int cnt = 0;
  for (uint32_t i : rem) {
    const FOp& f = ops[i];
    if (b.free(f.all, f.mixb) && (f.active & ~S) == 0) {
      ++cnt;
    } else {
      b.add(f.all, f.mixb);
      if ((b.mix & allq) == allq) break;
    }
  }
  return cnt;
}

inline int loc_of(const uint8_t* spos, int k, int q) {
  for (int j = 0; j < k; ++j)
    if (spos[j] == q) return j;
  return -1;
}

uint32_t local_bits(uint64_t mask, const uint8_t* spos, int k) {
  uint32_t r = 0;
  for (int j = 0; j < k; ++j)
    if (mask & bit(spos[j])) r |= 1u << j;
  return r;
}

void set_ins(TileOp* op, uint32_t bits) {
  int cnt = __builtin_popcount(bits);
  if (cnt > 4) {
    op->nins = 0xFF;
    return;
  }
  op->nins = 0;
  for (int j = 0; j < 32; ++j)
    if (bits & (1u << j)) op->ins[op->nins++] = (uint8_t)j;
}

// Shape of a 2x2 matrix c[0..7] = m00, m01, m10, m11 (re, im) for TK_M:
// kM2Real (every imaginary part 0), kM2Rx (m00, m11 real, m01, m10
// imaginary), else kM2General. The register kernels skip the products by
// those zeros: the sums they leave are the products' sums exactly, up to the
// sign of a zero result.
uint8_t m2_shape(const double* c) {
  if (c[1] == 0.0 && c[3] == 0.0 && c[5] == 0.0 && c[7] == 0.0) return kM2Real;
  if (c[1] == 0.0 && c[7] == 0.0 && c[2] == 0.0 && c[4] == 0.0) return kM2Rx;
  return kM2General;
}

// Physical positions of the logical qubits in mask m (pos: logical qubit ->
// physical position; nullptr: the identity).
inline uint64_t phys_mask(uint64_t m, const uint8_t* pos) {
  if (!pos) return m;
  uint64_t r = 0;
  for (; m; m &= m - 1) r |= bit(pos[__builtin_ctzll(m)]);
  return r;
}

// spos: the tile's logical qubits in local bit order; tile conditions on
// qubits outside the tile are physical masks (through pos).
void emit(const FOp& f, uint64_t S, const uint8_t* spos, int k, std::vector<TileOp>* out,
          std::vector<double>* pool, bool noise, const uint8_t* pos = nullptr) {
  TileOp op{};
  op.gi = f.gi;
  if (noise) {
    op.nq = f.nq;
    op.psel = f.psel;
    bool outside = (f.pmask & ~S) != 0 || (f.rzq >= 0 && !(S & bit(f.rzq)));
    for (int j = 0; j < 3; ++j) {
      op.qphys[j] = f.qphys[j];
      if (j < f.nq && !(S & bit(f.qphys[j]))) outside = true;
    }
    // with noise, an X error on an outside qubit changes the tile's frame, so
    // ops that read or hit outside qubits keep their relative order
    op.outside = outside ? 1 : 0;
  }
  const uint64_t pin = f.pmask & S, pout = f.pmask & ~S;
  op.lmask = local_bits(pin, spos, k);
  op.gmask = phys_mask(pout, pos);
  op.gval = op.gmask;
  std::copy(f.c, f.c + 8, op.c);
  switch (f.kind) {
    case FK_H: case FK_M: case FK_X: case FK_Y:
      op.kind = f.kind == FK_H ? TK_H : f.kind == FK_M ? TK_M : f.kind == FK_X ? TK_X : TK_Y;
      op.t = (uint8_t)loc_of(spos, k, f.t);
      if (op.kind == TK_M) op.t2 = m2_shape(op.c);
      set_ins(&op, op.lmask | (1u << op.t));
      break;
    case FK_DNEG: case FK_DPHASE:
      op.kind = f.kind == FK_DNEG ? TK_DNEG : TK_DPHASE;
      set_ins(&op, op.lmask);
      break;
    case FK_RZ:
      if (S & bit(f.rzq)) {
        op.kind = TK_RZ;
        op.t = (uint8_t)loc_of(spos, k, f.rzq);
        op.nins = 0;
      } else {  // bit constant per tile: two complementary tile-wide factors
        TileOp o0 = op, o1 = op;
        o0.kind = o1.kind = TK_DPHASE;
        o0.lmask = o1.lmask = 0;
        o0.nins = o1.nins = 0;
        o0.gmask = o1.gmask = phys_mask(bit(f.rzq), pos);
        o0.gval = 0;
        o1.gval = o0.gmask;
        o1.c[0] = f.c[2];
        o1.c[1] = f.c[3];
        o0.nq = 0;  // the gate's noise is drawn once, after the second op
        out->push_back(o0);
        out->push_back(o1);
        return;
      }
      break;
    case FK_SWAP:
      op.kind = TK_SWAP;
      op.t = (uint8_t)loc_of(spos, k, f.t);
      op.t2 = (uint8_t)loc_of(spos, k, f.t2);
      set_ins(&op, op.lmask | (1u << op.t) | (1u << op.t2));
      break;
    case FK_MAT4:
      op.kind = TK_MAT4;
      op.t = (uint8_t)loc_of(spos, k, f.t);
      op.t2 = (uint8_t)loc_of(spos, k, f.t2);
      op.aux = (uint32_t)pool->size();
      pool->insert(pool->end(), f.m4, f.m4 + 32);
      set_ins(&op, (1u << op.t) | (1u << op.t2));
      break;
  }
  out->push_back(op);
}

// ---- register segments inside a pass ------------------------------------------
// Each thread of the register-resident kernel holds the kRegAmps amplitudes
// whose local indices differ in the register bits R. Ops that mix only bits
// of R (and every diagonal op) run in registers; moving to a new R costs
// TK_LAYOUT ops (LDS exchanges in 2^reg_xchg_split(k, batched) rounds, or swap
// lists). The same greedy list scheduling as the pass planner picks R: ops
// on disjoint bits commute, so the emitted order stays a valid reordering. The pass starts in the layout that holds the top
// kRegBits local bits in registers (coalesced HBM access) and ends in one
// whose lowest local bits are thread bits.
struct SegOp {
  uint32_t act, all, mixb;
  uint32_t ctl;  // controls of a 2x2, SWAP or diagonal op (cheaper as register bits)
};

inline SegOp seg_masks_local(const TileOp& op) {
  switch (op.kind) {
    case TK_H: case TK_M: case TK_X: case TK_Y:
      return {1u << op.t, (1u << op.t) | op.lmask, 1u << op.t, op.lmask};
    case TK_SWAP: case TK_MAT4: {
      const uint32_t a = (1u << op.t) | (1u << op.t2);
      return {a, a | op.lmask, a, op.kind == TK_SWAP ? op.lmask : 0u};
    }
    case TK_RZ: return {0u, 1u << op.t, 0u, 0u};
    default: return {0u, op.lmask, 0u, op.lmask};
  }
}

inline SegOp seg_masks(const TileOp& op) {
  SegOp m = seg_masks_local(op);
  if (op.outside) m.all |= 1u << 31;  // ordered against other outside-qubit ops
  if (op.nq || op.outside) m.mixb = m.all;  // trajectory noise follows: keep the order
  return m;
}

int seg_sim(const std::vector<SegOp>& m, const std::vector<uint32_t>& rem, uint32_t R,
            uint32_t allb) {
  BlockerT<uint32_t> b;
  int cnt = 0;
  for (uint32_t i : rem) {
    if (b.free(m[i].all, m[i].mixb) && (m[i].act & ~R) == 0) {
      ++cnt;
    } else {
      b.add(m[i].all, m[i].mixb);
      if ((b.mix & allb) == allb) break;
    }
  }
  return cnt;
}

// Ops of a sweep over R whose controls all lie in R (they then need no
// per-lane selects).
int seg_bonus(const std::vector<SegOp>& m, const std::vector<uint32_t>& rem, uint32_t R,
              uint32_t allb) {
  BlockerT<uint32_t> b;
  int cnt = 0;
  for (uint32_t i : rem) {
    if (b.free(m[i].all, m[i].mixb) && (m[i].act & ~R) == 0) {
      if (m[i].ctl && (m[i].ctl & ~R) == 0) ++cnt;
    } else {
      b.add(m[i].all, m[i].mixb);
      if ((b.mix & allb) == allb) break;
    }
  }
  return cnt;
}

// Greedy choice of the next register set: starting from `seed`, add the bit
// (or pair of bits) that lets the most ops run; among bits that tie, the one
// that puts the most controls of those ops in registers. The set keeps at least `keep`
// bits of the current set `cur` (a layout change then takes one exchange).
uint32_t seg_choose(const std::vector<SegOp>& m, const std::vector<uint32_t>& rem, uint32_t seed,
                    uint32_t k, uint32_t allb, uint32_t cur, int keep) {
  uint32_t R = seed;
  int score = seg_sim(m, rem, R, allb);
  auto room = [&](uint32_t X) {  // X can still be completed with enough bits of cur
    const int have = __builtin_popcount(X & cur);
    return have >= keep || kRegBits - __builtin_popcount(X) >= keep - have;
  };
  while (__builtin_popcount(R) < kRegBits) {
    int best = -1, bests = -1, bestb = -1;
    for (uint32_t b = 0; b < k; ++b) {
      if ((R & (1u << b)) || !room(R | (1u << b))) continue;
      const int sc = seg_sim(m, rem, R | (1u << b), allb);
      if (sc < bests) continue;
      const int bo = seg_bonus(m, rem, R | (1u << b), allb);
      if (sc > bests || bo > bestb) { bests = sc; bestb = bo; best = (int)b; }
    }
    if (bests <= score && __builtin_popcount(R) + 2 <= kRegBits) {
      int b1 = -1, b2 = -1, bp = score;
      for (uint32_t b = 0; b < k; ++b) {
        if (R & (1u << b)) continue;
        for (uint32_t c = b + 1; c < k; ++c) {
          if ((R & (1u << c)) || !room(R | (1u << b) | (1u << c))) continue;
          const int sc = seg_sim(m, rem, R | (1u << b) | (1u << c), allb);
          if (sc > bp) { bp = sc; b1 = (int)b; b2 = (int)c; }
        }
      }
      if (b1 >= 0) {
        R |= (1u << b1) | (1u << b2);
        score = bp;
        continue;
      }
    }
    R |= 1u << best;
    score = bests;
  }
  return R;
}

// Register layout of a segment: slot j holds local bit s[j]; set = the bits.
// Thread index bit m stands for local bit tb[m] (nt = k - kRegBits of them).
struct RegSlots {
  uint8_t s[kRegBits];
  uint32_t set;
  uint8_t tb[16];
  uint32_t nt;
};

inline uint8_t slot_of(const RegSlots& L, uint32_t bitpos) {
  for (int j = 0; j < kRegBits; ++j)
    if (L.s[j] == bitpos) return (uint8_t)j;
  return 0xFF;
}

inline int tpos_of(const RegSlots& L, uint32_t bitpos) {
  for (uint32_t m = 0; m < L.nt; ++m)
    if (L.tb[m] == bitpos) return (int)m;
  return -1;
}

// A TK_LAYOUT op is either a whole-tile exchange (aux bit 31 clear: ins =
// the new slot bits, t < t2 = slots that keep their bit, the thread index
// bits then take the other local bits in ascending order) or a list of up to
// four swaps applied in order (aux bit 31 set): bits 28..30 = count, swap i
// in bits 7i..7i+6 = slot j (2 bits), operand (4 bits: a thread index bit m,
// or a slot), kind (1 bit: 0 = slot j <-> thread bit m, 1 = slot j <->
// slot); ins = the slot bits after the op.
struct SwapList {
  uint32_t packed = 0, n = 0;
  void add(uint32_t j, uint32_t operand, uint32_t kind) {
    packed |= (j | operand << 2 | kind << 6) << (7 * n);
    ++n;
  }
};

void emit_swaps(const RegSlots& after, SwapList* sl, std::vector<TileOp>* out) {
  if (sl->n == 0) return;
  TileOp op{};
  op.kind = TK_LAYOUT;
  for (int j = 0; j < kRegBits; ++j) op.ins[j] = after.s[j];
  op.aux = sl->packed | sl->n << 28 | 1u << 31;
  out->push_back(op);
  *sl = SwapList{};
}

// Slot j <-> thread index bit m, run through slot 3: slot j takes slot 3's
// bit, slot 3 takes thread bit m's, thread bit m takes slot j's.
void swap_rt(RegSlots* L, uint32_t j, uint32_t m, SwapList* sl, std::vector<TileOp>* out) {
  if (sl->n == 4) emit_swaps(*L, sl, out);
  const uint8_t b = L->s[j];
  L->set = (L->set & ~(1u << b)) | (1u << L->tb[m]);
  L->s[j] = L->s[3];
  L->s[3] = L->tb[m];
  L->tb[m] = b;
  sl->add(j, m, 0);
}

// Slot i <-> slot j.
void swap_rr(RegSlots* L, uint32_t i, uint32_t j, SwapList* sl, std::vector<TileOp>* out) {
  if (sl->n == 4) emit_swaps(*L, sl, out);
  std::swap(L->s[i], L->s[j]);
  sl->add(i, j, 1);
}

// Layout moves by swaps for tiles of up to 10 qubits when the kernel has
// them (the thread index bits are then lane bits: DPP and v_permlane swaps),
// else by whole-tile exchanges (measured faster for 12-qubit tiles: a swap
// there costs 52-155 us at 28 qubits against 124 us for a whole exchange).
inline bool use_swaps(uint32_t k, bool swaps) { return swaps && k <= 10; }

// Exchange buffer byte offset of local index l (the kernels' squeeze and
// swz): l without the bits at the q split positions pa (and pb), bits 4..7
// XORed into bits 0..3, times 16 bytes.
uint32_t xchg_offset(uint32_t l, uint32_t q, uint32_t pa, uint32_t pb) {
  auto squeeze1 = [](uint32_t x, uint32_t p) { return ((x >> (p + 1)) << p) | (x & ((1u << p) - 1u)); };
  if (q == 1) l = squeeze1(l, pa);
  if (q == 2) l = squeeze1(squeeze1(l, pa > pb ? pa : pb), pa > pb ? pb : pa);
  return (l ^ ((l >> 4) & 15u)) << 4;
}

// Whole-tile exchange op from layout `from` to `to`: ins = the new slot
// bits, t < t2 = slots whose bit is the same in both (the exchange of a 2^k
// tile runs in 2^q rounds keyed by the first q of those slots' bits,
// q = reg_xchg_split(k, batched)). The coefficient words carry what the
// kernel would otherwise derive per exchange (kXchgSorted, kXchgOld,
// kXchgNew): the new register bit positions in ascending order and every
// slot's buffer offset under both layouts.
TileOp exchange_op(const RegSlots& from, const RegSlots& to, uint32_t q) {
  TileOp op{};
  op.kind = TK_LAYOUT;
  int kept[kRegBits], nk = 0;
  for (int j = 0; j < kRegBits; ++j) {
    op.ins[j] = to.s[j];
    if (from.s[j] == to.s[j]) kept[nk++] = j;
  }
  op.t = (uint8_t)(nk > 0 ? kept[0] : 0);
  op.t2 = (uint8_t)(nk > 1 ? kept[1] : op.t + 1);
  uint32_t w[kXchgWords] = {};
  uint8_t sorted[kRegBits];
  for (int j = 0; j < kRegBits; ++j) sorted[j] = to.s[j];
  std::sort(sorted, sorted + kRegBits);
  for (int j = 0; j < kRegBits; ++j) w[kXchgSorted] |= (uint32_t)sorted[j] << (8 * j);
  const uint32_t pa = from.s[op.t], pb = from.s[op.t2 < kRegBits ? op.t2 : op.t];
  for (int j = 0; j < kRegBits; ++j) {
    const bool split = (q >= 1 && j == op.t) || (q == 2 && j == op.t2);
    w[kXchgOld + j] = split ? 0u : xchg_offset(1u << from.s[j], q, pa, pb);
    w[kXchgNew + j] = split ? 0u : xchg_offset(1u << to.s[j], q, pa, pb);
  }
  std::memcpy(op.c, w, sizeof(w));
  return op;
}

// Whole-tile exchange to the register set nr (bits in both keep their
// slots); the thread index bits then hold the other local bits in ascending
// order. One op changes at most kRegBits - reg_xchg_split(k, batched) slots.
void exchange_layout(RegSlots* cur, uint32_t nr, uint32_t k, bool batched,
                     std::vector<TileOp>* out) {
  const int max_changes = kRegBits - (int)reg_xchg_split(k, batched);
  while (cur->set != nr) {
    RegSlots to = *cur;
    uint32_t in = nr & ~cur->set;
    int changed = 0;
    for (int j = 0; j < kRegBits && in && changed < max_changes; ++j) {
      if (nr & (1u << to.s[j])) continue;
      const uint32_t b = (uint32_t)__builtin_ctz(in);
      in &= in - 1u;
      to.set = (to.set & ~(1u << to.s[j])) | (1u << b);
      to.s[j] = (uint8_t)b;
      ++changed;
    }
    out->push_back(exchange_op(*cur, to, reg_xchg_split(k, batched)));
    uint32_t m = 0;
    for (uint32_t b = 0; b < k; ++b)
      if (!(to.set & (1u << b))) to.tb[m++] = (uint8_t)b;
    *cur = to;
  }
}

// Moves layout *cur to the register set nr, by swaps (bits in both keep
// their slots; each entering bit trades places with a leaving one) or by
// whole-tile exchanges, whichever costs less.
void move_layout(RegSlots* cur, uint32_t nr, uint32_t k, bool swaps, std::vector<TileOp>* out) {
  if (cur->set == nr) return;
  if (!use_swaps(k, swaps)) {
    exchange_layout(cur, nr, k, swaps, out);
    return;
  }
  SwapList sl;
  uint32_t in = nr & ~cur->set;
  while (in) {  // a slot whose bit leaves takes an entering bit (through slot 3)
    int j = 0;
    while (nr & (1u << cur->s[j])) ++j;
    const uint32_t b = (uint32_t)__builtin_ctz(in);
    in &= in - 1u;
    swap_rt(cur, (uint32_t)j, (uint32_t)tpos_of(*cur, b), &sl, out);
  }
  emit_swaps(*cur, &sl, out);
}

// Puts register bits lo and hi into slots 2 and 3 (where the kernels apply
// 4x4 ops) by slot swaps: one swap list when the kernel takes swap lists,
// else one whole-tile exchange per swap (two slots stay put; the thread
// bits do not change; q: the exchange split).
void place_pair(RegSlots* cur, uint32_t lo, uint32_t hi, bool swaps, uint32_t q,
                std::vector<TileOp>* out) {
  SwapList sl;
  const uint32_t want[2] = {lo, hi};
  for (int d = 0; d < 2; ++d) {
    const int dst = 2 + d;
    if (cur->s[dst] == want[d]) continue;
    const uint32_t src = slot_of(*cur, want[d]);
    if (swaps) {
      swap_rr(cur, src, (uint32_t)dst, &sl, out);
    } else {
      RegSlots to = *cur;
      std::swap(to.s[src], to.s[dst]);
      out->push_back(exchange_op(*cur, to, q));
      *cur = to;
    }
  }
  emit_swaps(*cur, &sl, out);
}

// The pass stores from its final layout: the local bits wst stored at the
// lowest physical positions (the lowest local bits, up to three, unless the
// pass relabels qubits) must then sit on thread index bits below three, so
// eight consecutive lanes write 128 contiguous bytes.
void store_layout(RegSlots* cur, uint32_t k, bool swaps, std::vector<TileOp>* out,
                  uint32_t wst) {
  const uint32_t nlow = std::min<uint32_t>(3u, k - kRegBits);
  if (!use_swaps(k, swaps)) {
    // Exchanges leave the thread bits in ascending order: wst out of the
    // registers, and every local bit below wst's highest that is not in wst
    // in them (at most nlow such bits).
    const uint32_t below = wst ? ((2u << (31 - __builtin_clz(wst))) - 1u) & ~wst : 0u;
    if ((cur->set & wst) || (below & ~cur->set)) {
      uint32_t target = below;
      for (int j = 0; j < kRegBits && __builtin_popcount(target) < kRegBits; ++j)
        if (!(wst & (1u << cur->s[j]))) target |= 1u << cur->s[j];
      for (int b = (int)k - 1; b >= 0 && __builtin_popcount(target) < kRegBits; --b)
        if (!(wst & (1u << b))) target |= 1u << b;
      exchange_layout(cur, target, k, swaps, out);
    }
    return;
  }
  SwapList sl;
  for (uint32_t b = 0; b < nlow; ++b) {
    const int m = tpos_of(*cur, b);
    if (m >= 0 && (uint32_t)m < nlow) continue;
    int p = -1;  // a low thread index bit holding a bit that is not low
    for (uint32_t q = 0; q < nlow && p < 0; ++q)
      if (cur->tb[q] >= nlow) p = (int)q;
    uint32_t j = slot_of(*cur, b);
    if (j == 0xFF) {  // b sits on a higher thread index bit: bring it in first
      j = 0;
      while (cur->s[j] < nlow) ++j;
      swap_rt(cur, j, (uint32_t)m, &sl, out);
      j = slot_of(*cur, b);
    }
    swap_rt(cur, j, (uint32_t)p, &sl, out);
  }
  emit_swaps(*cur, &sl, out);
}

// Splits an op's controls into register slots and thread bits for layout L.
void reg_fields(TileOp* op, const RegSlots& L) {
  op->rsel = 0;
  for (int j = 0; j < kRegBits; ++j)
    if (op->lmask & (1u << L.s[j])) op->rsel |= (uint8_t)(1u << j);
  op->tl = op->lmask & ~L.set;
  op->rt = 0xFF;
  if (op->kind == TK_RZ && (L.set & (1u << op->t))) op->rt = slot_of(L, op->t);
}

// wst: the local bits the store puts at the lowest physical positions.
void segment_pass(const std::vector<TileOp>& ops, const std::vector<uint32_t>& gis, uint32_t k,
                  bool swaps, std::vector<TileOp>* out, std::vector<uint32_t>* order,
                  uint32_t wst) {
  const uint32_t L0 = ((1u << kRegBits) - 1u) << (k - kRegBits);
  const uint32_t allb = (1u << k) - 1u;
  std::vector<SegOp> m(ops.size());
  for (size_t i = 0; i < ops.size(); ++i) m[i] = seg_masks(ops[i]);
  std::vector<uint32_t> rem(ops.size());
  for (uint32_t i = 0; i < rem.size(); ++i) rem[i] = i;
  RegSlots cur;
  for (int j = 0; j < kRegBits; ++j) cur.s[j] = (uint8_t)(k - kRegBits + j);
  cur.set = L0;
  cur.nt = k - kRegBits;
  for (uint32_t m = 0; m < cur.nt; ++m) cur.tb[m] = (uint8_t)m;
  uint32_t R = L0;
  // slots an exchange keeps (plans with swap lists are the batched kernels')
  const int keep = use_swaps(k, swaps) ? 0 : (int)reg_xchg_split(k, swaps);
  uint32_t last_gi = 0xFFFFFFFFu;
  while (!rem.empty()) {
    const int ready = seg_sim(m, rem, R, allb);
    const uint32_t seed = ready == 0 ? m[rem[0]].act : 0u;
    const uint32_t nr = seg_choose(m, rem, seed, k, allb, R, keep);
    if (nr != R && (ready == 0 || seg_sim(m, rem, nr, allb) > ready)) {
      R = nr;
      move_layout(&cur, R, k, swaps, out);
    }
    BlockerT<uint32_t> blk;
    std::vector<uint32_t> next;
    next.reserve(rem.size());
    for (uint32_t i : rem) {
      if (blk.free(m[i].all, m[i].mixb) && (m[i].act & ~R) == 0) {
        TileOp op = ops[i];
        if (op.kind == TK_MAT4) {  // the lower local bit in slot 2, the higher in slot 3
          place_pair(&cur, op.t < op.t2 ? op.t : op.t2, op.t < op.t2 ? op.t2 : op.t,
                     use_swaps(k, swaps), reg_xchg_split(k, swaps), out);
          op.ins[0] = 2;
          op.ins[1] = 3;
        } else if (m[i].act) {
          op.ins[0] = slot_of(cur, op.t);
          if (op.kind == TK_SWAP) op.ins[1] = slot_of(cur, op.t2);
        }
        reg_fields(&op, cur);
        out->push_back(op);
        if (gis[i] != last_gi) {
          order->push_back(gis[i]);
          last_gi = gis[i];
        }
      } else {
        blk.add(m[i].all, m[i].mixb);
        next.push_back(i);
      }
    }
    rem.swap(next);
  }
  store_layout(&cur, k, swaps, out, wst);
}

// ---- factored passes (QOA_FUSE_FACTOR) ----------------------------------------
// An uncontrolled H is nf * [[1, 1], [1, -1]], an RY or RX (m00 = m11 = c) is
// c * [[1, m01/c], [m10/c, 1]], an RZ on a register slot is m00 * diag(1,
// m11/m00). Inside a pass these ops run without their scalars, and the
// product z of the scalars multiplies the coefficients of one op kept in its
// plain form (an RZ when an RZ was factored, z then being complex, else any
// candidate): linear ops commute with a scalar, so the pass's result is
// unchanged up to rounding. Without the scalars the amplitudes grow by 1/|z|
// inside the pass; factors are taken while that stays below 2^kFactorGrowth.
constexpr double kFactorGrowth = 64.0;

inline bool uncontrolled(const TileOp& op) {
  return op.tl == 0 && op.rsel == 0 && op.gmask == 0 && op.gval == 0;
}

// 0: not a candidate, 1: H, 2: RY-shaped, 3: RX-shaped, 4: RZ on a register slot
int factor_kind(const TileOp& op) {
  if (!uncontrolled(op)) return 0;
  if (op.kind == TK_H && op.c[1] == 0.0 && op.c[0] != 0.0) return 1;
  if (op.kind == TK_M && op.c[0] == op.c[6] && op.c[0] != 0.0) {
    if (op.t2 == kM2Real) return 2;
    if (op.t2 == kM2Rx) return 3;
  }
  if (op.kind == TK_RZ && op.rt != 0xFF && (op.c[0] != 0.0 || op.c[1] != 0.0)) return 4;
  return 0;
}

inline void cmul_into(double* a, const double* z) {  // a *= z
  const double r = a[0] * z[0] - a[1] * z[1], i = a[0] * z[1] + a[1] * z[0];
  a[0] = r;
  a[1] = i;
}

void factor_pass(std::vector<TileOp>* ops, size_t begin) {
  TileOp* sink = nullptr;
  for (size_t i = begin; i < ops->size() && !sink; ++i)
    if (factor_kind((*ops)[i]) == 4) sink = &(*ops)[i];
  for (size_t i = begin; i < ops->size() && !sink; ++i)
    if (factor_kind((*ops)[i]) != 0) sink = &(*ops)[i];
  if (!sink) return;
  double z[2] = {1.0, 0.0}, growth = 0.0;
  for (size_t i = begin; i < ops->size(); ++i) {
    TileOp& op = (*ops)[i];
    const int fk = factor_kind(op);
    // a complex factor needs the RZ sink (there is one whenever fk == 4 occurs)
    if (fk == 0 || &op == sink) continue;
    const double f = fk == 4 ? 1.0 : std::fabs(op.c[0]);
    if (growth - std::log2(f) > kFactorGrowth) continue;
    growth -= std::log2(f);
    if (fk == 1) {  // nf [[1, 1], [1, -1]]
      z[0] *= op.c[0];
      z[1] *= op.c[0];
      op.c[0] = 1.0;
    } else if (fk == 2 || fk == 3) {  // c [[1, m01 / c], [m10 / c, 1]]
      const double c = op.c[0];
      z[0] *= c;
      z[1] *= c;
      for (int w = 2; w < 6; ++w) op.c[w] /= c;
      op.c[0] = op.c[6] = 1.0;
      op.t2 = fk == 2 ? kM2RealU : kM2RxU;
    } else {  // m00 diag(1, m11 / m00): a phase on the registers of slot rt
      const double m00[2] = {op.c[0], op.c[1]};
      const double d = m00[0] * m00[0] + m00[1] * m00[1];
      const double inv[2] = {m00[0] / d, -m00[1] / d};
      double ph[2] = {op.c[2], op.c[3]};
      cmul_into(ph, inv);
      cmul_into(z, m00);
      op.kind = TK_DPHASE;
      op.rsel = (uint8_t)(1u << op.rt);
      op.lmask = 1u << op.t;
      op.rt = 0xFF;
      op.c[0] = ph[0];
      op.c[1] = ph[1];
      op.c[2] = op.c[3] = 0.0;
    }
  }
  if (z[0] == 1.0 && z[1] == 0.0) return;
  switch (sink->kind) {
    case TK_RZ:
      cmul_into(sink->c, z);
      cmul_into(sink->c + 2, z);
      break;
    case TK_H:  // z is real here (no RZ was factored)
      sink->c[0] *= z[0];
      break;
    default:  // TK_M of shape kM2Real / kM2Rx (z real): every stored part
      for (int w = 0; w < 8; ++w) sink->c[w] *= z[0];
      break;
  }
}

// ---- merged diagonal runs (QOA_FUSE_MERGE_DIAG) -------------------------------
// VALU instructions of a register op on the 16 amplitudes of a thread (a
// complex product is 6, a select 2) and of a TK_DIAGN op; the op dispatch
// (about 40 scalar instructions, issued beside the vector work) is counted as
// 20 vector instructions.
constexpr int kDispatchCost = 20;
inline int regs_of(uint32_t rs) { return kRegAmps >> __builtin_popcount(rs & 15u); }
int diag_cost(const TileOp& op) {
  switch (op.kind) {
    case TK_DNEG: return op.tl ? 6 * regs_of(op.rsel) + 2 : 2 * regs_of(op.rsel);
    case TK_DPHASE: return 6 * regs_of(op.rsel) + (op.tl ? 2 : 0);
    default: return 6 * kRegAmps + (op.rt == 0xFF ? 2 : 0);  // TK_RZ
  }
}

void diag_terms(const TileOp& op, std::vector<DiagTerm>* out) {
  DiagTerm t{};
  t.gmask = op.gmask;
  t.gval = op.gval;
  if (op.kind == TK_RZ) {  // m00 everywhere, then m11 / m00 where the bit is 1
    t.rs = 0;
    t.tl = 0;
    t.c[0] = op.c[0];
    t.c[1] = op.c[1];
    out->push_back(t);
    // m11 / m00 (= m11 * conj(m00) for a plain RZ; a factored pass's RZ
    // carries the pass's scalar in both entries)
    const double d = op.c[0] * op.c[0] + op.c[1] * op.c[1];
    t.c[0] = (op.c[2] * op.c[0] + op.c[3] * op.c[1]) / d;
    t.c[1] = (op.c[3] * op.c[0] - op.c[2] * op.c[1]) / d;
    if (op.rt != 0xFF) {
      t.rs = 1u << op.rt;
    } else {
      t.tl = 1u << op.t;
    }
    out->push_back(t);
    return;
  }
  t.rs = op.rsel;
  t.tl = op.tl;
  if (op.kind == TK_DNEG) {
    t.c[0] = -1.0;
    t.c[1] = 0.0;
  } else {
    t.c[0] = op.c[0];
    t.c[1] = op.c[1];
  }
  out->push_back(t);
}

// Replaces runs of consecutive diagonal ops in ops[begin, end) by TK_DIAGN
// ops where that costs fewer instructions. Every op of a run has the same
// register layout (a layout op ends a run), so its rsel/tl fields hold.
void merge_diag_runs(std::vector<TileOp>* ops, size_t begin, std::vector<double>* pool) {
  auto diag = [](const TileOp& op) {
    return op.kind == TK_DNEG || op.kind == TK_DPHASE || op.kind == TK_RZ;
  };
  std::vector<TileOp> out(ops->begin(), ops->begin() + (long)begin);
  size_t i = begin;
  while (i < ops->size()) {
    if (!diag((*ops)[i])) {
      out.push_back((*ops)[i++]);
      continue;
    }
    size_t j = i;
    int plain = 0;
    std::vector<DiagTerm> terms;
    while (j < ops->size() && diag((*ops)[j])) {
      plain += diag_cost((*ops)[j]) + kDispatchCost;
      diag_terms((*ops)[j], &terms);
      ++j;
    }
    std::stable_sort(terms.begin(), terms.end(),
                     [](const DiagTerm& a, const DiagTerm& b) { return a.rs < b.rs; });
    int merged = kDispatchCost;
    for (size_t t = 0; t < terms.size(); ++t) {
      merged += terms[t].tl ? 8 : 6;
      if (t == 0 || terms[t].rs != terms[t - 1].rs) merged += 6 * regs_of(terms[t].rs) + 4;
    }
    if (j - i < 2 || merged >= plain) {
      for (size_t t = i; t < j; ++t) out.push_back((*ops)[t]);
    } else {
      TileOp m{};
      m.kind = TK_DIAGN;
      m.aux = (uint32_t)pool->size();
      m.lmask = (uint32_t)terms.size();  // the term count
      m.rt = 0xFF;
      for (size_t t = 0; t < terms.size(); ++t)
        if (t + 1 == terms.size() || terms[t + 1].rs != terms[t].rs) terms[t].rs |= kDiagGroupEnd;
      // the terms, then one spare zero term: the kernel reads a term ahead
      const size_t at = pool->size();
      pool->resize(at + kDiagTermDoubles * (terms.size() + 1), 0.0);
      std::memcpy(pool->data() + at, terms.data(), sizeof(DiagTerm) * terms.size());
      out.push_back(m);
    }
    i = j;
  }
  ops->swap(out);
}

}  // namespace

int canonical_gate(uint32_t n, const qoa_gate& g, const double* mats, uint32_t flags,
                   std::vector<FOp>* out, uint32_t gi) {
  return canonical(n, g, mats, flags, out, gi);
}

namespace {

// The pass's tile qubits: from `seed`, greedily the qubit (or, when no single
// one helps, the pair) that lets the most ops of `rem` run, up to k qubits.
uint64_t choose_set(const std::vector<FOp>& ops, const std::vector<uint32_t>& rem, uint64_t seed,
                    uint32_t n, uint32_t k, uint64_t allq, uint64_t prefer = 0) {
  uint64_t S = seed;
  int cur = simulate(ops, rem, S, allq);
  while ((uint32_t)__builtin_popcountll(S) < k) {
    int best = -1, bests = -1;
    for (uint32_t q = 0; q < n; ++q) {
      if (S & bit(q)) continue;
      const int sc = simulate(ops, rem, S | bit(q), allq);
      // ties go to the preferred qubits (relabelled plans: qubits away from
      // their own position, so later passes can put them back)
      if (sc > bests || (sc == bests && (prefer & bit(q)) && !(prefer & bit((uint32_t)best)))) {
        bests = sc;
        best = (int)q;
      }
    }
    if (bests <= cur && (uint32_t)__builtin_popcountll(S) + 2 <= k) {
      // no single qubit helps: try the best pair (two-qubit gates need both)
      int b1 = -1, b2 = -1, bp = cur;
      for (uint32_t q = 0; q < n; ++q) {
        if (S & bit(q)) continue;
        for (uint32_t r = q + 1; r < n; ++r) {
          if (S & bit(r)) continue;
          const int sc = simulate(ops, rem, S | bit(q) | bit(r), allq);
          if (sc > bp) {
            bp = sc;
            b1 = (int)q;
            b2 = (int)r;
          }
        }
      }
      if (b1 >= 0) {
        S |= bit(b1) | bit(b2);
        cur = bp;
        continue;
      }
    }
    S |= bit(best);
    cur = bests;
  }
  return S;
}

}  // namespace

int build_plan(uint32_t n, const qoa_gate* g, uint32_t ng, const double* mats,
               const PlanOptions& opt, FusedPlan* plan) {
  if (n < 1 || n > 36) return set_error(QOA_ERR_ARG, "n_qubits %u out of range", n);
  // Default tile: 12 qubits. A single state of 13..17 qubits takes 11 (twice
  // the workgroups of a small grid: measured 227 against 244 us for a
  // 16-qubit, 470-gate circuit); batched plans keep 12 (their jobs fill the
  // GPU).
  const bool small = n > 12 && n < 18 && !(opt.flags & kPlanForceReg);
  uint32_t k = opt.tile_qubits ? opt.tile_qubits : (small ? 11 : 12);
  if (k > 13) return set_error(QOA_ERR_ARG, "tile_qubits %u exceeds 13 (128 KiB of LDS)", k);
  if (k < 4) k = 4;
  if (k > n) k = n;
  uint32_t m = opt.low_qubits ? opt.low_qubits : 3;
  if (m + 2 > k) m = k >= 2 ? k - 2 : 0;

  std::vector<FOp> ops;
  ops.reserve(ng);
  plan->n = n;
  plan->k = k;
  plan->passes.clear();
  plan->ops.clear();
  plan->pool.clear();
  plan->order.clear();
  for (uint32_t i = 0; i < ng; ++i) {
    size_t before = ops.size();
    int rc = canonical(n, g[i], mats, opt.flags, &ops, i);
    if (rc) return rc;
    if (ops.size() == before) plan->order.push_back(i);  // a no-op: position is irrelevant
  }
  if (opt.flags & kPlanNoise)  // errors follow every gate: no gate passes another
    for (FOp& f : ops) f.mixb = f.all;
  const uint64_t allq = n >= 64 ? ~0ULL : ((1ULL << n) - 1);
  std::vector<uint32_t> rem(ops.size());
  for (uint32_t i = 0; i < ops.size(); ++i) rem[i] = i;
  const bool reg_ok = (k >= 8 || (opt.flags & kPlanForceReg)) && k >= (uint32_t)kRegBits &&
                      !(opt.flags & QOA_FUSE_LDS_ONLY);
  // (not with swap-list layouts, nor with noise: trajectory errors name
  // their qubits by position)
  const bool relabel = (opt.flags & QOA_FUSE_RELABEL) && !(opt.flags & kPlanNoise) &&
                       !use_swaps(k, (opt.flags & kPlanSwaps) != 0) && reg_ok && k < n && m > 0 &&
                       k >= (uint32_t)kRegBits + 2 * m;
  // logical qubit q sits at physical position pos[q]; inv is the inverse
  uint8_t pos[64], inv[64];
  for (uint32_t q = 0; q < 64; ++q) pos[q] = inv[q] = (uint8_t)q;
  auto low_now = [&]() {  // logical qubits at the m lowest positions
    uint64_t w = 0;
    for (uint32_t p = 0; p < m; ++p) w |= bit(inv[p]);
    return w;
  };

  // One pass over the tile qubits S (logical), running the ops `take` (in
  // order); wn: the logical qubits the pass stores at the m lowest positions.
  auto add_pass = [&](uint64_t S, const std::vector<uint32_t>& take, uint64_t wn) {
    PassPlan pp{};
    pp.k = k;
    pp.reg_mode = reg_ok ? 1u : 0u;
    pp.op_begin = (uint32_t)plan->ops.size();
    uint8_t lq[16];  // logical qubit of local bit j
    uint32_t wst = (1u << std::min<uint32_t>(3u, k - kRegBits)) - 1u;
    uint8_t npos[64];
    std::memcpy(npos, pos, sizeof(npos));
    if (!relabel) {
      int j = 0;
      for (uint32_t q = 0; q < n; ++q)
        if (S & bit(q)) lq[j++] = (uint8_t)q;
    } else {
      // local bits: the current low qubits (by position), then the other
      // qubits of wn, then the rest of S by position
      const uint64_t w = low_now();
      int j = 0;
      for (uint32_t p = 0; p < m; ++p) lq[j++] = inv[p];
      for (uint32_t q = 0; q < n; ++q)
        if ((wn & bit(q)) && !(w & bit(q))) lq[j++] = (uint8_t)q;
      for (uint32_t p = 0; p < n; ++p)
        if ((S & bit(inv[p])) && !((w | wn) & bit(inv[p]))) lq[j++] = inv[p];
      // new positions: wn to 0..m-1 (lanes 0..m-1 store them; any order
      // within the 128-byte run), its own home first; the others to their
      // home position when it is in the tile, else where they are, else any
      // free tile position
      uint64_t freep = 0;  // free physical positions (of the tile's)
      for (uint32_t i = 0; i < k; ++i) freep |= bit(pos[lq[i]]);
      const uint64_t lowp = (1ULL << m) - 1;
      wst = 0;
      bool placed[16] = {};
      for (uint32_t i = 0; i < k; ++i)
        if ((wn & bit(lq[i])) && lq[i] < m) {
          npos[lq[i]] = lq[i];
          freep &= ~bit(lq[i]);
          placed[i] = true;
        }
      for (uint32_t i = 0; i < k; ++i)
        if (wn & bit(lq[i])) {
          wst |= 1u << i;
          if (placed[i]) continue;
          const uint32_t p = (uint32_t)__builtin_ctzll(freep & lowp);
          npos[lq[i]] = (uint8_t)p;
          freep &= ~bit(p);
          placed[i] = true;
        }
      for (int pass = 0; pass < 3; ++pass)
        for (uint32_t i = 0; i < k; ++i) {
          if (placed[i]) continue;
          const uint32_t q = lq[i];
          uint32_t want = 64;
          if (pass == 0 && (freep & bit(q))) want = q;
          if (pass == 1 && (freep & bit(pos[q]))) want = pos[q];
          if (pass == 2) want = (uint32_t)__builtin_ctzll(freep);
          if (want == 64) continue;
          npos[q] = (uint8_t)want;
          freep &= ~bit(want);
          placed[i] = true;
        }
    }
    for (uint32_t i = 0; i < k; ++i) {
      pp.spos[i] = pos[lq[i]];
      pp.spos_st[i] = npos[lq[i]];
    }
    std::vector<TileOp> pass_ops;
    std::vector<uint32_t> pass_gi;
    for (uint32_t i : take) {
      const FOp& f = ops[i];
      const size_t before = pass_ops.size();
      emit(f, S, lq, (int)k, &pass_ops, &plan->pool, (opt.flags & kPlanNoise) != 0,
           relabel ? pos : nullptr);
      for (size_t e = before; e < pass_ops.size(); ++e) pass_gi.push_back(f.gi);
    }
    if (pp.reg_mode) {
      segment_pass(pass_ops, pass_gi, k, (opt.flags & kPlanSwaps) != 0, &plan->ops, &plan->order,
                   wst);
      if (opt.flags & QOA_FUSE_FACTOR)  // scalars commute with the Pauli errors too
        factor_pass(&plan->ops, pp.op_begin);
      if ((opt.flags & QOA_FUSE_MERGE_DIAG) && !(opt.flags & kPlanNoise))
        merge_diag_runs(&plan->ops, pp.op_begin, &plan->pool);
    } else {
      plan->ops.insert(plan->ops.end(), pass_ops.begin(), pass_ops.end());
      for (size_t e = 0; e < pass_gi.size(); ++e)
        if (e == 0 || pass_gi[e] != pass_gi[e - 1]) plan->order.push_back(pass_gi[e]);
    }
    pp.op_end = (uint32_t)plan->ops.size();
    plan->passes.push_back(pp);
    std::memcpy(pos, npos, sizeof(pos));
    for (uint32_t q = 0; q < n; ++q) inv[pos[q]] = (uint8_t)q;
  };

  while (!rem.empty()) {
    uint64_t mis = 0;
    if (relabel)
      for (uint32_t q = 0; q < n; ++q)
        if (pos[q] != q) mis |= bit(q);
    uint64_t S =
        k >= n ? allq : choose_set(ops, rem, low_now() | ops[rem[0]].active, n, k, allq, mis);
    if (relabel) {
      // The last pass: when the qubits its ops mix, the low ones and every
      // qubit away from its own position fit in one tile, take that tile, so
      // the pass itself puts the order back (no extra pass).
      Blocker b0;
      bool last = true;
      uint64_t need = low_now() | ((1ULL << m) - 1);
      for (uint32_t i : rem) {
        const FOp& f = ops[i];
        if (!b0.free(f.all, f.mixb) || (f.active & ~S)) {
          last = false;
          break;
        }
        need |= f.active;
      }
      for (uint32_t q = 0; q < n; ++q)
        if (pos[q] != q) need |= bit(q);
      if (last && (uint32_t)__builtin_popcountll(need) <= k) {
        for (uint32_t q = 0; q < n && (uint32_t)__builtin_popcountll(need) < k; ++q) need |= bit(q);
        S = need;
      }
    }
    Blocker blk;
    std::vector<uint32_t> take, next;
    next.reserve(rem.size());
    for (uint32_t i : rem) {
      const FOp& f = ops[i];
      if (blk.free(f.all, f.mixb) && (f.active & ~S) == 0) {
        take.push_back(i);
      } else {
        blk.add(f.all, f.mixb);
        next.push_back(i);
      }
    }
    uint64_t wn = low_now();
    if (relabel) {
      // the next pass's low qubits: qubits of S the next pass would choose
      // anyway (its choice without the low constraint), the current low ones
      // first; after the last pass the home qubits 0..m-1 where possible
      const uint64_t home = (1ULL << m) - 1;
      const uint64_t pref = next.empty() ? home : choose_set(ops, next, ops[next[0]].active, n, k, allq);
      uint64_t cand = S & pref, w = 0;
      for (uint64_t c = cand & wn; c && (uint32_t)__builtin_popcountll(w) < m; c &= c - 1)
        w |= c & (~c + 1);
      for (uint64_t c = cand & ~wn; c && (uint32_t)__builtin_popcountll(w) < m; c &= c - 1)
        w |= c & (~c + 1);
      for (uint64_t c = wn; c && (uint32_t)__builtin_popcountll(w) < m; c &= c - 1)
        w |= c & (~c + 1);
      wn = w;
    }
    add_pass(S, take, wn);
    rem.swap(next);
  }
  if (relabel) {
    // Put every qubit back at its own position: passes without ops, each
    // over the low qubits and whole cycles of the position permutation (a
    // qubit reaches its home when the qubit there is in the same tile) as
    // far as they fit, else the longest chain that does.
    const uint64_t home_low = (1ULL << m) - 1;
    for (int guard = 0;; ++guard) {
      uint64_t mis = 0;
      for (uint32_t q = 0; q < n; ++q)
        if (pos[q] != q) mis |= bit(q);
      if (!mis) break;
      if (guard == 8) {  // not converging: plan without relabelling
        PlanOptions o2 = opt;
        o2.flags &= ~QOA_FUSE_RELABEL;
        return build_plan(n, g, ng, mats, o2, plan);
      }
      uint64_t S = low_now() | home_low;
      for (uint64_t c = mis; c; c &= c - 1) {  // whole cycles
        const uint32_t q0 = (uint32_t)__builtin_ctzll(c);
        uint64_t cyc = 0;
        for (uint32_t q = q0; !(cyc & bit(q)); q = inv[q]) cyc |= bit(q);
        if ((uint32_t)__builtin_popcountll(S | cyc) <= k) S |= cyc;
      }
      for (uint64_t c = S & mis; c && (uint32_t)__builtin_popcountll(S) < k; c &= c - 1) {
        // chains: the qubit sitting at a member's home joins while room is left
        for (uint32_t q = (uint32_t)__builtin_ctzll(c);
             (uint32_t)__builtin_popcountll(S) < k && !(S & bit(inv[q])); q = inv[q])
          S |= bit(inv[q]);
      }
      for (uint32_t q = 0; q < n && (uint32_t)__builtin_popcountll(S) < k; ++q) S |= bit(q);
      add_pass(S, {}, home_low);
    }
  }
  return QOA_OK;
}

}  // namespace qoa

extern "C" int qoa_tile_op_size(void) { return (int)sizeof(qoa::TileOp); }

extern "C" int qoa_plan_program(uint32_t n, const qoa_gate* g, uint32_t ng, const double* mats,
                                const qoa_fuse_opts* opts, void* ops_out, uint32_t max_ops,
                                uint8_t* spos_out, uint32_t* range_out, uint32_t max_passes,
                                uint32_t* n_ops, uint32_t* n_passes, uint32_t* tile_k) {
  using namespace qoa;
  if (!g && ng) return set_error(QOA_ERR_ARG, "gates is NULL");
  PlanOptions po{opts ? opts->tile_qubits : 0u, opts ? opts->low_qubits : 0u,
                 opts ? opts->flags : 0u};
  std::vector<double> zeros;
  if (!mats) {
    uint64_t maxaux = 0;
    for (uint32_t i = 0; i < ng; ++i)
      if (g[i].op == QOA_OP_U2 || g[i].op == QOA_OP_U4)
        maxaux = std::max<uint64_t>(maxaux, g[i].aux + 1);
    zeros.assign(32 * maxaux + 32, 0.0);
    mats = zeros.data();
  }
  FusedPlan plan;
  int rc = build_plan(n, g, ng, mats, po, &plan);
  if (rc) return rc;
  if (n_ops) *n_ops = (uint32_t)plan.ops.size();
  if (n_passes) *n_passes = (uint32_t)plan.passes.size();
  if (tile_k) *tile_k = plan.k;
  if (ops_out && max_ops < plan.ops.size())
    return set_error(QOA_ERR_ARG, "%zu ops do not fit in %u", plan.ops.size(), max_ops);
  if ((spos_out || range_out) && max_passes < plan.passes.size())
    return set_error(QOA_ERR_ARG, "%zu passes do not fit in %u", plan.passes.size(), max_passes);
  if (ops_out) std::memcpy(ops_out, plan.ops.data(), sizeof(TileOp) * plan.ops.size());
  if (spos_out && range_out) {
    for (size_t p = 0; p < plan.passes.size(); ++p) {
      std::memcpy(spos_out + 16 * p, plan.passes[p].spos, 16);
      range_out[3 * p] = plan.passes[p].op_begin;
      range_out[3 * p + 1] = plan.passes[p].op_end;
      range_out[3 * p + 2] = plan.passes[p].reg_mode;
    }
  }
  return QOA_OK;
}

extern "C" int qoa_plan_store_positions(uint32_t n, const qoa_gate* g, uint32_t ng,
                                        const double* mats, const qoa_fuse_opts* opts,
                                        uint8_t* spos_st_out, uint32_t max_passes,
                                        uint32_t* n_passes) {
  using namespace qoa;
  if (!g && ng) return set_error(QOA_ERR_ARG, "gates is NULL");
  if (!n_passes) return set_error(QOA_ERR_ARG, "n_passes is NULL");
  PlanOptions po{opts ? opts->tile_qubits : 0u, opts ? opts->low_qubits : 0u,
                 opts ? opts->flags : 0u};
  std::vector<double> zeros;
  if (!mats) {
    uint64_t maxaux = 0;
    for (uint32_t i = 0; i < ng; ++i)
      if (g[i].op == QOA_OP_U2 || g[i].op == QOA_OP_U4)
        maxaux = std::max<uint64_t>(maxaux, g[i].aux + 1);
    zeros.assign(32 * maxaux + 32, 0.0);
    mats = zeros.data();
  }
  FusedPlan plan;
  int rc = build_plan(n, g, ng, mats, po, &plan);
  if (rc) return rc;
  *n_passes = (uint32_t)plan.passes.size();
  if (spos_st_out && max_passes < plan.passes.size())
    return set_error(QOA_ERR_ARG, "%zu passes do not fit in %u", plan.passes.size(), max_passes);
  if (spos_st_out)
    for (size_t p = 0; p < plan.passes.size(); ++p)
      std::memcpy(spos_st_out + 16 * p, plan.passes[p].spos_st, 16);
  return QOA_OK;
}

extern "C" int qoa_plan_pool(uint32_t n, const qoa_gate* g, uint32_t ng, const double* mats,
                             const qoa_fuse_opts* opts, double* pool_out, uint32_t max_doubles,
                             uint32_t* n_doubles) {
  using namespace qoa;
  if (!g && ng) return set_error(QOA_ERR_ARG, "gates is NULL");
  if (!n_doubles) return set_error(QOA_ERR_ARG, "n_doubles is NULL");
  PlanOptions po{opts ? opts->tile_qubits : 0u, opts ? opts->low_qubits : 0u,
                 opts ? opts->flags : 0u};
  FusedPlan plan;
  int rc = build_plan(n, g, ng, mats, po, &plan);
  if (rc) return rc;
  *n_doubles = (uint32_t)plan.pool.size();
  if (pool_out && max_doubles < plan.pool.size())
    return set_error(QOA_ERR_ARG, "%zu pool doubles do not fit in %u", plan.pool.size(),
                     max_doubles);
  if (pool_out && !plan.pool.empty())
    std::memcpy(pool_out, plan.pool.data(), sizeof(double) * plan.pool.size());
  return QOA_OK;
}

namespace qoa {

int plan_order(uint32_t n, const qoa_gate* g, uint32_t ng, const PlanOptions& opt,
               std::vector<uint32_t>* order, uint32_t* n_passes) {
  FusedPlan plan;
  // matrices are only needed for coefficients, not for the schedule
  std::vector<double> zeros;
  uint64_t maxaux = 0;
  for (uint32_t i = 0; i < ng; ++i)
    if (g[i].op == QOA_OP_U2 || g[i].op == QOA_OP_U4) maxaux = std::max<uint64_t>(maxaux, g[i].aux + 1);
  zeros.assign(32 * maxaux + 32, 0.0);
  int rc = build_plan(n, g, ng, zeros.data(), opt, &plan);
  if (rc) return rc;
  if (order) *order = plan.order;
  if (n_passes) *n_passes = (uint32_t)plan.passes.size();
  return QOA_OK;
}

}  // namespace qoa
