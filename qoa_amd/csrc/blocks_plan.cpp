// blocks_plan.cpp — the dense-block fuser (host code, no device code).
//
// A greedy fuser packs gates into blocks that act on at most five qubits:
// the block's qubit set grows from the first pending gate by the qubit that
// lets it absorb the most gates, and a gate joins when its qubits lie in the
// set and it commutes with every gate left for a later block (they share no
// qubit that either mixes), so the emitted order is a valid reordering. Each
// block's 32x32 unitary is the FP64 product of its gates.
#include "blocks_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "gates.hpp"
#include "qoa_internal.hpp"

namespace qoa {
namespace {

// Applies f to the 32-vector v (local bits b <-> physical qubits q[b]).
void apply_local(const FOp& f, const int* loc, cd* v) {
  auto lm = [&](uint64_t m) {
    uint32_t r = 0;
    for (int q = 0; q < 64; ++q)
      if ((m >> q) & 1ULL) r |= 1u << loc[q];
    return r;
  };
  const uint32_t pm = lm(f.pmask);
  const cd c0(f.c[0], f.c[1]), c1(f.c[2], f.c[3]), c2(f.c[4], f.c[5]), c3(f.c[6], f.c[7]);
  switch (f.kind) {
    case FK_H: case FK_M: case FK_X: case FK_Y: {
      const uint32_t tb = 1u << loc[f.t];
      for (uint32_t i = 0; i < (uint32_t)kDim; ++i) {
        if ((i & tb) || (i & pm) != pm) continue;
        const cd a = v[i], b = v[i | tb];
        if (f.kind == FK_H) {
          v[i] = c0 * (a + b);
          v[i | tb] = c0 * (a - b);
        } else if (f.kind == FK_M) {
          v[i] = c0 * a + c1 * b;
          v[i | tb] = c2 * a + c3 * b;
        } else if (f.kind == FK_X) {
          v[i] = b;
          v[i | tb] = a;
        } else {
          v[i] = b * cd(0.0, -1.0);
          v[i | tb] = a * cd(0.0, 1.0);
        }
      }
    } break;
    case FK_DNEG:
      for (uint32_t i = 0; i < (uint32_t)kDim; ++i)
        if ((i & pm) == pm) v[i] = -v[i];
      break;
    case FK_DPHASE:
      for (uint32_t i = 0; i < (uint32_t)kDim; ++i)
        if ((i & pm) == pm) v[i] *= c0;
      break;
    case FK_RZ: {
      const uint32_t tb = 1u << loc[f.rzq];
      for (uint32_t i = 0; i < (uint32_t)kDim; ++i) v[i] *= (i & tb) ? c1 : c0;
    } break;
    case FK_SWAP: {
      const uint32_t m1 = 1u << loc[f.t], m2 = 1u << loc[f.t2];
      for (uint32_t i = 0; i < (uint32_t)kDim; ++i) {
        if ((i & (m1 | m2)) || (i & pm) != pm) continue;
        std::swap(v[i | m1], v[i | m2]);
      }
    } break;
    case FK_MAT4: {
      // basis (t, t2) = (q1, q2) ordered by (higher, lower) physical qubit,
      // as the per-gate 4x4 sweep
      const int lo = std::min(loc[f.t], loc[f.t2]), hi = std::max(loc[f.t], loc[f.t2]);
      const uint32_t ml = 1u << lo, mh = 1u << hi;
      for (uint32_t i = 0; i < (uint32_t)kDim; ++i) {
        if (i & (ml | mh)) continue;
        const uint32_t idx[4] = {i, i | ml, i | mh, i | ml | mh};
        cd w[4], o[4];
        for (int q = 0; q < 4; ++q) w[q] = v[idx[q]];
        for (int s = 0; s < 4; ++s) {
          o[s] = 0.0;
          for (int q = 0; q < 4; ++q) o[s] += cd(f.m4[8 * s + 2 * q], f.m4[8 * s + 2 * q + 1]) * w[q];
        }
        for (int s = 0; s < 4; ++s) v[idx[s]] = o[s];
      }
    } break;
  }
}

}  // namespace

int build_blocks(uint32_t n, const qoa_gate* g, uint32_t ng, const double* mats, uint32_t flags,
                 uint32_t bq, std::vector<Block>* out) {
  if (n < (uint32_t)kBlockQ || n > 36)
    return set_error(QOA_ERR_ARG, "dense blocks need 5..36 qubits (state has %u)", n);
  if (bq == 0) bq = kBlockQ;
  if (bq > (uint32_t)kBlockQ) return set_error(QOA_ERR_ARG, "block_qubits %u exceeds 5", bq);
  std::vector<FOp> ops;
  ops.reserve(ng);
  for (uint32_t i = 0; i < ng; ++i) {
    const uint32_t op = g[i].op;
    if (op == QOA_OP_MEASURE || op == QOA_OP_RESET || op == QOA_OP_RESET_ALL)
      return set_error(QOA_ERR_UNSUPPORTED, "gate %u (op 0x%02x) is not unitary", i, op);
    int rc = canonical_gate(n, g[i], mats, flags, &ops, i);
    if (rc) return rc;
  }
  for (const FOp& f : ops)
    if ((uint32_t)__builtin_popcountll(f.all) > bq)
      return set_error(QOA_ERR_UNSUPPORTED, "gate %u touches more than %u qubits", f.gi, bq);
  std::vector<uint32_t> rem(ops.size());
  for (uint32_t i = 0; i < rem.size(); ++i) rem[i] = i;
  // gates a block over qubit set S would take (scan stops once S is blocked)
  // (a gate passes a gate left for later when they share no qubit that
  // either of them mixes: then they commute)
  struct Blk {
    uint64_t any = 0, mix = 0;
    bool free(const FOp& f) const { return (f.mixb & any) == 0 && ((f.all & ~f.mixb) & mix) == 0; }
    void add(const FOp& f) {
      any |= f.all;
      mix |= f.mixb;
    }
  };
  auto count = [&](uint64_t S) {
    Blk b;
    int cnt = 0;
    for (uint32_t i : rem) {
      const FOp& f = ops[i];
      if (b.free(f) && (f.all & ~S) == 0) {
        ++cnt;
      } else {
        b.add(f);
        if ((b.mix & S) == S) break;
      }
    }
    return cnt;
  };
  out->clear();
  while (!rem.empty()) {
    // grow the qubit set from the first pending gate, one qubit at a time,
    // taking the candidate (a qubit of the next pending gates) that lets the
    // block absorb the most gates
    uint64_t Q = ops[rem[0]].all;
    while ((uint32_t)__builtin_popcountll(Q) < bq) {
      uint64_t cand = 0;
      for (size_t r = 0; r < rem.size() && r < 64; ++r) cand |= ops[rem[r]].all;
      cand &= ~Q;
      if (!cand) break;
      int best_cnt = -1;
      uint64_t best = 0;
      for (int q = 0; q < 64; ++q) {
        if (!((cand >> q) & 1ULL)) continue;
        const int c = count(Q | (1ULL << q));
        if (c > best_cnt) {
          best_cnt = c;
          best = 1ULL << q;
        }
      }
      Q |= best;
    }
    Blk blk;
    std::vector<uint32_t> take, next;
    for (uint32_t i : rem) {
      const FOp& f = ops[i];
      if (blk.free(f) && (f.all & ~Q) == 0) {
        take.push_back(i);
      } else {
        blk.add(f);
        next.push_back(i);
      }
    }
    // pad to five qubits with the lowest others (identity on them)
    for (uint32_t q = 0; q < n && __builtin_popcountll(Q) < kBlockQ; ++q) Q |= 1ULL << q;
    Block b;
    int loc[64];
    for (int q = 0; q < 64; ++q) loc[q] = -1;
    int j = 0;
    for (uint32_t q = 0; q < n; ++q)
      if ((Q >> q) & 1ULL) {
        b.q[j] = (uint8_t)q;
        loc[q] = j++;
      }
    b.u.assign(kDim * kDim, cd(0.0, 0.0));
    for (int i = 0; i < kDim; ++i) b.u[i * kDim + i] = 1.0;
    cd col[kDim];
    for (int k = 0; k < kDim; ++k) {
      for (int i = 0; i < kDim; ++i) col[i] = b.u[i * kDim + k];
      for (uint32_t t : take) apply_local(ops[t], loc, col);
      for (int i = 0; i < kDim; ++i) b.u[i * kDim + k] = col[i];
    }
    out->push_back(std::move(b));
    rem.swap(next);
  }
  return QOA_OK;
}

// A operands, step-major: w[s*64 + l] for K-step s of tile s / 32 and lane l
// (row i = l % 32, K index 2*(s%32) + l/32 of [Ur -Ui] or [Ui Ur]).
void pack_operands(const Block& b, float* w) {
  for (int s = 0; s < 64; ++s) {
    const int tile = s / 32, st = s % 32;
    for (int l = 0; l < 64; ++l) {
      const int i = l % 32, h = l / 32;
      const int k = 2 * (st % 16) + h;  // column of U
      const cd u = b.u[i * kDim + k];
      float val;
      if (tile == 0) val = st < 16 ? (float)u.real() : (float)(-u.imag());
      else val = st < 16 ? (float)u.imag() : (float)u.real();
      w[s * 64 + l] = val;
    }
  }
}

}  // namespace qoa

using namespace qoa;

extern "C" int qoa_plan_blocks(uint32_t n, const qoa_gate* g, uint32_t ng, const double* mats,
                               const qoa_block_opts* opts, uint32_t* qubits_out,
                               double* unitaries_out, uint32_t max_blocks,
                               uint32_t* n_blocks_out) {
  if (!g && ng) return set_error(QOA_ERR_ARG, "gates is NULL");
  if (!n_blocks_out) return set_error(QOA_ERR_ARG, "n_blocks_out is NULL");
  std::vector<Block> blocks;
  int rc = build_blocks(n, g, ng, mats, opts ? opts->flags : 0, opts ? opts->block_qubits : 0,
                        &blocks);
  if (rc) return rc;
  *n_blocks_out = (uint32_t)blocks.size();
  if (!qubits_out && !unitaries_out) return QOA_OK;
  if (blocks.size() > max_blocks)
    return set_error(QOA_ERR_ARG, "%zu blocks do not fit in %u", blocks.size(), max_blocks);
  for (size_t b = 0; b < blocks.size(); ++b) {
    if (qubits_out)
      for (int j = 0; j < kBlockQ; ++j) qubits_out[kBlockQ * b + j] = blocks[b].q[j];
    if (unitaries_out)
      for (int e = 0; e < kDim * kDim; ++e) {
        unitaries_out[2 * (kDim * kDim * b + e)] = blocks[b].u[e].real();
        unitaries_out[2 * (kDim * kDim * b + e) + 1] = blocks[b].u[e].imag();
      }
  }
  return QOA_OK;
}
