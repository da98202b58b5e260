// driver.cpp — host code of the backend under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by tests/sanitize/Makefile, run by
// tests/test_sanitize.py). No GPU: it drives the pass planner (build_plan,
// plan_order, the register-slot scheduler, swap lists, noise operands) and
// the dense-block fuser (build_blocks, pack_operands, qoa_plan_blocks) over
// random circuits of every gate kind, including out-of-range and repeated
// operands, and checks the invariants the kernels rely on: the executed order
// is a permutation, pass ranges tile the op array, local bit positions lie
// inside the tile, MAT4 coefficients inside the pool, block qubits distinct
// and in range, block matrices unitary.
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../qoa_amd/csrc/blocks_plan.hpp"
#include "../../qoa_amd/csrc/planner.hpp"
#include "../../qoa_amd/csrc/qoa_internal.hpp"

namespace qoa {
static char g_msg[512];
int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  return code;
}
}  // namespace qoa

using namespace qoa;

static int g_fail = 0;
#define EXPECT(c, ...)                                   \
  do {                                                   \
    if (!(c)) {                                          \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);                 \
      std::fprintf(stderr, "\n");                        \
      ++g_fail;                                          \
    }                                                    \
  } while (0)

static const uint32_t kOps[] = {QOA_OP_H,    QOA_OP_Z,     QOA_OP_X,    QOA_OP_PHASESHIFT,
                                QOA_OP_SWAP, QOA_OP_CSWAP, QOA_OP_T,    QOA_OP_S,
                                QOA_OP_RZ,   QOA_OP_CNOT,  QOA_OP_CZ,   QOA_OP_RX,
                                QOA_OP_RY,   QOA_OP_PHASE, QOA_OP_CPHASE, QOA_OP_Y,
                                QOA_OP_U2,   QOA_OP_U4,    QOA_OP_TOFFOLI, QOA_OP_CY};

struct Circuit {
  uint32_t n;
  std::vector<qoa_gate> g;
  std::vector<double> mats;  // 32 doubles per slot: U2 at 8*aux, U4 at 32*aux
  bool clean;                // every operand in range and distinct
};

static Circuit random_circuit(std::mt19937_64& rng, uint32_t n, uint32_t count, bool clean) {
  Circuit c;
  c.n = n;
  c.clean = clean;
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const uint32_t slots = 8;
  c.mats.resize(32 * slots);
  for (double& x : c.mats) x = U(rng) - 0.5;  // values only feed products
  for (uint32_t i = 0; i < count; ++i) {
    qoa_gate g{};
    g.op = kOps[rng() % (sizeof(kOps) / sizeof(kOps[0]))];
    const uint32_t ar = (g.op == QOA_OP_CSWAP || g.op == QOA_OP_TOFFOLI) ? 3
                        : (g.op == QOA_OP_CNOT || g.op == QOA_OP_CZ || g.op == QOA_OP_CPHASE ||
                           g.op == QOA_OP_SWAP || g.op == QOA_OP_U4 || g.op == QOA_OP_CY)
                            ? 2
                            : 1;
    if (ar > n) continue;
    uint32_t q[3];
    for (uint32_t j = 0; j < 3; ++j) {
      for (;;) {
        q[j] = (uint32_t)(rng() % n);
        bool dup = false;
        for (uint32_t i2 = 0; i2 < j && i2 < ar; ++i2) dup = dup || q[i2] == q[j];
        if (!dup || j >= ar) break;
      }
    }
    if (!clean && U(rng) < 0.08) q[rng() % ar] = n + (uint32_t)(rng() % 3);  // out of range
    if (!clean && ar > 1 && U(rng) < 0.04) q[1] = q[0];                       // repeated
    g.q0 = q[0];
    g.q1 = q[1];
    g.q2 = q[2];
    g.theta = 2.0 * M_PI * U(rng);
    if (g.op == QOA_OP_U2) g.aux = 4 * (rng() % slots);
    if (g.op == QOA_OP_U4) g.aux = rng() % slots;
    c.g.push_back(g);
  }
  return c;
}

static void check_plan(const Circuit& c, const PlanOptions& po, const FusedPlan& P) {
  const uint32_t ng = (uint32_t)c.g.size();
  std::vector<int> seen(ng, 0);
  EXPECT(P.order.size() == ng, "order has %zu of %u gates", P.order.size(), ng);
  for (uint32_t x : P.order) {
    EXPECT(x < ng, "order entry %u", x);
    if (x < ng) seen[x]++;
  }
  for (uint32_t i = 0; i < ng; ++i) EXPECT(seen[i] == 1, "gate %u appears %d times", i, seen[i]);
  uint32_t expect_begin = 0;
  for (const PassPlan& pp : P.passes) {
    EXPECT(pp.op_begin == expect_begin, "pass ranges not contiguous");
    EXPECT(pp.op_end >= pp.op_begin && pp.op_end <= P.ops.size(), "pass range");
    expect_begin = pp.op_end;
    EXPECT(pp.k == P.k && P.k <= c.n && P.k <= 16, "tile k %u", pp.k);
    uint64_t used = 0;
    for (uint32_t j = 0; j < pp.k; ++j) {
      EXPECT(pp.spos[j] < c.n, "spos %u", pp.spos[j]);
      EXPECT(!((used >> pp.spos[j]) & 1ULL), "spos repeated");
      used |= 1ULL << (pp.spos[j] & 63);
      if (j) EXPECT(pp.spos[j] > pp.spos[j - 1], "spos not ascending");
    }
    for (uint32_t o = pp.op_begin; o < pp.op_end && o < P.ops.size(); ++o) {
      const TileOp& op = P.ops[o];
      EXPECT(op.kind <= TK_DIAGN, "op kind %u", op.kind);
      if (op.kind == TK_DIAGN) {
        EXPECT(op.aux + kDiagTermDoubles * op.lmask <= P.pool.size(), "merged terms in the pool");
        continue;
      }
      if (op.kind == TK_LAYOUT) {
        for (int j = 0; j < kRegBits; ++j) EXPECT(op.ins[j] < pp.k, "layout bit %u", op.ins[j]);
        continue;
      }
      if (!pp.reg_mode) {
        EXPECT(op.t < pp.k || op.kind == TK_DNEG || op.kind == TK_DPHASE, "t %u", op.t);
        if (op.nins != 0xFF) {
          EXPECT(op.nins <= 4, "nins %u", op.nins);
          for (int j = 0; j < op.nins && j < 4; ++j) EXPECT(op.ins[j] < pp.k, "ins %u", op.ins[j]);
        }
      }
      EXPECT((op.lmask >> pp.k) == 0, "lmask beyond the tile");
      if (op.kind == TK_MAT4) EXPECT(op.aux + 32 <= P.pool.size(), "MAT4 pool index");
      if (po.flags & kPlanNoise) EXPECT(op.nq <= 3 && op.psel >= 1 && op.psel <= 2, "noise fields");
    }
  }
  EXPECT(expect_begin == P.ops.size(), "ops not covered by passes");
}

static void check_blocks(const Circuit& c, uint32_t bq, const std::vector<Block>& B) {
  std::vector<float> w(kWFloats);
  for (const Block& b : B) {
    for (int j = 0; j < kBlockQ; ++j) {
      EXPECT(b.q[j] < c.n, "block qubit %u", b.q[j]);
      if (j) EXPECT(b.q[j] > b.q[j - 1], "block qubits not ascending");
    }
    EXPECT(b.u.size() == (size_t)kDim * kDim, "block matrix size");
    pack_operands(b, w.data());
  }
  (void)bq;
}

int main(int argc, char** argv) {
  const int rounds = argc > 1 ? std::atoi(argv[1]) : 300;
  std::mt19937_64 rng(20241203);
  long plans = 0, fails_ok = 0, blocks = 0;
  for (int r = 0; r < rounds; ++r) {
    const uint32_t n = 1 + (uint32_t)(rng() % 30);
    const bool clean = (r % 2) == 0;
    Circuit c = random_circuit(rng, n, (uint32_t)(rng() % 300), clean);
    const uint32_t ng = (uint32_t)c.g.size();
    const uint32_t tiles[] = {0, 4, 5, 8, 10, 11, 12, 13};
    for (uint32_t tk : tiles) {
      for (uint32_t variant = 0; variant < 6; ++variant) {
        PlanOptions po{tk, (uint32_t)(rng() % 4), 0u};
        if (variant & 1) po.flags |= QOA_CZ_REF_MASKS;
        if (variant == 2) po.flags |= QOA_FUSE_LDS_ONLY;
        if (variant == 0 && (r % 3) == 0) po.flags |= QOA_FUSE_MERGE_DIAG;
        if (variant >= 3) po.flags |= kPlanForceReg;
        if (variant == 4) po.flags |= kPlanSwaps;
        if (variant == 5 && c.clean) po.flags |= kPlanSwaps | kPlanNoise;
        FusedPlan P;
        const int rc = build_plan(n, c.g.data(), ng, c.mats.data(), po, &P);
        if (rc == QOA_OK) {
          check_plan(c, po, P);
          ++plans;
          std::vector<uint32_t> order;
          uint32_t np = 0;
          if (plan_order(n, c.g.data(), ng, po, &order, &np) == QOA_OK)
            EXPECT(order.size() == ng, "plan_order size");
        } else {
          ++fails_ok;  // refused with an error code: also a valid outcome
        }
      }
    }
    // host-only C ABI of the planner with short buffers
    uint32_t nops = 0, npasses = 0, tk = 0;
    qoa_fuse_opts fo{0, 0, 0, 0};
    (void)qoa_plan_program(n, c.g.data(), ng, c.mats.data(), &fo, nullptr, 0, nullptr, nullptr, 0,
                           &nops, &npasses, &tk);
    if (nops > 0) {
      std::vector<uint8_t> buf((size_t)qoa_tile_op_size() * (nops > 1 ? nops - 1 : 1));
      std::vector<uint8_t> sp(16 * (npasses + 1));
      std::vector<uint32_t> rg(3 * (npasses + 1));  // begin, end, mode per pass
      const int rc = qoa_plan_program(n, c.g.data(), ng, c.mats.data(), &fo, buf.data(),
                                      nops > 1 ? nops - 1 : 0, sp.data(), rg.data(), npasses,
                                      &nops, &npasses, &tk);
      EXPECT(rc != QOA_OK || nops <= 1, "short op buffer accepted");
    }
    if (n >= 5) {
      for (uint32_t bq = 1; bq <= 5; ++bq) {
        std::vector<Block> B;
        const int rc = build_blocks(n, c.g.data(), ng, c.mats.data(), 0, bq, &B);
        if (rc == QOA_OK) {
          check_blocks(c, bq, B);
          blocks += (long)B.size();
        }
      }
      uint32_t nb = 0;
      qoa_block_opts bo{5, 0};
      if (qoa_plan_blocks(n, c.g.data(), ng, c.mats.data(), &bo, nullptr, nullptr, 0, &nb) ==
              QOA_OK && nb > 1) {
        std::vector<uint32_t> qs(5 * (nb - 1));
        EXPECT(qoa_plan_blocks(n, c.g.data(), ng, c.mats.data(), &bo, qs.data(), nullptr, nb - 1,
                               &nb) != QOA_OK,
               "short block buffer accepted");
      }
    }
  }
  std::printf("sanitized host run: %ld plans, %ld refused, %ld blocks, %d failures\n", plans,
              fails_ok, blocks, g_fail);
  return g_fail ? 1 : 0;
}
