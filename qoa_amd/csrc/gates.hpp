// gates.hpp — one gate of the C ABI as a canonical action (host side).
//
// Shared by the pass planner and the dense-block fuser: every gate becomes
// at most one action on the state with the layer-1 semantics of the per-gate
// entry points (out-of-range qubits are no-ops, RZ on an out-of-range qubit
// is the global factor m00, CZ/CPHASE masks per QOA_CZ_REF_MASKS).
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/qoa_hip.h"

namespace qoa {

enum FK { FK_H, FK_M, FK_X, FK_Y, FK_DNEG, FK_DPHASE, FK_RZ, FK_SWAP, FK_MAT4 };

struct FOp {
  int kind = 0;
  int t = -1, t2 = -1;  // mixed (active) qubits
  uint64_t pmask = 0;   // read-only qubits that must be 1
  int rzq = -1;         // RZ: read-only qubit whose bit picks the factor
  double c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double* m4 = nullptr;
  uint32_t gi = 0;
  uint64_t active = 0, all = 0;
  // Qubits on which this action does not commute with diagonal actions
  // (its mixed qubits; every qubit when trajectory noise follows the gate).
  // Two actions that overlap only on qubits where both act diagonally
  // commute, which is what the schedulers use to let gates pass each other.
  uint64_t mixb = 0;
  uint8_t nq = 0, psel = 1;  // trajectory-noise operands (the gate's qubits)
  uint8_t qphys[3] = {0, 0, 0};
};

// Appends the action of gate g (index gi) to out; a no-op gate appends
// nothing. Returns QOA_OK or an error code (message set).
int canonical_gate(uint32_t n, const qoa_gate& g, const double* mats, uint32_t flags,
                   std::vector<FOp>* out, uint32_t gi);

}  // namespace qoa
