// blocks_plan.hpp — the dense-block fuser (host side of the 5-qubit blocks).
#pragma once
#include <complex>
#include <cstdint>
#include <vector>

#include "../../include/qoa_hip.h"

namespace qoa {

using cd = std::complex<double>;
constexpr int kBlockQ = 5;
constexpr int kDim = 1 << kBlockQ;         // 32
constexpr int kWFloats = 64 * 64;          // A operands per block

struct Block {
  uint8_t q[kBlockQ];              // physical qubits, ascending (local bit b = q[b])
  std::vector<cd> u;               // 32x32 row-major, u[i*32+k]
};

// Packs the gates into blocks acting on at most bq (<= 5) qubits each, in an
// order that is a valid reordering of the circuit.
int build_blocks(uint32_t n, const qoa_gate* g, uint32_t ng, const double* mats, uint32_t flags,
                 uint32_t bq, std::vector<Block>* out);
// The block's MFMA A operands (kWFloats floats, step-major).
void pack_operands(const Block& b, float* w);

}  // namespace qoa
