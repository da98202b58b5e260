// traj_noise.hpp — the trajectory-noise draw shared by the pass kernels, the
// draw-table kernel and the host's plain per-gate trajectory loop.
#pragma once
#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#endif

#include <cstdint>

#include "../../include/qoa_hip.h"

namespace qoa {

// Philox4x32-10 (Salmon et al., SC'11) with counter (gi, q, traj, const) and
// key = seed: 0 = no error, else 1 + floor(3 * o2 / 2^32) picks X, Y, Z.
__host__ __device__ inline uint32_t traj_pauli(uint64_t seed, uint32_t traj, uint32_t gi,
                                               uint32_t q, double p) {
  uint32_t c0 = gi, c1 = q, c2 = traj, c3 = 0x7A3C5E1Fu;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  const double u = (double)((((uint64_t)c1 << 32) | c0) >> 11) * 0x1.0p-53;
  if (!(u < p)) return 0u;
  return 1u + (uint32_t)(((uint64_t)c2 * 3ULL) >> 32);
}

// Operands that take errors after a gate: one-qubit gates use p1 on q0,
// two-qubit gates p2 on (q0, q1), three-qubit gates p2 on (q0, q1, q2).
__host__ __device__ inline uint32_t traj_arity(uint32_t op) {
  switch (op) {
    case QOA_OP_CNOT: case QOA_OP_CZ: case QOA_OP_CPHASE: case QOA_OP_SWAP:
    case QOA_OP_U4: case QOA_OP_CY: return 2;
    case QOA_OP_CSWAP: case QOA_OP_TOFFOLI: return 3;
    default: return 1;
  }
}

}  // namespace qoa
