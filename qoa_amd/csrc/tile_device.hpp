// tile_device.hpp — device code shared by the fused and batched pass kernels.
//
// Arithmetic follows the per-gate sweeps exactly (num-complex products, no
// FMA: every includer is built with -ffp-contract=off), so a gate applied in
// a tile and by its own sweep give the same bits.
#pragma once
#include "planner.hpp"
#include "qoa_internal.hpp"

#pragma clang fp contract(off)

namespace qoa {
namespace tile {

constexpr int kThreads = 256;

struct Geo {
  int k;
  uint8_t spos[16];
};

__device__ __forceinline__ double2 cmul(double2 x, double2 y) {
  return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}

// insert zero bits at the op's sorted local positions
__device__ __forceinline__ uint32_t dep(uint32_t x, const TileOp& op) {
  for (int j = 0; j < op.nins; ++j) {
    const int b = op.ins[j];
    x = ((x >> b) << (b + 1)) | (x & ((1u << b) - 1u));
  }
  return x;
}

template <typename I>
__device__ __forceinline__ I scatter(uint32_t l, const Geo& g) {
  I off = 0;
  for (int j = 0; j < g.k; ++j) off |= (I)((l >> j) & 1u) << g.spos[j];
  return off;
}


// ---- register-resident passes -------------------------------------------------
// Thread t of a 2^(k-4)-thread block holds the 16 amplitudes whose local
// indices are thr_l(t) | regl[r], r = 0..15: four "register bits" of the
// local index come from r, the others from t. Ops whose mixed bits are
// register bits (and every diagonal op) run on registers; a TK_LAYOUT op
// changes the register bits through one LDS round trip.
constexpr int kR = kRegAmps;

__device__ __forceinline__ uint32_t lpad(uint32_t l) { return l + (l >> 4); }

template <int J, int KIND>
__device__ __forceinline__ void reg_pair(double2 (&v)[kR], const TileOp& op, uint32_t thr_l,
                                         const uint32_t (&regl)[kR]) {
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    if (((thr_l | regl[r]) & op.lmask) != op.lmask) continue;
    const double2 x = v[r], y = v[r1];
    if (KIND == TK_H) {
      const double2 nf = make_double2(op.c[0], op.c[1]);
      v[r] = cmul(nf, cadd(x, y));
      v[r1] = cmul(nf, csub(x, y));
    } else if (KIND == TK_M) {
      v[r] = cadd(cmul(make_double2(op.c[0], op.c[1]), x), cmul(make_double2(op.c[2], op.c[3]), y));
      v[r1] = cadd(cmul(make_double2(op.c[4], op.c[5]), x), cmul(make_double2(op.c[6], op.c[7]), y));
    } else if (KIND == TK_X) {
      v[r] = y;
      v[r1] = x;
    } else {  // TK_Y
      v[r] = cmul(y, make_double2(0.0, -1.0));
      v[r1] = cmul(x, make_double2(0.0, 1.0));
    }
  }
}

template <int KIND>
__device__ __forceinline__ void reg_pair_j(double2 (&v)[kR], const TileOp& op, uint32_t thr_l,
                                           const uint32_t (&regl)[kR]) {
  switch (op.ins[0]) {
    case 0: reg_pair<0, KIND>(v, op, thr_l, regl); break;
    case 1: reg_pair<1, KIND>(v, op, thr_l, regl); break;
    case 2: reg_pair<2, KIND>(v, op, thr_l, regl); break;
    default: reg_pair<3, KIND>(v, op, thr_l, regl); break;
  }
}

template <int A, int B>
__device__ __forceinline__ void reg_swap(double2 (&v)[kR], const TileOp& op, uint32_t thr_l,
                                         const uint32_t (&regl)[kR]) {
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & ((1 << A) | (1 << B))) continue;
    if (((thr_l | regl[r]) & op.lmask) != op.lmask) continue;
    const double2 u = v[r | (1 << B)];
    v[r | (1 << B)] = v[r | (1 << A)];
    v[r | (1 << A)] = u;
  }
}

// A < B are the register slots of the lower and higher local bit.
template <int A, int B>
__device__ __forceinline__ void reg_mat4(double2 (&v)[kR], const double* u) {
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & ((1 << A) | (1 << B))) continue;
    const int idx[4] = {r, r | (1 << A), r | (1 << B), r | (1 << A) | (1 << B)};
    double2 w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = v[idx[q]];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      double2 acc = cmul(make_double2(u[8 * s], u[8 * s + 1]), w[0]);
#pragma unroll
      for (int q = 1; q < 4; ++q)
        acc = cadd(acc, cmul(make_double2(u[8 * s + 2 * q], u[8 * s + 2 * q + 1]), w[q]));
      v[idx[s]] = acc;
    }
  }
}

// Local index bits of thread t for register bits R (t's bits fill the other
// local positions in ascending order) and the register offsets regl[r].
__device__ __forceinline__ void make_layout(uint32_t tid, int k, const uint8_t* rpos,
                                            uint32_t& thr_l, uint32_t (&regl)[kR]) {
  uint32_t rmask = 0;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) rmask |= 1u << rpos[j];
  uint32_t l = 0, src = 0;
  for (int b = 0; b < k; ++b) {
    if (rmask & (1u << b)) continue;
    l |= ((tid >> src) & 1u) << b;
    ++src;
  }
  thr_l = l;
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < kRegBits; ++j) o |= (uint32_t)((r >> j) & 1) << rpos[j];
    regl[r] = o;
  }
}

// Trajectory noise state of one tile pass (on == 0: no noise).
struct NoiseArgs {
  uint64_t seed;
  double p1, p2;
  uint32_t traj;
  uint32_t on;
  uint64_t frame;  // X frame entering the pass: logical index = physical ^ frame
};

// Philox4x32-10 (Salmon et al., SC'11) with counter (gi, q, traj, const) and
// key = seed: 0 = no error, else 1 + floor(3 * o2 / 2^32) picks X, Y, Z.
__device__ __forceinline__ uint32_t traj_pauli(uint64_t seed, uint32_t traj, uint32_t gi,
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

template <int J, bool Y>
__device__ __forceinline__ void reg_flip(double2 (&v)[kR]) {
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    const double2 x = v[r], y = v[r1];
    if (Y) {
      v[r] = cmul(y, make_double2(0.0, -1.0));
      v[r1] = cmul(x, make_double2(0.0, 1.0));
    } else {
      v[r] = y;
      v[r1] = x;
    }
  }
}

// Pauli `code` (1 X, 2 Y, 3 Z) on local bit b of a register-resident tile,
// with the arithmetic of apply_{x,y,z}_simd (vectorization.rs:942-999).
__device__ __forceinline__ void reg_pauli(double2 (&v)[kR], uint32_t code, uint32_t b,
                                          const uint8_t* rpos, uint32_t thr_l,
                                          const uint32_t (&regl)[kR], double2* lds, bool act) {
  if (code == 3) {
#pragma unroll
    for (int r = 0; r < kR; ++r)
      if (((thr_l | regl[r]) >> b) & 1u) v[r] = make_double2(v[r].x * -1.0, v[r].y * -1.0);
    return;
  }
  int slot = -1;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j)
    if (rpos[j] == b) slot = j;
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
  if (act) {
#pragma unroll
    for (int r = 0; r < kR; ++r) lds[lpad(thr_l | regl[r])] = v[r];
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int r = 0; r < kR; ++r) v[r] = lds[lpad((thr_l | regl[r]) ^ (1u << b))];
  }
  __syncthreads();
  if (code == 2) {
#pragma unroll
    for (int r = 0; r < kR; ++r)
      v[r] = (((thr_l | regl[r]) >> b) & 1u) ? cmul(v[r], make_double2(0.0, 1.0))
                                             : cmul(v[r], make_double2(0.0, -1.0));
  }
}

// One register-resident pass over the tile with base `base` (bits of S clear).
// Threads with act == false (beyond 2^(k-4)) only take part in barriers.
// With trajectory noise, errors on qubits of the tile are applied in place;
// an X or Y on a qubit outside the tile flips that bit of the X frame
// (its phase, a tile constant, is applied at once), and the frame bits that
// fall inside the tile are applied when the tile is loaded. Returns the frame
// leaving the pass.
template <typename I>
__device__ __forceinline__ uint64_t reg_tile(double2* __restrict__ a, I base, const Geo& geo,
                                             const TileOp* __restrict__ ops, int nops,
                                             const double* __restrict__ pool, double2* lds,
                                             bool act, bool init_zero, const NoiseArgs& nz) {
  const int k = geo.k;
  const uint32_t tid = threadIdx.x;
  uint8_t rpos[kRegBits];
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) rpos[j] = (uint8_t)(k - kRegBits + j);
  uint32_t thr_l;
  uint32_t regl[kR];
  make_layout(tid, k, rpos, thr_l, regl);
  const I gthr = scatter<I>(thr_l, geo);
  double2 v[kR];
  if (init_zero) {  // the state starts as |0...0>: nothing to read
#pragma unroll
    for (int r = 0; r < kR; ++r)
      v[r] = make_double2((base | gthr | scatter<I>(regl[r], geo)) == 0 ? 1.0 : 0.0, 0.0);
  } else if (act) {
#pragma unroll
    for (int r = 0; r < kR; ++r) v[r] = a[base | gthr | scatter<I>(regl[r], geo)];
  }
  uint64_t frame = nz.on ? nz.frame : 0ULL;
  if (frame) {
    uint64_t smask = 0;
    uint32_t xl = 0;
    for (int j = 0; j < k; ++j) {
      smask |= 1ULL << geo.spos[j];
      if ((frame >> geo.spos[j]) & 1ULL) xl |= 1u << j;
    }
    if (xl) {  // frame bits inside the tile: apply those X flips now
      if (act) {
#pragma unroll
        for (int r = 0; r < kR; ++r) lds[lpad(thr_l | regl[r])] = v[r];
      }
      __syncthreads();
      if (act) {
#pragma unroll
        for (int r = 0; r < kR; ++r) v[r] = lds[lpad((thr_l | regl[r]) ^ xl)];
      }
      __syncthreads();
      frame &= ~smask;
    }
  }

  for (int o = 0; o < nops; ++o) {
    const TileOp& op = ops[o];
    if ((((uint64_t)base ^ frame) & op.gmask) == op.gval) {  // uniform over the tile
      switch (op.kind) {
        case TK_LAYOUT: {
          if (act) {
#pragma unroll
            for (int r = 0; r < kR; ++r) lds[lpad(thr_l | regl[r])] = v[r];
          }
          __syncthreads();
#pragma unroll
          for (int j = 0; j < kRegBits; ++j) rpos[j] = op.ins[j];
          make_layout(tid, k, rpos, thr_l, regl);
          if (act) {
#pragma unroll
            for (int r = 0; r < kR; ++r) v[r] = lds[lpad(thr_l | regl[r])];
          }
          __syncthreads();
        } break;
        case TK_H: reg_pair_j<TK_H>(v, op, thr_l, regl); break;
        case TK_M: reg_pair_j<TK_M>(v, op, thr_l, regl); break;
        case TK_X: reg_pair_j<TK_X>(v, op, thr_l, regl); break;
        case TK_Y: reg_pair_j<TK_Y>(v, op, thr_l, regl); break;
        case TK_DNEG: {
#pragma unroll
          for (int r = 0; r < kR; ++r)
            if (((thr_l | regl[r]) & op.lmask) == op.lmask)
              v[r] = make_double2(v[r].x * -1.0, v[r].y * -1.0);
        } break;
        case TK_DPHASE: {
          const double2 f = make_double2(op.c[0], op.c[1]);
#pragma unroll
          for (int r = 0; r < kR; ++r)
            if (((thr_l | regl[r]) & op.lmask) == op.lmask) v[r] = cmul(v[r], f);
        } break;
        case TK_RZ: {
          const double2 m00 = make_double2(op.c[0], op.c[1]);
          const double2 m11 = make_double2(op.c[2], op.c[3]);
#pragma unroll
          for (int r = 0; r < kR; ++r)
            v[r] = (((thr_l | regl[r]) >> op.t) & 1u) ? cmul(v[r], m11) : cmul(v[r], m00);
        } break;
        case TK_SWAP: {
          const int a0 = op.ins[0] < op.ins[1] ? op.ins[0] : op.ins[1];
          const int b0 = op.ins[0] < op.ins[1] ? op.ins[1] : op.ins[0];
          switch (a0 * 4 + b0) {
            case 1: reg_swap<0, 1>(v, op, thr_l, regl); break;
            case 2: reg_swap<0, 2>(v, op, thr_l, regl); break;
            case 3: reg_swap<0, 3>(v, op, thr_l, regl); break;
            case 6: reg_swap<1, 2>(v, op, thr_l, regl); break;
            case 7: reg_swap<1, 3>(v, op, thr_l, regl); break;
            default: reg_swap<2, 3>(v, op, thr_l, regl); break;
          }
        } break;
        case TK_MAT4: {
          const int a0 = op.ins[0] < op.ins[1] ? op.ins[0] : op.ins[1];
          const int b0 = op.ins[0] < op.ins[1] ? op.ins[1] : op.ins[0];
          const double* u = pool + op.aux;
          switch (a0 * 4 + b0) {
            case 1: reg_mat4<0, 1>(v, u); break;
            case 2: reg_mat4<0, 2>(v, u); break;
            case 3: reg_mat4<0, 3>(v, u); break;
            case 6: reg_mat4<1, 2>(v, u); break;
            case 7: reg_mat4<1, 3>(v, u); break;
            default: reg_mat4<2, 3>(v, u); break;
          }
        } break;
        default:
          break;
      }
    }
    if (nz.on && op.nq) {  // trajectory errors after the gate, operand order
      const double p = op.psel == 1 ? nz.p1 : nz.p2;
      for (int j = 0; j < op.nq; ++j) {
        const uint32_t q = op.qphys[j];
        const uint32_t code = traj_pauli(nz.seed, nz.traj, op.gi, q, p);
        if (!code) continue;
        int lb = -1;
        for (int s = 0; s < k; ++s)
          if (geo.spos[s] == q) lb = s;
        if (lb >= 0) {
          reg_pauli(v, code, (uint32_t)lb, rpos, thr_l, regl, lds, act);
        } else if (code == 3) {  // Z on an outside qubit: a tile-wide sign
          if ((((uint64_t)base ^ frame) >> q) & 1ULL) {
#pragma unroll
            for (int r = 0; r < kR; ++r) v[r] = make_double2(v[r].x * -1.0, v[r].y * -1.0);
          }
        } else {
          frame ^= 1ULL << q;
          if (code == 2) {  // Y: phase -i / +i by the new logical bit
            const double2 f = ((((uint64_t)base ^ frame) >> q) & 1ULL) ? make_double2(0.0, 1.0)
                                                                        : make_double2(0.0, -1.0);
#pragma unroll
            for (int r = 0; r < kR; ++r) v[r] = cmul(v[r], f);
          }
        }
      }
    }
  }
  if (act) {
#pragma unroll
    for (int r = 0; r < kR; ++r) a[base | gthr | scatter<I>(regl[r], geo)] = v[r];
  }
  return frame;
}

}  // namespace tile
}  // namespace qoa
