// tile_device.hpp — device code shared by the fused and batched pass kernels.
//
// Arithmetic follows the per-gate sweeps exactly (num-complex products, no
// FMA: every includer is built with -ffp-contract=off), so a gate applied in
// a tile and by its own sweep give the same bits. The register-resident pass
// also has an FMA instantiation (F = true) whose complex products use fused
// multiply-adds; it is selected per call and is not bit-identical.
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

__device__ __forceinline__ uint32_t lpad(uint32_t l) { return l + (l >> 4); }

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

template <int J, int KIND, bool F>
__device__ __forceinline__ void reg_pair(double2 (&v)[kR], const TileOp& op) {
  const uint32_t rsel = op.rsel;
  const double2 c0 = make_double2(op.c[0], op.c[1]), c1 = make_double2(op.c[2], op.c[3]);
  const double2 c2 = make_double2(op.c[4], op.c[5]), c3 = make_double2(op.c[6], op.c[7]);
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & (1 << J)) continue;
    if (((uint32_t)r & rsel) != rsel) continue;  // uniform
    const int r1 = r | (1 << J);
    const double2 x = v[r], y = v[r1];
    if (KIND == kHReal) {
      // equals cmul((nf, 0), .) up to the sign of zero terms
      const double nf = c0.x;
      v[r] = make_double2(nf * (x.x + y.x), nf * (x.y + y.y));
      v[r1] = make_double2(nf * (x.x - y.x), nf * (x.y - y.y));
    } else if (KIND == TK_H) {
      v[r] = mulc<F>(c0, cadd(x, y));
      v[r1] = mulc<F>(c0, csub(x, y));
    } else if (KIND == TK_M) {
      v[r] = mul2c<F>(c0, x, c1, y);
      v[r1] = mul2c<F>(c2, x, c3, y);
    } else if (KIND == TK_X) {
      v[r] = y;
      v[r1] = x;
    } else {  // TK_Y
      v[r] = cmul(y, make_double2(0.0, -1.0));
      v[r1] = cmul(x, make_double2(0.0, 1.0));
    }
  }
}

template <int KIND, bool F>
__device__ __forceinline__ void reg_pair_j(double2 (&v)[kR], const TileOp& op) {
  switch (op.ins[0]) {
    case 0: reg_pair<0, KIND, F>(v, op); break;
    case 1: reg_pair<1, KIND, F>(v, op); break;
    case 2: reg_pair<2, KIND, F>(v, op); break;
    default: reg_pair<3, KIND, F>(v, op); break;
  }
}

template <int A, int B>
__device__ __forceinline__ void reg_swap(double2 (&v)[kR], uint32_t rsel) {
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & ((1 << A) | (1 << B))) continue;
    if (((uint32_t)r & rsel) != rsel) continue;
    const double2 u = v[r | (1 << B)];
    v[r | (1 << B)] = v[r | (1 << A)];
    v[r | (1 << A)] = u;
  }
}

// A < B are the register slots of the lower and higher local bit.
template <int A, int B, bool F>
__device__ __forceinline__ void reg_mat4(double2 (&v)[kR], const double* u) {
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r & ((1 << A) | (1 << B))) continue;
    const double2 w0 = v[r], w1 = v[r | (1 << A)], w2 = v[r | (1 << B)];
    const double2 w3 = v[r | (1 << A) | (1 << B)];
    double2 o[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      double2 acc = mulc<F>(make_double2(u[8 * s], u[8 * s + 1]), w0);
      acc = macc<F>(acc, make_double2(u[8 * s + 2], u[8 * s + 3]), w1);
      acc = macc<F>(acc, make_double2(u[8 * s + 4], u[8 * s + 5]), w2);
      acc = macc<F>(acc, make_double2(u[8 * s + 6], u[8 * s + 7]), w3);
      o[s] = acc;
    }
    v[r] = o[0];
    v[r | (1 << A)] = o[1];
    v[r | (1 << B)] = o[2];
    v[r | (1 << A) | (1 << B)] = o[3];
  }
}

// Writes the registers to LDS under layout L (padded rows: no bank conflicts
// for the strided layouts).
__device__ __forceinline__ void regs_to_lds(const double2 (&v)[kR], const RegLayout& L,
                                            double2* lds, uint32_t xl = 0) {
#pragma unroll
  for (int r = 0; r < kR; ++r) lds[lpad((L.thr_l | regl(L, r)) ^ xl)] = v[r];
}

__device__ __forceinline__ void lds_to_regs(double2 (&v)[kR], const RegLayout& L,
                                            const double2* lds) {
#pragma unroll
  for (int r = 0; r < kR; ++r) v[r] = lds[lpad(L.thr_l | regl(L, r))];
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
                                          const RegLayout& L, double2* lds, bool act) {
  if (code == 3) {
#pragma unroll
    for (int r = 0; r < kR; ++r)
      if (((L.thr_l | regl(L, r)) >> b) & 1u) v[r] = make_double2(v[r].x * -1.0, v[r].y * -1.0);
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
  if (act) regs_to_lds(v, L, lds, 1u << b);
  __syncthreads();
  if (act) lds_to_regs(v, L, lds);
  __syncthreads();
  if (code == 2) {
    const double2 f = ((L.thr_l >> b) & 1u) ? make_double2(0.0, 1.0) : make_double2(0.0, -1.0);
#pragma unroll
    for (int r = 0; r < kR; ++r) v[r] = cmul(v[r], f);
  }
}

// Applies one non-layout op whose tile condition holds to this thread's
// registers (the thread-bit controls have been checked by the caller).
template <bool F>
__device__ __forceinline__ void reg_op(double2 (&v)[kR], const TileOp& op, const RegLayout& L,
                                       const double* __restrict__ pool) {
  switch (op.kind) {
    case TK_H:
      if (op.c[1] == 0.0) reg_pair_j<kHReal, F>(v, op);
      else reg_pair_j<TK_H, F>(v, op);
      break;
    case TK_M: reg_pair_j<TK_M, F>(v, op); break;
    case TK_X: reg_pair_j<TK_X, F>(v, op); break;
    case TK_Y: reg_pair_j<TK_Y, F>(v, op); break;
    case TK_DNEG: {
      const uint32_t rsel = op.rsel;
#pragma unroll
      for (int r = 0; r < kR; ++r)
        if (((uint32_t)r & rsel) == rsel) v[r] = make_double2(v[r].x * -1.0, v[r].y * -1.0);
    } break;
    case TK_DPHASE: {
      const uint32_t rsel = op.rsel;
      const double2 f = make_double2(op.c[0], op.c[1]);
#pragma unroll
      for (int r = 0; r < kR; ++r)
        if (((uint32_t)r & rsel) == rsel) v[r] = mulc<F>(v[r], f);
    } break;
    case TK_RZ: {
      const double2 m00 = make_double2(op.c[0], op.c[1]);
      const double2 m11 = make_double2(op.c[2], op.c[3]);
      if (op.rt != 0xFF) {  // target held in register slot rt: the factor is uniform per r
        const uint32_t j = op.rt;
#pragma unroll
        for (int r = 0; r < kR; ++r) v[r] = mulc<F>(v[r], (((uint32_t)r >> j) & 1u) ? m11 : m00);
      } else {
        const double2 f = ((L.thr_l >> op.t) & 1u) ? m11 : m00;
#pragma unroll
        for (int r = 0; r < kR; ++r) v[r] = mulc<F>(v[r], f);
      }
    } break;
    case TK_SWAP: {
      const int a0 = op.ins[0] < op.ins[1] ? op.ins[0] : op.ins[1];
      const int b0 = op.ins[0] < op.ins[1] ? op.ins[1] : op.ins[0];
      const uint32_t rsel = op.rsel;
      switch (a0 * 4 + b0) {
        case 1: reg_swap<0, 1>(v, rsel); break;
        case 2: reg_swap<0, 2>(v, rsel); break;
        case 3: reg_swap<0, 3>(v, rsel); break;
        case 6: reg_swap<1, 2>(v, rsel); break;
        case 7: reg_swap<1, 3>(v, rsel); break;
        default: reg_swap<2, 3>(v, rsel); break;
      }
    } break;
    case TK_MAT4: {
      const int a0 = op.ins[0] < op.ins[1] ? op.ins[0] : op.ins[1];
      const int b0 = op.ins[0] < op.ins[1] ? op.ins[1] : op.ins[0];
      const double* u = pool + op.aux;
      switch (a0 * 4 + b0) {
        case 1: reg_mat4<0, 1, F>(v, u); break;
        case 2: reg_mat4<0, 2, F>(v, u); break;
        case 3: reg_mat4<0, 3, F>(v, u); break;
        case 6: reg_mat4<1, 2, F>(v, u); break;
        case 7: reg_mat4<1, 3, F>(v, u); break;
        default: reg_mat4<2, 3, F>(v, u); break;
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
// leaving the pass.
template <typename I, bool F>
__device__ __forceinline__ uint64_t reg_tile(double2* __restrict__ a, I base, const Geo& geo,
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
  // HBM addresses: base | thread part | register part (uniform per r)
  I gthr = 0;
  for (int j = 0; j < k - kRegBits; ++j) gthr |= (I)((tid >> j) & 1u) << geo.spos[j];
  I gb[kRegBits];
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) gb[j] = (I)1 << geo.spos[k - kRegBits + j];
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
    for (int r = 0; r < kR; ++r) v[r] = a[gaddr(r)];
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
      if (act) regs_to_lds(v, L, lds, xl);
      __syncthreads();
      if (act) lds_to_regs(v, L, lds);
      __syncthreads();
      frame &= ~smask;
    }
  }

  for (int o = 0; o < nops; ++o) {
    const TileOp& op = ops[o];
    if (op.kind == TK_LAYOUT) {
      if (act) regs_to_lds(v, L, lds);
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kRegBits; ++j) rpos[j] = op.ins[j];
      set_layout(L, tid, k, rpos);
      if (act) lds_to_regs(v, L, lds);
      __syncthreads();
    } else if ((((uint64_t)base ^ frame) & op.gmask) == op.gval) {  // uniform over the tile
      if ((L.thr_l & op.tl) == op.tl) reg_op<F>(v, op, L, pool);
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
          reg_pauli(v, code, (uint32_t)lb, L, lds, act);
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
    for (int r = 0; r < kR; ++r) a[gaddr(r)] = v[r];
  }
  return frame;
}

}  // namespace tile
}  // namespace qoa
