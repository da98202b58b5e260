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

// One register-resident pass over the tile with base `base` (bits of S clear).
// Threads with act == false (beyond 2^(k-4)) only take part in barriers.
template <typename I>
__device__ __forceinline__ void reg_tile(double2* __restrict__ a, I base, const Geo& geo,
                                         const TileOp* __restrict__ ops, int nops,
                                         const double* __restrict__ pool, double2* lds, bool act,
                                         bool init_zero) {
  const int k = geo.k;
  const uint32_t tid = threadIdx.x;
  // Start layout: the top 4 local bits in registers, the thread index on the
  // low local bits (HBM runs stay contiguous across lanes).
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

  for (int o = 0; o < nops; ++o) {
    const TileOp& op = ops[o];
    if (((uint64_t)base & op.gmask) != op.gval) continue;  // uniform over the tile
    switch (op.kind) {
      case TK_LAYOUT: {
        if (act) {
#pragma unroll
          for (int r = 0; r < kR; ++r) lds[lpad(thr_l | regl[r])] = v[r];
        }
        __syncthreads();
        uint8_t np[kRegBits];
#pragma unroll
        for (int j = 0; j < kRegBits; ++j) np[j] = op.ins[j];
        make_layout(tid, k, np, thr_l, regl);
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
        const double2 m00 = make_double2(op.c[0], op.c[1]), m11 = make_double2(op.c[2], op.c[3]);
#pragma unroll
        for (int r = 0; r < kR; ++r)
          v[r] = (((thr_l | regl[r]) >> op.t) & 1u) ? cmul(v[r], m11) : cmul(v[r], m00);
      } break;
      case TK_SWAP: {
        const int a0 = op.ins[0] < op.ins[1] ? op.ins[0] : op.ins[1];
        const int b0 = op.ins[0] < op.ins[1] ? op.ins[1] : op.ins[0];
        const int key = a0 * 4 + b0;
        switch (key) {
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
  if (act) {
#pragma unroll
    for (int r = 0; r < kR; ++r) a[base | gthr | scatter<I>(regl[r], geo)] = v[r];
  }
}


}  // namespace tile
}  // namespace qoa
