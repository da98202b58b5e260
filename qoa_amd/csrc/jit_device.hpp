// jit_device.hpp — the register pass of tile_device.hpp unrolled over one
// pass's op program, for kernels compiled at run time (hipRTC) per circuit.
//
// A specialised kernel is a straight sequence of calls, one per op: the
// primitive reg_op (or the exchange) would reach for that op's words, with
// the words as template arguments and literal constants, so reg_tile's op
// dispatch, loop and op-word prefetch are gone while every op runs the same
// arithmetic as in the op-loop kernel (the results are bit-identical to it).
// Coefficients stay in the op records and the pool.
#pragma once
#include "tile_device.hpp"

namespace qoa {
namespace tile {

// The tile a block holds: its amplitudes, register layout and base.
template <typename I>
struct SpecTile {
  double2 v[kR];
  RegLayout L;
  uint8_t rpos[kRegBits];
  I base;
  uint32_t tid;
  int k;
  Xchg X;
};

// The tile a block takes: the hardware hands consecutive blocks to the eight
// XCDs in turn, so with a multiple of eight tiles block b takes tile
// (b mod 8) * tiles/8 + b/8: each XCD (and its L2) works through one
// contiguous eighth of the tiles in order (measured 2 % faster per pass on
// the 28-qubit layered circuit than tile b).
__device__ __forceinline__ uint32_t spec_tile_index() {
  const uint32_t b = blockIdx.x, g = gridDim.x;
  return (g & 7u) ? b : (b & 7u) * (g >> 3) + (b >> 3);
}

// Loads the tile of block `tile` in the initial layout (the top kRegBits
// local bits in registers), as reg_tile does.
template <typename I>
__device__ __forceinline__ void spec_load(SpecTile<I>& S, double2* __restrict__ a, const Geo& geo,
                                          double2* lds, I tile) {
  const int k = geo.k;
  S.k = k;
  S.tid = threadIdx.x;
  S.X.lds = lds;
  const I base = tile_base(tile, geo);
  S.base = base;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) S.rpos[j] = (uint8_t)(k - kRegBits + j);
  set_layout(S.L, S.tid, k, S.rpos);
  I gthr = 0;
  for (int j = 0; j < k - kRegBits; ++j) gthr |= (I)((S.tid >> j) & 1u) << spos_at(geo, j);
  I gb[kRegBits];
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) gb[j] = (I)1 << spos_at(geo, k - kRegBits + j);
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    I o = base | gthr;
#pragma unroll
    for (int j = 0; j < kRegBits; ++j)
      if ((r >> j) & 1) o |= gb[j];
    S.v[r] = ld_stream(a + o);
  }
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
}

// Stores the tile from its final layout (logical register r in physical
// register pm(pi, r)). The thread's local bits pass through an empty asm
// statement: in a pass without layout changes the store addresses equal the
// load addresses, and the compiler would otherwise keep those 16 addresses
// live (32 VGPRs) through the whole pass.
template <typename I>
__device__ __forceinline__ void spec_store(SpecTile<I>& S, double2* __restrict__ a,
                                           const Geo& geo, uint64_t pi) {
  const int k = S.k;
  uint32_t thr = S.L.thr_l;
  asm volatile("" : "+v"(thr));
  I gthr = 0;
  for (int j = 0; j < k; ++j) gthr |= (I)((thr >> j) & 1u) << stpos_at(geo, j);
  I gb[kRegBits];
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) gb[j] = (I)1 << stpos_at(geo, S.rpos[j]);
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    I o = S.base | gthr;
#pragma unroll
    for (int j = 0; j < kRegBits; ++j)
      if ((r >> j) & 1) o |= gb[j];
    st_stream(a + o, S.v[pm(pi, r)]);
  }
}

// The calls a generated kernel makes per op are reg_op's, chosen on the host
// from the op's words (specialize.cpp mirrors reg_op's decisions), with these
// helpers for what they read at run time.

// Coefficients of op record p (c[0..7] as four complex values).
__device__ __forceinline__ Coef4 spec_coef4(const TileOp* __restrict__ p) {
  Coef4 k;
#pragma unroll
  for (int i = 0; i < 4; ++i) k.c[i] = make_double2(p->c[2 * i], p->c[2 * i + 1]);
  return k;
}
__device__ __forceinline__ double2 spec_c2(const TileOp* __restrict__ p, int i) {
  return make_double2(p->c[2 * i], p->c[2 * i + 1]);
}

// This thread's local bits hold tl (the op's thread-bit controls).
template <typename I>
__device__ __forceinline__ bool spec_tok(const SpecTile<I>& S, uint32_t tl) {
  return (S.L.thr_l & tl) == tl;
}

// RZ whose target is thread bit t: m11 where this thread's bit t is 1.
template <bool F, typename I>
__device__ __forceinline__ void spec_rz_thread(SpecTile<I>& S, double2 m00, double2 m11,
                                               uint32_t t) {
  const double2 f = sel2(((S.L.thr_l >> t) & 1u) != 0, m11, m00);
#pragma unroll
  for (int r = 0; r < kR; ++r) ip_mulc<F>(S.v[r], f);
}

// Whole-tile exchange op p to the register bits in w1 (a byte per slot),
// rounds keyed by slots A (and B), Q = reg_xchg_split(k, false); logical
// register r is physical register pm(pi, r) on both sides.
template <int Q, int A, int B, typename I>
__device__ __forceinline__ void spec_layout(SpecTile<I>& S, const TileOp* __restrict__ p,
                                            uint32_t w1, uint64_t pi) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(p) + kCoefWord;
#pragma unroll
  for (int j = 0; j < kRegBits; ++j) S.rpos[j] = (uint8_t)(w1 >> (8 * j));
  RegLayout N;
  set_layout_sorted(N, S.tid, S.rpos, w[kXchgSorted]);
  xchg_op<Q, A, B>(S.v, S.L, N, w, S.X, true, pi);
  S.L = N;
}

// A merged diagonal run term by term: the factor at c applies where the tile
// condition (base & gm) == gv holds and the thread's local bits hold tl; a
// group's product then multiplies the registers whose slots rs are all 1.
// A term with a tile condition multiplies by (1, 0) where it does not hold
// (exactly the value the op loop's skipped product leaves; no branch, so the
// straight-line kernel keeps its register assignment across the term).
template <typename I>
__device__ __forceinline__ void spec_term(SpecTile<I>& S, double2& acc,
                                          const double* __restrict__ c, const uint32_t tl,
                                          const uint64_t gm, const uint64_t gv) {
  const bool t = gm == 0 || ((uint64_t)S.base & gm) == gv;  // uniform
  const double cr = t ? c[0] : 1.0, ci = t ? c[1] : 0.0;
  if (tl == 0) {
    ip_mulc<true, true>(acc, make_double2(cr, ci));
  } else {
    const bool on = (S.L.thr_l & tl) == tl;
    ip_mulc<true>(acc, make_double2(on ? cr : 1.0, on ? ci : 0.0));
  }
}

template <int RS, typename I>
__device__ __forceinline__ void spec_group(SpecTile<I>& S, double2& acc, uint64_t pi) {
  reg_phase<RS, true, false>(S.v, acc, pi);
  acc = make_double2(1.0, 0.0);
}

}  // namespace tile
}  // namespace qoa
