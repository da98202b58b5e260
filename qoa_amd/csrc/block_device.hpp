// block_device.hpp — a dense 5-qubit block applied on FP32 matrix cores
// (device side, shared by the single-state and the batched block kernels).
//
// The state viewed as 32 x 2^(n-5) (rows: the block's qubits, columns: the
// other bits) is multiplied in place by U. The complex product is the real
// GEMM [Yr; Yi] = [Ur -Ui; Ui Ur] [Xr; Xi] (M = K = 64) on
// v_mfma_f32_32x32x2f32: one wave owns 32 columns; lane l holds column
// l % 32 and the rows of parity l / 32, so its 16 complex amplitudes are
// exactly the B operands of the 32 K-steps. Each workgroup stages the block's
// A operands (16 KiB, laid out step-major so LDS reads are conflict-free).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "blocks_plan.hpp"

namespace qoa {

constexpr int kColGroups = 4;              // 32-column groups per wave
constexpr int kBlockThreads = 256;         // 4 waves
constexpr uint64_t kColsPerWg = (uint64_t)(kBlockThreads / 64) * 32 * kColGroups;  // 512

typedef float f32x16 __attribute__((ext_vector_type(16)));

struct BlockArgs {
  uint64_t ncols;  // 2^(n-5)
  uint8_t q[kBlockQ];
};

template <typename I>
__device__ __forceinline__ I deposit5(I x, const BlockArgs& ba) {
#pragma unroll
  for (int j = 0; j < kBlockQ; ++j) {
    const int b = ba.q[j];
    x = ((x >> b) << (b + 1)) | (x & ((I(1) << b) - I(1)));
  }
  return x;
}

template <typename I>
__device__ __forceinline__ I rowoff(uint32_t r, const BlockArgs& ba) {
  I o = 0;
#pragma unroll
  for (int j = 0; j < kBlockQ; ++j) o |= (I)((r >> j) & 1u) << ba.q[j];
  return o;
}

// One workgroup's share of a block application: stages the block's A
// operands w in LDS (kWFloats floats), then its waves multiply columns
// wg * 512 .. wg * 512 + 511 of the state a (32 x ncols view). Every thread
// of the workgroup calls it (it has a barrier).
template <typename I>
__device__ __forceinline__ void block_apply(float2* __restrict__ a, const float* __restrict__ w,
                                            const BlockArgs& ba, uint64_t wg) {
  __shared__ float sw[kWFloats];
  for (int i = threadIdx.x; i < kWFloats / 4; i += kBlockThreads)
    reinterpret_cast<float4*>(sw)[i] = reinterpret_cast<const float4*>(w)[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, j = lane & 31;
  const uint64_t wave_col0 =
      (wg * (kBlockThreads / 64) + wave) * (uint64_t)(32 * kColGroups);
  // the column groups of this wave, loads of group g+1 issued before the
  // matrix products of group g
  auto load_group = [&](int g, float2 (&x)[16], I& base, bool& ok) {
    const uint64_t c = wave_col0 + (uint64_t)g * 32 + j;
    ok = c < ba.ncols;
    base = deposit5<I>((I)(ok ? c : 0), ba);
#pragma unroll
    for (int s = 0; s < 16; ++s)
      x[s] = ok ? a[base | rowoff<I>((uint32_t)(2 * s + h), ba)] : make_float2(0.f, 0.f);
  };
  if (wave_col0 >= ba.ncols) return;  // uniform over the wave (no barrier follows)
  float2 cur[16], nxt[16];
  I base, nbase = 0;
  bool ok, nok = false;
  load_group(0, cur, base, ok);
  for (int g = 0; g < kColGroups; ++g) {
    // The A operands are read from LDS again for every column group: hoisted
    // out of this loop they take 64 more VGPRs per lane (190 instead of 130,
    // two waves per SIMD instead of three), and the block took 3821 instead
    // of 3585 us at 30 qubits.
    asm volatile("" ::: "memory");
    const bool more = g + 1 < kColGroups && wave_col0 + (uint64_t)(g + 1) * 32 < ba.ncols;
    if (more) load_group(g + 1, nxt, nbase, nok);
    f32x16 acc0 = {}, acc1 = {};
#pragma unroll
    for (int s = 0; s < 32; ++s) {
      const float b = s < 16 ? cur[s].x : cur[s - 16].y;
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(sw[s * 64 + lane], b, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(sw[(32 + s) * 64 + lane], b, acc1, 0, 0, 0);
    }
    if (ok) {
#pragma unroll
      for (int r = 0; r < 16; ++r)
        a[base | rowoff<I>((uint32_t)((r & 3) + 8 * (r >> 2) + 4 * h), ba)] =
            make_float2(acc0[r], acc1[r]);
    }
    if (!more) break;
#pragma unroll
    for (int s = 0; s < 16; ++s) cur[s] = nxt[s];
    base = nbase;
    ok = nok;
  }
}

}  // namespace qoa
