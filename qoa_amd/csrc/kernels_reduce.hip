// kernels_reduce.hip — norm / probability reductions, renormalisation,
// measurement collapse and reset for FP64 states on gfx950.
//
// Reductions use a fixed grid (kRedBlocks blocks of 256) and a fixed tree:
// per-thread strided sums, a 64-lane shuffle tree, then the 4 wavefront sums
// in order; a single-block kernel adds the block partials in a fixed order.
// The result therefore depends only on the state length, never on timing or
// on what else runs on the device.
#include "qoa_internal.hpp"

#pragma clang fp contract(off)

namespace qoa {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off);
  return v;
}

// Returns the block total in thread 0 (fixed order).
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) t = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  __syncthreads();
  return t;
}

// NaN / Inf anywhere in the state: bit 0 / bit 1 of *flags (cleared first by
// k_clear_flags; one atomic OR per wave that saw any).
__global__ void k_clear_flags(unsigned int* flags) { *flags = 0u; }

template <typename I>
__global__ __launch_bounds__(kBlock) void k_nonfinite(const double2* __restrict__ a, I len,
                                                      unsigned int* __restrict__ flags) {
  unsigned int f = 0u;
  for (I i = (I)blockIdx.x * kBlock + threadIdx.x; i < len; i += (I)gridDim.x * kBlock) {
    const double2 v = a[i];
    if (__builtin_isnan(v.x) || __builtin_isnan(v.y)) f |= 1u;
    if (__builtin_isinf(v.x) || __builtin_isinf(v.y)) f |= 2u;
  }
  const bool any_nan = __ballot((f & 1u) != 0u) != 0ULL, any_inf = __ballot((f & 2u) != 0u) != 0ULL;
  if ((threadIdx.x & 63u) == 0u && (any_nan || any_inf))
    atomicOr(flags, (any_nan ? 1u : 0u) | (any_inf ? 2u : 0u));
}

// sum |a_i|^2 split by bit `selbit` (selbit == 0: everything goes to acc0).
template <typename I>
__global__ __launch_bounds__(kBlock) void k_norm_partials(const double2* __restrict__ a, I len,
                                                          I selbit, double* __restrict__ part) {
  __shared__ double sh[4];
  double acc0 = 0.0, acc1 = 0.0;
  const I stride = (I)gridDim.x * (I)(kBlock * 4);
  for (I base = (I)blockIdx.x * (I)(kBlock * 4) + threadIdx.x; base < len; base += stride) {
    double2 v[4];
    bool ok[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const I i = base + (I)(j * kBlock);
      ok[j] = i < len;
      if (ok[j]) v[j] = a[i];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (ok[j]) {
        const I i = base + (I)(j * kBlock);
        const double p = v[j].x * v[j].x + v[j].y * v[j].y;
        if ((i & selbit) == 0) acc0 += p;
        else acc1 += p;
      }
    }
  }
  const double s0 = block_sum(acc0, sh);
  const double s1 = block_sum(acc1, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[kRedBlocks + blockIdx.x] = s1;
  }
}

__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ part, int nb,
                                                     double* __restrict__ out0,
                                                     double* __restrict__ out1) {
  __shared__ double sh[4];
  double a0 = 0.0, a1 = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) {
    a0 += part[i];
    a1 += part[kRedBlocks + i];
  }
  const double s0 = block_sum(a0, sh);
  const double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) {
    *out0 = s0;
    if (out1) *out1 = s1;
  }
}

// normalize (quantum_state.rs:888-907): divide by sqrt(s) or reset to |0...0>.
template <typename I>
__global__ __launch_bounds__(kBlock) void k_normalize_apply(double2* __restrict__ a, I len,
                                                            const double* __restrict__ slot) {
  const double s = *slot;
  const I base = (I)blockIdx.x * (I)(kBlock * 4) + threadIdx.x;
  if (s > 1e-12) {
    const double nrm = sqrt(s);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const I i = base + (I)(j * kBlock);
      if (i < len) {
        double2 v = a[i];
        v.x = v.x / nrm;
        v.y = v.y / nrm;
        a[i] = v;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const I i = base + (I)(j * kBlock);
      if (i < len) a[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
    }
  }
}

// measure collapse (quantum_state.rs:445-452).
template <typename I>
__global__ __launch_bounds__(kBlock) void k_collapse(double2* __restrict__ a, I len, I bit,
                                                     int outcome, double f) {
  const I base = (I)blockIdx.x * (I)(kBlock * 4) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const I i = base + (I)(j * kBlock);
    if (i < len) {
      const int set = (i & bit) != 0;
      if ((outcome == 0 && set) || (outcome == 1 && !set)) {
        a[i] = make_double2(0.0, 0.0);
      } else {
        double2 v = a[i];
        v.x *= f;
        v.y *= f;
        a[i] = v;
      }
    }
  }
}

// apply_reset_simd second pass (vectorization.rs:1037-1055), with the bit-0
// norm read from the device slot.
template <typename I>
__global__ __launch_bounds__(kBlock) void k_reset_apply(double2* __restrict__ a, I len, I bit,
                                                        const double* __restrict__ slot) {
  const double ns = *slot;
  const I base = (I)blockIdx.x * (I)(kBlock * 4) + threadIdx.x;
  if (ns > 1e-12) {
    const double scale = 1.0 / sqrt(ns);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const I i = base + (I)(j * kBlock);
      if (i < len) {
        if ((i & bit) == 0) {
          double2 v = a[i];
          v.x *= scale;
          v.y *= scale;
          a[i] = v;
        } else {
          a[i] = make_double2(0.0, 0.0);
        }
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const I i = base + (I)(j * kBlock);
      if (i < len) a[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
    }
  }
}

template <typename I>
__global__ __launch_bounds__(kBlock) void k_probs(const double2* __restrict__ a, I len,
                                                  double* __restrict__ out) {
  const I base = (I)blockIdx.x * (I)(kBlock * 4) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const I i = base + (I)(j * kBlock);
    if (i < len) {
      const double2 v = a[i];
      out[i] = v.x * v.x + v.y * v.y;
    }
  }
}

// c128 <-> c64 conversions (device to device).
template <typename I>
__global__ __launch_bounds__(kBlock) void k_to_c64(const double2* __restrict__ src,
                                                   float2* __restrict__ dst, I len) {
  const I i = (I)blockIdx.x * kBlock + threadIdx.x;
  if (i < len) {
    const double2 v = src[i];
    dst[i] = make_float2((float)v.x, (float)v.y);
  }
}
template <typename I>
__global__ __launch_bounds__(kBlock) void k_to_c128(const float2* __restrict__ src,
                                                    double2* __restrict__ dst, I len) {
  const I i = (I)blockIdx.x * kBlock + threadIdx.x;
  if (i < len) {
    const float2 v = src[i];
    dst[i] = make_double2((double)v.x, (double)v.y);
  }
}

inline unsigned grid4(uint64_t count) {
  uint64_t g = (count + (uint64_t)(kBlock * 4) - 1) / (uint64_t)(kBlock * 4);
  return (unsigned)(g ? g : 1);
}
inline unsigned grid1(uint64_t count) {
  uint64_t g = (count + (uint64_t)kBlock - 1) / (uint64_t)kBlock;
  return (unsigned)(g ? g : 1);
}

}  // namespace

int launch_nonfinite(qoa_state* s, int slot) {
  unsigned int* flags = reinterpret_cast<unsigned int*>(s->red + kSlotBase + slot);
  const double2* a = (const double2*)s->amps;
  const uint64_t nb64 = (s->len + kBlock - 1) / kBlock;
  const unsigned nb = (unsigned)(nb64 < (uint64_t)kRedBlocks ? nb64 : kRedBlocks);
  hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(1), 0, s->stream, flags);
  timing_begin(s, KK_REDUCE, 16.0 * (double)s->len);
  if (use32(s))
    hipLaunchKernelGGL((k_nonfinite<uint32_t>), dim3(nb), dim3(kBlock), 0, s->stream, a,
                       (uint32_t)s->len, flags);
  else
    hipLaunchKernelGGL((k_nonfinite<uint64_t>), dim3(nb), dim3(kBlock), 0, s->stream, a,
                       (uint64_t)s->len, flags);
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_norm_sum(qoa_state* s, uint64_t selbit, bool two_way, int slot) {
  const double2* a = (const double2*)s->amps;
  uint64_t nb64 = (s->len + (uint64_t)(kBlock * 4) - 1) / (uint64_t)(kBlock * 4);
  const int nb = (int)(nb64 < (uint64_t)kRedBlocks ? (nb64 ? nb64 : 1) : kRedBlocks);
  double* part = s->red;
  timing_begin(s, KK_REDUCE, 16.0 * (double)s->len);
  if (use32(s))
    hipLaunchKernelGGL((k_norm_partials<uint32_t>), dim3(nb), dim3(kBlock), 0, s->stream, a,
                       (uint32_t)s->len, (uint32_t)selbit, part);
  else
    hipLaunchKernelGGL((k_norm_partials<uint64_t>), dim3(nb), dim3(kBlock), 0, s->stream, a,
                       (uint64_t)s->len, (uint64_t)selbit, part);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, s->stream, part, nb,
                     s->red + kSlotBase + slot, two_way ? s->red + kSlotBase + slot + 1 : nullptr);
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_normalize_apply(qoa_state* s, int slot) {
  double2* a = (double2*)s->amps;
  timing_begin(s, KK_REDUCE, 32.0 * (double)s->len);
  if (use32(s))
    hipLaunchKernelGGL((k_normalize_apply<uint32_t>), dim3(grid4(s->len)), dim3(kBlock), 0,
                       s->stream, a, (uint32_t)s->len, s->red + kSlotBase + slot);
  else
    hipLaunchKernelGGL((k_normalize_apply<uint64_t>), dim3(grid4(s->len)), dim3(kBlock), 0,
                       s->stream, a, (uint64_t)s->len, s->red + kSlotBase + slot);
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_collapse(qoa_state* s, uint64_t bit, int outcome, double f) {
  double2* a = (double2*)s->amps;
  timing_begin(s, KK_REDUCE, 24.0 * (double)s->len);
  if (use32(s))
    hipLaunchKernelGGL((k_collapse<uint32_t>), dim3(grid4(s->len)), dim3(kBlock), 0, s->stream, a,
                       (uint32_t)s->len, (uint32_t)bit, outcome, f);
  else
    hipLaunchKernelGGL((k_collapse<uint64_t>), dim3(grid4(s->len)), dim3(kBlock), 0, s->stream, a,
                       (uint64_t)s->len, (uint64_t)bit, outcome, f);
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_reset_apply(qoa_state* s, uint64_t bit, int slot) {
  double2* a = (double2*)s->amps;
  timing_begin(s, KK_REDUCE, 24.0 * (double)s->len);
  if (use32(s))
    hipLaunchKernelGGL((k_reset_apply<uint32_t>), dim3(grid4(s->len)), dim3(kBlock), 0, s->stream,
                       a, (uint32_t)s->len, (uint32_t)bit, s->red + kSlotBase + slot);
  else
    hipLaunchKernelGGL((k_reset_apply<uint64_t>), dim3(grid4(s->len)), dim3(kBlock), 0, s->stream,
                       a, (uint64_t)s->len, (uint64_t)bit, s->red + kSlotBase + slot);
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_probabilities(qoa_state* s, double* out) {
  const double2* a = (const double2*)s->amps;
  timing_begin(s, KK_REDUCE, 24.0 * (double)s->len);
  if (use32(s))
    hipLaunchKernelGGL((k_probs<uint32_t>), dim3(grid4(s->len)), dim3(kBlock), 0, s->stream, a,
                       (uint32_t)s->len, out);
  else
    hipLaunchKernelGGL((k_probs<uint64_t>), dim3(grid4(s->len)), dim3(kBlock), 0, s->stream, a,
                       (uint64_t)s->len, out);
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_convert(qoa_state* dst, const qoa_state* src) {
  const uint64_t len = dst->len;
  if (src->precision == QOA_C128 && dst->precision == QOA_C64) {
    hipLaunchKernelGGL((k_to_c64<uint64_t>), dim3(grid1(len)), dim3(kBlock), 0, dst->stream,
                       (const double2*)src->amps, (float2*)dst->amps, len);
  } else if (src->precision == QOA_C64 && dst->precision == QOA_C128) {
    hipLaunchKernelGGL((k_to_c128<uint64_t>), dim3(grid1(len)), dim3(kBlock), 0, dst->stream,
                       (const float2*)src->amps, (double2*)dst->amps, len);
  } else {
    return set_error(QOA_ERR_ARG, "launch_convert: precisions must differ");
  }
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_upload_c64(qoa_state* s, const double* dev_src) {
  hipLaunchKernelGGL((k_to_c64<uint64_t>), dim3(grid1(s->len)), dim3(kBlock), 0, s->stream,
                     (const double2*)dev_src, (float2*)s->amps, s->len);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_download_c64(qoa_state* s, double* dev_dst) {
  hipLaunchKernelGGL((k_to_c128<uint64_t>), dim3(grid1(s->len)), dim3(kBlock), 0, s->stream,
                     (const float2*)s->amps, (double2*)dev_dst, s->len);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

}  // namespace qoa
