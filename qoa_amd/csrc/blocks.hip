// blocks.hip — dense 5-qubit blocks on complex64 states (FP32 MFMA, gfx950).
//
// Host: the greedy fuser of blocks_plan.cpp packs gates into blocks that
// act on at most five qubits, each with its 32x32 unitary (the FP64 product
// of its gates, rounded once to FP32 here).
//
// Device: the state viewed as 32 x 2^(n-5) (rows: the block's qubits,
// columns: the other bits) is multiplied in place by U. The complex product
// is the real GEMM [Yr; Yi] = [Ur -Ui; Ui Ur] [Xr; Xi] (M = K = 64) on
// v_mfma_f32_32x32x2f32: one wave owns 32 columns; lane l holds column
// l % 32 and the rows of parity l / 32, so its 16 complex amplitudes are
// exactly the B operands of the 32 K-steps. Each workgroup stages the block's
// A operands (16 KiB, laid out step-major so LDS reads are conflict-free).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "block_device.hpp"
#include "blocks_plan.hpp"
#include "gates.hpp"
#include "qoa_internal.hpp"

namespace qoa {
namespace {

// ---- device -------------------------------------------------------------------
template <typename I>
__global__ __launch_bounds__(kBlockThreads) void k_block5(float2* __restrict__ a,
                                                         const float* __restrict__ w,
                                                         BlockArgs ba) {
  block_apply<I>(a, w, ba, blockIdx.x);
}

// Per-block partial sums of |a|^2 (fixed grid, fixed order).
__global__ __launch_bounds__(256) void k_c64_sumsq(const float2* __restrict__ a, uint64_t len,
                                                   double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * 256) {
    const float2 x = a[i];
    s += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_c64_scale(float2* __restrict__ a, uint64_t len,
                                                   double f) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * 256) {
    const float2 x = a[i];
    a[i] = make_float2((float)((double)x.x * f), (float)((double)x.y * f));
  }
}

template <typename T>
__device__ __forceinline__ double2 load_cd(const void* p, uint64_t i) {
  if constexpr (sizeof(T) == 16) {
    return reinterpret_cast<const double2*>(p)[i];
  } else {
    const float2 x = reinterpret_cast<const float2*>(p)[i];
    return make_double2((double)x.x, (double)x.y);
  }
}

// Partials per block: Re, Im of sum conj(a) b, sum |a|^2, sum |b|^2, max |a-b|.
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_compare(const void* __restrict__ a,
                                                 const void* __restrict__ b, uint64_t len,
                                                 double* __restrict__ part) {
  __shared__ double red[5][256];
  double re = 0, im = 0, na = 0, nb = 0, mx = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * 256) {
    const double2 x = load_cd<TA>(a, i), y = load_cd<TB>(b, i);
    re += x.x * y.x + x.y * y.y;
    im += x.x * y.y - x.y * y.x;
    na += x.x * x.x + x.y * x.y;
    nb += y.x * y.x + y.y * y.y;
    const double dx = x.x - y.x, dy = x.y - y.y;
    mx = fmax(mx, sqrt(dx * dx + dy * dy));
  }
  const int t = threadIdx.x;
  red[0][t] = re;
  red[1][t] = im;
  red[2][t] = na;
  red[3][t] = nb;
  red[4][t] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + w];
      red[4][t] = fmax(red[4][t], red[4][t + w]);
    }
    __syncthreads();
  }
  if (t < 5) part[5 * blockIdx.x + t] = red[t][0];
}

constexpr int kRedGrid = 1024;

struct BlockCache {
  uint64_t key = 0;
  std::vector<Block> blocks;
  float* d_w = nullptr;
};

uint64_t fnv(const void* p, size_t n, uint64_t h) {
  const unsigned char* c = (const unsigned char*)p;
  for (size_t i = 0; i < n; ++i) {
    h ^= c[i];
    h *= 0x100000001B3ULL;
  }
  return h;
}

int renormalize_c64(qoa_state* s) {
  double* part = nullptr;
  if (hipMalloc(&part, sizeof(double) * kRedGrid) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(QOA_ERR_OOM, "reduction buffer allocation failed");
  }
  std::vector<double> h(kRedGrid);
  int rc = QOA_OK;
  timing_begin(s, KK_REDUCE, 8.0 * (double)s->len);
  hipLaunchKernelGGL(k_c64_sumsq, dim3(kRedGrid), dim3(256), 0, s->stream,
                     (const float2*)s->amps, s->len, part);
  timing_end(s);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(h.data(), part, sizeof(double) * kRedGrid, hipMemcpyDeviceToHost,
                     s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    rc = set_error(QOA_ERR_HIP, "norm reduction failed");
  hipFree(part);
  if (rc) return rc;
  double sum = 0.0;
  for (double x : h) sum += x;
  if (!(sum > 1e-12)) return set_error(QOA_ERR_ARG, "state norm %g is too small to renormalise", sum);
  timing_begin(s, KK_REDUCE, 16.0 * (double)s->len);
  hipLaunchKernelGGL(k_c64_scale, dim3(kRedGrid), dim3(256), 0, s->stream, (float2*)s->amps,
                     s->len, 1.0 / std::sqrt(sum));
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

}  // namespace

void block_cache_free(qoa_state* s) {
  BlockCache* c = (BlockCache*)s->block_cache;
  if (!c) return;
  if (c->d_w) hipFree(c->d_w);
  delete c;
  s->block_cache = nullptr;
}

}  // namespace qoa

using namespace qoa;

extern "C" int qoa_run_blocks(qoa_state* s, const qoa_gate* g, uint32_t ng, const double* mats,
                              const qoa_block_opts* opts, uint32_t* n_blocks_out) {
  if (!valid(s)) return set_error(QOA_ERR_HANDLE, "invalid handle");
  if (s->precision != QOA_C64)
    return set_error(QOA_ERR_UNSUPPORTED, "dense blocks run on QOA_C64 states");
  if (!g && ng) return set_error(QOA_ERR_ARG, "gates is NULL");
  const uint32_t flags = opts ? opts->flags : 0, bq = opts ? opts->block_qubits : 0;
  uint64_t key = 0xCBF29CE484222325ULL;
  key = fnv(&s->n, sizeof(s->n), key);
  key = fnv(&flags, sizeof(flags), key);
  key = fnv(&bq, sizeof(bq), key);
  key = fnv(&ng, sizeof(ng), key);
  if (ng) key = fnv(g, sizeof(qoa_gate) * ng, key);
  for (uint32_t i = 0; i < ng; ++i) {
    if (g[i].op == QOA_OP_U2 && mats) key = fnv(mats + 8 * g[i].aux, 64, key);
    if (g[i].op == QOA_OP_U4 && mats) key = fnv(mats + 32 * g[i].aux, 256, key);
  }
  BlockCache* c = (BlockCache*)s->block_cache;
  if (!c || c->key != key) {
    block_cache_free(s);
    c = new BlockCache();
    s->block_cache = c;
    int rc = build_blocks(s->n, g, ng, mats, flags, bq, &c->blocks);
    if (rc) {
      block_cache_free(s);
      return rc;
    }
    std::vector<float> w((size_t)kWFloats * (c->blocks.size() + 1));
    for (size_t b = 0; b < c->blocks.size(); ++b) pack_operands(c->blocks[b], &w[kWFloats * b]);
    if (hipMalloc(&c->d_w, sizeof(float) * w.size()) != hipSuccess) {
      (void)hipGetLastError();
      block_cache_free(s);
      return set_error(QOA_ERR_OOM, "block operand upload failed");
    }
    QOA_HIP_CHECK(hipMemcpyAsync(c->d_w, w.data(), sizeof(float) * w.size(),
                                 hipMemcpyHostToDevice, s->stream));
    QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
    c->key = key;
  }
  const uint64_t ncols = s->len >> kBlockQ;
  const uint64_t cols_per_wg = (uint64_t)(kBlockThreads / 64) * 32 * kColGroups;
  const uint64_t grid = (ncols + cols_per_wg - 1) / cols_per_wg;
  if (grid > 0x7FFFFFFFULL) return set_error(QOA_ERR_ARG, "state too large for one launch");
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    BlockArgs ba{};
    ba.ncols = ncols;
    std::memcpy(ba.q, c->blocks[b].q, kBlockQ);
    const float* w = c->d_w + (size_t)kWFloats * b;
    timing_begin(s, KK_BLOCK, 16.0 * (double)s->len);
    if (use32(s))
      hipLaunchKernelGGL(k_block5<uint32_t>, dim3((unsigned)grid), dim3(kBlockThreads), 0,
                         s->stream, (float2*)s->amps, w, ba);
    else
      hipLaunchKernelGGL(k_block5<uint64_t>, dim3((unsigned)grid), dim3(kBlockThreads), 0,
                         s->stream, (float2*)s->amps, w, ba);
    timing_end(s);
    QOA_HIP_CHECK(hipGetLastError());
  }
  if (n_blocks_out) *n_blocks_out = (uint32_t)c->blocks.size();
  if (!(flags & QOA_BLOCK_NO_RENORM)) return renormalize_c64(s);
  return QOA_OK;
}

extern "C" int qoa_state_compare(qoa_state* a, qoa_state* b, double* max_abs_diff,
                                 double* fidelity) {
  if (!valid(a) || !valid(b)) return set_error(QOA_ERR_HANDLE, "invalid handle");
  if (a->n != b->n) return set_error(QOA_ERR_ARG, "qubit counts differ");
  QOA_HIP_CHECK(hipStreamSynchronize(b->stream));
  double* part = nullptr;
  if (hipMalloc(&part, sizeof(double) * 5 * kRedGrid) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(QOA_ERR_OOM, "comparison buffer allocation failed");
  }
  const bool a128 = a->precision == QOA_C128, b128 = b->precision == QOA_C128;
  const dim3 grid(kRedGrid), block(256);
  if (a128 && b128)
    hipLaunchKernelGGL((k_compare<double2, double2>), grid, block, 0, a->stream, a->amps, b->amps, a->len, part);
  else if (a128)
    hipLaunchKernelGGL((k_compare<double2, float2>), grid, block, 0, a->stream, a->amps, b->amps, a->len, part);
  else if (b128)
    hipLaunchKernelGGL((k_compare<float2, double2>), grid, block, 0, a->stream, a->amps, b->amps, a->len, part);
  else
    hipLaunchKernelGGL((k_compare<float2, float2>), grid, block, 0, a->stream, a->amps, b->amps, a->len, part);
  std::vector<double> h(5 * kRedGrid);
  int rc = QOA_OK;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                     a->stream) != hipSuccess ||
      hipStreamSynchronize(a->stream) != hipSuccess)
    rc = set_error(QOA_ERR_HIP, "comparison failed");
  hipFree(part);
  if (rc) return rc;
  double re = 0, im = 0, na = 0, nb = 0, mx = 0;
  for (int i = 0; i < kRedGrid; ++i) {
    re += h[5 * i];
    im += h[5 * i + 1];
    na += h[5 * i + 2];
    nb += h[5 * i + 3];
    mx = std::max(mx, h[5 * i + 4]);
  }
  if (max_abs_diff) *max_abs_diff = mx;
  if (fidelity) *fidelity = (na > 0 && nb > 0) ? (re * re + im * im) / (na * nb) : 0.0;
  return QOA_OK;
}
