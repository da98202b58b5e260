// sample.hip — Z-basis sampling of states on the device.
//
// The cumulative distribution is walked in a fixed order so that the result
// depends only on the amplitudes and the uniform: the state is cut into
// chunks of C = min(2^n, 4096) amplitudes; a chunk is cut into T_c =
// min(256, C) runs of C/T_c consecutive amplitudes; every sum below is a
// plain left-to-right sum (|a|^2 = re*re + im*im, no FMA):
//   run sums, chunk sum = sum of its run sums, total = sum of chunk sums.
// A sample with uniform u picks the first chunk, then run, then amplitude at
// which the running sum exceeds u * total (if rounding leaves none, the last
// one with non-zero weight). The CPU oracle mirrors this order exactly.
#include <mutex>

#include "qoa_internal.hpp"

#pragma clang fp contract(off)

namespace qoa {
namespace {

constexpr int kSampThreads = 256;

__device__ __forceinline__ double p2(double2 v) { return v.x * v.x + v.y * v.y; }

// Uniform for sample s of trajectory t: Philox4x32-10 with counter
// (0xFFFFFFFF, s, t, 0x5A3E1D2C), key = seed, u = (o1:o0 >> 11) * 2^-53.
__device__ __forceinline__ double sample_uniform(uint64_t seed, uint32_t traj, uint32_t s) {
  uint32_t c0 = 0xFFFFFFFFu, c1 = s, c2 = traj, c3 = 0x5A3E1D2Cu;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t q0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t q1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(q1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(q0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)q1;
    c3 = (uint32_t)q0;
    c0 = n0;
    c2 = n2;
  }
  return (double)((((uint64_t)c1 << 32) | c0) >> 11) * 0x1.0p-53;
}

// One block per (chunk, trajectory): chunk sums.
__global__ __launch_bounds__(kSampThreads) void k_chunk_sums(const double2* __restrict__ states,
                                                             uint64_t stride, uint32_t C,
                                                             uint32_t nch, uint64_t cs_stride,
                                                             double* __restrict__ cs) {
  __shared__ double part[kSampThreads];
  const uint32_t c = blockIdx.x, t = blockIdx.y;
  const uint32_t runs = C < kSampThreads ? C : kSampThreads, per = C / runs;
  const double2* a = states + t * stride + (uint64_t)c * C;
  if (threadIdx.x < runs) {
    double s = 0.0;
    for (uint32_t j = 0; j < per; ++j) s += p2(a[threadIdx.x * per + j]);
    part[threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (uint32_t r = 0; r < runs; ++r) s += part[r];
    cs[(uint64_t)t * cs_stride + c] = s;
  }
}

// One block per trajectory: replaces the chunk sums by their running sums
// (acc += cs[c], left to right: the same operations, in the same order, as
// the walk over chunks, so prefix[c] is exactly the walk's acc + cs[c]), and
// stores after them the last chunk with a non-zero sum and the running sum
// before it (the walk's fallback when rounding leaves no chunk above the
// target). Chunk sums pass through LDS in tiles; one thread adds.
constexpr uint32_t kPrefixTile = 2048;
__global__ __launch_bounds__(kSampThreads) void k_prefix(double* __restrict__ cs, uint32_t nch,
                                                         uint64_t cs_stride) {
  __shared__ double tile[kPrefixTile];
  double* c = cs + (uint64_t)blockIdx.x * cs_stride;
  double acc = 0.0, acc_last = 0.0;
  int64_t lastnz = -1;
  for (uint32_t base = 0; base < nch; base += kPrefixTile) {
    const uint32_t m = nch - base < kPrefixTile ? nch - base : kPrefixTile;
    for (uint32_t i = threadIdx.x; i < m; i += kSampThreads) tile[i] = c[base + i];
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t i = 0; i < m; ++i) {
        const double v = tile[i];
        if (v > 0.0) {
          lastnz = (int64_t)(base + i);
          acc_last = acc;
        }
        acc += v;
        tile[i] = acc;
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < m; i += kSampThreads) c[base + i] = tile[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    c[nch] = (double)lastnz;
    c[nch + 1] = acc_last;
  }
}

// Grid (trajectory, sample group): samples s = blockIdx.y, + gridDim.y, ...
// of trajectory blockIdx.x; out[t*count + s] = index ^ frame. The chunk is
// the first one whose running sum exceeds u * total (a binary search: the
// running sums never decrease); inside it the runs and then the amplitudes
// are walked left to right.
__global__ __launch_bounds__(kSampThreads) void k_walk(const double2* __restrict__ states,
                                                       uint64_t stride, uint32_t C, uint32_t nch,
                                                       const double* __restrict__ pre_all,
                                                       uint64_t cs_stride, uint64_t seed,
                                                       uint32_t count,
                                                       const uint64_t* __restrict__ frames,
                                                       uint32_t frame_stride,
                                                       uint64_t* __restrict__ out) {
  __shared__ double part[kSampThreads];
  __shared__ uint32_t sel_chunk;
  __shared__ double acc_s;
  const uint32_t t = blockIdx.x;
  const uint32_t runs = C < kSampThreads ? C : kSampThreads, per = C / runs;
  const double* pre = pre_all + (uint64_t)t * cs_stride;
  const double2* a = states + t * stride;
  const uint64_t frame = frames ? frames[(uint64_t)t * frame_stride] : 0ULL;
  const double total = pre[nch - 1];
  for (uint32_t s = blockIdx.y; s < count; s += gridDim.y) {
    const double target = sample_uniform(seed, t, s) * total;
    if (threadIdx.x == 0) {
      uint32_t lo = 0, hi = nch;  // first c with pre[c] > target
      while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (pre[mid] > target) hi = mid;
        else lo = mid + 1;
      }
      if (lo < nch) {
        sel_chunk = lo;
        acc_s = lo ? pre[lo - 1] : 0.0;
      } else {
        const int64_t lastnz = (int64_t)pre[nch];
        sel_chunk = lastnz < 0 ? 0u : (uint32_t)lastnz;
        acc_s = lastnz < 0 ? 0.0 : pre[nch + 1];
      }
    }
    __syncthreads();
    const double2* ch = a + (uint64_t)sel_chunk * C;
    if (threadIdx.x < runs) {
      double r = 0.0;
      for (uint32_t j = 0; j < per; ++j) r += p2(ch[threadIdx.x * per + j]);
      part[threadIdx.x] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double acc = acc_s;
      int chosen = -1, lastnz = -1;
      double acc_last = acc;
      for (uint32_t r = 0; r < runs; ++r) {
        if (part[r] > 0.0) {
          lastnz = (int)r;
          acc_last = acc;
        }
        if (acc + part[r] > target) {
          chosen = (int)r;
          break;
        }
        acc += part[r];
      }
      if (chosen < 0) {
        chosen = lastnz < 0 ? 0 : lastnz;
        acc = lastnz < 0 ? acc_s : acc_last;
      }
      const double2* run = ch + (uint64_t)chosen * per;
      int pick = -1, lastj = -1;
      for (uint32_t j = 0; j < per; ++j) {
        const double p = p2(run[j]);
        if (p > 0.0) lastj = (int)j;
        if (acc + p > target) {
          pick = (int)j;
          break;
        }
        acc += p;
      }
      if (pick < 0) pick = lastj < 0 ? 0 : lastj;
      const uint64_t idx = (uint64_t)sel_chunk * C + (uint64_t)chosen * per + (uint64_t)pick;
      out[(uint64_t)t * count + s] = idx ^ frame;
    }
    __syncthreads();
  }
}

}  // namespace

// Samples `count` outcomes from each of `ntraj` states of n qubits laid out
// `stride` amplitudes apart. `cs` is device scratch of
// sample_scratch_doubles(n, ntraj) doubles.
int launch_sample(hipStream_t st, const double2* states, uint64_t stride, uint32_t n,
                  uint32_t ntraj, uint64_t seed, uint32_t count, const uint64_t* frames,
                  uint32_t frame_stride, double* cs, uint64_t* out) {
  const uint64_t len = 1ULL << n;
  const uint32_t C = (uint32_t)(len < 4096 ? len : 4096);
  const uint32_t nch = (uint32_t)(len / C);
  const uint64_t cs_stride = (uint64_t)nch + 2;
  hipLaunchKernelGGL(k_chunk_sums, dim3(nch, ntraj), dim3(kSampThreads), 0, st, states, stride,
                     C, nch, cs_stride, cs);
  hipLaunchKernelGGL(k_prefix, dim3(ntraj), dim3(kSampThreads), 0, st, cs, nch, cs_stride);
  const uint32_t groups = count < 1024u ? count : 1024u;
  hipLaunchKernelGGL(k_walk, dim3(ntraj, groups), dim3(kSampThreads), 0, st, states, stride, C,
                     nch, cs, cs_stride, seed, count, frames, frame_stride, out);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

}  // namespace qoa

extern "C" int qoa_sample(qoa_state* s, uint64_t seed, uint32_t count, uint64_t* out) {
  using namespace qoa;
  if (!valid(s)) return set_error(QOA_ERR_HANDLE, "invalid handle");
  if (s->precision != QOA_C128) return set_error(QOA_ERR_UNSUPPORTED, "sampling needs QOA_C128");
  if (!out && count) return set_error(QOA_ERR_ARG, "out is NULL");
  if (count == 0) return QOA_OK;
  double* cs = nullptr;
  uint64_t* dout = nullptr;
  if (hipMalloc(&cs, sizeof(double) * sample_scratch_doubles(s->n, 1)) != hipSuccess ||
      hipMalloc(&dout, sizeof(uint64_t) * count) != hipSuccess) {
    (void)hipGetLastError();
    if (cs) hipFree(cs);
    return set_error(QOA_ERR_OOM, "sampling scratch");
  }
  int rc = launch_sample(s->stream, (const double2*)s->amps, 0, s->n, 1, seed, count, nullptr, 0,
                         cs, dout);
  if (rc == QOA_OK && hipMemcpyAsync(out, dout, sizeof(uint64_t) * count, hipMemcpyDeviceToHost,
                                     s->stream) != hipSuccess)
    rc = set_error(QOA_ERR_HIP, "sample copy failed");
  if (hipStreamSynchronize(s->stream) != hipSuccess && rc == QOA_OK)
    rc = set_error(QOA_ERR_HIP, "sampling failed");
  hipFree(cs);
  hipFree(dout);
  return rc;
}

// ---- allocation-free sampling ------------------------------------------------
// qoa_sample_noalloc keeps its device scratch in one file-scope array per
// device instead of allocating: the chunk running sums of a state of up to
// kNoAllocMaxQubits qubits (2^(n-12) + 2 doubles) and up to kNoAllocMaxCount
// samples. Every call on a device shares that device's copy; nothing in it
// has to start at zero (each call writes every slot it reads).
namespace qoa {
namespace {
constexpr uint32_t kNoAllocMaxQubits = 32;
constexpr uint32_t kNoAllocMaxCount = 65536;
constexpr uint64_t kNoAllocSumDoubles = (1ULL << (kNoAllocMaxQubits - 12)) + 2;
constexpr uint64_t kNoAllocSampleOff = (kNoAllocSumDoubles + 31) / 32 * 32;  // 256-B aligned
constexpr uint64_t kNoAllocDoubles = kNoAllocSampleOff + kNoAllocMaxCount;   // about 8.5 MiB
}  // namespace
}  // namespace qoa

static __device__ __align__(256) double scratch[qoa::kNoAllocDoubles];

namespace qoa {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (looked up once per device,
// with that device current). In host code the array's name is only a
// stand-in; kernels and copies take this address.
int scratch_for(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return set_error(QOA_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    QOA_HIP_CHECK(hipGetDevice(&cur));
    if (cur != device) QOA_HIP_CHECK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return set_error(QOA_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return QOA_OK;
}
}  // namespace
}  // namespace qoa

extern "C" int qoa_sample_noalloc(qoa_state* s, uint64_t seed, uint32_t count, uint64_t* out) {
  using namespace qoa;
  if (!valid(s)) return set_error(QOA_ERR_HANDLE, "invalid handle");
  if (s->precision != QOA_C128) return set_error(QOA_ERR_UNSUPPORTED, "sampling needs QOA_C128");
  if (!out && count) return set_error(QOA_ERR_ARG, "out is NULL");
  if (s->n > kNoAllocMaxQubits)
    return set_error(QOA_ERR_ARG, "qoa_sample_noalloc: %u qubits exceed its scratch (at most %u)",
                     s->n, kNoAllocMaxQubits);
  if (count > kNoAllocMaxCount)
    return set_error(QOA_ERR_ARG, "qoa_sample_noalloc: %u samples exceed its scratch (at most %u)",
                     count, kNoAllocMaxCount);
  if (count == 0) return QOA_OK;
  double* base = nullptr;
  int rc = scratch_for(s->device, &base);
  if (rc) return rc;
  uint64_t* dout = reinterpret_cast<uint64_t*>(base + kNoAllocSampleOff);
  rc = launch_sample(s->stream, (const double2*)s->amps, 0, s->n, 1, seed, count, nullptr, 0, base,
                     dout);
  if (rc == QOA_OK && hipMemcpyAsync(out, dout, sizeof(uint64_t) * count, hipMemcpyDeviceToHost,
                                     s->stream) != hipSuccess)
    rc = set_error(QOA_ERR_HIP, "sample copy failed");
  if (hipStreamSynchronize(s->stream) != hipSuccess && rc == QOA_OK)
    rc = set_error(QOA_ERR_HIP, "sampling failed");
  return rc;
}
