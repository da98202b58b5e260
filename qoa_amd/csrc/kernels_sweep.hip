// kernels_sweep.hip — per-gate FP64 amplitude sweeps for gfx950.
//
// One kernel per gate class. Every kernel does the reference's arithmetic
// (num-complex products, no FMA: this file is built with -ffp-contract=off)
// so a single gate here and in the CPU path agree bit for bit.
//
// Index mapping: a gate whose target is t and whose controls are C visits
// the 2^(n-1-|C|) pairs (i, i | 1<<t) with bit t clear and all C bits set.
// Pair p is turned into i by inserting a zero bit at each fixed position in
// ascending order and OR-ing the control mask ("deposit"). Consecutive lanes
// get consecutive p, so for t >= 6 a wavefront streams two contiguous 1 KiB
// runs per load. For a (compressed) target below 6 the partner sits in the
// same wavefront: each lane loads one amplitude of a fully contiguous run and
// fetches its partner with a cross-lane shuffle instead.
//
// Every sweep reads and writes each amplitude it visits once, so its loads
// and stores carry the nontemporal hint (ld_stream / st_stream; measured on
// the 28-qubit layered circuit: dense sweeps 5.60 -> 6.04 TB/s, diagonal
// 5.25 -> 5.38, permutations 5.25 -> 5.43).
#include "qoa_internal.hpp"
#include "tile_device.hpp"
using qoa::tile::ld_stream;
using qoa::tile::st_stream;

#pragma clang fp contract(off)

namespace qoa {
namespace {

constexpr int kBlock = 256;  // 4 wavefronts of 64
// Items per thread. One item per thread (the wave's loads one contiguous
// 1 KiB run per stream) measured fastest at 28 qubits: dense sweeps 5.62 TB/s,
// diagonal 5.33, permutations 5.23, against 5.46 / 5.00 / 4.99 with four
// items a thread 256 elements apart and 5.44 / 5.12 / 5.05 with two.
#ifndef QOA_SWEEP_ITEMS
#define QOA_SWEEP_ITEMS 1
#endif
constexpr int kItems = QOA_SWEEP_ITEMS;

// Workgroups walk the blocks of kBlock * kItems elements in strides of the
// grid (uniform per workgroup, so cross-lane shuffles stay in step): the
// grid is capped below so that the dispatch's work-item count fits 32 bits
// (a 32-qubit whole-state sweep has 2^24 blocks of 256).
#define FOR_EACH_BLOCK(count)                                                       \
  for (I blk_ = (I)blockIdx.x, nblk_ = ((count) + (I)(kBlock * kItems) - 1) /      \
                                          (I)(kBlock * kItems);                   \
       blk_ < nblk_; blk_ += (I)gridDim.x)

struct Ins {                 // fixed bit positions, ascending
  int n;
  uint8_t pos[44];
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

template <typename I>
__device__ __forceinline__ I deposit(I x, const Ins& ins) {
  for (int k = 0; k < ins.n; ++k) {
    const int b = ins.pos[k];
    const I lo = x & ((I(1) << b) - I(1));
    x = ((x >> b) << (b + 1)) | lo;
  }
  return x;
}

// ---- 2x2 operators: out0/out1 given a = amp(bit 0), b = amp(bit 1) ----------
struct OpH {  // apply_hadamard_simd (vectorization.rs:933-934)
  double2 nf;
  __device__ double2 out0(double2 a, double2 b) const { return cmul(nf, cadd(a, b)); }
  __device__ double2 out1(double2 a, double2 b) const { return cmul(nf, csub(a, b)); }
};
struct OpM {  // apply_rx_simd (vectorization.rs:1081-1082)
  double2 m00, m01, m10, m11;
  __device__ double2 out0(double2 a, double2 b) const { return cadd(cmul(m00, a), cmul(m01, b)); }
  __device__ double2 out1(double2 a, double2 b) const { return cadd(cmul(m10, a), cmul(m11, b)); }
};
struct OpX {  // apply_x_simd (vectorization.rs:950-957)
  __device__ double2 out0(double2, double2 b) const { return b; }
  __device__ double2 out1(double2 a, double2) const { return a; }
};
struct OpY {  // apply_y_simd (vectorization.rs:980-981)
  __device__ double2 out0(double2, double2 b) const { return cmul(b, make_double2(0.0, -1.0)); }
  __device__ double2 out1(double2 a, double2) const { return cmul(a, make_double2(0.0, 1.0)); }
};

template <typename I, class Op>
__global__ __launch_bounds__(kBlock) void k_pair(double2* __restrict__ a, I npairs, Ins ins,
                                                 I tbit, I cmask, Op op) {
  FOR_EACH_BLOCK(npairs) {
    const I base = blk_ * (I)(kBlock * kItems) + (I)threadIdx.x;
    I i0[kItems];
    double2 x[kItems], y[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I p = base + (I)(j * kBlock);
      i0[j] = deposit(p, ins) | cmask;
      if (p < npairs) {
        x[j] = ld_stream(a + i0[j]);
        y[j] = ld_stream(a + (i0[j] | tbit));
      }
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I p = base + (I)(j * kBlock);
      if (p < npairs) {
        const double2 o0 = op.out0(x[j], y[j]);
        const double2 o1 = op.out1(x[j], y[j]);
        st_stream(a + i0[j], o0);
        st_stream(a + (i0[j] | tbit), o1);
      }
    }
  }
}

// Element mode: virtual target tv < 6, partner in lane ^ (1 << tv).
template <typename I, class Op>
__global__ __launch_bounds__(kBlock) void k_shfl(double2* __restrict__ a, I nelem, Ins ins,
                                                 I cmask, int tv, Op op) {
  FOR_EACH_BLOCK(nelem) {
    const I base = blk_ * (I)(kBlock * kItems) + (I)threadIdx.x;
    const int hi = (threadIdx.x >> tv) & 1;
    const int lm = 1 << tv;
    I e[kItems];
    double2 v[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I p = base + (I)(j * kBlock);
      e[j] = deposit(p, ins) | cmask;
      v[j] = make_double2(0.0, 0.0);
      if (p < nelem) v[j] = ld_stream(a + e[j]);
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      double2 o;
      o.x = __shfl_xor(v[j].x, lm);
      o.y = __shfl_xor(v[j].y, lm);
      const double2 r = hi ? op.out1(o, v[j]) : op.out0(v[j], o);
      const I p = base + (I)(j * kBlock);
      if (p < nelem) st_stream(a + e[j], r);
    }
  }
}

template <typename I, bool NEG>
__global__ __launch_bounds__(kBlock) void k_diag(double2* __restrict__ a, I count, Ins ins,
                                                 I mask, double2 f) {
  FOR_EACH_BLOCK(count) {
    const I base = blk_ * (I)(kBlock * kItems) + (I)threadIdx.x;
    I e[kItems];
    double2 v[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I p = base + (I)(j * kBlock);
      e[j] = deposit(p, ins) | mask;
      if (p < count) v[j] = ld_stream(a + e[j]);
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I p = base + (I)(j * kBlock);
      if (p < count) {
        double2 r;
        if (NEG) r = make_double2(v[j].x * -1.0, v[j].y * -1.0);  // Complex::scale(-1.0)
        else r = cmul(v[j], f);
        st_stream(a + e[j], r);
      }
    }
  }
}

template <typename I>
__global__ __launch_bounds__(kBlock) void k_rz(double2* __restrict__ a, I len, I bit, double2 m00,
                                               double2 m11) {
  FOR_EACH_BLOCK(len) {
    const I base = blk_ * (I)(kBlock * kItems) + (I)threadIdx.x;
    double2 v[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I i = base + (I)(j * kBlock);
      if (i < len) v[j] = ld_stream(a + i);
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I i = base + (I)(j * kBlock);
      if (i < len) st_stream(a + i, (i & bit) ? cmul(v[j], m11) : cmul(v[j], m00));
    }
  }
}

// SWAP / CSWAP: exchange (base|c|m2, base|c|m1) for base with q1, q2 (and c) clear.
template <typename I>
__global__ __launch_bounds__(kBlock) void k_swap2(double2* __restrict__ a, I count, Ins ins,
                                                  I cmask, I m1, I m2) {
  FOR_EACH_BLOCK(count) {
    const I base = blk_ * (I)(kBlock * kItems) + (I)threadIdx.x;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I p = base + (I)(j * kBlock);
      if (p < count) {
        const I b = deposit(p, ins) | cmask;
        const double2 u = ld_stream(a + (b | m2)), w = ld_stream(a + (b | m1));
        st_stream(a + (b | m2), w);
        st_stream(a + (b | m1), u);
      }
    }
  }
}

struct Mat4 {
  double m[32];
};

// apply_two_qubit_unitary (quantum_state.rs:1064-1127): row s of u times the
// four amplitudes (00, 01, 10, 11) of the (high, low) qubit pair.
template <typename I>
__global__ __launch_bounds__(kBlock) void k_mat4(double2* __restrict__ a, I count, Ins ins, I ml,
                                                 I mh, Mat4 u) {
  FOR_EACH_BLOCK(count) {
    const I base = blk_ * (I)(kBlock * kItems) + (I)threadIdx.x;
#pragma unroll 1
    for (int j = 0; j < kItems; ++j) {
      const I p = base + (I)(j * kBlock);
      if (p < count) {
        const I b = deposit(p, ins);
        const I idx[4] = {b, b | ml, b | mh, b | ml | mh};
        double2 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = a[idx[k]];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          double2 acc = cmul(make_double2(u.m[8 * s], u.m[8 * s + 1]), v[0]);
#pragma unroll
          for (int k = 1; k < 4; ++k)
            acc = cadd(acc, cmul(make_double2(u.m[8 * s + 2 * k], u.m[8 * s + 2 * k + 1]), v[k]));
          a[idx[s]] = acc;
        }
      }
    }
  }
}

struct PauliList {
  int n;
  uint8_t q[64];
  uint8_t code[64];  // 1 X, 2 Y, 3 Z
};

// Follows one amplitude through the ordered X/Y/Z sweeps of
// apply_depolarizing_noise (quantum_state.rs:872-885), doing each sweep's own
// arithmetic, and returns where it ends up.
template <typename I>
__device__ __forceinline__ I pauli_chain(I idx, double2& v, const PauliList& pl) {
  for (int k = 0; k < pl.n; ++k) {
    const int q = pl.q[k];
    const I bit = I(1) << q;
    const int set = (idx & bit) != 0;
    const int c = pl.code[k];
    if (c == 1) {
      idx ^= bit;
    } else if (c == 2) {
      v = set ? cmul(v, make_double2(0.0, -1.0)) : cmul(v, make_double2(0.0, 1.0));
      idx ^= bit;
    } else if (set) {
      v = make_double2(v.x * -1.0, v.y * -1.0);
    }
  }
  return idx;
}

template <typename I>
__global__ __launch_bounds__(kBlock) void k_pauli_pairs(double2* __restrict__ a, I npairs, int low,
                                                        I xmask, PauliList pl) {
  FOR_EACH_BLOCK(npairs) {
    const I base = blk_ * (I)(kBlock * kItems) + (I)threadIdx.x;
#pragma unroll 1
    for (int j = 0; j < kItems; ++j) {
      const I p = base + (I)(j * kBlock);
      if (p < npairs) {
        const I lo = p & ((I(1) << low) - I(1));
        const I i = ((p >> low) << (low + 1)) | lo;
        const I k = i ^ xmask;
        double2 vi = a[i], vk = a[k];
        const I di = pauli_chain(i, vi, pl);
        const I dk = pauli_chain(k, vk, pl);
        a[di] = vi;
        a[dk] = vk;
      }
    }
  }
}

template <typename I>
__global__ __launch_bounds__(kBlock) void k_pauli_diag(double2* __restrict__ a, I len,
                                                       PauliList pl) {
  FOR_EACH_BLOCK(len) {
    const I base = blk_ * (I)(kBlock * kItems) + (I)threadIdx.x;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I i = base + (I)(j * kBlock);
      if (i < len) {
        double2 v = a[i];
        pauli_chain(i, v, pl);
        a[i] = v;
      }
    }
  }
}

template <typename I>
__global__ __launch_bounds__(kBlock) void k_set_basis0(double2* __restrict__ a, I len) {
  FOR_EACH_BLOCK(len) {
    const I base = blk_ * (I)(kBlock * kItems) + (I)threadIdx.x;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const I i = base + (I)(j * kBlock);
      if (i < len) a[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
    }
  }
}

// ---- host helpers -----------------------------------------------------------
constexpr uint64_t kMaxGrid = 1ULL << 20;  // 2^28 work-items per dispatch at most
inline unsigned grid_for(uint64_t count) {
  uint64_t g = (count + (uint64_t)(kBlock * kItems) - 1) / (uint64_t)(kBlock * kItems);
  if (g > kMaxGrid) g = kMaxGrid;
  return (unsigned)(g ? g : 1);
}

Ins make_ins(uint64_t bits) {
  Ins ins{};
  ins.n = 0;
  for (int b = 0; b < 64; ++b)
    if (bits & (1ULL << b)) ins.pos[ins.n++] = (uint8_t)b;
  return ins;
}

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

// Algorithmic HBM bytes of a sweep that reads and writes the amplitudes whose
// bits in `forced` are fixed, at 64-byte-line granularity (4 amplitudes):
// fixed bits below position 2 do not reduce the lines touched.
inline double sweep_bytes(const qoa_state* s, uint64_t forced) {
  return 32.0 * (double)s->len / (double)(1ULL << popc(forced & ~3ULL));
}

inline double2 d2(Cx c) { return make_double2(c.re, c.im); }

template <class Op>
int launch_2x2(qoa_state* s, uint32_t t, uint64_t cmask, Op op, int kind) {
  if (cmask & (1ULL << t)) return QOA_OK;  // target among controls: nothing visits a pair
  double2* a = (double2*)s->amps;
  const int nc = popc(cmask);
  // compressed target position: controls below t are removed from the element index
  const int tv = (int)t - popc(cmask & ((1ULL << t) - 1));
  const uint64_t nelem = s->len >> nc;
  timing_begin(s, kind, sweep_bytes(s, cmask));
  if (tv < 6) {
    Ins ins = make_ins(cmask);
    if (use32(s))
      hipLaunchKernelGGL((k_shfl<uint32_t, Op>), dim3(grid_for(nelem)), dim3(kBlock), 0, s->stream,
                         a, (uint32_t)nelem, ins, (uint32_t)cmask, tv, op);
    else
      hipLaunchKernelGGL((k_shfl<uint64_t, Op>), dim3(grid_for(nelem)), dim3(kBlock), 0, s->stream,
                         a, (uint64_t)nelem, ins, (uint64_t)cmask, tv, op);
  } else {
    Ins ins = make_ins(cmask | (1ULL << t));
    const uint64_t np = nelem >> 1;
    if (use32(s))
      hipLaunchKernelGGL((k_pair<uint32_t, Op>), dim3(grid_for(np)), dim3(kBlock), 0, s->stream, a,
                         (uint32_t)np, ins, (uint32_t)(1u << t), (uint32_t)cmask, op);
    else
      hipLaunchKernelGGL((k_pair<uint64_t, Op>), dim3(grid_for(np)), dim3(kBlock), 0, s->stream, a,
                         (uint64_t)np, ins, (uint64_t)(1ULL << t), (uint64_t)cmask, op);
  }
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

}  // namespace

int launch_hadamard(qoa_state* s, Cx nf, uint32_t t, uint64_t cmask) {
  return launch_2x2(s, t, cmask, OpH{d2(nf)}, KK_DENSE);
}
int launch_matrix2(qoa_state* s, const Cx m[4], uint32_t t, uint64_t cmask) {
  return launch_2x2(s, t, cmask, OpM{d2(m[0]), d2(m[1]), d2(m[2]), d2(m[3])}, KK_DENSE);
}
int launch_pauli_x(qoa_state* s, uint32_t t, uint64_t cmask) {
  return launch_2x2(s, t, cmask, OpX{}, KK_PERM);
}
int launch_pauli_y(qoa_state* s, uint32_t t, uint64_t cmask) {
  return launch_2x2(s, t, cmask, OpY{}, KK_DENSE);
}

int launch_diag_mask(qoa_state* s, uint64_t mask, bool neg, Cx f) {
  double2* a = (double2*)s->amps;
  const uint64_t count = s->len >> popc(mask);
  Ins ins = make_ins(mask);
  timing_begin(s, KK_DIAG, sweep_bytes(s, mask) * (mask ? 1.0 : 1.0));
  if (use32(s)) {
    if (neg)
      hipLaunchKernelGGL((k_diag<uint32_t, true>), dim3(grid_for(count)), dim3(kBlock), 0,
                         s->stream, a, (uint32_t)count, ins, (uint32_t)mask, d2(f));
    else
      hipLaunchKernelGGL((k_diag<uint32_t, false>), dim3(grid_for(count)), dim3(kBlock), 0,
                         s->stream, a, (uint32_t)count, ins, (uint32_t)mask, d2(f));
  } else {
    if (neg)
      hipLaunchKernelGGL((k_diag<uint64_t, true>), dim3(grid_for(count)), dim3(kBlock), 0,
                         s->stream, a, (uint64_t)count, ins, (uint64_t)mask, d2(f));
    else
      hipLaunchKernelGGL((k_diag<uint64_t, false>), dim3(grid_for(count)), dim3(kBlock), 0,
                         s->stream, a, (uint64_t)count, ins, (uint64_t)mask, d2(f));
  }
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_rz(qoa_state* s, uint64_t bit, Cx m00, Cx m11) {
  double2* a = (double2*)s->amps;
  timing_begin(s, KK_DIAG, 32.0 * (double)s->len);
  if (use32(s))
    hipLaunchKernelGGL((k_rz<uint32_t>), dim3(grid_for(s->len)), dim3(kBlock), 0, s->stream, a,
                       (uint32_t)s->len, (uint32_t)bit, d2(m00), d2(m11));
  else
    hipLaunchKernelGGL((k_rz<uint64_t>), dim3(grid_for(s->len)), dim3(kBlock), 0, s->stream, a,
                       (uint64_t)s->len, (uint64_t)bit, d2(m00), d2(m11));
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_swap2(qoa_state* s, uint32_t q1, uint32_t q2, uint64_t cmask) {
  if (q1 == q2) return QOA_OK;
  double2* a = (double2*)s->amps;
  const uint64_t m1 = 1ULL << q1, m2 = 1ULL << q2;
  if (cmask & (m1 | m2)) return QOA_OK;
  Ins ins = make_ins(cmask | m1 | m2);
  const uint64_t count = s->len >> ins.n;
  // touched lines: those where the two bits differ (and controls set)
  double bytes = 32.0 * (double)s->len / (double)(1ULL << popc(cmask & ~3ULL));
  if (q1 >= 2 && q2 >= 2) bytes *= 0.5;
  timing_begin(s, KK_PERM, bytes);
  if (use32(s))
    hipLaunchKernelGGL((k_swap2<uint32_t>), dim3(grid_for(count)), dim3(kBlock), 0, s->stream, a,
                       (uint32_t)count, ins, (uint32_t)cmask, (uint32_t)m1, (uint32_t)m2);
  else
    hipLaunchKernelGGL((k_swap2<uint64_t>), dim3(grid_for(count)), dim3(kBlock), 0, s->stream, a,
                       (uint64_t)count, ins, (uint64_t)cmask, m1, m2);
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_matrix4(qoa_state* s, uint32_t q1, uint32_t q2, const double* m32) {
  double2* a = (double2*)s->amps;
  const uint32_t lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
  Mat4 u;
  for (int k = 0; k < 32; ++k) u.m[k] = m32[k];
  Ins ins = make_ins((1ULL << lo) | (1ULL << hi));
  const uint64_t count = s->len >> 2;
  timing_begin(s, KK_DENSE, 32.0 * (double)s->len);
  if (use32(s))
    hipLaunchKernelGGL((k_mat4<uint32_t>), dim3(grid_for(count)), dim3(kBlock), 0, s->stream, a,
                       (uint32_t)count, ins, (uint32_t)(1u << lo), (uint32_t)(1u << hi), u);
  else
    hipLaunchKernelGGL((k_mat4<uint64_t>), dim3(grid_for(count)), dim3(kBlock), 0, s->stream, a,
                       (uint64_t)count, ins, (uint64_t)(1ULL << lo), (uint64_t)(1ULL << hi), u);
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_pauli_codes(qoa_state* s, const uint8_t* codes, uint32_t n_codes) {
  PauliList pl{};
  uint64_t xmask = 0;
  for (uint32_t q = 0; q < n_codes && q < s->n; ++q) {
    if (codes[q] < 1 || codes[q] > 3) continue;
    pl.q[pl.n] = (uint8_t)q;
    pl.code[pl.n] = codes[q];
    pl.n++;
    if (codes[q] != 3) xmask |= 1ULL << q;
  }
  if (pl.n == 0) return QOA_OK;
  double2* a = (double2*)s->amps;
  timing_begin(s, KK_PERM, 32.0 * (double)s->len);
  if (xmask) {
    const int low = __builtin_ctzll(xmask);
    const uint64_t np = s->len >> 1;
    if (use32(s))
      hipLaunchKernelGGL((k_pauli_pairs<uint32_t>), dim3(grid_for(np)), dim3(kBlock), 0, s->stream,
                         a, (uint32_t)np, low, (uint32_t)xmask, pl);
    else
      hipLaunchKernelGGL((k_pauli_pairs<uint64_t>), dim3(grid_for(np)), dim3(kBlock), 0, s->stream,
                         a, (uint64_t)np, low, xmask, pl);
  } else {
    if (use32(s))
      hipLaunchKernelGGL((k_pauli_diag<uint32_t>), dim3(grid_for(s->len)), dim3(kBlock), 0,
                         s->stream, a, (uint32_t)s->len, pl);
    else
      hipLaunchKernelGGL((k_pauli_diag<uint64_t>), dim3(grid_for(s->len)), dim3(kBlock), 0,
                         s->stream, a, (uint64_t)s->len, pl);
  }
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

int launch_set_basis0(qoa_state* s) {
  double2* a = (double2*)s->amps;
  timing_begin(s, KK_OTHER, 16.0 * (double)s->len);
  if (use32(s))
    hipLaunchKernelGGL((k_set_basis0<uint32_t>), dim3(grid_for(s->len)), dim3(kBlock), 0,
                       s->stream, a, (uint32_t)s->len);
  else
    hipLaunchKernelGGL((k_set_basis0<uint64_t>), dim3(grid_for(s->len)), dim3(kBlock), 0,
                       s->stream, a, (uint64_t)s->len);
  timing_end(s);
  QOA_HIP_CHECK(hipGetLastError());
  return QOA_OK;
}

}  // namespace qoa
