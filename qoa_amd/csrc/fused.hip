// fused.hip — LDS-tiled multi-gate passes (FP64) for gfx950.
//
// One workgroup per tile: the 2^k amplitudes whose bits outside the pass's
// qubit set S equal the tile's base. The tile is read from HBM once (S holds
// the low qubits, so loads come in contiguous runs), every gate of the pass
// is applied in LDS with the same per-amplitude arithmetic as the per-gate
// sweeps (this file is built with -ffp-contract=off), and the tile is written
// back once. Gates that only read a qubit outside S see that bit as a
// constant of the tile and become per-tile conditions or scalars.
#include <cstring>
#include <mutex>

#include "planner.hpp"
#include "qoa_internal.hpp"

#pragma clang fp contract(off)

namespace qoa {
namespace {

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

template <typename I>
__global__ __launch_bounds__(kThreads) void k_fused(double2* __restrict__ a,
                                                   const TileOp* __restrict__ ops, int nops,
                                                   const double* __restrict__ pool, Geo geo) {
  extern __shared__ double2 lds[];
  const uint32_t T = 1u << geo.k;
  const uint32_t tid = threadIdx.x;
  // tile base: the block index spread over the bits outside S
  I base = (I)blockIdx.x;
  for (int j = 0; j < geo.k; ++j) {
    const int b = geo.spos[j];
    base = ((base >> b) << (b + 1)) | (base & ((I(1) << b) - I(1)));
  }
  for (uint32_t l = tid; l < T; l += kThreads) lds[l] = a[base | scatter<I>(l, geo)];
  __syncthreads();

  for (int o = 0; o < nops; ++o) {
    const TileOp& op = ops[o];
    if (((uint64_t)base & op.gmask) != op.gval) continue;  // uniform over the tile
    const uint32_t cnt = T >> (op.nins == 0xFF ? 0 : op.nins);
    switch (op.kind) {
      case TK_H: {
        const double2 nf = make_double2(op.c[0], op.c[1]);
        const uint32_t tb = 1u << op.t;
        for (uint32_t p = tid; p < cnt; p += kThreads) {
          const uint32_t l0 = dep(p, op) | op.lmask, l1 = l0 | tb;
          const double2 x = lds[l0], y = lds[l1];
          lds[l0] = cmul(nf, cadd(x, y));
          lds[l1] = cmul(nf, csub(x, y));
        }
      } break;
      case TK_M: {
        const double2 m00 = make_double2(op.c[0], op.c[1]), m01 = make_double2(op.c[2], op.c[3]);
        const double2 m10 = make_double2(op.c[4], op.c[5]), m11 = make_double2(op.c[6], op.c[7]);
        const uint32_t tb = 1u << op.t;
        for (uint32_t p = tid; p < cnt; p += kThreads) {
          const uint32_t l0 = dep(p, op) | op.lmask, l1 = l0 | tb;
          const double2 x = lds[l0], y = lds[l1];
          lds[l0] = cadd(cmul(m00, x), cmul(m01, y));
          lds[l1] = cadd(cmul(m10, x), cmul(m11, y));
        }
      } break;
      case TK_X: {
        const uint32_t tb = 1u << op.t;
        for (uint32_t p = tid; p < cnt; p += kThreads) {
          const uint32_t l0 = dep(p, op) | op.lmask, l1 = l0 | tb;
          const double2 x = lds[l0], y = lds[l1];
          lds[l0] = y;
          lds[l1] = x;
        }
      } break;
      case TK_Y: {
        const uint32_t tb = 1u << op.t;
        for (uint32_t p = tid; p < cnt; p += kThreads) {
          const uint32_t l0 = dep(p, op) | op.lmask, l1 = l0 | tb;
          const double2 x = lds[l0], y = lds[l1];
          lds[l0] = cmul(y, make_double2(0.0, -1.0));
          lds[l1] = cmul(x, make_double2(0.0, 1.0));
        }
      } break;
      case TK_DNEG:
      case TK_DPHASE: {
        const bool neg = op.kind == TK_DNEG;
        const double2 f = make_double2(op.c[0], op.c[1]);
        if (op.nins == 0xFF) {
          for (uint32_t l = tid; l < T; l += kThreads) {
            if ((l & op.lmask) != op.lmask) continue;
            const double2 v = lds[l];
            lds[l] = neg ? make_double2(v.x * -1.0, v.y * -1.0) : cmul(v, f);
          }
        } else {
          for (uint32_t p = tid; p < cnt; p += kThreads) {
            const uint32_t l = dep(p, op) | op.lmask;
            const double2 v = lds[l];
            lds[l] = neg ? make_double2(v.x * -1.0, v.y * -1.0) : cmul(v, f);
          }
        }
      } break;
      case TK_RZ: {
        const double2 m00 = make_double2(op.c[0], op.c[1]), m11 = make_double2(op.c[2], op.c[3]);
        for (uint32_t l = tid; l < T; l += kThreads) {
          const double2 v = lds[l];
          lds[l] = ((l >> op.t) & 1u) ? cmul(v, m11) : cmul(v, m00);
        }
      } break;
      case TK_SWAP: {
        const uint32_t m1 = 1u << op.t, m2 = 1u << op.t2;
        for (uint32_t p = tid; p < cnt; p += kThreads) {
          const uint32_t l = dep(p, op) | op.lmask;
          const double2 u = lds[l | m2], w = lds[l | m1];
          lds[l | m2] = w;
          lds[l | m1] = u;
        }
      } break;
      case TK_MAT4: {
        // local (t, t2) = (q1, q2); the reference orders the 4x4 basis by the
        // (higher, lower) physical qubit, and local order follows physical order.
        const uint32_t lo = op.t < op.t2 ? op.t : op.t2, hi = op.t < op.t2 ? op.t2 : op.t;
        const uint32_t ml = 1u << lo, mh = 1u << hi;
        const double* u = pool + op.aux;
        for (uint32_t p = tid; p < cnt; p += kThreads) {
          const uint32_t b = dep(p, op);
          const uint32_t idx[4] = {b, b | ml, b | mh, b | ml | mh};
          double2 v[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) v[q] = lds[idx[q]];
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            double2 acc = cmul(make_double2(u[8 * s], u[8 * s + 1]), v[0]);
#pragma unroll
            for (int q = 1; q < 4; ++q)
              acc = cadd(acc, cmul(make_double2(u[8 * s + 2 * q], u[8 * s + 2 * q + 1]), v[q]));
            lds[idx[s]] = acc;
          }
        }
      } break;
      default:
        break;
    }
    __syncthreads();
  }

  for (uint32_t l = tid; l < T; l += kThreads) a[base | scatter<I>(l, geo)] = lds[l];
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

template <typename I>
__global__ __launch_bounds__(512) void k_fused_reg(double2* __restrict__ a,
                                                  const TileOp* __restrict__ ops, int nops,
                                                  const double* __restrict__ pool, Geo geo) {
  extern __shared__ double2 lds[];
  const int k = geo.k;
  const uint32_t tid = threadIdx.x;
  I base = (I)blockIdx.x;
  for (int j = 0; j < k; ++j) {
    const int b = geo.spos[j];
    base = ((base >> b) << (b + 1)) | (base & ((I(1) << b) - I(1)));
  }
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
#pragma unroll
  for (int r = 0; r < kR; ++r) v[r] = a[base | gthr | scatter<I>(regl[r], geo)];

  for (int o = 0; o < nops; ++o) {
    const TileOp& op = ops[o];
    if (((uint64_t)base & op.gmask) != op.gval) continue;  // uniform over the tile
    switch (op.kind) {
      case TK_LAYOUT: {
#pragma unroll
        for (int r = 0; r < kR; ++r) lds[lpad(thr_l | regl[r])] = v[r];
        __syncthreads();
        uint8_t np[kRegBits];
#pragma unroll
        for (int j = 0; j < kRegBits; ++j) np[j] = op.ins[j];
        make_layout(tid, k, np, thr_l, regl);
#pragma unroll
        for (int r = 0; r < kR; ++r) v[r] = lds[lpad(thr_l | regl[r])];
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
#pragma unroll
  for (int r = 0; r < kR; ++r) a[base | gthr | scatter<I>(regl[r], geo)] = v[r];
}

// FNV-1a over the bytes that determine a plan.
uint64_t fnv(const void* p, size_t n, uint64_t h) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 0x100000001B3ULL;
  }
  return h;
}

}  // namespace

// Device copy of the last plan run on a handle (keyed by a hash of its inputs).
struct FusedCache {
  uint64_t key = 0;
  FusedPlan plan;
  TileOp* d_ops = nullptr;
  double* d_pool = nullptr;
};

void fused_cache_free(qoa_state* s) {
  FusedCache* c = (FusedCache*)s->fused_cache;
  if (!c) return;
  if (c->d_ops) hipFree(c->d_ops);
  if (c->d_pool) hipFree(c->d_pool);
  delete c;
  s->fused_cache = nullptr;
}

int run_fused(qoa_state* s, const qoa_gate* g, uint32_t ng, const double* mats,
              const PlanOptions& opt) {
  uint64_t key = 0xCBF29CE484222325ULL;
  key = fnv(&s->n, sizeof(s->n), key);
  key = fnv(&opt, sizeof(opt), key);
  key = fnv(&ng, sizeof(ng), key);
  key = fnv(g, sizeof(qoa_gate) * ng, key);
  for (uint32_t i = 0; i < ng; ++i) {
    if (g[i].op == QOA_OP_U2 && mats) key = fnv(mats + 8 * g[i].aux, 64, key);
    if (g[i].op == QOA_OP_U4 && mats) key = fnv(mats + 32 * g[i].aux, 256, key);
  }
  FusedCache* c = (FusedCache*)s->fused_cache;
  if (!c || c->key != key) {
    fused_cache_free(s);
    c = new FusedCache();
    s->fused_cache = c;
    int rc = build_plan(s->n, g, ng, mats, opt, &c->plan);
    if (rc) {
      fused_cache_free(s);
      return rc;
    }
    const size_t ob = sizeof(TileOp) * (c->plan.ops.size() + 1);
    const size_t pb = sizeof(double) * (c->plan.pool.size() + 1);
    if (hipMalloc(&c->d_ops, ob) != hipSuccess || hipMalloc(&c->d_pool, pb) != hipSuccess) {
      (void)hipGetLastError();
      fused_cache_free(s);
      return set_error(QOA_ERR_OOM, "fused plan upload failed");
    }
    if (!c->plan.ops.empty())
      QOA_HIP_CHECK(hipMemcpyAsync(c->d_ops, c->plan.ops.data(), sizeof(TileOp) * c->plan.ops.size(),
                                   hipMemcpyHostToDevice, s->stream));
    if (!c->plan.pool.empty())
      QOA_HIP_CHECK(hipMemcpyAsync(c->d_pool, c->plan.pool.data(),
                                   sizeof(double) * c->plan.pool.size(), hipMemcpyHostToDevice,
                                   s->stream));
    QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
    c->key = key;
  }
  const FusedPlan& P = c->plan;
  const uint32_t k = P.k;
  const size_t lds_bytes = sizeof(double2) << k;
  static std::once_flag once;
  std::call_once(once, [] {
    (void)hipFuncSetAttribute((const void*)k_fused<uint32_t>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
    (void)hipFuncSetAttribute((const void*)k_fused<uint64_t>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
    (void)hipFuncSetAttribute((const void*)k_fused_reg<uint32_t>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
    (void)hipFuncSetAttribute((const void*)k_fused_reg<uint64_t>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
  });
  const uint64_t tiles = 1ULL << (s->n - k);
  for (const PassPlan& pp : P.passes) {
    Geo geo{};
    geo.k = (int)k;
    std::memcpy(geo.spos, pp.spos, sizeof(geo.spos));
    const int nops = (int)(pp.op_end - pp.op_begin);
    const TileOp* ops = c->d_ops + pp.op_begin;
    double2* a = (double2*)s->amps;
    timing_begin(s, KK_FUSED, 32.0 * (double)s->len);
    if (pp.reg_mode) {
      // 16 amplitudes per thread; LDS holds the padded tile for layout changes
      const unsigned threads = 1u << (k - kRegBits);
      const size_t lbytes = sizeof(double2) * ((1u << k) + (1u << k) / 16);
      if (use32(s))
        hipLaunchKernelGGL((k_fused_reg<uint32_t>), dim3((unsigned)tiles), dim3(threads), lbytes,
                           s->stream, a, ops, nops, c->d_pool, geo);
      else
        hipLaunchKernelGGL((k_fused_reg<uint64_t>), dim3((unsigned)tiles), dim3(threads), lbytes,
                           s->stream, a, ops, nops, c->d_pool, geo);
    } else if (use32(s)) {
      hipLaunchKernelGGL((k_fused<uint32_t>), dim3((unsigned)tiles), dim3(kThreads), lds_bytes,
                         s->stream, a, ops, nops, c->d_pool, geo);
    } else {
      hipLaunchKernelGGL((k_fused<uint64_t>), dim3((unsigned)tiles), dim3(kThreads), lds_bytes,
                         s->stream, a, ops, nops, c->d_pool, geo);
    }
    timing_end(s);
    QOA_HIP_CHECK(hipGetLastError());
  }
  return QOA_OK;
}

}  // namespace qoa
