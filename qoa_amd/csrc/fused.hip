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
#include "specialize.hpp"
#include "tile_device.hpp"

#pragma clang fp contract(off)

namespace qoa {
namespace {

using namespace tile;

template <typename I>
__global__ __launch_bounds__(kThreads) void k_fused(double2* __restrict__ a,
                                                   const TileOp* __restrict__ ops, int nops,
                                                   const double* __restrict__ pool, Geo geo) {
  extern __shared__ double2 lds[];
  const uint32_t T = 1u << geo.k;
  const uint32_t tid = threadIdx.x;
  // tile base: the block index spread over the bits outside S (LDS passes
  // load and store at the same positions, spw)
  const I base = tile_base((I)blockIdx.x, geo);
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

// At most 96 VGPRs: five waves per SIMD (five 256-thread blocks, 32 KiB of
// LDS each, fill a CU).
template <typename I, bool F>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5))) void k_fused_reg(double2* __restrict__ a,
                                                  const TileOp* __restrict__ ops, int nops,
                                                  const double* __restrict__ pool, Geo geo) {
  extern __shared__ double2 lds[];
  const I base = tile_base((I)blockIdx.x, geo);
  reg_tile<I, F, false, false, false>(a, base, geo, ops, nops, pool, lds, true, false, NoiseArgs{});
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
  // QOA_FUSE_SPECIALIZE: one compiled kernel per pass, for the arithmetic
  // (spec_mode bit 0: FMA) and index width (bit 1: 64-bit) they were made for
  SpecKernels spec;
  int spec_mode = -1;
};

void fused_cache_free(qoa_state* s) {
  FusedCache* c = (FusedCache*)s->fused_cache;
  if (!c) return;
  if (c->d_ops) hipFree(c->d_ops);
  if (c->d_pool) hipFree(c->d_pool);
  delete c;
  s->fused_cache = nullptr;
}

Geo pass_geo(const PassPlan& pp, uint32_t k) {
  Geo g{};
  g.k = (int)k;
  pass_geo_words(pp, g.spw, g.stw, g.srt);
  return g;
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
  // merged diagonal runs exist in the FMA instantiation only
  const bool fma = (opt.flags & (QOA_FUSE_FMA | QOA_FUSE_MERGE_DIAG | QOA_FUSE_FACTOR)) != 0;
  bool spec = false;
  if (opt.flags & QOA_FUSE_SPECIALIZE) {
    spec = true;
    for (const PassPlan& pp : P.passes) spec = spec && pp.reg_mode;
    const int mode = (fma ? 1 : 0) | (use32(s) ? 0 : 2);
    if (spec && c->spec_mode != mode) {
      c->spec_mode = -1;
      int rc = specialize_plan(P, fma, !use32(s), &c->spec);
      if (rc) return rc;
      c->spec_mode = mode;
    }
  }
  const size_t lds_bytes = sizeof(double2) << k;
  static std::once_flag once;
  std::call_once(once, [] {
    (void)hipFuncSetAttribute((const void*)k_fused<uint32_t>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
    (void)hipFuncSetAttribute((const void*)k_fused<uint64_t>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
    (void)hipFuncSetAttribute((const void*)k_fused_reg<uint32_t, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
    (void)hipFuncSetAttribute((const void*)k_fused_reg<uint64_t, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
    (void)hipFuncSetAttribute((const void*)k_fused_reg<uint32_t, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
    (void)hipFuncSetAttribute((const void*)k_fused_reg<uint64_t, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
  });
  const uint64_t tiles = 1ULL << (s->n - k);
  for (const PassPlan& pp : P.passes) {
    const Geo geo = pass_geo(pp, k);
    const int nops = (int)(pp.op_end - pp.op_begin);
    const TileOp* ops = c->d_ops + pp.op_begin;
    double2* a = (double2*)s->amps;
    timing_begin(s, KK_FUSED, 32.0 * (double)s->len);
    if (pp.reg_mode) {
      // 16 amplitudes per thread; LDS holds the exchange buffer for layout
      // changes (a quarter of a 12-qubit tile)
      const unsigned threads = 1u << (k - kRegBits);
      const size_t lbytes = sizeof(double2) * reg_xchg_elems(k, false);
      const dim3 grid((unsigned)tiles), block(threads);
      if (spec) {
        const double* pool = c->d_pool;
        void* args[] = {&a, &ops, &pool};
        QOA_HIP_CHECK(hipModuleLaunchKernel(c->spec.fn[&pp - P.passes.data()], (unsigned)tiles, 1,
                                            1, threads, 1, 1, (unsigned)lbytes, s->stream, args,
                                            nullptr));
      } else if (use32(s)) {
        if (fma)
          hipLaunchKernelGGL((k_fused_reg<uint32_t, true>), grid, block, lbytes, s->stream, a, ops,
                             nops, c->d_pool, geo);
        else
          hipLaunchKernelGGL((k_fused_reg<uint32_t, false>), grid, block, lbytes, s->stream, a,
                             ops, nops, c->d_pool, geo);
      } else if (fma) {
        hipLaunchKernelGGL((k_fused_reg<uint64_t, true>), grid, block, lbytes, s->stream, a, ops,
                           nops, c->d_pool, geo);
      } else {
        hipLaunchKernelGGL((k_fused_reg<uint64_t, false>), grid, block, lbytes, s->stream, a, ops,
                           nops, c->d_pool, geo);
      }
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
