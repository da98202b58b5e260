// batch_kernel.hpp — the work-item pass kernel shared by batches and the server.
//
// One launch runs one pass for many states: block b finds its work item (the
// last item whose first block is <= b) and runs a register-resident pass on
// tile b - item.block_begin of that item's state. Items carry their own
// trajectory-noise key and X-frame slots.
#pragma once
#include "planner.hpp"
#include "qoa_internal.hpp"
#include "tile_device.hpp"

#include <algorithm>
#include <vector>

namespace qoa {
namespace tile {

constexpr uint32_t kBatchMagic = 0x51424154u;  // "QBAT"
constexpr int kBatchThreads = 256;              // tiles of up to 12 qubits

// Threads per block for a tile of 2^k amplitudes: one wave up to k = 10
// (2^(k-4) <= 64 threads), else 2^(k-4).
constexpr uint32_t batch_threads(uint32_t k) { return k <= 10 ? 64u : 1u << (k - kRegBits); }

struct WorkItem {
  uint64_t state_off;    // first amplitude of the instance's state
  uint64_t codes_off;    // noise draws of op i of the op buffer: codes[codes_off + i]
                         // (the sum taken mod 2^64)
  uint32_t ops_begin, nops;
  uint32_t block_begin;  // first block of this item in its round
  uint32_t traj;
  uint32_t frame_in, frame_out;  // slots of the X frame before / after the pass
  uint8_t k, init, noise;
  uint8_t fma;      // the item's plan has merged diagonal runs: FMA kernels
  uint32_t spw[4];  // tile qubit load positions as in Geo (whole words: scalar loads)
  uint32_t stw[4];  // and store positions (spw unless the plan relabels qubits)
  uint32_t srt[4];  // the positions ascending (the tile base)
  uint8_t pad1[4];
};
static_assert(sizeof(WorkItem) == 96, "WorkItem layout");

// NOISE = false: no item of the launch has trajectory noise. SWAPS: the
// items' plans may hold swap-list layout ops (tiles of up to 10 qubits,
// planned with kPlanSwaps); plans of larger tiles hold whole-tile exchanges
// only, which either instantiation runs. With noise or swaps at most 168
// VGPRs (three waves per SIMD); the noise-free swap-free kernel, whose blocks
// have no idle lanes, at most 128 (it takes 108: four waves per SIMD, measured
// 1% faster than a 96-VGPR five-wave build with two spilled registers).
// F: FMA arithmetic with merged diagonal runs (noise-free batches planned
// with QOA_FUSE_MERGE_DIAG).
template <typename I, bool NOISE, bool SWAPS, bool F = false>
__global__ __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(NOISE || SWAPS ? 3 : 4)))
void k_batch_reg(double2* __restrict__ states,
                                                            const WorkItem* __restrict__ items,
                                                            int nitems,
                                                            const TileOp* __restrict__ ops,
                                                            const double* __restrict__ pool,
                                                            uint64_t* __restrict__ frames,
                                                            const uint32_t* __restrict__ codes) {
  extern __shared__ double2 lds[];
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {  // last item whose first block is <= blockIdx.x (uniform)
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].block_begin <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const WorkItem& it = items[lo];
  // Batch plans do not relabel (stw = spw, srt = spw). The noisy kernel reads
  // the positions from the item (measured: 161 -> 152 VGPRs, scratch 44 -> 24
  // bytes per lane, serving 423 -> 445 jobs/s); the noise-free ones keep
  // their five words in registers (pointer reads measured 6 % slower there).
  std::conditional_t<NOISE, GeoItem, GeoSame> geo;
  geo.k = it.k;
  if constexpr (NOISE) {
    geo.spw = it.spw;
    geo.stw = it.stw;
    geo.srt = it.srt;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) geo.spw[j] = it.spw[j];
  }
  const I base = tile_base((I)(blockIdx.x - it.block_begin), geo);
  // Groups of 11+-qubit tiles (the only ones the swap-free kernel runs) have
  // exactly 2^(k-4) threads per block; one-wave blocks of smaller tiles have
  // idle lanes.
  const bool act = !SWAPS || threadIdx.x < (1u << (it.k - kRegBits));
  NoiseArgs nz;
  nz.on = it.noise;
  nz.codes = codes + (it.codes_off + it.ops_begin);
  nz.frame = (it.noise && !it.init) ? frames[it.frame_in] : 0ULL;
  const uint64_t fout = reg_tile<I, F, NOISE, true, SWAPS>(states + it.state_off, base, geo, ops + it.ops_begin,
                                    (int)it.nops, pool, lds, act, it.init != 0, nz);
  // every tile of the instance computes the same frame; the first one stores it
  if (NOISE && it.noise && blockIdx.x == it.block_begin && threadIdx.x == 0)
    frames[it.frame_out] = fout;
}

// Items of one launch: all have the same block size.
struct ItemGroup {
  uint32_t first = 0, count = 0;  // range in the grouped item array
  uint32_t threads = 0, kmax = kRegBits;
  uint64_t blocks = 0;
  bool noise = false;
  bool fma = false;
};

// Orders items (with nblk[i] tiles each) into groups of equal block size,
// stable within a group, and numbers block_begin within each group.
inline void group_items(const std::vector<WorkItem>& in, const std::vector<uint64_t>& nblk,
                        std::vector<WorkItem>* out, std::vector<ItemGroup>* groups) {
  out->clear();
  groups->clear();
  for (uint32_t th = 64; th <= (uint32_t)kBatchThreads; th *= 2) {
    ItemGroup g;
    g.first = (uint32_t)out->size();
    g.threads = th;
    for (size_t i = 0; i < in.size(); ++i) {
      if (batch_threads(in[i].k) != th) continue;
      WorkItem w = in[i];
      w.block_begin = (uint32_t)g.blocks;
      g.blocks += nblk[i];
      g.kmax = std::max<uint32_t>(g.kmax, w.k);
      g.noise = g.noise || w.noise != 0;
      g.fma = g.fma || w.fma != 0;
      out->push_back(w);
    }
    g.count = (uint32_t)out->size() - g.first;
    if (g.count) groups->push_back(g);
  }
}

// Lets every instantiation take dynamic LDS beyond the default limit (the
// exchange buffer holds 2^(k - reg_xchg_split(k, true)) amplitudes).
inline void batch_kernel_attrs() {
  const void* fns[5] = {(const void*)k_batch_reg<uint32_t, false, false>,
                        (const void*)k_batch_reg<uint32_t, false, true>,
                        (const void*)k_batch_reg<uint32_t, true, true>,
                        (const void*)k_batch_reg<uint32_t, false, false, true>,
                        (const void*)k_batch_reg<uint32_t, false, true, true>};
  for (const void* f : fns)
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 139264);
}

// Launches one group (items already on the device at d_items + g.first).
// Noise-free one-wave groups (tiles of up to 10 qubits) run the swap-list
// kernel, larger noise-free tiles the kernel without it. Noisy groups always
// run the swap-list kernel: without the swap code the noisy kernel's register
// allocation spills 24 VGPRs instead of 3, and its serving throughput
// measured 125 jobs/s against 190 with it.
inline void launch_group(hipStream_t st, const ItemGroup& g, double2* states,
                         const WorkItem* d_items, const TileOp* ops, const double* pool,
                         uint64_t* frames, const uint32_t* codes) {
  const size_t lbytes = sizeof(double2) * reg_xchg_elems(g.kmax, true);
  // (merged plans exist only in noise-free batches)
  auto fn = g.noise ? k_batch_reg<uint32_t, true, true>
            : g.fma ? (g.kmax <= 10 ? k_batch_reg<uint32_t, false, true, true>
                                    : k_batch_reg<uint32_t, false, false, true>)
                    : (g.kmax <= 10 ? k_batch_reg<uint32_t, false, true>
                                    : k_batch_reg<uint32_t, false, false>);
  hipLaunchKernelGGL(fn, dim3((unsigned)g.blocks), dim3(g.threads), lbytes, st, states,
                     d_items + g.first, (int)g.count, ops, pool, frames, codes);
}

// Trajectory draws of one job: word t * nops + o holds, 2 bits per operand
// j of op o, traj_pauli(seed, t, gi, qphys[j], p1 or p2) for trajectory t
// (the counters of the in-kernel channel, drawn once instead of in every
// tile).
static __global__ __launch_bounds__(256) void k_draw_codes(const TileOp* __restrict__ ops, uint32_t nops,
                                                    uint64_t seed, double p1, double p2,
                                                    uint32_t ntraj, uint32_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (uint64_t)ntraj * nops) return;
  const uint32_t t = (uint32_t)(i / nops), o = (uint32_t)(i % nops);
  const TileOp& op = ops[o];
  uint32_t w = 0;
  for (uint32_t j = 0; j < op.nq && j < 3; ++j)
    w |= traj_pauli(seed, t, op.gi, op.qphys[j], op.psel == 1 ? p1 : p2) << (2 * j);
  out[i] = w;
}

inline void draw_codes(hipStream_t st, const TileOp* ops, uint32_t nops, uint64_t seed, double p1,
                       double p2, uint32_t ntraj, uint32_t* out) {
  const uint64_t n = (uint64_t)ntraj * nops;
  if (n == 0) return;
  hipLaunchKernelGGL(k_draw_codes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ops, nops,
                     seed, p1, p2, ntraj, out);
}

}  // namespace tile
}  // namespace qoa
