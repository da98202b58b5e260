// batch.hip — many independent circuits of uneven size in one launch per round.
//
// Every job instance (job x trajectory) owns a state in one device buffer and
// a fused plan (register-resident passes). Round r runs pass r of every
// instance that has one: the launch's blocks are the concatenated tiles of
// those passes, found through a per-round work-item table (prefix sums of
// tile counts). The number of launches is the largest pass count in the
// batch, not the number of jobs. Pass 0 synthesises |0...0> instead of
// reading the state.
//
// The FP32-fused form (QOA_C64 jobs) runs dense 5-qubit blocks instead:
// round r applies block r of every job on FP32 matrix cores, each workgroup
// finding its job and 512-column share through the round's item table.
//
// A batch's whole launch sequence is captured once as a HIP graph and
// replayed (no host launch cost per round). QOA_BATCH_AUTOTUNE picks each
// qubit count's tile by timing one of its jobs with each candidate.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <array>
#include <map>
#include <tuple>
#include <mutex>
#include <thread>

#include "batch_kernel.hpp"
#include "block_device.hpp"
#include "blocks_plan.hpp"
#include "planner.hpp"
#include "qoa_internal.hpp"

namespace qoa {
namespace {

using namespace tile;


struct Instance {
  uint32_t job, traj, n;
  uint64_t off;  // amplitude offset in the batch buffer
};

struct Round {
  WorkItem* d_items = nullptr;
  std::vector<ItemGroup> groups;  // one launch each
};

// One job instance's block of a C64 round.
struct BlkItem {
  uint64_t state_off;  // first amplitude (float2) of the instance
  uint64_t w_off;      // the block's A operands: w + w_off (kWFloats floats)
  uint64_t ncols;      // 2^(n-5)
  uint32_t wg_begin;   // first workgroup of this item in its round
  uint8_t q[kBlockQ];  // the block's qubits, ascending
  uint8_t pad[11];
};
static_assert(sizeof(BlkItem) == 48, "BlkItem layout");

struct BlkRound {
  BlkItem* d_items = nullptr;
  uint32_t nitems = 0;
  uint64_t wgs = 0;
};

// Workgroup b of a C64 round: its item (the last whose first workgroup is
// <= b) and that item's 512-column share.
__global__ __launch_bounds__(kBlockThreads) void k_block_batch(float2* __restrict__ states,
                                                               const float* __restrict__ w,
                                                               const BlkItem* __restrict__ items,
                                                               int nitems) {
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].wg_begin <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const BlkItem& it = items[lo];
  BlockArgs ba;
  ba.ncols = it.ncols;
#pragma unroll
  for (int j = 0; j < kBlockQ; ++j) ba.q[j] = it.q[j];
  block_apply<uint32_t>(states + it.state_off, w + it.w_off, ba, blockIdx.x - it.wg_begin);
}

// |0...0> in every C64 instance: zero everything, then 1 at each start.
__global__ __launch_bounds__(256) void k_zero_c64(float2* __restrict__ a, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    a[i] = make_float2(0.f, 0.f);
}
__global__ __launch_bounds__(256) void k_starts_c64(float2* __restrict__ a,
                                                    const uint64_t* __restrict__ starts,
                                                    uint32_t count) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < count) a[starts[i]] = make_float2(1.f, 0.f);
}

}  // namespace
}  // namespace qoa

struct qoa_batch {
  uint32_t magic;
  int device;
  hipStream_t stream;
  double2* states;
  uint64_t total_amps;
  uint32_t max_n;
  std::vector<qoa_job_desc> jobs;
  std::vector<qoa::Instance> inst;
  std::vector<uint32_t> first_inst;  // per job: index of its trajectory 0
  std::vector<qoa::FusedPlan> plans;  // per job
  qoa::TileOp* d_ops;
  double* d_pool;
  std::vector<qoa::Round> rounds;
  uint64_t* frames;  // 2 X-frame slots per instance (noisy trajectories)
  uint32_t* codes;   // noise draws: per instance one word per op of its job's plan
  bool noisy;
  bool merged;       // noise-free plans with merged diagonal runs (FMA kernels)
  double alg_bytes;  // algorithmic HBM bytes of one run
  hipEvent_t ev0, ev1;
  float last_ms;
  uint32_t precision;              // QOA_C128: register passes; QOA_C64: 5-qubit blocks
  std::vector<uint32_t> job_tile;  // C128: a job's tile qubits; C64: its block count
  // C64 (FP32 5-qubit blocks)
  float2* states64;
  float* d_w;                      // every block's A operands, job after job
  uint64_t* d_starts;              // each instance's first amplitude
  std::vector<qoa::BlkRound> brounds;
  // the launch sequence as a graph (captured at the first run)
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace qoa {

static bool batch_valid(const qoa_batch* b) { return b && b->magic == kBatchMagic; }

static void batch_free(qoa_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  for (auto& r : b->rounds)
    if (r.d_items) hipFree(r.d_items);
  for (auto& r : b->brounds)
    if (r.d_items) hipFree(r.d_items);
  if (b->states64) hipFree(b->states64);
  if (b->d_w) hipFree(b->d_w);
  if (b->d_starts) hipFree(b->d_starts);
  if (b->d_ops) hipFree(b->d_ops);
  if (b->d_pool) hipFree(b->d_pool);
  if (b->frames) hipFree(b->frames);
  if (b->codes) hipFree(b->codes);
  if (b->states) hipFree(b->states);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->stream) hipStreamDestroy(b->stream);
  b->magic = 0;
  delete b;
}

// ---- autotuning ------------------------------------------------------------------
// Tile per qubit count, chosen by timing one job of that count (the one with
// the most gates) with each candidate tile, as a batch of copies large enough
// to fill the device (2^23 amplitudes, at most 256 copies): the measure is a
// job's share of batched throughput, not its latency alone. Cached for the
// process by (qubit count, gate count, mask flags).
std::mutex g_tune_mu;
std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t> g_tuned;

int autotune_tiles(const qoa_job_desc* jobs, uint32_t n_jobs, const qoa_gate* gates,
                   uint64_t n_gates_total, const double* mats, const qoa_fuse_opts* opts,
                   uint32_t tile, std::vector<uint32_t>* tile_of) {
  tile_of->assign(n_jobs, tile);
  std::map<uint32_t, uint32_t> rep;  // qubit count -> representative job
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const uint32_t n = jobs[j].n_qubits;
    if (n <= 10) continue;  // a whole-state tile; nothing to choose
    auto it = rep.find(n);
    if (it == rep.end() || jobs[j].n_gates > jobs[it->second].n_gates) rep[n] = j;
  }
  // trials in the arithmetic the batch will run
  const uint32_t fl =
      opts ? (opts->flags & (QOA_CZ_REF_MASKS | QOA_FUSE_FACTOR | QOA_FUSE_MERGE_DIAG)) : 0u;
  std::map<uint32_t, uint32_t> best_of;
  for (const auto& kv : rep) {
    const uint32_t n = kv.first, j = kv.second;
    const auto key = std::make_tuple(n, jobs[j].n_gates, fl);
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto c = g_tuned.find(key);
      if (c != g_tuned.end()) {
        best_of[n] = c->second;
        continue;
      }
    }
    double best_ms = 1e300;
    uint32_t best_k = std::min(tile, n);
    for (uint32_t k = 10; k <= 12 && k <= n; ++k) {
      qoa_job_desc d = jobs[j];
      d.trajectories = n >= 23 ? 1u : std::min(256u, 1u << (23 - n));  // copies of the job
      d.p1 = d.p2 = 0.0;
      qoa_fuse_opts o{k, opts ? opts->low_qubits : 0u, fl, 0u};
      qoa_batch* tb = nullptr;
      int rc = qoa_batch_create(&d, 1, gates, n_gates_total, mats, &o, &tb);
      if (rc) return rc;
      double ms = 1e300, t = 0;
      for (int rep_i = 0; rep_i < 3 && rc == QOA_OK; ++rep_i) {
        rc = qoa_batch_run(tb);
        if (rc == QOA_OK) rc = qoa_batch_synchronize(tb, &t);
        if (rep_i > 0) ms = std::min(ms, t);  // the first run warms up (and captures)
      }
      qoa_batch_destroy(tb);
      if (rc) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best_k = k;
      }
    }
    best_of[n] = best_k;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = best_k;
  }
  for (uint32_t j = 0; j < n_jobs; ++j) {
    auto it = best_of.find(jobs[j].n_qubits);
    if (it != best_of.end()) (*tile_of)[j] = it->second;
  }
  return QOA_OK;
}

// ---- FP32-fused form: 5-qubit blocks --------------------------------------------
// Fuses every job into blocks (host threads), uploads their A operands and one
// item table per round.
int create_c64(qoa_batch* b, const qoa_gate* gates, const double* mats, uint32_t flags) {
  const uint32_t n_jobs = (uint32_t)b->jobs.size();
  std::vector<std::vector<float>> w(n_jobs);
  std::vector<std::vector<std::array<uint8_t, kBlockQ>>> qs(n_jobs);
  std::vector<int> rcs(n_jobs, 0);
  {
    const unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    std::atomic<uint32_t> next{0};
    for (unsigned t = 0; t < nt; ++t) {
      pool.emplace_back([&]() {
        for (;;) {
          const uint32_t j = next.fetch_add(1);
          if (j >= n_jobs) break;
          const qoa_job_desc& d = b->jobs[j];
          std::vector<Block> blocks;
          rcs[j] = build_blocks(d.n_qubits, gates + d.gate_offset, d.n_gates, mats,
                                flags & QOA_CZ_REF_MASKS, kBlockQ, &blocks);
          if (rcs[j]) continue;
          w[j].resize((size_t)kWFloats * blocks.size());
          qs[j].resize(blocks.size());
          for (size_t k = 0; k < blocks.size(); ++k) {
            pack_operands(blocks[k], &w[j][(size_t)kWFloats * k]);
            std::memcpy(qs[j][k].data(), blocks[k].q, kBlockQ);
          }
        }
      });
    }
    for (auto& th : pool) th.join();
  }
  for (uint32_t j = 0; j < n_jobs; ++j)
    if (rcs[j]) return set_error(rcs[j], "job %u: block fusion failed: %s", j, qoa_last_error());
  std::vector<uint64_t> wbase(n_jobs);
  uint64_t wtot = 0;
  b->job_tile.resize(n_jobs);
  for (uint32_t j = 0; j < n_jobs; ++j) {
    wbase[j] = wtot;
    wtot += w[j].size();
    b->job_tile[j] = (uint32_t)qs[j].size();
  }
  uint64_t off = 0;
  std::vector<uint64_t> starts;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    b->first_inst[j] = (uint32_t)b->inst.size();
    for (uint32_t t = 0; t < b->jobs[j].trajectories; ++t) {
      b->inst.push_back(Instance{j, t, b->jobs[j].n_qubits, off});
      starts.push_back(off);
      off += 1ULL << b->jobs[j].n_qubits;  // >= 32 amplitudes = 256 bytes
    }
  }
  b->total_amps = off;
  if (hipMalloc(&b->states64, sizeof(float2) * off) != hipSuccess ||
      hipMalloc(&b->d_w, sizeof(float) * (wtot + 4)) != hipSuccess ||
      hipMalloc(&b->d_starts, sizeof(uint64_t) * starts.size()) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(QOA_ERR_OOM, "C64 batch: %.2f GiB of states and %.2f GiB of operands",
                     8.0 * off / 1073741824.0, 4.0 * wtot / 1073741824.0);
  }
  QOA_HIP_CHECK(hipMemcpy(b->d_starts, starts.data(), sizeof(uint64_t) * starts.size(),
                          hipMemcpyHostToDevice));
  for (uint32_t j = 0; j < n_jobs; ++j) {
    if (!w[j].empty())
      QOA_HIP_CHECK(hipMemcpy(b->d_w + wbase[j], w[j].data(), sizeof(float) * w[j].size(),
                              hipMemcpyHostToDevice));
    std::vector<float>().swap(w[j]);
  }
  uint32_t max_blocks = 0;
  for (const auto& q : qs) max_blocks = std::max<uint32_t>(max_blocks, (uint32_t)q.size());
  b->alg_bytes = 8.0 * (double)off;  // the |0...0> fill
  for (uint32_t r = 0; r < max_blocks; ++r) {
    std::vector<BlkItem> items;
    BlkRound rd;
    for (const Instance& in : b->inst) {
      if (r >= qs[in.job].size()) continue;
      BlkItem it{};
      it.state_off = in.off;
      it.w_off = wbase[in.job] + (uint64_t)kWFloats * r;
      it.ncols = 1ULL << (in.n - kBlockQ);
      it.wg_begin = (uint32_t)rd.wgs;
      std::memcpy(it.q, qs[in.job][r].data(), kBlockQ);
      rd.wgs += (it.ncols + kColsPerWg - 1) / kColsPerWg;
      items.push_back(it);
      b->alg_bytes += 16.0 * (double)(1ULL << in.n);
    }
    if (rd.wgs >= (1ULL << 31)) return set_error(QOA_ERR_ARG, "C64 round %u too large", r);
    rd.nitems = (uint32_t)items.size();
    if (hipMalloc(&rd.d_items, sizeof(BlkItem) * items.size()) != hipSuccess) {
      (void)hipGetLastError();
      return set_error(QOA_ERR_OOM, "C64 round items");
    }
    b->brounds.push_back(rd);
    QOA_HIP_CHECK(hipMemcpy(rd.d_items, items.data(), sizeof(BlkItem) * items.size(),
                            hipMemcpyHostToDevice));
  }
  return QOA_OK;
}

// Enqueues one run's launches on the batch's stream.
int enqueue_run(qoa_batch* b) {
  if (b->precision == QOA_C64) {
    const uint64_t zg = std::min<uint64_t>((b->total_amps + 255) / 256, 8192);
    hipLaunchKernelGGL(k_zero_c64, dim3((unsigned)zg), dim3(256), 0, b->stream, b->states64,
                       b->total_amps);
    const uint32_t ni = (uint32_t)b->inst.size();
    hipLaunchKernelGGL(k_starts_c64, dim3((ni + 255) / 256), dim3(256), 0, b->stream, b->states64,
                       b->d_starts, ni);
    QOA_HIP_CHECK(hipGetLastError());
    for (const BlkRound& rd : b->brounds) {
      hipLaunchKernelGGL(k_block_batch, dim3((unsigned)rd.wgs), dim3(kBlockThreads), 0, b->stream,
                         b->states64, b->d_w, rd.d_items, (int)rd.nitems);
      QOA_HIP_CHECK(hipGetLastError());
    }
    return QOA_OK;
  }
  for (const Round& rd : b->rounds) {
    for (const ItemGroup& g : rd.groups) {
      launch_group(b->stream, g, b->states, rd.d_items, b->d_ops, b->d_pool, b->frames, b->codes);
      QOA_HIP_CHECK(hipGetLastError());
    }
  }
  return QOA_OK;
}

}  // namespace qoa

using namespace qoa;

extern "C" {

int qoa_batch_create(const qoa_job_desc* jobs, uint32_t n_jobs, const qoa_gate* gates,
                     uint64_t n_gates_total, const double* mats, const qoa_fuse_opts* opts,
                     qoa_batch** out) {
  if (!out) return set_error(QOA_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!jobs || n_jobs == 0) return set_error(QOA_ERR_ARG, "no jobs");
  const uint32_t tile = opts && opts->tile_qubits ? opts->tile_qubits : 12;
  if (tile < 8 || tile > 12) return set_error(QOA_ERR_ARG, "batch tile_qubits must be 8..12");
  const uint32_t prec = jobs[0].precision;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const qoa_job_desc& d = jobs[j];
    if (d.n_qubits < kRegBits || d.n_qubits > 30)
      return set_error(QOA_ERR_ARG, "job %u: n_qubits %u outside 4..30", j, d.n_qubits);
    if (d.gate_offset + d.n_gates > n_gates_total)
      return set_error(QOA_ERR_ARG, "job %u: gates beyond the gate array", j);
    if (d.trajectories < 1) return set_error(QOA_ERR_ARG, "job %u: trajectories < 1", j);
    if (d.precision != QOA_C128 && d.precision != QOA_C64)
      return set_error(QOA_ERR_ARG, "job %u: unknown precision %u", j, d.precision);
    if (d.precision != prec)
      return set_error(QOA_ERR_UNSUPPORTED, "job %u: a batch's jobs share one precision", j);
    if (d.p1 < 0.0 || d.p1 > 1.0 || d.p2 < 0.0 || d.p2 > 1.0)
      return set_error(QOA_ERR_ARG, "job %u: noise probabilities outside [0,1]", j);
    if (prec == QOA_C64 && (d.p1 > 0.0 || d.p2 > 0.0))
      return set_error(QOA_ERR_UNSUPPORTED, "job %u: QOA_C64 batches are noise-free", j);
    if (prec == QOA_C64 && d.n_qubits < (uint32_t)kBlockQ)
      return set_error(QOA_ERR_ARG, "job %u: QOA_C64 jobs need at least 5 qubits", j);
    if (d.p1 > 0.0 || d.p2 > 0.0) {  // trajectories need well-formed gates
      for (uint64_t i = d.gate_offset; i < d.gate_offset + d.n_gates; ++i) {
        const qoa_gate& g = gates[i];
        const bool two = g.op == QOA_OP_CNOT || g.op == QOA_OP_CZ || g.op == QOA_OP_CPHASE ||
                         g.op == QOA_OP_SWAP || g.op == QOA_OP_U4 || g.op == QOA_OP_CY;
        const bool three = g.op == QOA_OP_CSWAP || g.op == QOA_OP_TOFFOLI;
        bool ok = g.q0 < d.n_qubits;
        if (two || three) ok = ok && g.q1 < d.n_qubits && g.q1 != g.q0;
        if (three) ok = ok && g.q2 < d.n_qubits && g.q2 != g.q0 && g.q2 != g.q1;
        if (!ok && g.op != QOA_OP_RESET_ALL)
          return set_error(QOA_ERR_ARG, "job %u gate %llu: noisy trajectories need distinct "
                           "in-range operands", j, (unsigned long long)(i - d.gate_offset));
      }
    }
  }
  qoa_batch* b = new qoa_batch();
  b->magic = kBatchMagic;
  b->stream = nullptr;
  b->states = nullptr;
  b->d_ops = nullptr;
  b->d_pool = nullptr;
  b->frames = nullptr;
  b->codes = nullptr;
  b->noisy = false;
  b->merged = false;
  b->ev0 = b->ev1 = nullptr;
  b->last_ms = 0.f;
  b->precision = prec;
  b->states64 = nullptr;
  b->d_w = nullptr;
  b->d_starts = nullptr;
  b->graph = nullptr;
  b->exec = nullptr;
  b->jobs.assign(jobs, jobs + n_jobs);
  if (prec == QOA_C64) {
    b->first_inst.resize(n_jobs);
    b->max_n = 0;
    for (uint32_t j = 0; j < n_jobs; ++j) b->max_n = std::max(b->max_n, jobs[j].n_qubits);
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    int rc = e == hipSuccess ? create_c64(b, gates, mats, opts ? opts->flags : 0u)
                             : set_error(QOA_ERR_HIP, "batch stream/event creation failed: %s",
                                         hipGetErrorString(e));
    if (rc) {
      batch_free(b);
      return rc;
    }
    *out = b;
    return QOA_OK;
  }
  b->plans.resize(n_jobs);
  {  // tiles: the requested one, or per qubit count as timed
    int rc = QOA_OK;
    if (opts && (opts->flags & QOA_BATCH_AUTOTUNE))
      rc = autotune_tiles(jobs, n_jobs, gates, n_gates_total, mats, opts, tile, &b->job_tile);
    else
      b->job_tile.assign(n_jobs, tile);
    if (rc) {
      batch_free(b);
      return rc;
    }
    for (uint32_t j = 0; j < n_jobs; ++j) b->job_tile[j] = std::min(b->job_tile[j], jobs[j].n_qubits);
  }

  // plan every job (host threads); merged diagonal runs only when no job is
  // noisy (their FMA kernels have no noise code)
  bool any_noisy = false;
  for (uint32_t j = 0; j < n_jobs; ++j) any_noisy = any_noisy || jobs[j].p1 > 0.0 || jobs[j].p2 > 0.0;
  const bool merge = opts && (opts->flags & QOA_FUSE_MERGE_DIAG) && !any_noisy;
  b->merged = merge;
  std::vector<int> rcs(n_jobs, 0);
  {
    const unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    std::atomic<uint32_t> next{0};
    for (unsigned t = 0; t < nt; ++t) {
      pool.emplace_back([&]() {
        for (;;) {
          const uint32_t j = next.fetch_add(1);
          if (j >= n_jobs) break;
          const bool noisy = jobs[j].p1 > 0.0 || jobs[j].p2 > 0.0;
          PlanOptions po{b->job_tile[j], opts ? opts->low_qubits : 0u,  // noisy jobs too
                         (opts ? (opts->flags & (QOA_CZ_REF_MASKS | QOA_FUSE_FACTOR)) : 0u) |
                             (merge ? QOA_FUSE_MERGE_DIAG : 0u) | kPlanForceReg | kPlanSwaps |
                             (noisy ? kPlanNoise : 0u)};
          rcs[j] = build_plan(jobs[j].n_qubits, gates + jobs[j].gate_offset, jobs[j].n_gates,
                              mats, po, &b->plans[j]);
        }
      });
    }
    for (auto& th : pool) th.join();
  }
  for (uint32_t j = 0; j < n_jobs; ++j) {
    if (rcs[j]) {
      batch_free(b);
      return set_error(rcs[j], "job %u: planning failed: %s", j, qoa_last_error());
    }
  }
  // flatten op programs
  std::vector<TileOp> all_ops;
  std::vector<double> all_pool;
  std::vector<uint32_t> ops_base(n_jobs), pool_base(n_jobs);
  for (uint32_t j = 0; j < n_jobs; ++j) {
    ops_base[j] = (uint32_t)all_ops.size();
    pool_base[j] = (uint32_t)all_pool.size();
    for (TileOp op : b->plans[j].ops) {
      if (op.kind == TK_MAT4 || op.kind == TK_DIAGN) op.aux += pool_base[j];
      all_ops.push_back(op);
    }
    all_pool.insert(all_pool.end(), b->plans[j].pool.begin(), b->plans[j].pool.end());
  }
  // instances and their offsets (each state starts on a 256-byte boundary)
  uint64_t off = 0;
  b->max_n = 0;
  b->first_inst.resize(n_jobs);
  for (uint32_t j = 0; j < n_jobs; ++j) {
    b->first_inst[j] = (uint32_t)b->inst.size();
    for (uint32_t t = 0; t < jobs[j].trajectories; ++t) {
      b->inst.push_back(Instance{j, t, jobs[j].n_qubits, off});
      off += std::max<uint64_t>(1ULL << jobs[j].n_qubits, 16);
    }
    b->max_n = std::max(b->max_n, jobs[j].n_qubits);
  }
  b->total_amps = off;
  int rc = QOA_OK;
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e != hipSuccess) {
    batch_free(b);
    return set_error(QOA_ERR_HIP, "batch stream/event creation failed: %s", hipGetErrorString(e));
  }
  if (hipMalloc(&b->states, sizeof(double2) * off) != hipSuccess) {
    (void)hipGetLastError();
    batch_free(b);
    return set_error(QOA_ERR_OOM, "batch states need %.2f GiB", 16.0 * off / 1073741824.0);
  }
  for (uint32_t j = 0; j < n_jobs; ++j) b->noisy |= jobs[j].p1 > 0.0 || jobs[j].p2 > 0.0;
  if (hipMalloc(&b->frames, sizeof(uint64_t) * 2 * b->inst.size()) != hipSuccess ||
      hipMalloc(&b->d_ops, sizeof(TileOp) * (all_ops.size() + 1)) != hipSuccess ||
      hipMalloc(&b->d_pool, sizeof(double) * (all_pool.size() + 1)) != hipSuccess) {
    (void)hipGetLastError();
    batch_free(b);
    return set_error(QOA_ERR_OOM, "batch op programs");
  }
  if (!all_ops.empty() &&
      hipMemcpy(b->d_ops, all_ops.data(), sizeof(TileOp) * all_ops.size(), hipMemcpyHostToDevice) !=
          hipSuccess)
    rc = set_error(QOA_ERR_HIP, "op upload failed");
  if (rc == QOA_OK && !all_pool.empty() &&
      hipMemcpy(b->d_pool, all_pool.data(), sizeof(double) * all_pool.size(),
                hipMemcpyHostToDevice) != hipSuccess)
    rc = set_error(QOA_ERR_HIP, "pool upload failed");
  // noise draws: instance ii's words start at codes_base[ii] (a job's
  // trajectories are consecutive instances, so its table is t * nops + o);
  // only noisy jobs have words (the kernel reads none for the others)
  std::vector<uint64_t> codes_base(b->inst.size(), 0);
  if (rc == QOA_OK && b->noisy) {
    uint64_t words = 0;
    for (size_t ii = 0; ii < b->inst.size(); ++ii) {
      const qoa_job_desc& jd = b->jobs[b->inst[ii].job];
      codes_base[ii] = words;
      if (jd.p1 > 0.0 || jd.p2 > 0.0) words += b->plans[b->inst[ii].job].ops.size();
    }
    if (hipMalloc(&b->codes, sizeof(uint32_t) * (words + 1)) != hipSuccess) {
      (void)hipGetLastError();
      rc = set_error(QOA_ERR_OOM, "noise draws");
    }
    for (uint32_t j = 0; rc == QOA_OK && j < n_jobs; ++j) {
      const qoa_job_desc& jd = b->jobs[j];
      if (!(jd.p1 > 0.0 || jd.p2 > 0.0)) continue;
      draw_codes(b->stream, b->d_ops + ops_base[j], (uint32_t)b->plans[j].ops.size(),
                 jd.noise_seed, jd.p1, jd.p2, jd.trajectories, b->codes + codes_base[b->first_inst[j]]);
      if (hipGetLastError() != hipSuccess) rc = set_error(QOA_ERR_HIP, "noise draw launch failed");
    }
  }
  // rounds
  uint32_t max_passes = 0;
  for (auto& p : b->plans) max_passes = std::max<uint32_t>(max_passes, (uint32_t)p.passes.size());
  b->alg_bytes = 0.0;
  for (uint32_t r = 0; rc == QOA_OK && r < max_passes; ++r) {
    std::vector<WorkItem> items;
    std::vector<uint64_t> nblk;
    Round rd;
    for (uint32_t ii = 0; ii < b->inst.size(); ++ii) {
      const Instance& in = b->inst[ii];
      const FusedPlan& P = b->plans[in.job];
      if (r >= P.passes.size()) continue;
      const PassPlan& pp = P.passes[r];
      const qoa_job_desc& jd = b->jobs[in.job];
      WorkItem w{};
      w.state_off = in.off;
      w.noise = (jd.p1 > 0.0 || jd.p2 > 0.0) ? 1 : 0;
      w.fma = b->merged ? 1 : 0;
      w.codes_off = codes_base[ii] - (uint64_t)ops_base[in.job];
      w.traj = in.traj;
      w.frame_in = 2 * ii + (r & 1);
      w.frame_out = 2 * ii + ((r + 1) & 1);
      w.ops_begin = ops_base[in.job] + pp.op_begin;
      w.nops = pp.op_end - pp.op_begin;
      w.k = (uint8_t)P.k;
      w.init = r == 0 ? 1 : 0;
      std::memcpy(w.spw, pp.spos, 16);
      std::memcpy(w.stw, pp.spos_st, 16);
      {
        uint32_t spw[4];
        pass_geo_words(pp, spw, w.stw, w.srt);
      }
      items.push_back(w);
      nblk.push_back(1ULL << (in.n - P.k));
      b->alg_bytes += (r == 0 ? 16.0 : 32.0) * (double)(1ULL << in.n);
    }
    std::vector<WorkItem> grouped;
    group_items(items, nblk, &grouped, &rd.groups);
    items.swap(grouped);
    bool too_many = false;
    for (const ItemGroup& g : rd.groups) too_many = too_many || g.blocks >= (1ULL << 31);
    if (too_many) {
      rc = set_error(QOA_ERR_ARG, "round %u has too many tiles", r);
      break;
    }
    if (hipMalloc(&rd.d_items, sizeof(WorkItem) * items.size()) != hipSuccess) {
      (void)hipGetLastError();
      rc = set_error(QOA_ERR_OOM, "work items");
      break;
    }
    b->rounds.push_back(rd);
    if (hipMemcpy(rd.d_items, items.data(), sizeof(WorkItem) * items.size(),
                  hipMemcpyHostToDevice) != hipSuccess)
      rc = set_error(QOA_ERR_HIP, "work item upload failed");
  }
  if (rc != QOA_OK) {
    batch_free(b);
    return rc;
  }
  static std::once_flag once;
  std::call_once(once, batch_kernel_attrs);
  *out = b;
  return QOA_OK;
}

int qoa_batch_run(qoa_batch* b) {
  if (!batch_valid(b)) return set_error(QOA_ERR_HANDLE, "invalid batch");
  QOA_HIP_CHECK(hipEventRecord(b->ev0, b->stream));
  if (g_batch_graphs) {
    if (!b->exec) {  // capture the launch sequence once (kernels only: no memset nodes)
      QOA_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_run(b);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return rc ? rc : set_error(QOA_ERR_HIP, "batch capture failed: %s", hipGetErrorString(e));
      }
      const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
      b->graph = g;
      if (ei != hipSuccess) {
        b->exec = nullptr;
        return set_error(QOA_ERR_HIP, "batch graph instantiation failed: %s", hipGetErrorString(ei));
      }
    }
    QOA_HIP_CHECK(hipGraphLaunch(b->exec, b->stream));
  } else {
    const int rc = enqueue_run(b);
    if (rc) return rc;
  }
  QOA_HIP_CHECK(hipEventRecord(b->ev1, b->stream));
  return QOA_OK;
}

int qoa_batch_synchronize(qoa_batch* b, double* run_ms) {
  if (!batch_valid(b)) return set_error(QOA_ERR_HANDLE, "invalid batch");
  QOA_HIP_CHECK(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  QOA_HIP_CHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  if (run_ms) *run_ms = ms;
  return QOA_OK;
}

int qoa_batch_info(qoa_batch* b, uint32_t* n_rounds, uint64_t* total_amps, double* alg_bytes) {
  if (!batch_valid(b)) return set_error(QOA_ERR_HANDLE, "invalid batch");
  if (n_rounds)
    *n_rounds = (uint32_t)(b->precision == QOA_C64 ? b->brounds.size() : b->rounds.size());
  if (total_amps) *total_amps = b->total_amps;
  if (alg_bytes) *alg_bytes = b->alg_bytes;
  return QOA_OK;
}

int qoa_batch_download(qoa_batch* b, uint32_t job, uint32_t traj, double* amps, uint64_t n_amps) {
  if (!batch_valid(b)) return set_error(QOA_ERR_HANDLE, "invalid batch");
  if (job >= b->jobs.size() || traj >= b->jobs[job].trajectories)
    return set_error(QOA_ERR_ARG, "no such job/trajectory");
  const Instance& in = b->inst[b->first_inst[job] + traj];
  if (!amps || n_amps != (1ULL << in.n)) return set_error(QOA_ERR_ARG, "output must hold 2^n amplitudes");
  if (b->precision == QOA_C64) {  // widen the FP32 amplitudes on the host
    std::vector<float> tmp(2 * n_amps);
    QOA_HIP_CHECK(hipMemcpyAsync(tmp.data(), b->states64 + in.off, sizeof(float2) * n_amps,
                                 hipMemcpyDeviceToHost, b->stream));
    QOA_HIP_CHECK(hipStreamSynchronize(b->stream));
    for (uint64_t i = 0; i < 2 * n_amps; ++i) amps[i] = (double)tmp[i];
    return QOA_OK;
  }
  QOA_HIP_CHECK(hipMemcpyAsync(amps, b->states + in.off, sizeof(double2) * n_amps,
                               hipMemcpyDeviceToHost, b->stream));
  uint64_t frame = 0;
  const uint32_t np = (uint32_t)b->plans[job].passes.size();
  const qoa_job_desc& jd = b->jobs[job];
  if ((jd.p1 > 0.0 || jd.p2 > 0.0) && np > 0)
    QOA_HIP_CHECK(hipMemcpyAsync(&frame, b->frames + 2 * (b->first_inst[job] + traj) + (np & 1),
                                 sizeof(uint64_t), hipMemcpyDeviceToHost, b->stream));
  QOA_HIP_CHECK(hipStreamSynchronize(b->stream));
  if (frame) {  // logical amplitude i is stored at physical index i ^ frame
    std::vector<double> tmp(amps, amps + 2 * n_amps);
    for (uint64_t i = 0; i < n_amps; ++i) {
      amps[2 * i] = tmp[2 * (i ^ frame)];
      amps[2 * i + 1] = tmp[2 * (i ^ frame) + 1];
    }
  }
  return QOA_OK;
}

int qoa_batch_order(qoa_batch* b, uint32_t job, uint32_t* order, uint32_t n_order) {
  if (!batch_valid(b)) return set_error(QOA_ERR_HANDLE, "invalid batch");
  if (job >= b->jobs.size()) return set_error(QOA_ERR_ARG, "no such job");
  if (b->precision == QOA_C64)
    return set_error(QOA_ERR_UNSUPPORTED, "block batches do not record a gate order");
  const auto& o = b->plans[job].order;
  if (!order || n_order != o.size()) return set_error(QOA_ERR_ARG, "order must hold n_gates entries");
  std::memcpy(order, o.data(), sizeof(uint32_t) * o.size());
  return QOA_OK;
}

int qoa_batch_job_plan(qoa_batch* b, uint32_t job, uint32_t* tile_qubits, uint32_t* n_steps) {
  if (!batch_valid(b)) return set_error(QOA_ERR_HANDLE, "invalid batch");
  if (job >= b->jobs.size()) return set_error(QOA_ERR_ARG, "no such job");
  if (b->precision == QOA_C64) {
    if (tile_qubits) *tile_qubits = kBlockQ;
    if (n_steps) *n_steps = b->job_tile[job];
  } else {
    if (tile_qubits) *tile_qubits = b->job_tile[job];
    if (n_steps) *n_steps = (uint32_t)b->plans[job].passes.size();
  }
  return QOA_OK;
}

int qoa_batch_destroy(qoa_batch* b) {
  if (!batch_valid(b)) return set_error(QOA_ERR_HANDLE, "invalid batch");
  batch_free(b);
  return QOA_OK;
}

}  // extern "C"
