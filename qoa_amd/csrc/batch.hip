// batch.hip — many independent circuits of uneven size in one launch per round.
//
// Every job instance (job x trajectory) owns a state in one device buffer and
// a fused plan (register-resident passes). Round r runs pass r of every
// instance that has one: the launch's blocks are the concatenated tiles of
// those passes, found through a per-round work-item table (prefix sums of
// tile counts). The number of launches is the largest pass count in the
// batch, not the number of jobs. Pass 0 synthesises |0...0> instead of
// reading the state.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

#include "planner.hpp"
#include "qoa_internal.hpp"
#include "batch_kernel.hpp"

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
  double alg_bytes;  // algorithmic HBM bytes of one run
  hipEvent_t ev0, ev1;
  float last_ms;
};

namespace qoa {

static bool batch_valid(const qoa_batch* b) { return b && b->magic == kBatchMagic; }

static void batch_free(qoa_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& r : b->rounds)
    if (r.d_items) hipFree(r.d_items);
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
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const qoa_job_desc& d = jobs[j];
    if (d.n_qubits < kRegBits || d.n_qubits > 30)
      return set_error(QOA_ERR_ARG, "job %u: n_qubits %u outside 4..30", j, d.n_qubits);
    if (d.gate_offset + d.n_gates > n_gates_total)
      return set_error(QOA_ERR_ARG, "job %u: gates beyond the gate array", j);
    if (d.trajectories < 1) return set_error(QOA_ERR_ARG, "job %u: trajectories < 1", j);
    if (d.precision != QOA_C128) return set_error(QOA_ERR_UNSUPPORTED, "job %u: only QOA_C128", j);
    if (d.p1 < 0.0 || d.p1 > 1.0 || d.p2 < 0.0 || d.p2 > 1.0)
      return set_error(QOA_ERR_ARG, "job %u: noise probabilities outside [0,1]", j);
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
  b->ev0 = b->ev1 = nullptr;
  b->last_ms = 0.f;
  b->jobs.assign(jobs, jobs + n_jobs);
  b->plans.resize(n_jobs);

  // plan every job (host threads)
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
          PlanOptions po{std::min(tile, jobs[j].n_qubits), opts ? opts->low_qubits : 0u,
                         (opts ? (opts->flags & QOA_CZ_REF_MASKS) : 0u) | kPlanForceReg | kPlanSwaps |
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
      w.codes_off = codes_base[ii] - (uint64_t)ops_base[in.job];
      w.traj = in.traj;
      w.frame_in = 2 * ii + (r & 1);
      w.frame_out = 2 * ii + ((r + 1) & 1);
      w.ops_begin = ops_base[in.job] + pp.op_begin;
      w.nops = pp.op_end - pp.op_begin;
      w.k = (uint8_t)P.k;
      w.init = r == 0 ? 1 : 0;
      std::memcpy(w.spos, pp.spos, 16);
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
  for (const Round& rd : b->rounds) {
    for (const ItemGroup& g : rd.groups) {
      launch_group(b->stream, g, b->states, rd.d_items, b->d_ops, b->d_pool, b->frames, b->codes);
      QOA_HIP_CHECK(hipGetLastError());
    }
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
  if (n_rounds) *n_rounds = (uint32_t)b->rounds.size();
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
  const auto& o = b->plans[job].order;
  if (!order || n_order != o.size()) return set_error(QOA_ERR_ARG, "order must hold n_gates entries");
  std::memcpy(order, o.data(), sizeof(uint32_t) * o.size());
  return QOA_OK;
}

int qoa_batch_destroy(qoa_batch* b) {
  if (!batch_valid(b)) return set_error(QOA_ERR_HANDLE, "invalid batch");
  batch_free(b);
  return QOA_OK;
}

}  // extern "C"
