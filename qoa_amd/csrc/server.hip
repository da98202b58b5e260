// server.hip — continuous-batching serving engine for circuit jobs.
//
// Clients submit jobs (a circuit, a qubit count, a number of noisy
// trajectories) with an arrival time on the server clock. An engine thread
//   1. hands arrived jobs to planner threads (fused register passes),
//   2. admits planned jobs in arrival order while their footprint (states,
//      op program, sampling scratch) fits the HBM arena,
//   3. launches one step at a time: a single batched kernel that runs the
//      next pass of every admitted job, each job at its own pass index, so
//      jobs that arrive while others run join the next step,
//   4. after a job's last pass, samples one outcome per trajectory on the
//      device, copies the samples back, and frees the job's memory once the
//      copy has landed.
// Every job's arithmetic depends only on its own inputs (plan, Philox keys,
// fixed summation orders), so its result is the same alone or in any mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "batch_kernel.hpp"
#include "planner.hpp"
#include "qoa_internal.hpp"

namespace qoa {
namespace {

using namespace tile;

constexpr uint32_t kServerMagic = 0x51535256u;  // "QSRV"
constexpr uint32_t kFrameBlock = 2048;           // frame slots per job (2 per trajectory)
constexpr uint32_t kMaxTraj = kFrameBlock / 2;
constexpr uint32_t kRing = 4;                    // step buffers in flight

// First-fit allocator over [0, capacity) in units of `align` bytes.
class Arena {
 public:
  void init(uint64_t cap, uint64_t align) {
    cap_ = cap / align * align;
    align_ = align;
    free_.clear();
    if (cap_) free_[0] = cap_;
    used_ = peak_ = 0;
  }
  bool alloc(uint64_t bytes, uint64_t* off) {
    bytes = (bytes + align_ - 1) / align_ * align_;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const uint64_t rest = it->second - bytes, start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = (bytes + align_ - 1) / align_ * align_;
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t capacity() const { return cap_; }
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t cap_ = 0, align_ = 256, used_ = 0, peak_ = 0;
};

struct SrvJob {
  uint64_t id = 0;
  qoa_job_desc desc{};
  std::vector<qoa_gate> gates;
  std::vector<double> mats;
  double t_arrival = 0, t_admit = -1, t_done = -1;
  FusedPlan plan;
  int status = QOA_OK;
  uint32_t samples_per_traj = 1;
  // device placement
  uint64_t state_off = 0, state_bytes = 0;  // arena bytes
  uint64_t aux_off = 0, aux_bytes = 0;      // arena bytes: cs + samples + op pool
  uint64_t ops_off = 0, nops = 0;           // op slots in the op pool
  uint64_t codes_word = 0;                  // noise draws (arena as 32-bit words)
  uint32_t frame_block = 0;
  uint32_t next_pass = 0;
  hipEvent_t done_ev = nullptr;
  std::vector<uint64_t> samples;  // host copy (pinned staging below)
  uint64_t* pinned = nullptr;
  uint32_t pinned_slot = 0;
  bool finishing = false;
};

}  // namespace
}  // namespace qoa

struct qoa_server {
  uint32_t magic;
  qoa_server_opts opts;
  hipStream_t stream;         // steps, sampling, sample copies
  hipStream_t upload_stream;  // op programs of admitted jobs
  uint64_t* pinned_pool;      // kMaxTraj sample slots per job block
  std::vector<uint32_t> free_pinned;
  uint8_t* arena_base;
  qoa::Arena arena;      // job states and scratch (bytes)
  qoa::TileOp* d_ops;
  qoa::Arena op_pool;    // TileOp slots
  uint64_t* d_frames;
  std::vector<uint32_t> free_frame_blocks;
  // step ring
  qoa::tile::WorkItem* h_items[qoa::kRing];
  qoa::tile::WorkItem* d_items[qoa::kRing];
  hipEvent_t step_ev[qoa::kRing];
  bool step_used[qoa::kRing];
  uint32_t ring_pos;
  uint32_t max_items;
  // jobs
  std::mutex mu;
  std::condition_variable cv;
  std::multimap<double, std::shared_ptr<qoa::SrvJob>> pending;  // by arrival time
  std::deque<std::shared_ptr<qoa::SrvJob>> ready;                // planned, waiting for memory
  std::vector<std::shared_ptr<qoa::SrvJob>> active;
  std::unordered_map<uint64_t, std::shared_ptr<qoa::SrvJob>> done;
  std::atomic<uint64_t> next_id;
  uint64_t submitted, completed, rejected, steps;
  double busy_ms;
  // planner threads
  std::deque<std::shared_ptr<qoa::SrvJob>> to_plan;
  std::vector<std::shared_ptr<qoa::SrvJob>> planned;
  std::vector<std::thread> planners;
  std::thread engine;
  bool stop;
  int fatal;
  std::chrono::steady_clock::time_point t0;
};

namespace qoa {
namespace {

double now_s(const qoa_server* s) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

uint64_t job_state_bytes(const qoa_job_desc& d) {
  return (uint64_t)d.trajectories * (1ULL << d.n_qubits) * 16ULL;
}

uint64_t job_aux_bytes(const SrvJob& j) {
  uint64_t b = sample_scratch_doubles(j.desc.n_qubits, j.desc.trajectories) * 8;  // chunk sums
  b = (b + 255) / 256 * 256;
  b += (uint64_t)j.desc.trajectories * j.samples_per_traj * 8;                // samples
  b = (b + 255) / 256 * 256;
  b += j.plan.pool.size() * 8 + 8;                                            // MAT4 pool
  if (j.desc.p1 > 0.0 || j.desc.p2 > 0.0) {                                   // noise draws
    b = (b + 255) / 256 * 256;
    b += (uint64_t)j.desc.trajectories * j.plan.ops.size() * 4;
  }
  return b;
}

void planner_main(qoa_server* s) {
  for (;;) {
    std::shared_ptr<SrvJob> j;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->to_plan.empty(); });
      if (s->stop) return;
      j = s->to_plan.front();
      s->to_plan.pop_front();
    }
    const bool noisy = j->desc.p1 > 0.0 || j->desc.p2 > 0.0;
    PlanOptions po{std::min(s->opts.tile_qubits, j->desc.n_qubits), 0u,
                   (s->opts.flags & (QOA_CZ_REF_MASKS | QOA_FUSE_FACTOR)) | kPlanForceReg | kPlanSwaps |
                       (noisy ? kPlanNoise : 0u)};
    const int rc = build_plan(j->desc.n_qubits, j->gates.data(), (uint32_t)j->gates.size(),
                              j->mats.empty() ? nullptr : j->mats.data(), po, &j->plan);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (rc) j->status = rc;
      s->planned.push_back(j);
    }
    s->cv.notify_all();
  }
}

void finish_job(qoa_server* s, const std::shared_ptr<SrvJob>& j, int status) {
  j->status = status;
  j->t_done = now_s(s);
  if (j->state_bytes) s->arena.release(j->state_off, j->state_bytes);
  if (j->aux_bytes) s->arena.release(j->aux_off, j->aux_bytes);
  if (j->nops) s->op_pool.release(j->ops_off, j->nops);
  if (j->state_bytes) s->free_frame_blocks.push_back(j->frame_block);
  j->state_bytes = j->aux_bytes = j->nops = 0;
  if (j->done_ev) {
    hipEventDestroy(j->done_ev);
    j->done_ev = nullptr;
  }
  if (j->pinned) {
    s->free_pinned.push_back(j->pinned_slot);
    j->pinned = nullptr;
  }
  j->gates.clear();
  j->gates.shrink_to_fit();
  j->plan = FusedPlan();
  std::lock_guard<std::mutex> lk(s->mu);
  s->done[j->id] = j;
  s->completed++;
  s->cv.notify_all();
}

// Tries to place a planned job; returns false when memory does not fit now.
bool admit(qoa_server* s, const std::shared_ptr<SrvJob>& j) {
  if (s->free_frame_blocks.empty() || s->free_pinned.empty()) return false;
  const uint64_t sb = job_state_bytes(j->desc), ab = job_aux_bytes(*j);
  const uint64_t nops = j->plan.ops.size() + 1;
  uint64_t so, ao, oo;
  if (!s->arena.alloc(sb, &so)) return false;
  if (!s->arena.alloc(ab, &ao)) {
    s->arena.release(so, sb);
    return false;
  }
  if (!s->op_pool.alloc(nops, &oo)) {
    s->arena.release(so, sb);
    s->arena.release(ao, ab);
    return false;
  }
  j->state_off = so;
  j->state_bytes = sb;
  j->aux_off = ao;
  j->aux_bytes = ab;
  j->ops_off = oo;
  j->nops = nops;
  j->frame_block = s->free_frame_blocks.back();
  s->free_frame_blocks.pop_back();
  j->pinned_slot = s->free_pinned.back();
  s->free_pinned.pop_back();
  j->pinned = s->pinned_pool + (uint64_t)j->pinned_slot * kMaxTraj;
  // op program (with MAT4 offsets into the job's own pool region)
  uint64_t pool_off = ao + sample_scratch_doubles(j->desc.n_qubits, j->desc.trajectories) * 8;
  pool_off = (pool_off + 255) / 256 * 256;
  pool_off += (uint64_t)j->desc.trajectories * j->samples_per_traj * 8;
  pool_off = (pool_off + 255) / 256 * 256;
  // MAT4 coefficients live in the arena; ops address them relative to the
  // arena base viewed as doubles (the kernel gets arena_base as the pool)
  if (!j->plan.pool.empty() && pool_off / 8 + j->plan.pool.size() >= (1ULL << 32)) {
    s->arena.release(so, sb);
    s->arena.release(ao, ab);
    s->op_pool.release(oo, nops);
    j->state_bytes = j->aux_bytes = j->nops = 0;
    s->free_frame_blocks.push_back(j->frame_block);
    s->free_pinned.push_back(j->pinned_slot);
    j->pinned = nullptr;
    j->status = QOA_ERR_UNSUPPORTED;  // 4x4 coefficients beyond 32-bit pool offsets
    return false;
  }
  std::vector<TileOp> ops = j->plan.ops;
  for (TileOp& op : ops)
    if (op.kind == TK_MAT4 || op.kind == TK_DIAGN) op.aux += (uint32_t)(pool_off / 8);
  if (!ops.empty())
    (void)hipMemcpyAsync(s->d_ops + oo, ops.data(), sizeof(TileOp) * ops.size(),
                         hipMemcpyHostToDevice, s->upload_stream);
  if (!j->plan.pool.empty())
    (void)hipMemcpyAsync(s->arena_base + pool_off, j->plan.pool.data(),
                         sizeof(double) * j->plan.pool.size(), hipMemcpyHostToDevice,
                         s->upload_stream);
  // host vectors are pageable: wait for the upload (not for running steps)
  (void)hipStreamSynchronize(s->upload_stream);
  if (j->desc.p1 > 0.0 || j->desc.p2 > 0.0) {  // trajectory draws, before the job's first step
    uint64_t codes_off = pool_off + j->plan.pool.size() * 8 + 8;
    codes_off = (codes_off + 255) / 256 * 256;
    j->codes_word = codes_off / 4;
    tile::draw_codes(s->stream, s->d_ops + oo, (uint32_t)j->plan.ops.size(), j->desc.noise_seed,
                     j->desc.p1, j->desc.p2, j->desc.trajectories,
                     reinterpret_cast<uint32_t*>(s->arena_base + codes_off));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {  // the engine stops; waiters see the fatal flag
      s->fatal = set_error(QOA_ERR_HIP, "noise draw launch failed: %s", hipGetErrorString(e));
    }
  }
  j->t_admit = now_s(s);
  return true;
}

int launch_step(qoa_server* s) {
  // one pass of every active job that has passes left
  const uint32_t slot = s->ring_pos;
  if (s->step_used[slot]) QOA_HIP_CHECK(hipEventSynchronize(s->step_ev[slot]));
  std::vector<tile::WorkItem> built;
  std::vector<uint64_t> nblk;
  uint32_t nitems = 0;
  std::vector<std::shared_ptr<SrvJob>> last;
  for (auto& j : s->active) {
    if (j->finishing || j->next_pass >= j->plan.passes.size()) continue;
    const uint32_t r = j->next_pass;
    const PassPlan& pp = j->plan.passes[r];
    const uint32_t T = j->desc.trajectories;
    if (nitems + T > s->max_items) break;  // the rest joins the next step
    const uint64_t len = 1ULL << j->desc.n_qubits;
    for (uint32_t t = 0; t < T; ++t) {
      tile::WorkItem w;
      std::memset(&w, 0, sizeof(w));
      w.state_off = j->state_off / 16 + (uint64_t)t * len;
      w.noise = (j->desc.p1 > 0.0 || j->desc.p2 > 0.0) ? 1 : 0;
      if (w.noise)  // the job's draw table in the arena (noise-free jobs have none)
        w.codes_off = j->codes_word + (uint64_t)t * j->plan.ops.size() - j->ops_off;
      w.traj = t;
      w.frame_in = j->frame_block * kFrameBlock + 2 * t + (r & 1);
      w.frame_out = j->frame_block * kFrameBlock + 2 * t + ((r + 1) & 1);
      w.ops_begin = (uint32_t)j->ops_off + pp.op_begin;
      w.nops = pp.op_end - pp.op_begin;
      w.k = (uint8_t)j->plan.k;
      w.init = r == 0 ? 1 : 0;
      std::memcpy(w.spw, pp.spos, 16);
      std::memcpy(w.stw, pp.spos_st, 16);
      {
        uint32_t spw[4];
        pass_geo_words(pp, spw, w.stw, w.srt);
      }
      built.push_back(w);
      nblk.push_back(len >> j->plan.k);
      ++nitems;
    }
    j->next_pass++;
    if (j->next_pass == j->plan.passes.size()) last.push_back(j);
  }
  if (nitems == 0) return QOA_OK;
  std::vector<tile::WorkItem> grouped;
  std::vector<tile::ItemGroup> groups;
  tile::group_items(built, nblk, &grouped, &groups);
  tile::WorkItem* items = s->h_items[slot];
  std::memcpy(items, grouped.data(), sizeof(tile::WorkItem) * nitems);
  QOA_HIP_CHECK(hipMemcpyAsync(s->d_items[slot], items, sizeof(tile::WorkItem) * nitems,
                               hipMemcpyHostToDevice, s->stream));
  for (const tile::ItemGroup& g : groups) {
    tile::launch_group(s->stream, g, (double2*)s->arena_base, s->d_items[slot], s->d_ops,
                       (const double*)s->arena_base, s->d_frames,
                       reinterpret_cast<const uint32_t*>(s->arena_base));
    QOA_HIP_CHECK(hipGetLastError());
  }
  for (auto& j : last) {  // sample each trajectory and copy the samples back
    const uint32_t T = j->desc.trajectories, P = (uint32_t)j->plan.passes.size();
    const uint64_t len = 1ULL << j->desc.n_qubits;
    double* cs = (double*)(s->arena_base + j->aux_off);
    uint64_t so = j->aux_off + sample_scratch_doubles(j->desc.n_qubits, T) * 8;
    so = (so + 255) / 256 * 256;
    uint64_t* dsamp = (uint64_t*)(s->arena_base + so);
    const bool noisy = j->desc.p1 > 0.0 || j->desc.p2 > 0.0;
    const uint64_t* fr = noisy ? s->d_frames + j->frame_block * kFrameBlock + (P & 1) : nullptr;
    int rc = launch_sample(s->stream, (const double2*)(s->arena_base + j->state_off), len,
                           j->desc.n_qubits, T, j->desc.noise_seed ^ 0x53414D504C45ULL,
                           j->samples_per_traj, fr, 2, cs, dsamp);
    if (rc) return rc;
    const size_t nb = sizeof(uint64_t) * T * j->samples_per_traj;
    QOA_HIP_CHECK(hipMemcpyAsync(j->pinned, dsamp, nb, hipMemcpyDeviceToHost, s->stream));
    QOA_HIP_CHECK(hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming));
    QOA_HIP_CHECK(hipEventRecord(j->done_ev, s->stream));
    j->finishing = true;
  }
  QOA_HIP_CHECK(hipEventRecord(s->step_ev[slot], s->stream));
  s->step_used[slot] = true;
  s->ring_pos = (slot + 1) % kRing;
  s->steps++;
  return QOA_OK;
}

void engine_main(qoa_server* s) {
  (void)hipSetDevice(0);
  for (;;) {
    std::vector<std::shared_ptr<SrvJob>> newly;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->stop) break;
      const double t = now_s(s);
      while (!s->pending.empty() && s->pending.begin()->first <= t) {
        s->to_plan.push_back(s->pending.begin()->second);
        s->pending.erase(s->pending.begin());
      }
      if (!s->to_plan.empty()) s->cv.notify_all();
      newly.swap(s->planned);
    }
    // planned jobs join the admission queue in arrival order
    std::sort(newly.begin(), newly.end(),
              [](const std::shared_ptr<SrvJob>& a, const std::shared_ptr<SrvJob>& b) {
                return a->t_arrival < b->t_arrival || (a->t_arrival == b->t_arrival && a->id < b->id);
              });
    for (auto& j : newly) {
      if (j->status != QOA_OK) finish_job(s, j, j->status);
      else s->ready.push_back(j);
    }
    while (!s->ready.empty()) {
      std::shared_ptr<SrvJob> j = s->ready.front();
      if (admit(s, j)) {
        s->active.push_back(j);
        s->ready.pop_front();
      } else if (j->status != QOA_OK) {
        s->ready.pop_front();
        finish_job(s, j, j->status);
      } else {
        break;  // arrival order: later jobs wait behind the head
      }
    }
    // after a fatal error every admitted job finishes with that error
    const int rc = s->fatal ? s->fatal : launch_step(s);
    if (rc) {
      s->fatal = rc;
      for (auto& j : s->active) finish_job(s, j, rc);
      s->active.clear();
    }
    // retire jobs whose samples have landed
    for (size_t i = 0; i < s->active.size();) {
      auto& j = s->active[i];
      if (j->finishing && hipEventQuery(j->done_ev) == hipSuccess) {
        j->samples.assign(j->pinned, j->pinned + (size_t)j->desc.trajectories * j->samples_per_traj);
        finish_job(s, j, QOA_OK);
        s->active[i] = s->active.back();
        s->active.pop_back();
      } else {
        ++i;
      }
    }
    bool idle;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      bool runnable = false;
      for (auto& j : s->active)
        if (!j->finishing) runnable = true;
      idle = !runnable && s->to_plan.empty() && s->planned.empty() &&
             (s->ready.empty() || s->active.empty());
    }
    if (idle) {
      std::unique_lock<std::mutex> lk(s->mu);
      double wait = 0.0002;
      if (!s->pending.empty()) wait = std::min(wait, std::max(0.0, s->pending.begin()->first - now_s(s)));
      bool finishing = false;
      for (auto& j : s->active) finishing |= j->finishing;
      if (finishing) wait = std::min(wait, 0.00002);
      s->cv.wait_for(lk, std::chrono::duration<double>(wait));
    }
  }
}

bool server_valid(const qoa_server* s) { return s && s->magic == kServerMagic; }

}  // namespace
}  // namespace qoa

using namespace qoa;

extern "C" {

int qoa_server_create(const qoa_server_opts* opts, qoa_server** out) {
  if (!out) return set_error(QOA_ERR_ARG, "out is NULL");
  *out = nullptr;
  qoa_server_opts o{};
  if (opts) o = *opts;
  if (!o.tile_qubits) o.tile_qubits = 12;
  if (o.tile_qubits < 8 || o.tile_qubits > 12) return set_error(QOA_ERR_ARG, "tile_qubits 8..12");
  if (!o.planner_threads) o.planner_threads = 4;
  if (!o.max_items) o.max_items = 65536;
  if (!o.max_ops) o.max_ops = 1u << 21;
  if (!o.max_jobs) o.max_jobs = 4096;
  size_t free_b = 0, total_b = 0;
  QOA_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  if (!o.budget_bytes) o.budget_bytes = (uint64_t)(0.8 * (double)free_b);
  if (o.budget_bytes > free_b)
    return set_error(QOA_ERR_OOM, "budget %.1f GiB exceeds free device memory %.1f GiB",
                     o.budget_bytes / 1073741824.0, free_b / 1073741824.0);
  qoa_server* s = new qoa_server();
  s->magic = kServerMagic;
  s->opts = o;
  s->stream = nullptr;
  s->upload_stream = nullptr;
  s->pinned_pool = nullptr;
  s->arena_base = nullptr;
  s->d_ops = nullptr;
  s->d_frames = nullptr;
  for (uint32_t i = 0; i < kRing; ++i) {
    s->h_items[i] = nullptr;
    s->d_items[i] = nullptr;
    s->step_ev[i] = nullptr;
    s->step_used[i] = false;
  }
  s->ring_pos = 0;
  s->max_items = o.max_items;
  s->next_id = 1;
  s->submitted = s->completed = s->rejected = s->steps = 0;
  s->busy_ms = 0;
  s->stop = false;
  s->fatal = 0;
  s->t0 = std::chrono::steady_clock::now();
  bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&s->upload_stream, hipStreamNonBlocking) == hipSuccess &&
            hipHostMalloc((void**)&s->pinned_pool, sizeof(uint64_t) * kMaxTraj * o.max_jobs,
                          hipHostMallocDefault) == hipSuccess &&
            hipMalloc(&s->arena_base, o.budget_bytes) == hipSuccess &&
            hipMalloc(&s->d_ops, sizeof(TileOp) * ((uint64_t)o.max_ops + 1)) == hipSuccess &&
            hipMalloc(&s->d_frames, sizeof(uint64_t) * (uint64_t)kFrameBlock * o.max_jobs) ==
                hipSuccess;
  for (uint32_t i = 0; ok && i < kRing; ++i) {
    ok = hipHostMalloc((void**)&s->h_items[i], sizeof(tile::WorkItem) * o.max_items,
                       hipHostMallocDefault) == hipSuccess &&
         hipMalloc(&s->d_items[i], sizeof(tile::WorkItem) * o.max_items) == hipSuccess &&
         hipEventCreateWithFlags(&s->step_ev[i], hipEventDisableTiming) == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    qoa_server_destroy(s);
    return set_error(QOA_ERR_OOM, "server allocation failed");
  }
  s->arena.init(o.budget_bytes, 256);
  s->op_pool.init(o.max_ops, 1);
  for (uint32_t b = 0; b < o.max_jobs; ++b) {
    s->free_frame_blocks.push_back(o.max_jobs - 1 - b);
    s->free_pinned.push_back(o.max_jobs - 1 - b);
  }
  static std::once_flag once;
  std::call_once(once, tile::batch_kernel_attrs);
  for (uint32_t t = 0; t < o.planner_threads; ++t) s->planners.emplace_back(planner_main, s);
  s->engine = std::thread(engine_main, s);
  *out = s;
  return QOA_OK;
}

int qoa_server_submit(qoa_server* s, const qoa_job_desc* job, const qoa_gate* gates,
                      const double* matrices, uint32_t n_matrices, double arrival_s,
                      uint64_t* job_id) {
  if (!server_valid(s)) return set_error(QOA_ERR_HANDLE, "invalid server");
  if (!job || (!gates && job->n_gates)) return set_error(QOA_ERR_ARG, "job or gates is NULL");
  if (job->n_qubits < kRegBits || job->n_qubits > 30)
    return set_error(QOA_ERR_ARG, "n_qubits %u outside 4..30", job->n_qubits);
  if (job->trajectories < 1 || job->trajectories > kMaxTraj)
    return set_error(QOA_ERR_ARG, "trajectories must be 1..%u", kMaxTraj);
  if (job->precision != QOA_C128) return set_error(QOA_ERR_UNSUPPORTED, "only QOA_C128 jobs");
  if (job->p1 < 0 || job->p1 > 1 || job->p2 < 0 || job->p2 > 1)
    return set_error(QOA_ERR_ARG, "noise probabilities outside [0,1]");
  auto j = std::make_shared<SrvJob>();
  j->id = s->next_id.fetch_add(1);
  j->desc = *job;
  j->gates.assign(gates, gates + job->n_gates);
  if (matrices && n_matrices) j->mats.assign(matrices, matrices + 8 * (size_t)n_matrices);
  j->t_arrival = arrival_s;
  j->samples_per_traj = 1;
  if (job_state_bytes(*job) > s->arena.capacity()) {
    std::lock_guard<std::mutex> lk(s->mu);
    j->status = QOA_ERR_ADMISSION;
    j->t_done = arrival_s;
    s->done[j->id] = j;
    s->rejected++;
    s->submitted++;
    if (job_id) *job_id = j->id;
    return set_error(QOA_ERR_ADMISSION, "job needs %.2f GiB, budget is %.2f GiB",
                     job_state_bytes(*job) / 1073741824.0, s->arena.capacity() / 1073741824.0);
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->pending.emplace(arrival_s, j);
    s->submitted++;
  }
  s->cv.notify_all();
  if (job_id) *job_id = j->id;
  return QOA_OK;
}

double qoa_server_now(qoa_server* s) { return server_valid(s) ? now_s(s) : -1.0; }

int qoa_server_wait(qoa_server* s, uint64_t job_id, double timeout_s, qoa_job_result* res,
                    uint64_t* samples, uint32_t max_samples) {
  if (!server_valid(s)) return set_error(QOA_ERR_HANDLE, "invalid server");
  std::shared_ptr<SrvJob> j;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    const bool got = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
      return s->done.count(job_id) > 0;
    });
    if (!got) return set_error(QOA_ERR_ARG, "job %llu not finished within the timeout",
                               (unsigned long long)job_id);
    j = s->done[job_id];
    s->done.erase(job_id);
  }
  if (res) {
    res->job_id = j->id;
    res->status = j->status;
    res->n_qubits = j->desc.n_qubits;
    res->trajectories = j->desc.trajectories;
    res->n_passes = (uint32_t)0;
    res->t_arrival = j->t_arrival;
    res->t_admit = j->t_admit;
    res->t_done = j->t_done;
    res->n_samples = (uint32_t)j->samples.size();
  }
  if (samples && max_samples)
    std::memcpy(samples, j->samples.data(),
                sizeof(uint64_t) * std::min<size_t>(max_samples, j->samples.size()));
  return j->status;
}

int qoa_server_drain(qoa_server* s, double timeout_s) {
  if (!server_valid(s)) return set_error(QOA_ERR_HANDLE, "invalid server");
  std::unique_lock<std::mutex> lk(s->mu);
  const bool ok = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                 [&] { return s->completed + s->rejected >= s->submitted; });
  return ok ? QOA_OK : set_error(QOA_ERR_ARG, "drain timed out");
}

int qoa_server_get_stats(qoa_server* s, qoa_server_stats* st) {
  if (!server_valid(s) || !st) return set_error(QOA_ERR_HANDLE, "invalid server");
  std::lock_guard<std::mutex> lk(s->mu);
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->rejected = s->rejected;
  st->steps = s->steps;
  st->arena_bytes = s->arena.capacity();
  st->arena_peak = s->arena.peak();
  st->fatal = s->fatal;
  return QOA_OK;
}

int qoa_server_destroy(qoa_server* s) {
  if (!server_valid(s)) return set_error(QOA_ERR_HANDLE, "invalid server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->engine.joinable()) s->engine.join();
  for (auto& t : s->planners) t.join();
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& j : s->active)
    if (j->done_ev) hipEventDestroy(j->done_ev);
  if (s->pinned_pool) hipHostFree(s->pinned_pool);
  if (s->upload_stream) hipStreamDestroy(s->upload_stream);
  for (uint32_t i = 0; i < kRing; ++i) {
    if (s->h_items[i]) hipHostFree(s->h_items[i]);
    if (s->d_items[i]) hipFree(s->d_items[i]);
    if (s->step_ev[i]) hipEventDestroy(s->step_ev[i]);
  }
  if (s->d_frames) hipFree(s->d_frames);
  if (s->d_ops) hipFree(s->d_ops);
  if (s->arena_base) hipFree(s->arena_base);
  if (s->stream) hipStreamDestroy(s->stream);
  s->magic = 0;
  delete s;
  return QOA_OK;
}

}  // extern "C"
