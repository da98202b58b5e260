// api.cpp — the extern "C" boundary (include/qoa_hip.h) of the gfx950 backend.
//
// Layer 1 (qoa_apply_*) keeps the parameter meaning and out-of-range
// behaviour of the src/vectorization.rs free functions; layer 2 (qoa_qs_*)
// keeps the QuantumState wrapper semantics of src/runtime/quantum_state.rs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "qoa_internal.hpp"
#include "traj_noise.hpp"

namespace qoa {

static thread_local char g_err[512] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

bool valid(const qoa_state* s) { return s != nullptr && s->magic == kMagic; }

bool g_force_index64 = false;
bool g_batch_graphs = true;

const char* kind_name(int k) {
  static const char* names[KK_COUNT] = {"sweep_dense", "sweep_diag", "sweep_perm",
                                        "fused_pass",  "reduce",     "other",
                                        "fused_block"};
  return (k >= 0 && k < KK_COUNT) ? names[k] : "?";
}

// ---- Philox4x32-10 host stream (Salmon et al., SC'11) ----------------------
static void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
  }
}

void HostRng::seed(uint64_t s) {
  key[0] = (uint32_t)s;
  key[1] = (uint32_t)(s >> 32);
  ctr = 0;
  have = 0;
}

// Draw k = u64 half (k & 1) of block (k >> 1), block counter {lo, hi, 0, 0}.
uint64_t HostRng::next_u64() {
  if (have == 0) {
    const uint64_t blk = ctr >> 1;
    buf[0] = (uint32_t)blk;
    buf[1] = (uint32_t)(blk >> 32);
    buf[2] = 0;
    buf[3] = 0;
    philox10(buf, key[0], key[1]);
    have = 2;
  }
  const int half = (int)(ctr & 1);
  ++ctr;
  --have;
  return ((uint64_t)buf[2 * half + 1] << 32) | buf[2 * half];
}

double HostRng::next_f64() { return (double)(next_u64() >> 11) * 0x1.0p-53; }
uint32_t HostRng::below3() { return (uint32_t)(((next_u64() >> 32) * 3ULL) >> 32); }

// ---- timing -------------------------------------------------------------------
static hipEvent_t pool_get(qoa_state* s) {
  if (!s->event_pool.empty()) {
    hipEvent_t e = s->event_pool.back();
    s->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void timing_begin(qoa_state* s, int kind, double bytes) {
  if (!s->timing) return;
  TimedLaunch t;
  t.start = pool_get(s);
  t.stop = pool_get(s);
  t.kind = kind;
  t.bytes = bytes;
  if (!t.start || !t.stop) return;
  (void)hipEventRecord(t.start, s->stream);
  s->pending.push_back(t);
}

void timing_end(qoa_state* s) {
  if (!s->timing || s->pending.empty()) return;
  (void)hipEventRecord(s->pending.back().stop, s->stream);
}

int timing_collect(qoa_state* s) {
  if (s->pending.empty()) return QOA_OK;
  QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
  for (auto& t : s->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
      s->kind_ms[t.kind] += ms;
      s->kind_bytes[t.kind] += t.bytes;
      s->kind_launches[t.kind] += 1;
    }
    s->event_pool.push_back(t.start);
    s->event_pool.push_back(t.stop);
  }
  s->pending.clear();
  return QOA_OK;
}

// ---- gate coefficients (host libm, like the reference) -------------------------
// cos/sin pairs come from one sincos call, as rustc/LLVM lowers the
// reference's side-by-side .cos()/.sin() on linux-gnu (sincos and separate
// cos/sin can differ in the last bit).
static inline void cis(double x, double* c, double* s) { ::sincos(x, s, c); }
static Cx from_polar(double r, double th) {
  double c, s;
  cis(th, &c, &s);
  return Cx{r * c, r * s};
}

static inline uint64_t bit_or_zero(uint32_t q, uint32_t n) { return q < n ? (1ULL << q) : 0ULL; }

// Layer-1 single gate dispatch (shared by qoa_run_gates and the QuantumState layer).
int apply_gate(qoa_state* s, const qoa_gate* g, const double* mats, uint32_t flags) {
  const uint32_t n = s->n;
  switch (g->op) {
    case QOA_OP_H:
      return qoa_apply_hadamard(s, 1.0 / M_SQRT2, 0.0, g->q0);
    case QOA_OP_X: return qoa_apply_x(s, g->q0);
    case QOA_OP_Y: return qoa_apply_y(s, g->q0);
    case QOA_OP_Z: return qoa_apply_z(s, g->q0);
    case QOA_OP_S: return qoa_apply_s(s, g->q0);
    case QOA_OP_T: return qoa_apply_t(s, g->q0);
    case QOA_OP_PHASESHIFT:
    case QOA_OP_PHASE: return qoa_apply_phaseshift(s, g->q0, g->theta);
    case QOA_OP_RX: return qoa_apply_rx(s, g->q0, g->theta);
    case QOA_OP_RY: return qoa_apply_ry(s, g->q0, g->theta);
    case QOA_OP_RZ: return qoa_apply_rz(s, g->q0, g->theta);
    case QOA_OP_CNOT: return qoa_apply_cnot(s, g->q0, g->q1);
    case QOA_OP_CZ:
    case QOA_OP_CPHASE: {
      uint64_t cm, tm;
      if (flags & QOA_CZ_REF_MASKS) {
        cm = g->q0;
        tm = g->q1;
      } else {
        if (g->q0 >= 64 || g->q1 >= 64)
          return set_error(QOA_ERR_ARG, "qubit index >= 64");
        cm = 1ULL << g->q0;
        tm = 1ULL << g->q1;
      }
      return g->op == QOA_OP_CZ ? qoa_apply_cz(s, cm, tm)
                                : qoa_apply_controlled_phase_rotation(s, cm, tm, g->theta);
    }
    case QOA_OP_SWAP: return qoa_apply_swap(s, g->q0, g->q1);
    case QOA_OP_CSWAP: return qoa_apply_controlled_swap(s, g->q0, g->q1, g->q2);
    case QOA_OP_U2:
      if (!mats) return set_error(QOA_ERR_ARG, "U2 gate without a matrix table");
      return qoa_apply_matrix2(s, g->q0, mats + 8 * g->aux);
    case QOA_OP_U4:
      if (!mats) return set_error(QOA_ERR_ARG, "U4 gate without a matrix table");
      return qoa_apply_matrix4(s, g->q0, g->q1, mats + 32 * g->aux);
    case QOA_OP_TOFFOLI: return qoa_apply_toffoli(s, g->q0, g->q1, g->q2);
    case QOA_OP_CY: return qoa_apply_controlled_y(s, g->q0, g->q1);
    case QOA_OP_RESET_ALL: return qoa_apply_reset_all(s);
    default:
      (void)n;
      return set_error(QOA_ERR_ARG, "op 0x%02x is not a unitary gate", g->op);
  }
}

}  // namespace qoa

using namespace qoa;

#define CHECK_STATE(s)                                                             \
  do {                                                                             \
    if (!::qoa::valid(s)) return ::qoa::set_error(QOA_ERR_HANDLE, "invalid handle"); \
  } while (0)
#define CHECK_C128(s)                                                               \
  do {                                                                              \
    if ((s)->precision != QOA_C128)                                                 \
      return ::qoa::set_error(QOA_ERR_UNSUPPORTED, "operation needs a QOA_C128 state"); \
  } while (0)
#define CHECK_Q64(q)                                                                \
  do {                                                                              \
    if ((q) >= 64) return ::qoa::set_error(QOA_ERR_ARG, "qubit index %u >= 64", (unsigned)(q)); \
  } while (0)

extern "C" {

int qoa_version(void) { return 100; }

const char* qoa_last_error(void) { return g_err; }

int qoa_device_info(char* name, size_t name_len, uint64_t* hbm_bytes, int* num_cus) {
  int dev = 0;
  QOA_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t p;
  QOA_HIP_CHECK(hipGetDeviceProperties(&p, dev));
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (hbm_bytes) *hbm_bytes = (uint64_t)p.totalGlobalMem;
  if (num_cus) *num_cus = p.multiProcessorCount;
  return QOA_OK;
}

int qoa_set_option(const char* key, int64_t value) {
  if (!key) return set_error(QOA_ERR_ARG, "key is NULL");
  if (std::strcmp(key, "index64") == 0) {
    g_force_index64 = value != 0;
    return QOA_OK;
  }
  if (std::strcmp(key, "batch_graphs") == 0) {
    g_batch_graphs = value != 0;
    return QOA_OK;
  }
  return set_error(QOA_ERR_ARG, "unknown option '%s'", key);
}

// ---- lifecycle -------------------------------------------------------------------
int qoa_state_create(uint32_t n_qubits, uint32_t precision, qoa_state** out) {
  if (!out) return set_error(QOA_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (precision != QOA_C128 && precision != QOA_C64)
    return set_error(QOA_ERR_ARG, "unknown precision %u", precision);
  const uint32_t n = n_qubits < 1 ? 1 : n_qubits;  // QuantumState::new: n.max(1)
  if (n > 36) return set_error(QOA_ERR_ARG, "n_qubits %u exceeds 36", n);
  qoa_state* s = new qoa_state();
  s->magic = kMagic;
  s->n = n;
  s->len = 1ULL << n;
  s->precision = precision;
  s->amps = nullptr;
  s->red = nullptr;
  s->fused_cache = nullptr;
  s->block_cache = nullptr;
  s->stream = nullptr;
  s->noise_kind = QOA_NOISE_NONE;
  s->noise_p = 0.0;
  s->flags = 0;
  s->timing = false;
  for (int k = 0; k < KK_COUNT; ++k) {
    s->kind_ms[k] = 0;
    s->kind_bytes[k] = 0;
    s->kind_launches[k] = 0;
  }
  // QuantumState::new seeds from ThreadRng (quantum_state.rs:76): not reproducible
  // unless seed_rng is called.
  std::random_device rd;
  s->rng.seed(((uint64_t)rd() << 32) ^ rd());
  hipError_t e = hipGetDevice(&s->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    int rc = set_error(QOA_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
    s->magic = 0;
    delete s;
    return rc;
  }
  const size_t bytes = (size_t)s->len * (precision == QOA_C128 ? 16 : 8);
  e = hipMalloc(&s->amps, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    hipStreamDestroy(s->stream);
    s->magic = 0;
    delete s;
    return set_error(QOA_ERR_OOM, "hipMalloc of %zu bytes for %u qubits failed", bytes, n);
  }
  e = hipMalloc(&s->red, sizeof(double) * kRedDoubles);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    hipFree(s->amps);
    hipStreamDestroy(s->stream);
    s->magic = 0;
    delete s;
    return set_error(QOA_ERR_OOM, "hipMalloc of reduction scratch failed");
  }
  int rc;
  if (precision == QOA_C128) {
    rc = launch_set_basis0(s);
  } else {
    rc = QOA_OK;
    if (hipMemsetAsync(s->amps, 0, bytes, s->stream) != hipSuccess)
      rc = set_error(QOA_ERR_HIP, "hipMemsetAsync failed");
    const float one[2] = {1.0f, 0.0f};
    if (rc == QOA_OK &&
        hipMemcpyAsync(s->amps, one, sizeof(one), hipMemcpyHostToDevice, s->stream) != hipSuccess)
      rc = set_error(QOA_ERR_HIP, "hipMemcpyAsync failed");
  }
  if (rc == QOA_OK && hipStreamSynchronize(s->stream) != hipSuccess)
    rc = set_error(QOA_ERR_HIP, "initialisation failed");
  if (rc != QOA_OK) {
    qoa_state_destroy(s);
    return rc;
  }
  *out = s;
  return QOA_OK;
}

int qoa_state_destroy(qoa_state* s) {
  CHECK_STATE(s);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& t : s->pending) {
    hipEventDestroy(t.start);
    hipEventDestroy(t.stop);
  }
  for (auto e : s->event_pool) hipEventDestroy(e);
  fused_cache_free(s);
  block_cache_free(s);
  if (s->amps) hipFree(s->amps);
  if (s->red) hipFree(s->red);
  if (s->stream) hipStreamDestroy(s->stream);
  s->magic = 0;
  delete s;
  return QOA_OK;
}

int qoa_state_num_qubits(const qoa_state* s, uint32_t* n_out) {
  CHECK_STATE(s);
  if (!n_out) return set_error(QOA_ERR_ARG, "n_out is NULL");
  *n_out = s->n;
  return QOA_OK;
}

int qoa_state_upload(qoa_state* s, const double* amps, uint64_t n_amps) {
  CHECK_STATE(s);
  if (!amps || n_amps != s->len)
    return set_error(QOA_ERR_ARG, "upload needs exactly %llu amplitudes", (unsigned long long)s->len);
  if (s->precision == QOA_C128) {
    QOA_HIP_CHECK(hipMemcpyAsync(s->amps, amps, s->len * 16, hipMemcpyHostToDevice, s->stream));
    QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
    return QOA_OK;
  }
  double* tmp = nullptr;
  if (hipMalloc(&tmp, s->len * 16) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(QOA_ERR_OOM, "staging buffer allocation failed");
  }
  int rc = QOA_OK;
  if (hipMemcpyAsync(tmp, amps, s->len * 16, hipMemcpyHostToDevice, s->stream) != hipSuccess)
    rc = set_error(QOA_ERR_HIP, "upload copy failed");
  if (rc == QOA_OK) rc = launch_upload_c64(s, tmp);
  if (hipStreamSynchronize(s->stream) != hipSuccess && rc == QOA_OK)
    rc = set_error(QOA_ERR_HIP, "upload failed");
  hipFree(tmp);
  return rc;
}

int qoa_state_download(qoa_state* s, double* amps, uint64_t n_amps) {
  CHECK_STATE(s);
  if (!amps || n_amps != s->len)
    return set_error(QOA_ERR_ARG, "download needs exactly %llu amplitudes",
                     (unsigned long long)s->len);
  if (s->precision == QOA_C128) {
    QOA_HIP_CHECK(hipMemcpyAsync(amps, s->amps, s->len * 16, hipMemcpyDeviceToHost, s->stream));
    QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
    return QOA_OK;
  }
  double* tmp = nullptr;
  if (hipMalloc(&tmp, s->len * 16) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(QOA_ERR_OOM, "staging buffer allocation failed");
  }
  int rc = launch_download_c64(s, tmp);
  if (rc == QOA_OK &&
      hipMemcpyAsync(amps, tmp, s->len * 16, hipMemcpyDeviceToHost, s->stream) != hipSuccess)
    rc = set_error(QOA_ERR_HIP, "download copy failed");
  if (hipStreamSynchronize(s->stream) != hipSuccess && rc == QOA_OK)
    rc = set_error(QOA_ERR_HIP, "download failed");
  hipFree(tmp);
  return rc;
}

int qoa_state_synchronize(qoa_state* s) {
  CHECK_STATE(s);
  QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
  return QOA_OK;
}

int qoa_state_copy(qoa_state* dst, const qoa_state* src) {
  CHECK_STATE(dst);
  CHECK_STATE(src);
  if (dst->n != src->n) return set_error(QOA_ERR_ARG, "qubit counts differ");
  QOA_HIP_CHECK(hipStreamSynchronize(src->stream));
  if (dst->precision == src->precision) {
    const size_t bytes = dst->len * (dst->precision == QOA_C128 ? 16 : 8);
    QOA_HIP_CHECK(hipMemcpyAsync(dst->amps, src->amps, bytes, hipMemcpyDeviceToDevice, dst->stream));
  } else {
    const int rc = launch_convert(dst, src);
    if (rc) return rc;
  }
  // The copy reads src on dst's stream: wait for it, so that work enqueued
  // next on src's own stream cannot overwrite src while it is being read.
  QOA_HIP_CHECK(hipStreamSynchronize(dst->stream));
  return QOA_OK;
}

// ---- layer 1 ------------------------------------------------------------------------
int qoa_apply_hadamard(qoa_state* s, double nf_re, double nf_im, uint32_t q) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  if (q >= s->n) return QOA_OK;  // j = i|bit >= len for every i: no pair visited
  return launch_hadamard(s, Cx{nf_re, nf_im}, q, 0);
}

int qoa_apply_x(qoa_state* s, uint32_t q) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  if (q >= s->n) return QOA_OK;
  return launch_pauli_x(s, q, 0);
}

int qoa_apply_y(qoa_state* s, uint32_t q) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  if (q >= s->n) return QOA_OK;
  return launch_pauli_y(s, q, 0);
}

int qoa_apply_z(qoa_state* s, uint32_t q) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  if (q >= s->n) return QOA_OK;  // (i & bit) != 0 never holds
  return launch_diag_mask(s, 1ULL << q, true, Cx{0, 0});
}

static int phase_gate(qoa_state* s, uint32_t q, Cx f) {
  if (q >= s->n) return QOA_OK;
  return launch_diag_mask(s, 1ULL << q, false, f);
}

int qoa_apply_t(qoa_state* s, uint32_t q) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  return phase_gate(s, q, from_polar(1.0, M_PI / 4.0));  // vectorization.rs:218
}

int qoa_apply_s(qoa_state* s, uint32_t q) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  return phase_gate(s, q, from_polar(1.0, M_PI / 2.0));  // vectorization.rs:263
}

int qoa_apply_phaseshift(qoa_state* s, uint32_t q, double angle) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  return phase_gate(s, q, from_polar(1.0, angle));  // vectorization.rs:308
}

int qoa_apply_reset(qoa_state* s, uint32_t q) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  const uint64_t bit = bit_or_zero(q, s->n);  // q >= n: every index counts as bit-0
  int rc = launch_norm_sum(s, bit, false, 4);
  if (rc) return rc;
  return launch_reset_apply(s, bit, 4);
}

int qoa_apply_swap(qoa_state* s, uint32_t q1, uint32_t q2) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q1); CHECK_Q64(q2);
  if (q1 >= s->n || q2 >= s->n || q1 == q2) return QOA_OK;
  return launch_swap2(s, q1, q2, 0);
}

int qoa_apply_controlled_swap(qoa_state* s, uint32_t c, uint32_t t1, uint32_t t2) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(c); CHECK_Q64(t1); CHECK_Q64(t2);
  if (c >= s->n || t1 >= s->n || t2 >= s->n || t1 == t2) return QOA_OK;
  // With c equal to one target the reference loop degenerates into a
  // controlled flip of the other target (vectorization.rs:1209-1212).
  if (c == t1) return launch_pauli_x(s, t2, 1ULL << c);
  if (c == t2) return launch_pauli_x(s, t1, 1ULL << c);
  return launch_swap2(s, t1, t2, 1ULL << c);
}

int qoa_apply_rx(qoa_state* s, uint32_t q, double angle) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  if (q >= s->n) return QOA_OK;
  double c, sn;
  cis(angle / 2.0, &c, &sn);  // vectorization.rs:529-535
  const Cx m[4] = {{c, 0.0}, {0.0, -sn}, {0.0, -sn}, {c, 0.0}};
  return launch_matrix2(s, m, q, 0);
}

int qoa_apply_ry(qoa_state* s, uint32_t q, double angle) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  if (q >= s->n) return QOA_OK;
  double c, sn;
  cis(angle / 2.0, &c, &sn);  // vectorization.rs:589-595
  const Cx m[4] = {{c, 0.0}, {-sn, 0.0}, {sn, 0.0}, {c, 0.0}};
  return launch_matrix2(s, m, q, 0);
}

int qoa_apply_rz(qoa_state* s, uint32_t q, double angle) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  const Cx m00 = from_polar(1.0, -angle / 2.0), m11 = from_polar(1.0, angle / 2.0);  // :650-651
  return launch_rz(s, bit_or_zero(q, s->n), m00, m11);  // q >= n: global phase m00
}

int qoa_apply_cnot(qoa_state* s, uint32_t c, uint32_t t) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(c); CHECK_Q64(t);
  if (c >= s->n || t >= s->n || c == t) return QOA_OK;
  return launch_pauli_x(s, t, 1ULL << c);
}

static inline bool mask_in_range(uint64_t m, uint32_t n) { return n >= 64 || (m >> n) == 0; }

int qoa_apply_cz(qoa_state* s, uint64_t cm, uint64_t tm) {
  CHECK_STATE(s); CHECK_C128(s);
  const uint64_t m = cm | tm;
  if (!mask_in_range(m, s->n)) return QOA_OK;  // (i & m) == m never holds
  return launch_diag_mask(s, m, true, Cx{0, 0});
}

int qoa_apply_controlled_phase_rotation(qoa_state* s, uint64_t cm, uint64_t tm, double angle) {
  CHECK_STATE(s); CHECK_C128(s);
  const uint64_t m = cm | tm;
  if (!mask_in_range(m, s->n)) return QOA_OK;
  return launch_diag_mask(s, m, false, from_polar(1.0, angle));  // vectorization.rs:794
}

int qoa_apply_reset_all(qoa_state* s) {
  CHECK_STATE(s); CHECK_C128(s);
  return launch_set_basis0(s);
}

int qoa_apply_matrix2(qoa_state* s, uint32_t q, const double* m) {
  CHECK_STATE(s); CHECK_C128(s); CHECK_Q64(q);
  if (!m) return set_error(QOA_ERR_ARG, "matrix is NULL");
  if (q >= s->n) return QOA_OK;
  const Cx c[4] = {{m[0], m[1]}, {m[2], m[3]}, {m[4], m[5]}, {m[6], m[7]}};
  return launch_matrix2(s, c, q, 0);
}

int qoa_apply_matrix4(qoa_state* s, uint32_t q1, uint32_t q2, const double* m) {
  CHECK_STATE(s); CHECK_C128(s);
  if (!m) return set_error(QOA_ERR_ARG, "matrix is NULL");
  if (q1 >= s->n || q2 >= s->n)
    return set_error(QOA_ERR_QUBIT, "qubit indices (%u, %u) out of bounds for %u-qubit state", q1,
                     q2, s->n);
  if (q1 == q2) return set_error(QOA_ERR_ARG, "qubits cannot be the same for two-qubit unitary");
  return launch_matrix4(s, q1, q2, m);
}

int qoa_apply_toffoli(qoa_state* s, uint32_t c1, uint32_t c2, uint32_t t) {
  CHECK_STATE(s); CHECK_C128(s);
  if (c1 >= s->n || c2 >= s->n || t >= s->n)
    return set_error(QOA_ERR_QUBIT, "toffoli qubit out of bounds");
  if (c1 == c2 || c1 == t || c2 == t)
    return set_error(QOA_ERR_ARG, "control and target qubits must be distinct for toffoli gate");
  return launch_pauli_x(s, t, (1ULL << c1) | (1ULL << c2));
}

int qoa_apply_controlled_y(qoa_state* s, uint32_t c, uint32_t t) {
  CHECK_STATE(s); CHECK_C128(s);
  if (c >= s->n || t >= s->n) return set_error(QOA_ERR_QUBIT, "controlled-y qubit out of bounds");
  if (c == t) return set_error(QOA_ERR_ARG, "control and target qubits cannot be the same");
  return launch_pauli_y(s, t, 1ULL << c);
}

// ---- reductions ---------------------------------------------------------------------
int qoa_validate(qoa_state* s, int* has_nan, int* has_inf, double* norm_sqr) {
  CHECK_STATE(s); CHECK_C128(s);
  int rc = launch_nonfinite(s, 6);
  if (rc == QOA_OK) rc = launch_norm_sum(s, 0, false, 0);
  if (rc) return rc;
  double slots[2];
  QOA_HIP_CHECK(hipMemcpyAsync(&slots[0], s->red + kSlotBase, sizeof(double), hipMemcpyDeviceToHost,
                               s->stream));
  QOA_HIP_CHECK(hipMemcpyAsync(&slots[1], s->red + kSlotBase + 6, sizeof(double),
                               hipMemcpyDeviceToHost, s->stream));
  QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
  unsigned int flags;
  std::memcpy(&flags, &slots[1], sizeof(flags));
  if (has_nan) *has_nan = (flags & 1u) ? 1 : 0;
  if (has_inf) *has_inf = (flags & 2u) ? 1 : 0;
  if (norm_sqr) *norm_sqr = slots[0];
  return QOA_OK;
}

int qoa_norm_sqr(qoa_state* s, double* out) {
  CHECK_STATE(s); CHECK_C128(s);
  if (!out) return set_error(QOA_ERR_ARG, "out is NULL");
  int rc = launch_norm_sum(s, 0, false, 0);
  if (rc) return rc;
  QOA_HIP_CHECK(hipMemcpyAsync(out, s->red + kSlotBase, sizeof(double), hipMemcpyDeviceToHost,
                               s->stream));
  QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
  return QOA_OK;
}

int qoa_normalize(qoa_state* s, double* norm_sqr_out, int* was_reset) {
  CHECK_STATE(s); CHECK_C128(s);
  int rc = launch_norm_sum(s, 0, false, 0);
  if (rc) return rc;
  rc = launch_normalize_apply(s, 0);
  if (rc) return rc;
  if (norm_sqr_out || was_reset) {
    double v = 0.0;
    QOA_HIP_CHECK(hipMemcpyAsync(&v, s->red + kSlotBase, sizeof(double), hipMemcpyDeviceToHost,
                                 s->stream));
    QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
    if (norm_sqr_out) *norm_sqr_out = v;
    if (was_reset) *was_reset = !(v > 1e-12);
  }
  return QOA_OK;
}

int qoa_measure(qoa_state* s, uint32_t q, double r, int* outcome, double* p0, double* p1) {
  CHECK_STATE(s); CHECK_C128(s);
  if (q >= s->n) return set_error(QOA_ERR_QUBIT, "qubit index %u out of bounds", q);
  const uint64_t bit = 1ULL << q;
  int rc = launch_norm_sum(s, bit, true, 2);
  if (rc) return rc;
  double p[2];
  QOA_HIP_CHECK(hipMemcpyAsync(p, s->red + kSlotBase + 2, sizeof(p), hipMemcpyDeviceToHost,
                               s->stream));
  QOA_HIP_CHECK(hipStreamSynchronize(s->stream));
  const int out = (r < p[0]) ? 0 : 1;  // quantum_state.rs:425-429
  double f;
  if (out == 0) f = p[0] > 1e-12 ? 1.0 / std::sqrt(p[0]) : 0.0;
  else f = p[1] > 1e-12 ? 1.0 / std::sqrt(p[1]) : 0.0;
  rc = launch_collapse(s, bit, out, f);
  if (rc) return rc;
  if (outcome) *outcome = out;
  if (p0) *p0 = p[0];
  if (p1) *p1 = p[1];
  return QOA_OK;
}

int qoa_probabilities(qoa_state* s, double* out, uint64_t n_amps) {
  CHECK_STATE(s); CHECK_C128(s);
  if (!out || n_amps != s->len) return set_error(QOA_ERR_ARG, "output must hold 2^n doubles");
  double* dev = nullptr;
  if (hipMalloc(&dev, s->len * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(QOA_ERR_OOM, "probability buffer allocation failed");
  }
  int rc = launch_probabilities(s, dev);
  if (rc == QOA_OK && hipMemcpyAsync(out, dev, s->len * sizeof(double), hipMemcpyDeviceToHost,
                                     s->stream) != hipSuccess)
    rc = set_error(QOA_ERR_HIP, "probability copy failed");
  if (hipStreamSynchronize(s->stream) != hipSuccess && rc == QOA_OK)
    rc = set_error(QOA_ERR_HIP, "probabilities failed");
  hipFree(dev);
  return rc;
}

int qoa_apply_pauli_codes(qoa_state* s, const uint8_t* codes, uint32_t n_codes) {
  CHECK_STATE(s); CHECK_C128(s);
  if (!codes && n_codes) return set_error(QOA_ERR_ARG, "codes is NULL");
  if (n_codes > 64) return set_error(QOA_ERR_ARG, "at most 64 codes");
  return launch_pauli_codes(s, codes, n_codes);
}

// ---- circuits ----------------------------------------------------------------------
int qoa_run_gates(qoa_state* s, const qoa_gate* gates, uint32_t n_gates, const double* mats,
                  uint32_t flags) {
  CHECK_STATE(s); CHECK_C128(s);
  if (!gates && n_gates) return set_error(QOA_ERR_ARG, "gates is NULL");
  for (uint32_t k = 0; k < n_gates; ++k) {
    int rc = apply_gate(s, &gates[k], mats, flags);
    if (rc) return rc;
  }
  return QOA_OK;
}

int qoa_run_trajectory(qoa_state* s, const qoa_gate* gates, uint32_t n_gates,
                       const double* mats, uint32_t flags, uint64_t seed, uint32_t traj,
                       double p1, double p2, const uint32_t* order) {
  CHECK_STATE(s); CHECK_C128(s);
  if (!gates && n_gates) return set_error(QOA_ERR_ARG, "gates is NULL");
  if (!(p1 >= 0.0 && p1 <= 1.0 && p2 >= 0.0 && p2 <= 1.0))
    return set_error(QOA_ERR_ARG, "noise probabilities outside [0,1]");
  for (uint32_t k = 0; k < n_gates; ++k) {  // validate before the first launch
    const uint32_t gi = order ? order[k] : k;
    if (gi >= n_gates) return set_error(QOA_ERR_ARG, "order[%u] = %u out of range", k, gi);
    const qoa_gate& g = gates[gi];
    const uint32_t ar = traj_arity(g.op);
    const uint32_t q[3] = {g.q0, g.q1, g.q2};
    for (uint32_t j = 0; j < ar; ++j) {
      if (q[j] >= s->n) return set_error(QOA_ERR_ARG, "gate %u: operand %u out of range", gi, q[j]);
      for (uint32_t i = 0; i < j; ++i)
        if (q[i] == q[j]) return set_error(QOA_ERR_ARG, "gate %u: repeated operand", gi);
    }
    if (g.op == QOA_OP_MEASURE || g.op == QOA_OP_RESET || g.op == QOA_OP_RESET_ALL)
      return set_error(QOA_ERR_ARG, "gate %u: op 0x%02x has no trajectory form", gi, g.op);
  }
  for (uint32_t k = 0; k < n_gates; ++k) {
    const uint32_t gi = order ? order[k] : k;
    const qoa_gate& g = gates[gi];
    int rc = apply_gate(s, &g, mats, flags);
    if (rc) return rc;
    const uint32_t ar = traj_arity(g.op);
    const double p = ar == 1 ? p1 : p2;
    if (!(p > 0.0)) continue;
    const uint32_t q[3] = {g.q0, g.q1, g.q2};
    for (uint32_t j = 0; j < ar; ++j) {  // errors after the gate, operand order
      const uint32_t code = traj_pauli(seed, traj, gi, q[j], p);
      if (code == 1) rc = launch_pauli_x(s, q[j], 0);
      else if (code == 2) rc = launch_pauli_y(s, q[j], 0);
      else if (code == 3) rc = launch_diag_mask(s, 1ULL << q[j], true, Cx{0.0, 0.0});
      if (rc) return rc;
    }
  }
  return QOA_OK;
}

int qoa_run_fused(qoa_state* s, const qoa_gate* gates, uint32_t n_gates, const double* mats,
                  const qoa_fuse_opts* opts) {
  CHECK_STATE(s); CHECK_C128(s);
  if (!gates && n_gates) return set_error(QOA_ERR_ARG, "gates is NULL");
  PlanOptions po{opts ? opts->tile_qubits : 0u, opts ? opts->low_qubits : 0u,
                 opts ? opts->flags : 0u};
  return run_fused(s, gates, n_gates, mats, po);
}

int qoa_plan_fused(uint32_t n_qubits, const qoa_gate* gates, uint32_t n_gates,
                   const qoa_fuse_opts* opts, uint32_t* order, uint32_t* n_passes) {
  if (!gates && n_gates) return set_error(QOA_ERR_ARG, "gates is NULL");
  PlanOptions po{opts ? opts->tile_qubits : 0u, opts ? opts->low_qubits : 0u,
                 opts ? opts->flags : 0u};
  std::vector<uint32_t> ord;
  uint32_t np = 0;
  int rc = plan_order(n_qubits, gates, n_gates, po, &ord, &np);
  if (rc) return rc;
  if (order) std::memcpy(order, ord.data(), sizeof(uint32_t) * ord.size());
  if (n_passes) *n_passes = np;
  return QOA_OK;
}

int qoa_timing_enable(qoa_state* s, int enable) {
  CHECK_STATE(s);
  int rc = timing_collect(s);
  s->timing = enable != 0;
  if (enable) {
    for (int k = 0; k < KK_COUNT; ++k) {
      s->kind_ms[k] = 0;
      s->kind_bytes[k] = 0;
      s->kind_launches[k] = 0;
    }
  }
  return rc;
}

int qoa_timing_read(qoa_state* s, const char* kind, double* total_ms, uint64_t* launches,
                    double* bytes) {
  CHECK_STATE(s);
  int rc = timing_collect(s);
  if (rc) return rc;
  for (int k = 0; k < KK_COUNT; ++k) {
    if (kind && std::strcmp(kind, kind_name(k)) == 0) {
      if (total_ms) *total_ms = s->kind_ms[k];
      if (launches) *launches = s->kind_launches[k];
      if (bytes) *bytes = s->kind_bytes[k];
      return QOA_OK;
    }
  }
  return set_error(QOA_ERR_ARG, "unknown kernel kind '%s'", kind ? kind : "(null)");
}

// ---- layer 2: QuantumState ---------------------------------------------------------
int qoa_qs_set_noise(qoa_state* s, int kind, double p) {
  CHECK_STATE(s);
  if (kind < QOA_NOISE_NONE || kind > QOA_NOISE_FIXED) return set_error(QOA_ERR_ARG, "noise kind");
  s->noise_kind = kind;
  s->noise_p = p;
  return QOA_OK;
}

int qoa_qs_seed_rng(qoa_state* s, uint64_t seed) {
  CHECK_STATE(s);
  s->rng.seed(seed);
  return QOA_OK;
}

int qoa_qs_set_flags(qoa_state* s, uint32_t flags) {
  CHECK_STATE(s);
  s->flags = flags;
  return QOA_OK;
}

// apply_noise (quantum_state.rs:839-869) then normalize (:888-907). The Pauli
// errors chosen for one call act on distinct qubits and are applied as one
// fused sweep with the per-qubit sweep arithmetic.
int qoa_qs_apply_noise(qoa_state* s) {
  CHECK_STATE(s); CHECK_C128(s);
  double level = -1.0;
  if (s->noise_kind == QOA_NOISE_RANDOM) level = 0.1 + (1.0 - 0.1) * s->rng.next_f64();
  else if (s->noise_kind == QOA_NOISE_FIXED && s->noise_p > 0.0) level = s->noise_p;
  if (level >= 0.0) {
    uint8_t codes[64] = {0};
    bool any = false;
    for (uint32_t q = 0; q < s->n; ++q) {
      const double r = s->rng.next_f64();
      if (r < level) {
        codes[q] = (uint8_t)(1 + s->rng.below3());
        any = true;
      }
    }
    if (any) {
      int rc = launch_pauli_codes(s, codes, s->n);
      if (rc) return rc;
    }
  }
  int rc = launch_norm_sum(s, 0, false, 0);
  if (rc) return rc;
  return launch_normalize_apply(s, 0);
}

static bool gate_in_range(uint32_t n, const qoa_gate* g) {
  switch (g->op) {
    case QOA_OP_RESET_ALL: return true;
    case QOA_OP_CNOT: case QOA_OP_CZ: case QOA_OP_CPHASE: case QOA_OP_SWAP:
    case QOA_OP_U4: case QOA_OP_CY:
      return g->q0 < n && g->q1 < n && g->q0 != g->q1;
    case QOA_OP_CSWAP: case QOA_OP_TOFFOLI:
      return g->q0 < n && g->q1 < n && g->q2 < n && g->q0 != g->q1 && g->q0 != g->q2 &&
             g->q1 != g->q2;
    default: return g->q0 < n;
  }
}

int qoa_qs_apply(qoa_state* s, const qoa_gate* g, const double* mats) {
  CHECK_STATE(s); CHECK_C128(s);
  if (!g) return set_error(QOA_ERR_ARG, "gate is NULL");
  if (g->op == QOA_OP_MEASURE) {
    int out;
    return qoa_qs_measure(s, g->q0, &out);
  }
  if (!gate_in_range(s->n, g))
    return set_error(QOA_ERR_QUBIT,
                     "error: qubit index out of bounds for %u-qubit state (op 0x%02x)", s->n, g->op);
  int rc = (g->op == QOA_OP_RESET) ? qoa_apply_reset(s, g->q0) : apply_gate(s, g, mats, s->flags);
  if (rc) return rc;
  return qoa_qs_apply_noise(s);
}

int qoa_qs_measure(qoa_state* s, uint32_t q, int* outcome) {
  CHECK_STATE(s); CHECK_C128(s);
  if (q >= s->n) return set_error(QOA_ERR_QUBIT, "qubit index %u out of bounds", q);
  const double r = s->rng.next_f64();  // quantum_state.rs:424
  int out = 0;
  int rc = qoa_measure(s, q, r, &out, nullptr, nullptr);
  if (rc) return rc;
  if (outcome) *outcome = out;
  return qoa_qs_apply_noise(s);
}

int qoa_qs_measure_all(qoa_state* s, uint8_t* out, uint32_t n_out) {
  CHECK_STATE(s); CHECK_C128(s);
  if (!out || n_out < s->n) return set_error(QOA_ERR_ARG, "output must hold n entries");
  qoa_state* tmp = nullptr;
  int rc = qoa_state_create(s->n, QOA_C128, &tmp);
  if (rc) return rc;
  rc = qoa_state_copy(tmp, s);
  if (rc == QOA_OK) {
    // measure_all clones the whole state, RNG included (quantum_state.rs:909-918)
    tmp->rng = s->rng;
    tmp->noise_kind = s->noise_kind;
    tmp->noise_p = s->noise_p;
    tmp->flags = s->flags;
    for (uint32_t q = 0; q < s->n && rc == QOA_OK; ++q) {
      int o = 0;
      rc = qoa_qs_measure(tmp, q, &o);
      out[q] = (uint8_t)o;
    }
  }
  qoa_state_destroy(tmp);
  return rc;
}

}  // extern "C"
