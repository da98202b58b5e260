// specialize.cpp — circuit-specialised register passes (QOA_FUSE_SPECIALIZE).
//
// The op-loop kernel (k_fused_reg) decodes every op of a pass in every wave:
// scalar loads of the op words, a tree of scalar branches down to the op's
// code, the loop itself. For a plan that runs more than once, this file
// writes one kernel per pass whose op words are literal constants (calls into
// jit_device.hpp, the same device code as the op loop), compiles the passes
// with hipRTC for the device's architecture (one program per plan), and keeps
// the code objects in a process-wide cache and an on-disk one, both keyed by
// the generated source, so a circuit structure compiles once.
// Coefficients stay in the plan's op records and pool, which the kernels read
// as the op loop does; the results are bit-identical to the op-loop kernel.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "planner.hpp"
#include "qoa_internal.hpp"
#include "specialize.hpp"

namespace qoa {

// Header texts that generated sources include (embedded by the build from
// jit_device.hpp, tile_device.hpp, planner.hpp, traj_noise.hpp, qoa_hip.h).
extern const int kRtcHeaderCount;
extern const char* const kRtcHeaderNames[];
extern const char* const kRtcHeaderTexts[];

namespace {

uint64_t fnv64(const std::string& s, uint64_t h = 0xCBF29CE484222325ULL) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 0x100000001B3ULL;
  }
  return h;
}

void append_u32(std::string* out, uint32_t x) {
  char b[16];
  std::snprintf(b, sizeof(b), "%uu", x);
  *out += b;
}

void append_u64(std::string* out, uint64_t x) {
  char b[32];
  std::snprintf(b, sizeof(b), "%lluull", (unsigned long long)x);
  *out += b;
}

// Logical register r of the tile lives in physical register perm[r]: X,
// CNOT and SWAP ops on register slots (without thread or tile conditions)
// only permute this map, and every other op is emitted with it.
struct Perm {
  uint8_t p[kRegAmps];
  Perm() {
    for (int r = 0; r < kRegAmps; ++r) p[r] = (uint8_t)r;
  }
  uint64_t word() const {
    uint64_t w = 0;
    for (int r = 0; r < kRegAmps; ++r) w |= (uint64_t)p[r] << (4 * r);
    return w;
  }
};

// Applies op to the permutation instead of emitting code, when it is a
// register permutation without thread or tile conditions; returns whether it
// did. (reg_pair<slot, rsel, TK_X> and reg_swap<A, B> move exactly these
// registers.)
bool permute(const TileOp& op, Perm* pm) {
  if (op.tl != 0 || op.gmask != 0 || op.gval != 0) return false;
  if (op.kind == TK_X) {
    const uint32_t j = std::min<uint32_t>(op.ins[0], 3u), rs = op.rsel & ~(1u << j);
    for (uint32_t r = 0; r < (uint32_t)kRegAmps; ++r)
      if (!(r & (1u << j)) && (r & rs) == rs) std::swap(pm->p[r], pm->p[r | (1u << j)]);
    return true;
  }
  if (op.kind == TK_SWAP) {
    const uint32_t a = std::min(op.ins[0], op.ins[1]), c = std::max(op.ins[0], op.ins[1]);
    if (a == c || c > 3) return false;
    const uint32_t ma = 1u << a, mc = 1u << c, rs = op.rsel;
    for (uint32_t r = 0; r < (uint32_t)kRegAmps; ++r)
      if (!(r & (ma | mc)) && (r & rs) == rs) std::swap(pm->p[r | ma], pm->p[r | mc]);
    return true;
  }
  return false;
}

// The call reg_op (or reg_tile's exchange) makes for op `op`, record index
// i of its pass, with everything that depends on the op's words resolved
// here as reg_op resolves it on the device (kind, 2x2 shape, register slot,
// register and thread-bit controls), on registers permuted by pi.
std::string op_call(const TileOp& op, uint32_t i, uint32_t k, bool fma, uint64_t pi) {
  char b[320];
  const char* F = fma ? "true" : "false";
  const uint32_t rsel = op.rsel, tl = op.tl;
  const uint32_t j = std::min<uint32_t>(op.ins[0], 3u);  // a 2x2 op's register slot
  switch (op.kind) {
    case TK_H: case TK_M: case TK_X: case TK_Y: {
      const char* kind;
      bool shaped = false;  // kMReal / kMRx: the general 2x2 code under controls
      if (op.kind == TK_H) {
        kind = op.c[1] != 0.0 ? "TK_H" : op.c[0] == 1.0 ? "kHRaw" : "kHReal";
      } else if (op.kind == TK_M) {
        shaped = op.t2 == kM2Real || op.t2 == kM2Rx;
        kind = op.t2 == kM2Real    ? "kMReal"
               : op.t2 == kM2Rx    ? "kMRx"
               : op.t2 == kM2RealU ? "kMRealU"
               : op.t2 == kM2RxU   ? "kMRxU"
                                   : "TK_M";
      } else {
        kind = op.kind == TK_X ? "TK_X" : "TK_Y";
      }
      if (tl == 0 && rsel == 0)
        std::snprintf(b, sizeof(b), "reg_pair<%u, 0, %s, %s>(S.v, spec_coef4(ops + %u), %lluull);", j,
                      kind, F, i, (unsigned long long)pi);
      else if (op.kind == TK_X && tl == 0)
        std::snprintf(b, sizeof(b), "reg_pair<%u, %u, TK_X, %s>(S.v, spec_coef4(ops + %u), %lluull);",
                      j, rsel, F, i, (unsigned long long)pi);
      else
        std::snprintf(b, sizeof(b),
                      "reg_pair_sel<%u, %s, %s>(S.v, spec_coef4(ops + %u), %uu, spec_tok(S, %uu),"
                      " %lluull);",
                      j, shaped ? "TK_M" : kind, F, i, rsel, tl, (unsigned long long)pi);
    } break;
    case TK_DNEG:
      if (tl == 0)
        std::snprintf(b, sizeof(b), "reg_neg<%u>(S.v, %lluull);", rsel, (unsigned long long)pi);
      else
        std::snprintf(b, sizeof(b), "reg_scale<%u>(S.v, spec_tok(S, %uu) ? -1.0 : 1.0, %lluull);",
                      rsel, tl, (unsigned long long)pi);
      break;
    case TK_DPHASE:
      if (tl == 0)
        std::snprintf(b, sizeof(b), "reg_phase<%u, %s, true>(S.v, spec_c2(ops + %u, 0), %lluull);",
                      rsel, F, i, (unsigned long long)pi);
      else
        std::snprintf(b, sizeof(b),
                      "reg_phase<%u, %s, false>(S.v, sel2(spec_tok(S, %uu), spec_c2(ops + %u, 0),"
                      " make_double2(1.0, 0.0)), %lluull);",
                      rsel, F, tl, i, (unsigned long long)pi);
      break;
    case TK_RZ:
      if (op.rt < kRegBits)
        std::snprintf(b, sizeof(b),
                      "reg_rz<%u, %s>(S.v, spec_c2(ops + %u, 0), spec_c2(ops + %u, 1), %lluull);",
                      op.rt, F, i, i, (unsigned long long)pi);
      else
        std::snprintf(b, sizeof(b),
                      "spec_rz_thread<%s>(S, spec_c2(ops + %u, 0), spec_c2(ops + %u, 1), %uu);", F,
                      i, i, (uint32_t)op.t);
      break;
    case TK_SWAP: {
      const uint32_t s0 = op.ins[0], s1 = op.ins[1];
      const uint32_t a0 = std::min(s0, s1), b0 = std::max(s0, s1);
      // reg_op's switch over the slot pair (its default case is slots 2, 3)
      const bool known = (a0 == 0 && b0 >= 1 && b0 <= 3) || (a0 == 1 && (b0 == 2 || b0 == 3));
      std::snprintf(b, sizeof(b), "reg_swap<%u, %u>(S.v, %uu, %uu, spec_tok(S, %uu), %lluull);",
                    known ? a0 : 2u, known ? b0 : 3u, rsel, tl, tl, (unsigned long long)pi);
    } break;
    case TK_MAT4:
      std::snprintf(b, sizeof(b), "reg_mat4<2, 3, %s>(S.v, pool + %u, %lluull);", F, op.aux,
                    (unsigned long long)pi);
      break;
    case TK_LAYOUT: {  // relayout's choice of the round-keying slots
      const uint32_t q = reg_xchg_split(k, false), a = op.t, c = op.t2;
      uint32_t A = 0, B = 1;
      if (q == 1) {
        A = a > 3 ? 3u : a;
        B = (A + 1) & 3u;
      } else if (q == 2) {
        const uint32_t ab = a * 4 + c;
        const bool known = ab == 1 || ab == 2 || ab == 3 || ab == 6 || ab == 7;
        A = known ? a : 2u;
        B = known ? c : 3u;
      }
      uint32_t w1;
      std::memcpy(&w1, op.ins, sizeof(w1));
      std::snprintf(b, sizeof(b), "spec_layout<%u, %u, %u>(S, ops + %u, %uu, %lluull);", q, A, B,
                    i, w1, (unsigned long long)pi);
    } break;
    default:
      b[0] = '\0';
      break;
  }
  return b;
}

// One pass as a kernel: load, the ops in order, store.
void pass_source(const FusedPlan& P, uint32_t p, bool fma, bool idx64, std::string* out) {
  const PassPlan& pp = P.passes[p];
  const uint32_t k = pp.k;
  const uint32_t threads = 1u << (k - kRegBits);
  const char* I = idx64 ? "uint64_t" : "uint32_t";
  uint32_t spw[4], stw[4], srt[4];
  pass_geo_words(pp, spw, stw, srt);
  char head[1024];
  std::snprintf(head, sizeof(head),
                "extern \"C\" __global__ __launch_bounds__(%u) __attribute__((amdgpu_waves_per_eu(5)))"
                " void qoa_spec_%u(double2* __restrict__ a, const TileOp* __restrict__ ops,"
                " const double* __restrict__ pool) {\n"
                "  extern __shared__ double2 lds[];\n"
                "  Geo geo; geo.k = %u;"
                " geo.spw[0] = %uu; geo.spw[1] = %uu; geo.spw[2] = %uu; geo.spw[3] = %uu;"
                " geo.stw[0] = %uu; geo.stw[1] = %uu; geo.stw[2] = %uu; geo.stw[3] = %uu;"
                " geo.srt[0] = %uu; geo.srt[1] = %uu; geo.srt[2] = %uu; geo.srt[3] = %uu;\n"
                "  SpecTile<%s> S;\n"
                "  spec_load(S, a, geo, lds, (%s)spec_tile_index());\n",
                threads, p, k, spw[0], spw[1], spw[2], spw[3], stw[0], stw[1], stw[2], stw[3],
                srt[0], srt[1], srt[2], srt[3], I, I);
  *out += head;
  Perm perm;
  for (uint32_t i = pp.op_begin; i < pp.op_end; ++i) {
    const TileOp& op = P.ops[i];
    if (permute(op, &perm)) continue;
    const uint64_t pi = perm.word();
    std::string call;
    if (op.kind == TK_DIAGN) {  // terms at pool[aux], lmask of them (FMA kernels only)
      if (!fma) continue;
      call = "{ double2 acc = make_double2(1.0, 0.0);\n";
      for (uint32_t t = 0; t < op.lmask; ++t) {
        DiagTerm d;
        std::memcpy(&d, P.pool.data() + op.aux + kDiagTermDoubles * t, sizeof(d));
        call += "    spec_term(S, acc, pool + ";
        append_u32(&call, op.aux + kDiagTermDoubles * t + 3);  // c[0], c[1]
        call += ", ";
        append_u32(&call, d.tl);
        call += ", ";
        append_u64(&call, d.gmask);
        call += ", ";
        append_u64(&call, d.gval);
        call += ");\n";
        if (d.rs & kDiagGroupEnd) {
          call += "    spec_group<" + std::to_string(d.rs & 15u) + ">(S, acc, ";
          append_u64(&call, pi);
          call += ");\n";
        }
      }
      call += "  }";
    } else {
      call = op_call(op, i - pp.op_begin, k, fma, pi);
    }
    if (op.kind != TK_LAYOUT && (op.gmask || op.gval)) {  // a tile condition (outside qubits)
      *out += "  if (((uint64_t)S.base & ";
      append_u64(out, op.gmask);
      *out += ") == ";
      append_u64(out, op.gval);
      *out += ") " + call + "\n";
    } else {
      *out += "  " + call + "\n";
    }
  }
  *out += "  spec_store(S, a, geo, ";
  append_u64(out, perm.word());
  *out += ");\n}\n";
}

std::string rtc_log(hiprtcProgram prog) {
  size_t n = 0;
  if (hiprtcGetProgramLogSize(prog, &n) != HIPRTC_SUCCESS || n == 0) return std::string();
  std::string log(n, '\0');
  if (hiprtcGetProgramLog(prog, &log[0]) != HIPRTC_SUCCESS) return std::string();
  return log;
}

// The library's device flags: unfused FP64 products, no packed FP32, and the
// two code generation options the register kernels rely on.
const char* const kRtcOptions[] = {"-O3", "-std=c++17", "-ffp-contract=off",
                                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                                   "-mllvm", "-simplifycfg-sink-common=false",
                                   "-mllvm", "-structurizecfg-skip-uniform-regions=true"};
constexpr int kRtcOptionCount = (int)(sizeof(kRtcOptions) / sizeof(kRtcOptions[0]));

// Hash of what a compiled program depends on besides its own source: the
// embedded device headers and the options (so a library built from changed
// device code never takes an older code object from the disk cache).
uint64_t toolchain_key() {
  static const uint64_t key = [] {
    uint64_t h = 0xCBF29CE484222325ULL;
    for (int i = 0; i < kRtcHeaderCount; ++i) h = fnv64(kRtcHeaderNames[i], fnv64(kRtcHeaderTexts[i], h));
    for (int i = 0; i < kRtcOptionCount; ++i) h = fnv64(kRtcOptions[i], h);
    return h;
  }();
  return key;
}

// Compiles one source; returns an error message or "".
std::string rtc_compile(const std::string& src, const std::string& arch, std::vector<char>* code) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "qoa_spec.hip", kRtcHeaderCount, kRtcHeaderTexts,
                          kRtcHeaderNames) != HIPRTC_SUCCESS)
    return "hiprtcCreateProgram failed";
  const std::string off = "--offload-arch=" + arch;
  std::vector<const char*> opts{off.c_str()};
  opts.insert(opts.end(), kRtcOptions, kRtcOptions + kRtcOptionCount);
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  std::string err;
  if (r != HIPRTC_SUCCESS) {
    err = std::string("hipRTC: ") + hiprtcGetErrorString(r) + "\n" + rtc_log(prog).substr(0, 4000);
  } else {
    size_t n = 0;
    if (hiprtcGetCodeSize(prog, &n) != HIPRTC_SUCCESS || n == 0) {
      err = "hiprtcGetCodeSize failed";
    } else {
      code->resize(n);
      if (hiprtcGetCode(prog, code->data()) != HIPRTC_SUCCESS) err = "hiprtcGetCode failed";
    }
  }
  hiprtcDestroyProgram(&prog);
  return err;
}

// Code objects by source hash (process-wide; the same circuit structure on
// any handle compiles once). Loaded modules are kept per device.
struct ModuleEntry {
  std::vector<char> code;
  std::unordered_map<int, hipModule_t> loaded;
};
std::mutex g_mu;
std::unordered_map<uint64_t, std::shared_ptr<ModuleEntry>> g_modules;
SpecStats g_stats;

}  // namespace

std::string specialized_source(const FusedPlan& P, uint32_t p0, uint32_t p1, bool fma,
                               bool idx64) {
  std::string src = "#include \"jit_device.hpp\"\nnamespace qoa {\nusing namespace tile;\n";
  for (uint32_t p = p0; p < p1; ++p) pass_source(P, p, fma, idx64, &src);
  src += "}  // namespace qoa\n";
  return src;
}

namespace {

// Directory of the on-disk code cache: QOA_SPEC_CACHE, else
// $HOME/.cache/qoa_amd/spec; "" when neither can be used.
std::string cache_dir() {
  const char* d = std::getenv("QOA_SPEC_CACHE");
  if (d) return d;
  const char* home = std::getenv("HOME");
  if (!home || !*home) return std::string();
  return std::string(home) + "/.cache/qoa_amd/spec";
}

bool read_file(const std::string& path, std::vector<char>* out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::vector<char> buf;
  char tmp[1 << 16];
  size_t n;
  while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  if (ok && !buf.empty()) *out = std::move(buf);
  return ok && !out->empty();
}

// Writes through a temporary name and renames, so a concurrent reader sees a
// whole file or none; failures only cost a later compile.
void write_file(const std::string& dir, const std::string& path, const std::vector<char>& code) {
  std::string p;
  for (size_t i = 1; i <= dir.size(); ++i)  // mkdir -p
    if (i == dir.size() || dir[i] == '/') {
      p = dir.substr(0, i);
      mkdir(p.c_str(), 0755);
    }
  const std::string tmp = path + ".tmp" + std::to_string((unsigned long)getpid());
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;
  const bool ok = std::fwrite(code.data(), 1, code.size(), f) == code.size();
  if (std::fclose(f) == 0 && ok) std::rename(tmp.c_str(), path.c_str());
  else std::remove(tmp.c_str());
}

// One program holds every pass of a plan (hipRTC compiles one program at a
// time in a process, and each program pays for parsing the device headers:
// the 28-qubit layered circuit's 35 passes compile in about 17 s as one
// program, 31 s as 16). Returns the program's cache entry.
int compile_plan(const FusedPlan& P, bool fma, bool idx64, const std::string& arch,
                 std::shared_ptr<ModuleEntry>* ent_out) {
  for (const PassPlan& pp : P.passes)
    if (!pp.reg_mode)
      return set_error(QOA_ERR_UNSUPPORTED, "specialised passes need register-mode plans");
  const std::string src = specialized_source(P, 0, (uint32_t)P.passes.size(), fma, idx64);
  const uint64_t key = fnv64(arch, fnv64(src, toolchain_key()));
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_modules.find(key);
    if (it != g_modules.end()) {
      *ent_out = it->second;
      return QOA_OK;
    }
  }
  std::vector<char> code;
  const std::string dir = cache_dir();
  char name[40];
  std::snprintf(name, sizeof(name), "/%016llx.co", (unsigned long long)key);
  const std::string path = dir.empty() ? std::string() : dir + name;
  if (path.empty() || !read_file(path, &code)) {
    const auto t0 = std::chrono::steady_clock::now();
    const std::string err = rtc_compile(src, arch, &code);
    if (!err.empty()) return set_error(QOA_ERR_HIP, "%s", err.c_str());
    std::lock_guard<std::mutex> lk(g_mu);
    g_stats.compiles += 1;
    g_stats.compile_seconds +=
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!path.empty()) write_file(dir, path, code);
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto& e = g_modules[key];
  if (!e) {
    e = std::make_shared<ModuleEntry>();
    e->code = std::move(code);
  }
  g_stats.programs = (uint32_t)g_modules.size();
  *ent_out = e;
  return QOA_OK;
}

}  // namespace

int specialize_compile_only(const FusedPlan& P, bool fma, bool idx64, const std::string& arch) {
  std::shared_ptr<ModuleEntry> ent;
  return P.passes.empty() ? QOA_OK : compile_plan(P, fma, idx64, arch, &ent);
}

int specialize_plan(const FusedPlan& P, bool fma, bool idx64, SpecKernels* out) {
  out->fn.clear();
  out->modules.clear();
  if (P.passes.empty()) return QOA_OK;
  int dev = 0;
  QOA_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  QOA_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  std::shared_ptr<ModuleEntry> ent;
  int rc = compile_plan(P, fma, idx64, prop.gcnArchName, &ent);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = ent->loaded.find(dev);
  if (it == ent->loaded.end()) {
    hipModule_t m;
    QOA_HIP_CHECK(hipModuleLoadData(&m, ent->code.data()));
    it = ent->loaded.emplace(dev, m).first;
  }
  for (uint32_t p = 0; p < (uint32_t)P.passes.size(); ++p) {
    char name[32];
    std::snprintf(name, sizeof(name), "qoa_spec_%u", p);
    hipFunction_t f;
    QOA_HIP_CHECK(hipModuleGetFunction(&f, it->second, name));
    out->fn.push_back(f);
  }
  out->modules.push_back(ent);
  return QOA_OK;
}

SpecStats specialize_stats() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_stats;
}

}  // namespace qoa

extern "C" int qoa_specialize_info(uint32_t* programs, uint32_t* compiles, double* compile_seconds) {
  const qoa::SpecStats s = qoa::specialize_stats();
  if (programs) *programs = s.programs;
  if (compiles) *compiles = s.compiles;
  if (compile_seconds) *compile_seconds = s.compile_seconds;
  return QOA_OK;
}

extern "C" int qoa_specialize_compile(uint32_t n_qubits, const qoa_gate* gates, uint32_t n_gates,
                                      const double* matrices, const qoa_fuse_opts* opts,
                                      const char* arch) {
  using namespace qoa;
  if (!gates && n_gates) return set_error(QOA_ERR_ARG, "gates is NULL");
  if (!arch || !*arch) return set_error(QOA_ERR_ARG, "arch is empty");
  const PlanOptions po{opts ? opts->tile_qubits : 0u, opts ? opts->low_qubits : 0u,
                       opts ? opts->flags : 0u};
  FusedPlan plan;
  int rc = build_plan(n_qubits, gates, n_gates, matrices, po, &plan);
  if (rc) return rc;
  const bool fma = (po.flags & (QOA_FUSE_FMA | QOA_FUSE_MERGE_DIAG)) != 0;
  return specialize_compile_only(plan, fma, n_qubits > 31 || g_force_index64, arch);
}
