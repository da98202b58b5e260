// specialize.hpp — circuit-specialised register passes (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "planner.hpp"

namespace qoa {

// One compiled kernel per pass of a plan (and the code objects they live in).
struct SpecKernels {
  std::vector<hipFunction_t> fn;
  std::vector<std::shared_ptr<void>> modules;
};

struct SpecStats {
  uint32_t programs = 0;         // code objects in the process-wide cache
  uint32_t compiles = 0;         // programs compiled so far
  double compile_seconds = 0.0;  // wall time spent compiling
};

// Source of the kernels for passes [p0, p1) of P (kernel qoa_spec_<p>).
std::string specialized_source(const FusedPlan& P, uint32_t p0, uint32_t p1, bool fma, bool idx64);
// Compiles (or finds in the cache) and loads the kernels of every pass of P
// on the current device. fma: the FMA instantiation (also merged diagonal
// runs); idx64: 64-bit index arithmetic.
int specialize_plan(const FusedPlan& P, bool fma, bool idx64, SpecKernels* out);
// Compiles P's kernels for `arch` without loading them (no device needed).
int specialize_compile_only(const FusedPlan& P, bool fma, bool idx64, const std::string& arch);
SpecStats specialize_stats();

}  // namespace qoa
