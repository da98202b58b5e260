// fused.hip — LDS-tiled multi-gate passes (placeholder until the planner lands).
#include "qoa_internal.hpp"

namespace qoa {

int plan_order(uint32_t, const qoa_gate*, uint32_t, const PlanOptions&, std::vector<uint32_t>*,
               uint32_t*) {
  return set_error(QOA_ERR_UNSUPPORTED, "fused execution is not built yet");
}

int run_fused(qoa_state*, const qoa_gate*, uint32_t, const double*, const PlanOptions&) {
  return set_error(QOA_ERR_UNSUPPORTED, "fused execution is not built yet");
}

}  // namespace qoa
