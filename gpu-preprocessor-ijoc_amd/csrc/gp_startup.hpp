// gp_startup.hpp -- the structural Startup rules (gp_startup.hip).
#pragma once

#include <functional>

#include "gp_common.hpp"

namespace gp {

struct StartupOutcome {
  bool applicable = false;
  int32_t count = 0;        // candidates accepted by the analysis
  int32_t primal_only = 0;  // GP_PRIMAL_ONLY_* bit of a rule written into a plan
};

// Supplies the plan an accepted rule writes into (allocating it on first
// use); nullptr = analysis only.
using PlanSource = std::function<gp_plan*()>;

// One rule (GP_STARTUP_*): the analysis reads `lp`, or the plan's vectors and
// masks when `pending` is given; when it accepts, the records and deletions go
// into get_plan().
StartupOutcome startup_rule(int32_t rule, const gp_device_lp& lp, const gp_params& p, gp_plan* pending,
                            const PlanSource& get_plan);

}  // namespace gp
