/*
 * pslp_driver.c -- CPU baseline only.  Runs the reference's bundled CPU
 * presolver (PSLP 0.0.11, third_party/pslp) on one CSR LP with its default
 * settings and reports the time the reference reports
 * (solver_interface/adapters/pslp.cpp:232-275: wall time of new_presolver +
 * run_presolver, and stats->time_presolve).  Linked with the PSLP sources by
 * ref.mk into _ref/libpslp.so; never part of the engine.
 */
#include <stddef.h>
#include <time.h>

#include "PSLP_API.h"
#include "PSLP_stats.h"

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* out[0] wall seconds (new_presolver + run_presolver), out[1] time_presolve,
 * out[2..4] reduced rows, cols, nnz; returns the PresolveStatus or -1. */
int pslp_presolve_timed(const double* Ax, const int* Ai, const int* Ap, size_t m, size_t n, size_t nnz,
                        const double* lhs, const double* rhs, const double* lbs, const double* ubs,
                        const double* c, double* out) {
  Settings* stgs = default_settings();
  if (!stgs) return -1;
  const double t0 = now_s();
  Presolver* pre = new_presolver(Ax, Ai, Ap, m, n, nnz, lhs, rhs, lbs, ubs, c, stgs);
  if (!pre) {
    free_settings(stgs);
    return -1;
  }
  const PresolveStatus status = run_presolver(pre);
  out[0] = now_s() - t0;
  out[1] = pre->stats ? pre->stats->time_presolve : -1.0;
  out[2] = pre->reduced_prob ? (double)pre->reduced_prob->m : -1.0;
  out[3] = pre->reduced_prob ? (double)pre->reduced_prob->n : -1.0;
  out[4] = pre->reduced_prob ? (double)pre->reduced_prob->nnz : -1.0;
  free_presolver(pre);
  free_settings(stgs);
  return (int)status;
}
