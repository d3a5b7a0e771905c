// gp_postsolve.hip -- postsolve: scatter the reduced solution to original
// positions, restore fixed columns, replay the tape in reverse (primal pass,
// then dual pass) and project column duals from the original model
// (gpu_postsolve.cu:535-670; record handlers postsolve/postsolve_<name>.cuh).
#include "gp_tape.hpp"

namespace gp {
namespace {

struct PsState {
  double* x;
  double* y;
  double* z;
  uint8_t* retrieved;
  gp_device_lp model;  // original model (valid when has_model)
  int has_model;
  double tol;
};

struct PsTape {
  const int32_t* types;
  const gp_offset* is;
  const gp_offset* vs;
  const int32_t* idx;
  const double* vals;
};

__device__ inline bool ps_finite(double x) { return x > -INFINITY && x < INFINITY; }
__device__ inline bool at_lower(double x, double l, double tol) { return ps_finite(l) && x <= l + tol; }
__device__ inline bool at_upper(double x, double u, double tol) { return ps_finite(u) && x >= u - tol; }
__device__ inline double clamp_inf(double x, double lo, double hi) {
  double y = x;
  if (ps_finite(lo)) y = cmax(y, lo);
  if (ps_finite(hi)) y = cmin(y, hi);
  return y;
}

__device__ void replay_fixed(PsState& s, const int32_t* ix, int64_t i0, int64_t i1, const double* v, int64_t v0,
                             bool primal, bool dual) {
  const int32_t col = ix[i0];
  if (primal) s.x[col] = v[v0];
  if (dual) {
    double zj = v[v0 + 1];
    for (int64_t p = i0 + 1; p <= i1; ++p) zj -= v[v0 + 2 + (p - (i0 + 1))] * s.y[ix[p]];
    s.z[col] = zj;
    s.retrieved[col] = 1;
  }
}

__device__ void replay_infinite(PsState& s, const int32_t* ix, int64_t i0, const double* v, int64_t v0, bool primal,
                                bool dual) {
  const bool pos = ix[i0] > 0;
  const int32_t col = ix[i0 + 1];
  const int32_t rows = ix[i0 + 2];
  double extreme = v[v0];
  int64_t ip = i0 + 3, vp = v0 + 1;
  if (primal) {
    for (int32_t r = 0; r < rows; ++r) {
      const int64_t len = ix[ip];
      double side = v[vp], coeff = 0.0;
      for (int64_t t = 0; t < len; ++t) {
        const int32_t c = ix[ip + 1 + t];
        const double a = v[vp + 1 + t];
        if (c == col) coeff = a;
        else side -= a * s.x[c];
      }
      const double cand = side / coeff;
      extreme = pos ? cmax(extreme, cand) : cmin(extreme, cand);
      ip += len + 1;
      vp += len + 1;
    }
    s.x[col] = extreme;
  }
  if (dual) {
    s.z[col] = 0.0;
    s.retrieved[col] = 1;
  }
}

__device__ void replay_singleton_column(PsState& s, const int32_t* ix, int64_t i0, const double* v, int64_t v0,
                                        bool primal, bool dual) {
  const int32_t ec = ix[i0], row = ix[i0 + 1], support = ix[i0 + 2];
  const double pivot = v[v0], rhs = v[v0 + 1], cost = v[v0 + 4];
  const bool deleted = v[v0 + 5] != 0.0;
  double partial = 0.0;
  for (int32_t t = 0; t < support; ++t) partial += v[v0 + 6 + t] * s.x[ix[i0 + 3 + t]];
  if (primal) s.x[ec] = (rhs - partial) / pivot;
  if (dual && row >= 0) {
    if (deleted) s.y[row] = cost / pivot;
    else s.y[row] += cost / pivot;
  }
  if (dual) {
    s.z[ec] = deleted || row < 0 ? 0.0 : cost - pivot * s.y[row];
    s.retrieved[ec] = 1;
  }
}

__device__ double merged_source_value(double lo, double hi, double tol) {
  if (isfinite(lo) && isfinite(hi)) return 0.5 * (lo + hi);
  if (isfinite(lo)) {
    if (0.0 > lo + tol) return 0.0;
    return lo + cmax(1.0, fabs(lo) * cmax(tol, 1.0e-12));
  }
  if (isfinite(hi)) {
    if (0.0 < hi - tol) return 0.0;
    return hi - cmax(1.0, fabs(hi) * cmax(tol, 1.0e-12));
  }
  return 0.0;
}

__device__ void replay_parallel_column(PsState& s, const int32_t* ix, int64_t i0, const double* v, int64_t v0,
                                       bool primal, bool dual) {
  const int32_t from = ix[i0], to = ix[i0 + 1];
  const double ratio = v[v0], lo0 = v[v0 + 1], hi0 = v[v0 + 2], tl = v[v0 + 3], tu = v[v0 + 4];
  const double tol = s.tol;
  if (primal) {
    const double yv = s.x[to];
    const double zto = s.z[to];
    double lo = lo0, hi = hi0;
    if (isfinite(tu)) {
      const double b = (yv - tu) / ratio;
      if (ratio > 0.0) lo = cmax(lo, b);
      else hi = cmin(hi, b);
    }
    if (isfinite(tl)) {
      const double b = (yv - tl) / ratio;
      if (ratio > 0.0) hi = cmin(hi, b);
      else lo = cmax(lo, b);
    }
    double lower = -INFINITY, upper = INFINITY;
    if (ratio > 0.0) {
      if (isfinite(tl) && isfinite(lo0)) lower = tl + ratio * lo0;
      if (isfinite(tu) && isfinite(hi0)) upper = tu + ratio * hi0;
    } else {
      if (isfinite(tl) && isfinite(hi0)) lower = tl + ratio * hi0;
      if (isfinite(tu) && isfinite(lo0)) upper = tu + ratio * lo0;
    }
    double xf;
    if (zto > tol && isfinite(lower) && fabs(yv - lower) <= tol) xf = ratio > 0.0 ? lo0 : hi0;
    else if (zto < -tol && isfinite(upper) && fabs(yv - upper) <= tol) xf = ratio > 0.0 ? hi0 : lo0;
    else xf = merged_source_value(lo, hi, tol);
    s.x[from] = xf;
    s.x[to] = yv - ratio * xf;
  }
  if (dual) {
    s.z[from] = ratio * s.z[to];
    s.retrieved[from] = 1;
  }
}

__device__ void replay_doubleton(PsState& s, const int32_t* ix, int64_t i0, int64_t i1, const double* v, int64_t v0,
                                 int64_t v1, bool primal, bool dual) {
  const int32_t ec = ix[i0], kc = ix[i0 + 1], row = ix[i0 + 2];
  const double ecoef = v[v0], kcoef = v[v0 + 1], rhs = v[v0 + 2], eobj = v[v0 + 9];
  const double alpha = -kcoef / ecoef, beta = rhs / ecoef;
  const double tol = s.tol;
  if (primal) s.x[ec] = beta + alpha * s.x[kc];
  if (dual && s.has_model) {
    const double okl = v[v0 + 5], oku = v[v0 + 6], nkl = v[v0 + 7], nku = v[v0 + 8];
    const double lsrc = v1 >= v0 + 10 ? v[v0 + 10] : 0.0;
    const double usrc = v1 >= v0 + 11 ? v[v0 + 11] : 0.0;
    double zk = 0.0, ze = 0.0;
    const double xk = s.x[kc], zh = s.z[kc];
    if (fabs(alpha) <= tol) {
      zk = zh;
    } else {
      const bool al = ps_finite(nkl) && xk <= nkl + tol;
      const bool au = ps_finite(nku) && xk >= nku - tol;
      if (al || au) {
        double src = al ? lsrc : usrc;
        if (src == 0.0) {
          if (al) src = (ps_finite(okl) && fabs(nkl - okl) <= tol) ? 1.0 : 2.0;
          else src = (ps_finite(oku) && fabs(nku - oku) <= tol) ? 1.0 : 2.0;
        }
        if (src == 2.0) {
          zk = 0.0;
          ze = zh / alpha;
        } else {
          zk = zh;
          ze = 0.0;
        }
      }
    }
    double sum = 0.0;
    if (i1 >= i0 + 3 && v1 >= v0 + 11) {
      const int32_t sc = ix[i0 + 3];
      for (int32_t t = 0; t < sc; ++t) sum += v[v0 + 12 + t] * s.y[ix[i0 + 4 + t]];
      s.y[row] = (eobj - sum - ze) / ecoef;
    } else {
      const gp_csr& AT = s.model.AT;
      for (gp_offset p = AT.row_ptr[ec]; p < AT.row_ptr[ec + 1]; ++p)
        if (AT.col_idx[p] != row) sum += AT.vals[p] * s.y[AT.col_idx[p]];
      s.y[row] = (s.model.c[ec] - sum - ze) / ecoef;
    }
    s.z[kc] = zk;
    s.z[ec] = ze;
    s.retrieved[kc] = 1;
    s.retrieved[ec] = 1;
  } else if (dual) {
    s.z[ec] = 0.0;
    s.retrieved[ec] = 1;
    if (row >= 0) s.y[row] = eobj / ecoef;
  }
}

__device__ int row_state(double act, double lo, double hi, double tol) {
  const bool al = ps_finite(lo) && act <= lo + tol;
  const bool au = ps_finite(hi) && act >= hi - tol;
  if (al && au) return 3;
  if (al) return 1;
  if (au) return 2;
  return 0;
}

__device__ double row_activity(const gp_csr& A, int32_t row, const double* x) {
  double a = 0.0;
  for (gp_offset p = A.row_ptr[row]; p < A.row_ptr[row + 1]; ++p) a += A.vals[p] * x[A.col_idx[p]];
  return a;
}

__device__ void replay_parallel_row(PsState& s, const int32_t* ix, int64_t i0, const double* v, int64_t v0) {
  const int32_t kept = ix[i0], del = ix[i0 + 1];
  const double ratio = v[v0], kal = v[v0 + 1], kau = v[v0 + 2], dal = v[v0 + 3], dau = v[v0 + 4];
  const double tol = s.tol;
  if (fabs(ratio) <= tol) return;
  const double rl = dal * ratio, ru = dau * ratio;
  const double sal = ratio >= 0.0 ? rl : ru, sau = ratio >= 0.0 ? ru : rl;
  if (sal <= kal + tol && kau <= sau + tol) {
    s.y[del] = 0.0;
    return;
  }
  if (!s.has_model) return;
  const double ap = row_activity(s.model.A, kept, s.x);
  const double ad = row_activity(s.model.A, del, s.x);
  const int kst = row_state(ap, kal, kau, tol);
  const int dst = row_state(ad, dal, dau, tol);
  double ydl, ydh, ypl, yph;
  if (dst == 3) { ydl = -INFINITY; ydh = INFINITY; }
  else if (dst == 1) { ydl = 0.0; ydh = INFINITY; }
  else if (dst == 2) { ydl = -INFINITY; ydh = 0.0; }
  else { ydl = 0.0; ydh = 0.0; }
  const double pivot = ratio * s.y[kept];
  if (kst == 3) { ypl = -INFINITY; yph = INFINITY; }
  else if (kst == 0) { ypl = pivot; yph = pivot; }
  else if (kst == 1) {
    if (ratio > 0.0) { ypl = -INFINITY; yph = pivot; }
    else { ypl = pivot; yph = INFINITY; }
  } else {
    if (ratio > 0.0) { ypl = pivot; yph = INFINITY; }
    else { ypl = -INFINITY; yph = pivot; }
  }
  const double sl = cmax(ydl, ypl), sh = cmin(ydh, yph);
  if (sl <= sh + tol) {
    const double ydn = clamp_inf(0.0, sl, sh);
    s.y[kept] = s.y[kept] - ydn / ratio;
    s.y[del] = ydn;
    return;
  }
  const double yb = s.y[kept];
  double yp = yb, yd = 0.0;
  bool ok = true;
  if (fabs(yb) <= tol) {
    yp = 0.0;
    yd = 0.0;
  } else {
    const double bl = cmax(kal, sal), bu = cmin(kau, sau);
    if (!(bl <= bu + tol)) {
      ok = false;
    } else {
      const bool lt = bl > kal + tol, ut = bu < kau - tol;
      if (lt && ut) ok = false;
      else if (!lt && !ut) { yp = yb; yd = 0.0; }
      else if (yb > 0.0) {
        if (lt) { yp = 0.0; yd = ratio * yb; }
        else { yp = yb; yd = 0.0; }
      } else {
        if (lt) { yp = yb; yd = 0.0; }
        else { yp = 0.0; yd = ratio * yb; }
      }
    }
  }
  if (ok) {
    s.y[kept] = yp;
    s.y[del] = yd;
  }
}

__device__ void replay_singleton_row(PsState& s, const int32_t* ix, int64_t i0, const double* v, int64_t v0) {
  const int32_t row = ix[i0], col = ix[i0 + 1];
  const double oal = v[v0], oau = v[v0 + 1], a = v[v0 + 2];
  const double tol = s.tol;
  if (fabs(a) <= tol) return;
  const gp_device_lp& M = s.model;
  double g = M.c[col];
  for (gp_offset p = M.AT.row_ptr[col]; p < M.AT.row_ptr[col + 1]; ++p)
    if (M.AT.col_idx[p] != row) g -= M.AT.vals[p] * s.y[M.AT.col_idx[p]];
  const double xj = s.x[col], zb = s.z[col];
  const double act = a * xj;
  const bool rl = isfinite(oal) && act <= oal + tol;
  const bool ru = isfinite(oau) && act >= oau - tol;
  double izl = -INFINITY, izh = INFINITY;
  if (!(rl && ru)) {
    if (rl) {
      if (a > 0.0) izh = g;
      else izl = g;
    } else if (ru) {
      if (a > 0.0) izl = g;
      else izh = g;
    } else {
      izl = g;
      izh = g;
    }
  }
  const bool al = at_lower(xj, M.l[col], tol), au = at_upper(xj, M.u[col], tol);
  double zl = izl, zh = izh;
  if (!al && !au) {
    if ((isfinite(izl) && izl > tol) || (isfinite(izh) && izh < -tol)) {
      zl = g;
      zh = g;
    } else {
      zl = 0.0;
      zh = 0.0;
    }
  } else if (al && !au) {
    zl = cmax(zl, 0.0);
  } else if (!al && au) {
    zh = cmin(zh, 0.0);
  }
  if (isfinite(zl) && isfinite(zh) && zl > zh + tol) return;
  const double zn = clamp_inf(zb, zl, zh);
  s.z[col] = zn;
  s.y[row] = (g - zn) / a;
  s.retrieved[col] = 1;
}

__device__ void replay_row_implied(PsState& s, const int32_t* ix, int64_t i0, const double* v, int64_t v0) {
  const int32_t col = ix[i0], row = ix[i0 + 1];
  const double tol = s.tol;
  if (!s.retrieved[col]) return;
  const double ol = v[v0], ou = v[v0 + 1], nl = v[v0 + 2], nu = v[v0 + 3];
  const bool lch = nl > ol + tol, uch = nu < ou - tol;
  if (lch == uch) return;
  const double implied = lch ? nl : nu;
  const double other = lch ? ou : ol;
  const bool other_lower = !lch;
  const double xj = s.x[col], zj = s.z[col];
  if (ps_finite(other) && fabs(xj - other) <= tol) {
    const bool ok = other_lower ? (zj >= -tol) : (zj <= tol);
    if (ok) return;
  }
  if (fabs(xj - implied) > tol) return;
  const gp_csr& A = s.model.A;
  double aij = 0.0;
  for (gp_offset p = A.row_ptr[row]; p < A.row_ptr[row + 1]; ++p)
    if (A.col_idx[p] == col) {
      aij = A.vals[p];
      break;
    }
  if (fabs(aij) <= tol) return;
  s.y[row] += zj / aij;
  for (gp_offset p = A.row_ptr[row]; p < A.row_ptr[row + 1]; ++p) {
    const int32_t k = A.col_idx[p];
    if (k == col || !s.retrieved[k]) continue;
    s.z[k] -= (A.vals[p] / aij) * zj;
  }
  s.z[col] = 0.0;
  s.retrieved[col] = 1;
}

__device__ void replay_record(PsState& s, const PsTape& t, int64_t k, bool primal, bool dual) {
  const int32_t type = t.types[k];
  const int64_t i0 = t.is[k], i1 = t.is[k + 1] - 1;
  const int64_t v0 = t.vs[k], v1 = t.vs[k + 1] - 1;
  switch (type) {
    case GP_RED_PARALLEL_COLUMN: replay_parallel_column(s, t.idx, i0, t.vals, v0, primal, dual); break;
    case GP_RED_DOUBLETON_EQUATION: replay_doubleton(s, t.idx, i0, i1, t.vals, v0, v1, primal, dual); break;
    case GP_RED_SINGLETON_COLUMN_SUBSTITUTION: replay_singleton_column(s, t.idx, i0, t.vals, v0, primal, dual); break;
    case GP_RED_FIXED_VARIABLE: replay_fixed(s, t.idx, i0, i1, t.vals, v0, primal, dual); break;
    case GP_RED_INFINITE_BOUND_DUAL_FIX: replay_infinite(s, t.idx, i0, t.vals, v0, primal, dual); break;
    case GP_RED_ROW_DUAL_SHIFT:
      if (dual) s.y[t.idx[i0]] += t.vals[v0];
      break;
    case GP_RED_PARALLEL_ROW:
      if (dual && i1 >= i0 + 1 && v1 >= v0 + 4) replay_parallel_row(s, t.idx, i0, t.vals, v0);
      break;
    case GP_RED_DELETED_ROW:
      if (dual && s.has_model && i1 >= i0 + 1 && v1 >= v0 + 2) replay_singleton_row(s, t.idx, i0, t.vals, v0);
      break;
    case GP_RED_ROW_IMPLIED_BOUND_CHANGE:
      if (dual && s.has_model && i1 >= i0 + 1) replay_row_implied(s, t.idx, i0, t.vals, v0);
      break;
    default: break;
  }
}

// Reverse replay of the whole tape, one pass (serial, one lane).
__global__ void k_replay_serial(PsState s, PsTape t, int64_t count, int primal, int dual) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int64_t k = count - 1; k >= 0; --k) replay_record(s, t, k, primal != 0, dual != 0);
}

// ---- dependency-ordered parallel replay ----------------------------------------
// Within one pass (primal or dual) a record reads and writes state entries:
// x / z / retrieved of a column, y of a row.  Listing, per entry, its
// accesses in replay order (reverse tape order) gives the only orderings the
// serial replay's result depends on: a write must follow every earlier access
// of the entry, a read must follow the entry's previous write.  The replay
// runs in rounds: a round replays, in parallel, every pending record whose
// accesses are all unblocked, then counts them as done.  Records replayed in
// one round touch no common entry that either of them writes, so the result
// is the serial replay's bit for bit, and the number of rounds is the depth
// of the dependency graph (a handful on presolve tapes), not the size of its
// connected pieces.  The reference instead joins conflicting records into
// components and replays each component serially (parallel_replay.cuh).
// A RowImpliedBoundChange that writes a whole row of z is pruned before the
// dual pass when its own test on the final x and its tape values already
// makes it a no-op.

template <class F>
__device__ void record_accesses(const PsState& s, const PsTape& t, int64_t k, bool primal, int32_t n0, F& f) {
  const int32_t* ix = t.idx;
  const int64_t i0 = t.is[k], i1 = t.is[k + 1] - 1;
  const int64_t v0 = t.vs[k], v1 = t.vs[k + 1] - 1;
  switch (t.types[k]) {
    case GP_RED_FIXED_VARIABLE:
      f(ix[i0], true);
      if (!primal)
        for (int64_t p = i0 + 1; p <= i1; ++p) f(n0 + ix[p], false);
      break;
    case GP_RED_INFINITE_BOUND_DUAL_FIX: {
      const int32_t col = ix[i0 + 1];
      f(col, true);
      if (primal) {
        const int32_t rows = ix[i0 + 2];
        int64_t ip = i0 + 3;
        for (int32_t r = 0; r < rows; ++r) {
          const int64_t len = ix[ip];
          for (int64_t q = 0; q < len; ++q)
            if (ix[ip + 1 + q] != col) f(ix[ip + 1 + q], false);
          ip += len + 1;
        }
      }
      break;
    }
    case GP_RED_SINGLETON_COLUMN_SUBSTITUTION: {
      f(ix[i0], true);
      const int32_t row = ix[i0 + 1], support = ix[i0 + 2];
      if (primal) {
        for (int32_t q = 0; q < support; ++q) f(ix[i0 + 3 + q], false);
      } else if (row >= 0) {
        f(n0 + row, true);
      }
      break;
    }
    case GP_RED_PARALLEL_COLUMN:
      f(ix[i0], true);
      f(ix[i0 + 1], primal);
      break;
    case GP_RED_DOUBLETON_EQUATION: {
      const int32_t ec = ix[i0], kc = ix[i0 + 1], row = ix[i0 + 2];
      f(ec, true);
      if (primal) {
        f(kc, false);
        break;
      }
      if (row >= 0) f(n0 + row, true);
      if (!s.has_model) break;
      f(kc, true);
      if (i1 >= i0 + 3 && v1 >= v0 + 11) {
        const int32_t sc = ix[i0 + 3];
        for (int32_t q = 0; q < sc; ++q) f(n0 + ix[i0 + 4 + q], false);
      } else {
        const gp_csr& AT = s.model.AT;
        for (gp_offset p = AT.row_ptr[ec]; p < AT.row_ptr[ec + 1]; ++p) f(n0 + AT.col_idx[p], false);
      }
      break;
    }
    case GP_RED_ROW_DUAL_SHIFT:
      if (!primal) f(n0 + ix[i0], true);
      break;
    case GP_RED_PARALLEL_ROW:
      if (!primal && i1 >= i0 + 1 && v1 >= v0 + 4) {
        f(n0 + ix[i0], true);
        f(n0 + ix[i0 + 1], true);
      }
      break;
    case GP_RED_DELETED_ROW:
      if (!primal && s.has_model && i1 >= i0 + 1 && v1 >= v0 + 2) {
        const int32_t row = ix[i0], col = ix[i0 + 1];
        f(col, true);
        f(n0 + row, true);
        const gp_csr& AT = s.model.AT;
        for (gp_offset p = AT.row_ptr[col]; p < AT.row_ptr[col + 1]; ++p) f(n0 + AT.col_idx[p], false);
      }
      break;
    case GP_RED_ROW_IMPLIED_BOUND_CHANGE:
      if (!primal && s.has_model && i1 >= i0 + 1) {
        const int32_t col = ix[i0], row = ix[i0 + 1];
        // the record's early exits that depend on tape values and final x only
        const double ol = t.vals[v0], ou = t.vals[v0 + 1], nl = t.vals[v0 + 2], nu = t.vals[v0 + 3];
        const bool lch = nl > ol + s.tol, uch = nu < ou - s.tol;
        if (lch == uch) break;
        if (fabs(s.x[col] - (lch ? nl : nu)) > s.tol) break;
        f(col, true);
        f(n0 + row, true);
        const gp_csr& A = s.model.A;
        for (gp_offset p = A.row_ptr[row]; p < A.row_ptr[row + 1]; ++p) f(A.col_idx[p], true);
      }
      break;
    default: break;
  }
}

// Access enumeration into flat arrays.
struct CountVisit {
  int32_t limit;
  int32_t n = 0;
  __device__ void operator()(int32_t e, bool) {
    if (e >= 0 && e < limit) ++n;
  }
};

struct FillVisit {
  uint64_t* keys;
  int32_t* ids;
  uint8_t* writes;
  int64_t a;
  uint32_t rank;
  int32_t limit;
  __device__ void operator()(int32_t e, bool write) {
    if (e < 0 || e >= limit) return;
    keys[a] = (static_cast<uint64_t>(e) << 32) | rank;
    ids[a] = static_cast<int32_t>(a);
    writes[a] = write ? 1 : 0;
    ++a;
  }
};

__global__ void k_acc_count(PsState s, PsTape t, int32_t count, int primal, int32_t n0, int32_t entities,
                            int64_t* acc_start) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  CountVisit f{entities};
  record_accesses(s, t, k, primal != 0, n0, f);
  acc_start[k] = f.n;
}

// Sort key of an access: (entry << 32 | replay rank), replay rank = count-1-k.
__global__ void k_acc_fill(PsState s, PsTape t, int32_t count, int primal, int32_t n0, int32_t entities,
                           const int64_t* acc_start, uint64_t* keys, int32_t* ids, uint8_t* writes) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  FillVisit f{keys, ids, writes, acc_start[k], static_cast<uint32_t>(count - 1 - k), entities};
  record_accesses(s, t, k, primal != 0, n0, f);
}

// Per sorted access: first of its (entry, record) run?  A record that lists
// one entry twice counts once, as a write if any of the copies writes.
__global__ void k_acc_runs(const uint64_t* keys, const int32_t* ids, const uint8_t* writes, int32_t total,
                           int32_t* first, int32_t* first_write) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const bool head = q == 0 || keys[q - 1] != keys[q];
  int32_t w = 0;
  if (head)
    for (int64_t r = q; r < total && keys[r] == keys[q]; ++r) w |= writes[ids[r]];
  first[q] = head ? 1 : 0;
  first_write[q] = head ? w : 0;
}

// Entry-relative positions: ent_base[e] = scan values at the entry's first access.
__global__ void k_acc_entry_base(const uint64_t* keys, const int32_t* pos, const int32_t* wpos, int32_t total,
                                 int32_t* ent_pos, int32_t* ent_wpos) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const uint32_t e = static_cast<uint32_t>(keys[q] >> 32);
  if (q == 0 || static_cast<uint32_t>(keys[q - 1] >> 32) != e) {
    ent_pos[e] = pos[q];
    ent_wpos[e] = wpos[q];
  }
}

// Per access (in record order): position among the entry's accesses, number
// of writes before it, write flag; position -1 marks a repeated access.
struct AccInfo {
  int32_t entry;
  int32_t position;
  int32_t writes_before;
  int32_t write;
};

__global__ void k_acc_info(const uint64_t* keys, const int32_t* ids, const int32_t* first, const int32_t* first_write,
                           const int32_t* pos, const int32_t* wpos, const int32_t* ent_pos, const int32_t* ent_wpos,
                           int32_t total, AccInfo* info) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int32_t e = static_cast<int32_t>(keys[q] >> 32);
  AccInfo a{e, -1, 0, 0};
  if (first[q]) {
    a.position = pos[q] - ent_pos[e];
    a.writes_before = wpos[q] - ent_wpos[e];
    a.write = first_write[q];
  }
  info[ids[q]] = a;
}

// state[k]: 0 pending, -1 no accesses in this pass, r > 0 replayed in round r.
__global__ void k_acc_state(const int64_t* acc_start, int32_t count, int32_t* state) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) state[k] = acc_start[k + 1] > acc_start[k] ? 0 : -1;
}

__global__ void k_round_replay(PsState s, PsTape t, int32_t count, int primal, int dual, const int64_t* acc_start,
                               const AccInfo* info, const int32_t* done, const int32_t* writes_done, int32_t* state,
                               int32_t round) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count || state[k] != 0) return;
  for (int64_t a = acc_start[k]; a < acc_start[k + 1]; ++a) {
    const AccInfo f = info[a];
    if (f.position < 0) continue;
    if (f.write ? done[f.entry] != f.position : writes_done[f.entry] < f.writes_before) return;
  }
  replay_record(s, t, k, primal != 0, dual != 0);
  state[k] = round;
}

__global__ void __launch_bounds__(kBlock) k_round_commit(int32_t count, const int64_t* acc_start,
                                                          const AccInfo* info, int32_t* done, int32_t* writes_done,
                                                          const int32_t* state, int32_t round,
                                                          unsigned long long* pending) {
  __shared__ unsigned long long lds[kBlock / 64];
  unsigned long long left = 0;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < count;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t st = state[k];
    if (st == 0) ++left;
    if (st != round) continue;
    for (int64_t a = acc_start[k]; a < acc_start[k + 1]; ++a) {
      const AccInfo f = info[a];
      if (f.position < 0) continue;
      atomicAdd(done + f.entry, 1);
      if (f.write) atomicAdd(writes_done + f.entry, 1);
    }
  }
  for (int o = 32; o > 0; o >>= 1) left += __shfl_xor(left, o, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = left;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tsum = 0;
    for (int w = 0; w < kBlock / 64; ++w) tsum += lds[w];
    if (tsum) atomicAdd(pending, tsum);
  }
}

// Finishes a pass whose dependency chain is long and thin: the pending
// records, serially in replay order, after every record already replayed.
__global__ void k_replay_pending(PsState s, PsTape t, int32_t count, int primal, int dual, const int32_t* state) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int64_t k = count - 1; k >= 0; --k)
    if (state[k] == 0) replay_record(s, t, k, primal != 0, dual != 0);
}

// One pass of the dependency-ordered replay; returns false (nothing replayed)
// when the access lists exceed the sort's int32 item limit.
bool replay_rounds(const PsState& s, const PsTape& t, int32_t count, int32_t m0, int32_t n0, bool primal) {
  const int32_t entities = n0 + m0;
  DVec<int64_t> acc_start(static_cast<size_t>(count) + 1);
  hipLaunchKernelGGL(k_acc_count, dim3(grid_for(count)), dim3(kBlock), 0, 0, s, t, count, primal ? 1 : 0, n0,
                     entities, acc_start.data());
  GP_LAUNCH_CHECK("postsolve access count");
  int64_t total64 = 0;
  exclusive_scan_i64(acc_start.data(), count, &total64);
  h2d(acc_start.data() + count, &total64, 1);
  if (total64 >= INT32_MAX) return false;
  const int32_t total = static_cast<int32_t>(total64);
  DVec<int32_t> state(static_cast<size_t>(count));
  hipLaunchKernelGGL(k_acc_state, dim3(grid_for(count)), dim3(kBlock), 0, 0, acc_start.data(), count, state.data());
  GP_LAUNCH_CHECK("postsolve access state");
  if (total == 0) return true;
  DVec<AccInfo> info(static_cast<size_t>(total));
  {
    DVec<uint64_t> keys(static_cast<size_t>(total));
    DVec<int32_t> ids(static_cast<size_t>(total));
    DVec<uint8_t> writes(static_cast<size_t>(total));
    hipLaunchKernelGGL(k_acc_fill, dim3(grid_for(count)), dim3(kBlock), 0, 0, s, t, count, primal ? 1 : 0, n0,
                       entities, acc_start.data(), keys.data(), ids.data(), writes.data());
    GP_LAUNCH_CHECK("postsolve access fill");
    sort_pairs_u64_i32(keys.data(), ids.data(), total, 64);
    DVec<int32_t> pos(static_cast<size_t>(total)), wpos(static_cast<size_t>(total));
    DVec<int32_t> first_write(static_cast<size_t>(total));
    hipLaunchKernelGGL(k_acc_runs, dim3(grid_for(total)), dim3(kBlock), 0, 0, keys.data(), ids.data(), writes.data(),
                       total, pos.data(), first_write.data());
    GP_LAUNCH_CHECK("postsolve access runs");
    DVec<int32_t> first(static_cast<size_t>(total));
    d2d(first.data(), pos.data(), static_cast<size_t>(total));
    d2d(wpos.data(), first_write.data(), static_cast<size_t>(total));
    exclusive_scan_i32_total(pos.data(), total);
    exclusive_scan_i32_total(wpos.data(), total);
    DVec<int32_t> ent_pos(static_cast<size_t>(entities)), ent_wpos(static_cast<size_t>(entities));
    hipLaunchKernelGGL(k_acc_entry_base, dim3(grid_for(total)), dim3(kBlock), 0, 0, keys.data(), pos.data(),
                       wpos.data(), total, ent_pos.data(), ent_wpos.data());
    hipLaunchKernelGGL(k_acc_info, dim3(grid_for(total)), dim3(kBlock), 0, 0, keys.data(), ids.data(), first.data(),
                       first_write.data(), pos.data(), wpos.data(), ent_pos.data(), ent_wpos.data(), total,
                       info.data());
    GP_LAUNCH_CHECK("postsolve access info");
  }
  DVec<int32_t> done(static_cast<size_t>(entities)), writes_done(static_cast<size_t>(entities));
  DVec<unsigned long long> pending(1);
  GP_CHECK(hipMemsetAsync(done.data(), 0, sizeof(int32_t) * static_cast<size_t>(entities), 0));
  GP_CHECK(hipMemsetAsync(writes_done.data(), 0, sizeof(int32_t) * static_cast<size_t>(entities), 0));
  unsigned long long before = ~0ull;
  for (int32_t round = 1;; ++round) {
    GP_CHECK(hipMemsetAsync(pending.data(), 0, sizeof(unsigned long long), 0));
    hipLaunchKernelGGL(k_round_replay, dim3(grid_for(count)), dim3(kBlock), 0, 0, s, t, count, primal ? 1 : 0,
                       primal ? 0 : 1, acc_start.data(), info.data(), done.data(), writes_done.data(), state.data(),
                       round);
    hipLaunchKernelGGL(k_round_commit, dim3(grid_capped(count, kBlock, 2048)), dim3(kBlock), 0, 0, count,
                       acc_start.data(), info.data(), done.data(), writes_done.data(), state.data(), round,
                       pending.data());
    GP_LAUNCH_CHECK("postsolve replay round");
    const unsigned long long left = read1(pending.data());
    if (left == 0) break;
    // every round replays at least the earliest pending record in replay order
    if (left >= before) throw Error(GP_ERR_RUNTIME, "postsolve replay made no progress");
    if (round >= 16 && before - left < 64) {
      hipLaunchKernelGGL(k_replay_pending, dim3(1), dim3(64), 0, 0, s, t, count, primal ? 1 : 0, primal ? 0 : 1,
                         state.data());
      GP_LAUNCH_CHECK("postsolve replay tail");
      break;
    }
    before = left;
  }
  return true;
}

// Tapes shorter than this replay serially (the component analysis costs a
// few launches and a sort); GP_POSTSOLVE_PARALLEL_MIN overrides it.
int64_t parallel_min_records() {
  const char* e = getenv("GP_POSTSOLVE_PARALLEL_MIN");
  if (e && *e) return strtoll(e, nullptr, 10);
  return 1024;
}

__global__ void k_scatter(double* dst, const double* src, const int32_t* map, int32_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[map[i]] = src[i];
}
__global__ void k_mark(uint8_t* mask, const int32_t* map, int32_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) mask[map[i]] = 1;
}
__global__ void k_restore_fixed(double* x, const int32_t* idx, const double* val, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[idx[i]] = val[i];
}
__global__ void k_protect(const int32_t* types, const gp_offset* is, const int32_t* idx, int64_t count, int32_t n,
                          uint8_t* mask) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int32_t t = types[k];
  const gp_offset p = is[k];
  int32_t a = -1, b = -1;
  if (t == GP_RED_SINGLETON_COLUMN_SUBSTITUTION || t == GP_RED_PARALLEL_COLUMN) a = idx[p];
  if (t == GP_RED_DOUBLETON_EQUATION) {
    a = idx[p];
    b = idx[p + 1];
  }
  if (t == GP_RED_INFINITE_BOUND_DUAL_FIX) a = idx[p + 1];
  if (a >= 0 && a < n) mask[a] = 1;
  if (b >= 0 && b < n) mask[b] = 1;
}
__global__ void k_project(PsState s, const uint8_t* protect, int32_t n) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n || protect[j]) return;
  const gp_device_lp& M = s.model;
  double aty = 0.0;
  for (gp_offset p = M.AT.row_ptr[j]; p < M.AT.row_ptr[j + 1]; ++p) aty += M.AT.vals[p] * s.y[M.AT.col_idx[p]];
  const double r = M.c[j] - aty;
  const double xj = s.x[j], lj = M.l[j], uj = M.u[j];
  const bool al = at_lower(xj, lj, s.tol), au = at_upper(xj, uj, s.tol);
  double zj;
  if (al && au) zj = r;
  else if (!al && !au) zj = 0.0;
  else if (al) zj = cmax(r, 0.0);
  else zj = cmin(r, 0.0);
  const double eps = cmin(s.tol, 1.0e-12);
  if (zj > eps && !isfinite(lj)) zj = 0.0;
  else if (zj < -eps && !isfinite(uj)) zj = 0.0;
  s.z[j] = zj;
}

// Tape validation (validate_postsolve_tape, gpu_postsolve.cu:400-437).
__global__ void k_validate(const int32_t* types, const gp_offset* is, const gp_offset* vs, const uint8_t* modes,
                           int64_t count, int64_t ni, int64_t nv, int32_t* bad) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  if (is[k] < 0 || is[k] > is[k + 1] || is[k + 1] > ni || vs[k] < 0 || vs[k] > vs[k + 1] || vs[k + 1] > nv ||
      types[k] < 0 || types[k] > GP_RED_FME_COL || modes[k] > GP_DUAL_MINIMAL)
    atomicExch(bad, 1);
  if (k == 0 && (is[0] != 0 || vs[0] != 0 || is[count] != ni || vs[count] != nv)) atomicExch(bad, 1);
}

}  // namespace



}  // namespace gp

using namespace gp;

extern "C" int gp_postprocess(const double* x, const double* y, const double* z, const gp_summary* summary,
                              const gp_device_lp* original, double* x_out, double* y_out, double* z_out) {
  return guarded([&] {
    if (!summary || !summary->record) throw std::invalid_argument("summary has no record");
    const RecordView r = record_view(summary->record);
    if (!summary->presolve_enabled) {
      if (x && x_out) d2h(x_out, x, static_cast<size_t>(r.n0));
      if (y && y_out) d2h(y_out, y, static_cast<size_t>(r.m0));
      if (z && z_out) d2h(z_out, z, static_cast<size_t>(r.n0));
      return;
    }
    const DeviceTape& t = *r.tape;
    if (t.count > 0) {
      DVec<int32_t> bad(1);
      GP_CHECK(hipMemsetAsync(bad.data(), 0, sizeof(int32_t), 0));
      hipLaunchKernelGGL(k_validate, dim3(grid_for(t.count)), dim3(kBlock), 0, 0, t.types.data(), t.istarts.data(),
                         t.vstarts.data(), t.modes.data(), t.count, t.icount, t.vcount, bad.data());
      GP_LAUNCH_CHECK("validate tape");
      if (read1(bad.data())) throw Error(GP_ERR_RUNTIME, "malformed postsolve tape");
    }
    const int32_t m0 = r.m0, n0 = r.n0;
    DVec<double> xo(n0 ? n0 : 1), yo(m0 ? m0 : 1), zo(n0 ? n0 : 1);
    DVec<uint8_t> ret(n0 ? n0 : 1);
    GP_CHECK(hipMemsetAsync(xo.data(), 0, sizeof(double) * n0, 0));
    GP_CHECK(hipMemsetAsync(zo.data(), 0, sizeof(double) * n0, 0));
    GP_CHECK(hipMemsetAsync(yo.data(), 0, sizeof(double) * m0, 0));
    GP_CHECK(hipMemsetAsync(ret.data(), 0, n0, 0));
    if (r.n1 > 0) {
      if (x) hipLaunchKernelGGL(k_scatter, dim3(grid_for(r.n1)), dim3(kBlock), 0, 0, xo.data(), x, r.cols, r.n1);
      if (z) hipLaunchKernelGGL(k_scatter, dim3(grid_for(r.n1)), dim3(kBlock), 0, 0, zo.data(), z, r.cols, r.n1);
      hipLaunchKernelGGL(k_mark, dim3(grid_for(r.n1)), dim3(kBlock), 0, 0, ret.data(), r.cols, r.n1);
    }
    if (r.m1 > 0 && y) hipLaunchKernelGGL(k_scatter, dim3(grid_for(r.m1)), dim3(kBlock), 0, 0, yo.data(), y, r.rows, r.m1);
    if (r.fixed_count > 0)
      hipLaunchKernelGGL(k_restore_fixed, dim3(grid_for(r.fixed_count)), dim3(kBlock), 0, 0, xo.data(), r.fixed_idx,
                         r.fixed_val, r.fixed_count);
    GP_LAUNCH_CHECK("postsolve scatter");
    PsState s{};
    s.x = xo.data();
    s.y = yo.data();
    s.z = zo.data();
    s.retrieved = ret.data();
    s.has_model = original != nullptr;
    if (original) s.model = *original;
    s.tol = 1e-7;
    PsTape pt{t.types.data(), t.istarts.data(), t.vstarts.data(), t.idx.data(), t.vals.data()};
    const bool rounds = t.count > 0 && t.count >= parallel_min_records() && t.count < INT32_MAX;
    // the dual pass is analysed after the primal replay: its pruning reads the final x
    if (t.count > 0 && !(rounds && replay_rounds(s, pt, static_cast<int32_t>(t.count), m0, n0, true)))
      hipLaunchKernelGGL(k_replay_serial, dim3(1), dim3(64), 0, 0, s, pt, t.count, 1, 0);
    if (t.count > 0 && !(rounds && replay_rounds(s, pt, static_cast<int32_t>(t.count), m0, n0, false)))
      hipLaunchKernelGGL(k_replay_serial, dim3(1), dim3(64), 0, 0, s, pt, t.count, 0, 1);
    GP_LAUNCH_CHECK("postsolve replay");
    if (original && n0 > 0) {
      DVec<uint8_t> protect(n0);
      GP_CHECK(hipMemsetAsync(protect.data(), 0, n0, 0));
      if (t.count > 0)
        hipLaunchKernelGGL(k_protect, dim3(grid_for(t.count)), dim3(kBlock), 0, 0, t.types.data(), t.istarts.data(),
                           t.idx.data(), t.count, n0, protect.data());
      hipLaunchKernelGGL(k_project, dim3(grid_for(n0)), dim3(kBlock), 0, 0, s, protect.data(), n0);
      GP_LAUNCH_CHECK("postsolve projection");
    }
    if (x_out) d2h(x_out, xo.data(), static_cast<size_t>(n0));
    if (y_out) d2h(y_out, yo.data(), static_cast<size_t>(m0));
    if (z_out) d2h(z_out, zo.data(), static_cast<size_t>(n0));
  });
}
