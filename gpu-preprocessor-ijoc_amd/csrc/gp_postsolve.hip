// gp_postsolve.hip -- postsolve: scatter the reduced solution to original
// positions, restore fixed columns, replay the tape in reverse (primal pass,
// then dual pass) and project column duals from the original model
// (gpu_postsolve.cu:535-670; record handlers postsolve/postsolve_<name>.cuh).
#include "gp_fold.hpp"
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

// Projected column (FmeCol; postsolve_projected_variables.cuh:14-68).
// indices: K, eliminated column, K deleted rows, N, L, (orient, len) x L,
// N support columns; values: (constant, len coefficients) x L.  The column
// takes the largest lower side, else the smallest upper side, else 0.
__device__ void replay_projected(PsState& s, const int32_t* ix, int64_t i0, const double* v, int64_t v0, bool primal,
                                 bool dual) {
  const int32_t K = ix[i0], col = ix[i0 + 1];
  const int32_t L = ix[i0 + 3 + K];
  if (primal) {
    int64_t meta = i0 + 4 + K, cp = meta + 2 * static_cast<int64_t>(L), vp = v0;
    double lo = -INFINITY, hi = INFINITY;
    bool has_lo = false, has_hi = false;
    for (int32_t side = 0; side < L; ++side) {
      const int32_t orient = ix[meta++];
      const int32_t len = ix[meta++];
      double cand = v[vp];
      for (int32_t t = 0; t < len; ++t) cand += v[vp + 1 + t] * s.x[ix[cp++]];
      if (orient > 0) {
        lo = cmax(lo, cand);
        has_lo = true;
      } else {
        hi = cmin(hi, cand);
        has_hi = true;
      }
      vp += 1 + len;
    }
    s.x[col] = has_lo ? lo : (has_hi ? hi : 0.0);
  }
  if (dual) {
    s.z[col] = 0.0;
    s.retrieved[col] = 1;
    for (int32_t t = 0; t < K; ++t) s.y[ix[i0 + 2 + t]] = 0.0;
  }
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
    case GP_RED_FME_COL: replay_projected(s, t.idx, i0, t.vals, v0, primal, dual); break;
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
    case GP_RED_FME_COL: {
      const int32_t K = ix[i0];
      f(ix[i0 + 1], true);
      if (primal) {
        const int32_t N = ix[i0 + 2 + K], L = ix[i0 + 3 + K];
        const int64_t cp = i0 + 4 + K + 2 * static_cast<int64_t>(L);
        for (int32_t q = 0; q < N; ++q) f(ix[cp + q], false);
      } else {
        for (int32_t q = 0; q < K; ++q) f(n0 + ix[i0 + 2 + q], true);
      }
      break;
    }
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

namespace {

// ---- structural primal recoveries (Startup rules' payloads) -----------------------
// postsolve_piecewise_linear_reformulation_absolute_value.cu:9-103 and
// postsolve_piecewise_linear_reformulation_variable_pair.cu:9-41.
__global__ void k_restore_splits(double* x, const int32_t* t_col, const int32_t* e_col, const double* rho,
                                 int32_t k) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const double t = x[t_col[i]], e = x[e_col[i]];
  x[t_col[i]] = t + e;
  x[e_col[i]] = (t - e) / rho[i];
}
__global__ void k_restore_pairs(double* x, const int32_t* bound_col, const int32_t* free_col, int32_t k) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < k) x[free_col[i]] = x[bound_col[i]];
}
__global__ void k_restore_linked(double* x, const int32_t* slack_col, const int32_t* t_col, const double* factor,
                                 int32_t k) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < k) x[slack_col[i]] = factor[i] * x[t_col[i]];
}
// one slack per lane: max of its lower bound (0 when absent or infinite) and
// factor * x over its columns
__global__ void k_restore_max(double* x, const int32_t* slack_col, const gp_offset* starts, const int32_t* t_cols,
                              const double* factors, int32_t k, const double* lower) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= k || starts[i + 1] == starts[i]) return;
  const int32_t sc = slack_col[i];
  double v = lower ? lower[sc] : 0.0;
  if (!isfinite(v)) v = 0.0;
  for (gp_offset q = starts[i]; q < starts[i + 1]; ++q) v = cmax(v, factors[q] * x[t_cols[q]]);
  x[sc] = v;
}
__global__ void k_cache_roots(const double* x, const int32_t* rp, const int32_t* rm, double* cp, double* cm,
                              int32_t k) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < k) {
    cp[i] = x[rp[i]];
    cm[i] = x[rm[i]];
  }
}
__global__ void k_restore_members(double* x, const int32_t* pc, const int32_t* mc, const double* ps, const double* ms,
                                  const double* cp, const double* cm, int32_t k) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < k) {
    x[pc[i]] = ps[i] + cp[i];
    x[mc[i]] = ms[i] + cm[i];
  }
}

template <class T>
DVec<T> upload(const T* src, int64_t count) {
  DVec<T> d(static_cast<size_t>(count > 0 ? count : 1));
  if (count > 0) h2d(d.data(), src, static_cast<size_t>(count));
  return d;
}

void restore_abs_value(double* x, const gp_abs_value_step& st, const double* lower) {
  if (st.split_count > 0) {
    auto t = upload(st.split_t_col, st.split_count), e = upload(st.split_e_col, st.split_count);
    auto r = upload(st.split_rho, st.split_count);
    hipLaunchKernelGGL(k_restore_splits, dim3(grid_for(st.split_count)), dim3(kBlock), 0, 0, x, t.data(), e.data(),
                       r.data(), st.split_count);
  }
  if (st.pair_count > 0) {
    auto b = upload(st.pair_bound_col, st.pair_count), f = upload(st.pair_free_col, st.pair_count);
    hipLaunchKernelGGL(k_restore_pairs, dim3(grid_for(st.pair_count)), dim3(kBlock), 0, 0, x, b.data(), f.data(),
                       st.pair_count);
  }
  if (st.linked_count > 0) {
    auto sc = upload(st.linked_slack_col, st.linked_count), t = upload(st.linked_t_col, st.linked_count);
    auto fa = upload(st.linked_factor, st.linked_count);
    hipLaunchKernelGGL(k_restore_linked, dim3(grid_for(st.linked_count)), dim3(kBlock), 0, 0, x, sc.data(), t.data(),
                       fa.data(), st.linked_count);
  }
  if (st.max_count > 0) {
    const int64_t total = st.max_starts[st.max_count];
    auto sc = upload(st.max_slack_col, st.max_count);
    auto starts = upload(st.max_starts, static_cast<int64_t>(st.max_count) + 1);
    auto t = upload(st.max_t_cols, total);
    auto fa = upload(st.max_factors, total);
    hipLaunchKernelGGL(k_restore_max, dim3(grid_for(st.max_count)), dim3(kBlock), 0, 0, x, sc.data(), starts.data(),
                       t.data(), fa.data(), st.max_count, lower);
  }
  GP_LAUNCH_CHECK("structural recovery (absolute value)");
}

void restore_var_pair(double* x, const gp_var_pair_step& st) {
  if (st.count <= 0) return;
  const int32_t k = st.count;
  auto pc = upload(st.plus_cols, k), mc = upload(st.minus_cols, k);
  auto rp = upload(st.root_plus_cols, k), rm = upload(st.root_minus_cols, k);
  auto ps = upload(st.plus_shift, k), ms = upload(st.minus_shift, k);
  DVec<double> cp(k), cm(k);
  hipLaunchKernelGGL(k_cache_roots, dim3(grid_for(k)), dim3(kBlock), 0, 0, x, rp.data(), rm.data(), cp.data(),
                     cm.data(), k);
  hipLaunchKernelGGL(k_restore_members, dim3(grid_for(k)), dim3(kBlock), 0, 0, x, pc.data(), mc.data(), ps.data(),
                     ms.data(), cp.data(), cm.data(), k);
  GP_LAUNCH_CHECK("structural recovery (variable pair)");
}

// Reverse replay of tape records [begin, end) in one pass.
void replay_range(const PsState& s, const PsTape& t, int64_t begin, int64_t end, int32_t m0, int32_t n0,
                  bool primal) {
  if (end <= begin) return;
  const PsTape v{t.types + begin, t.is + begin, t.vs + begin, t.idx, t.vals};
  const int64_t count = end - begin;
  const bool rounds = count >= parallel_min_records() && count < INT32_MAX;
  if (!(rounds && replay_rounds(s, v, static_cast<int32_t>(count), m0, n0, primal)))
    hipLaunchKernelGGL(k_replay_serial, dim3(1), dim3(64), 0, 0, s, v, count, primal ? 1 : 0, primal ? 0 : 1);
  GP_LAUNCH_CHECK("postsolve replay");
}

// Device-side inputs of one postsolve (maps, fixed columns, tape) plus the
// host-side structural timeline (may be empty).
struct PostsolveInputs {
  int32_t m0, n0, m1, n1;
  const int32_t* rows;
  const int32_t* cols;
  const int32_t* fixed_idx;
  const double* fixed_val;
  int64_t fixed_count;
  PsTape tape;
  int64_t count;
  const gp_host_record* timeline;  // checkpoints and payloads, or null
};

// postsolve_gpu (gpu_postsolve.cu:535-670): scatter, fixed restore, primal
// replay between structural checkpoints in reverse, dual replay, projection.
// Device copies of a postsolved point (filled instead of host arrays when
// the caller continues on the device, e.g. to unfold).
struct PostsolvedDevice {
  DVec<double> x, y, z;
};

void postsolve_run(const PostsolveInputs& in, const double* x, const double* y, const double* z,
                   const gp_device_lp* original, const double* original_l, double* x_out, double* y_out,
                   double* z_out, PostsolvedDevice* keep = nullptr) {
  const int32_t m0 = in.m0, n0 = in.n0;
  if (!original_l && original) original_l = original->l;
  DVec<double> xo(n0 ? n0 : 1), yo(m0 ? m0 : 1), zo(n0 ? n0 : 1);
  DVec<uint8_t> ret(n0 ? n0 : 1);
  GP_CHECK(hipMemsetAsync(xo.data(), 0, sizeof(double) * n0, 0));
  GP_CHECK(hipMemsetAsync(zo.data(), 0, sizeof(double) * n0, 0));
  GP_CHECK(hipMemsetAsync(yo.data(), 0, sizeof(double) * m0, 0));
  GP_CHECK(hipMemsetAsync(ret.data(), 0, n0, 0));
  if (in.n1 > 0) {
    if (x) hipLaunchKernelGGL(k_scatter, dim3(grid_for(in.n1)), dim3(kBlock), 0, 0, xo.data(), x, in.cols, in.n1);
    if (z) hipLaunchKernelGGL(k_scatter, dim3(grid_for(in.n1)), dim3(kBlock), 0, 0, zo.data(), z, in.cols, in.n1);
    hipLaunchKernelGGL(k_mark, dim3(grid_for(in.n1)), dim3(kBlock), 0, 0, ret.data(), in.cols, in.n1);
  }
  if (in.m1 > 0 && y)
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(in.m1)), dim3(kBlock), 0, 0, yo.data(), y, in.rows, in.m1);
  if (in.fixed_count > 0)
    hipLaunchKernelGGL(k_restore_fixed, dim3(grid_for(in.fixed_count)), dim3(kBlock), 0, 0, xo.data(), in.fixed_idx,
                       in.fixed_val, in.fixed_count);
  GP_LAUNCH_CHECK("postsolve scatter");
  PsState s{};
  s.x = xo.data();
  s.y = yo.data();
  s.z = zo.data();
  s.retrieved = ret.data();
  s.has_model = original != nullptr;
  if (original) s.model = *original;
  s.tol = 1e-7;
  // primal: the tape ranges between checkpoints, newest first, each followed
  // by its structural payload; the dual pass is analysed after the primal
  // replay (its pruning reads the final x)
  int64_t range_end = in.count;
  const gp_host_record* tl = in.timeline;
  for (int32_t c = tl ? tl->checkpoint_count - 1 : -1; c >= 0; --c) {
    const gp_checkpoint& cp = tl->checkpoints[c];
    replay_range(s, in.tape, cp.tape_position, range_end, m0, n0, true);
    if (cp.kind == GP_RECOVERY_ABS_VALUE) restore_abs_value(s.x, tl->abs_value[cp.payload_index], original_l);
    else restore_var_pair(s.x, tl->var_pair[cp.payload_index]);
    range_end = cp.tape_position;
  }
  replay_range(s, in.tape, 0, range_end, m0, n0, true);
  replay_range(s, in.tape, 0, in.count, m0, n0, false);
  // projection needs x, y and z (the reference has no y without rows) and the
  // model's AT, c, l and u
  if (original && n0 > 0 && m0 > 0 && original->AT.rows > 0 && original->AT.row_ptr && original->c && original->l &&
      original->u) {
    DVec<uint8_t> protect(n0);
    GP_CHECK(hipMemsetAsync(protect.data(), 0, n0, 0));
    if (in.count > 0)
      hipLaunchKernelGGL(k_protect, dim3(grid_for(in.count)), dim3(kBlock), 0, 0, in.tape.types, in.tape.is,
                         in.tape.idx, in.count, n0, protect.data());
    hipLaunchKernelGGL(k_project, dim3(grid_for(n0)), dim3(kBlock), 0, 0, s, protect.data(), n0);
    GP_LAUNCH_CHECK("postsolve projection");
  }
  if (keep) {
    keep->x = std::move(xo);
    keep->y = std::move(yo);
    keep->z = std::move(zo);
    return;
  }
  if (x_out) d2h(x_out, xo.data(), static_cast<size_t>(n0));
  if (y_out) d2h(y_out, yo.data(), static_cast<size_t>(m0));
  if (z_out) d2h(z_out, zo.data(), static_cast<size_t>(n0));
}

// ---- host-side validation of a caller-built record ---------------------------------
[[noreturn]] void bad_record(const char* what) { throw Error(GP_ERR_RUNTIME, what); }

// validate_postsolve_tape (gpu_postsolve.cu:400-437), plus the index ranges
// each record's replay reads, so a malformed record is an error, not a fault.
void validate_host_tape(const gp_host_record& r) {
  const int64_t count = r.record_count;
  if (count < 0 || r.index_count < 0 || r.value_count < 0) bad_record("malformed postsolve tape offsets");
  if (count == 0) {
    if ((r.index_starts && r.index_starts[0] != 0) || (r.value_starts && r.value_starts[0] != 0))
      bad_record("postsolve tape offsets must start at zero");
    if (r.index_count != 0 || r.value_count != 0) bad_record("postsolve tape terminal offsets do not match data");
    return;
  }
  if (!r.types || !r.index_starts || !r.value_starts || !r.dual_modes) bad_record("malformed postsolve tape offsets");
  if (r.index_starts[0] != 0 || r.value_starts[0] != 0) bad_record("postsolve tape offsets must start at zero");
  for (int64_t k = 0; k < count; ++k) {
    if (r.index_starts[k] < 0 || r.index_starts[k] > r.index_starts[k + 1] || r.value_starts[k] < 0 ||
        r.value_starts[k] > r.value_starts[k + 1])
      bad_record("postsolve tape offsets are not monotone");
    if (r.types[k] < GP_RED_FIXED_VARIABLE || r.types[k] > GP_RED_FME_COL)
      bad_record("postsolve tape contains an unknown reduction");
    if (r.dual_modes[k] > GP_DUAL_MINIMAL) bad_record("postsolve tape contains an unknown dual mode");
  }
  if (r.index_starts[count] != r.index_count || r.value_starts[count] != r.value_count)
    bad_record("postsolve tape terminal offsets do not match data");
  if ((r.index_count && !r.indices) || (r.value_count && !r.vals)) bad_record("postsolve tape data missing");
  const int32_t* ix = r.indices;
  auto col = [&](int64_t p) {
    if (ix[p] < 0 || ix[p] >= r.n0) bad_record("postsolve record names an invalid original column");
  };
  auto row = [&](int64_t p, bool allow_none) {
    if (allow_none && ix[p] < 0) return;
    if (ix[p] < 0 || ix[p] >= r.m0) bad_record("postsolve record names an invalid original row");
  };
  for (int64_t k = 0; k < count; ++k) {
    const int64_t i0 = r.index_starts[k], ni = r.index_starts[k + 1] - i0;
    const int64_t nv = r.value_starts[k + 1] - r.value_starts[k];
    auto need = [&](int64_t i, int64_t v) {
      if (ni < i || nv < v) bad_record("postsolve record payload too short");
    };
    switch (r.types[k]) {
      case GP_RED_FIXED_VARIABLE:
        need(1, 1 + ni);
        col(i0);
        for (int64_t p = 1; p < ni; ++p) row(i0 + p, false);
        break;
      case GP_RED_INFINITE_BOUND_DUAL_FIX: {
        need(3, 1);
        col(i0 + 1);
        int64_t ip = 3, vp = 1;
        for (int32_t q = 0; q < ix[i0 + 2]; ++q) {
          need(ip + 1, vp);
          const int64_t len = ix[i0 + ip];
          if (len < 0) bad_record("postsolve record payload too short");
          need(ip + 1 + len, vp + 1 + len);
          for (int64_t t = 0; t < len; ++t) col(i0 + ip + 1 + t);
          ip += len + 1;
          vp += len + 1;
        }
        break;
      }
      case GP_RED_SINGLETON_COLUMN_SUBSTITUTION: {
        need(3, 6);
        col(i0);
        row(i0 + 1, true);
        const int64_t sup = ix[i0 + 2];
        if (sup < 0) bad_record("postsolve record payload too short");
        need(3 + sup, 6 + sup);
        for (int64_t t = 0; t < sup; ++t) col(i0 + 3 + t);
        break;
      }
      case GP_RED_PARALLEL_COLUMN:
        need(2, 5);
        col(i0);
        col(i0 + 1);
        break;
      case GP_RED_PARALLEL_ROW:
        if (ni >= 2 && nv >= 5) {
          row(i0, false);
          row(i0 + 1, false);
        }
        break;
      case GP_RED_DELETED_ROW:
        if (ni >= 2 && nv >= 3) {
          row(i0, false);
          col(i0 + 1);
        }
        break;
      case GP_RED_ROW_DUAL_SHIFT:
        need(1, 1);
        row(i0, false);
        break;
      case GP_RED_ROW_IMPLIED_BOUND_CHANGE:
        if (ni >= 2) {
          need(2, 4);
          col(i0);
          row(i0 + 1, false);
        }
        break;
      case GP_RED_DOUBLETON_EQUATION: {
        need(3, 10);
        col(i0);
        col(i0 + 1);
        row(i0 + 2, true);
        if (ni >= 4 && nv >= 12) {
          const int64_t sc = ix[i0 + 3];
          if (sc < 0) bad_record("postsolve record payload too short");
          need(4 + sc, 12 + sc);
          for (int64_t t = 0; t < sc; ++t) row(i0 + 4 + t, false);
        }
        break;
      }
      case GP_RED_FME_COL: {
        need(2, 0);
        const int64_t K = ix[i0];
        if (K < 0) bad_record("postsolve record payload too short");
        need(4 + K, 0);
        col(i0 + 1);
        for (int64_t t = 0; t < K; ++t) row(i0 + 2 + t, false);
        const int64_t N = ix[i0 + 2 + K], L = ix[i0 + 3 + K];
        if (N < 0 || L < 0) bad_record("postsolve record payload too short");
        need(4 + K + 2 * L + N, 0);
        int64_t lens = 0;
        for (int64_t q = 0; q < L; ++q) {
          const int64_t len = ix[i0 + 4 + K + 2 * q + 1];
          if (len < 0) bad_record("postsolve record payload too short");
          lens += len;
        }
        if (lens > N) bad_record("postsolve record payload too short");
        need(4 + K + 2 * L + N, L + lens);
        for (int64_t t = 0; t < N; ++t) col(i0 + 4 + K + 2 * L + t);
        break;
      }
      default: break;  // row / variable bound changes replay nothing
    }
  }
}

void recovery_col(int32_t c, int32_t n0, const char* what) {
  if (c < 0 || c >= n0) throw Error(GP_ERR_RUNTIME, std::string(what) + " contains an invalid original column");
}

// validate_primal_recovery_timeline (gpu_postsolve.cu:439-503)
void validate_timeline(const gp_host_record& r) {
  const int32_t na = r.abs_value_count > 0 ? r.abs_value_count : 0;
  const int32_t nvp = r.var_pair_count > 0 ? r.var_pair_count : 0;
  if ((na && !r.abs_value) || (nvp && !r.var_pair) || (r.checkpoint_count > 0 && !r.checkpoints))
    bad_record("primal recovery payload missing");
  std::vector<uint8_t> seen_a(na, 0), seen_v(nvp, 0);
  int64_t prev = 0;
  for (int32_t c = 0; c < r.checkpoint_count; ++c) {
    const gp_checkpoint& cp = r.checkpoints[c];
    if (cp.tape_position < prev || cp.tape_position > r.record_count)
      bad_record("primal recovery checkpoints are not chronological");
    prev = cp.tape_position;
    if (cp.kind == GP_RECOVERY_ABS_VALUE) {
      if (cp.payload_index < 0 || cp.payload_index >= na || seen_a[cp.payload_index])
        bad_record("invalid or duplicate absolute-value recovery checkpoint");
      seen_a[cp.payload_index] = 1;
      const gp_abs_value_step& st = r.abs_value[cp.payload_index];
      for (int32_t i = 0; i < st.split_count; ++i) {
        recovery_col(st.split_t_col[i], r.n0, "absolute-value split recovery");
        recovery_col(st.split_e_col[i], r.n0, "absolute-value split recovery");
      }
      for (int32_t i = 0; i < st.pair_count; ++i) {
        recovery_col(st.pair_bound_col[i], r.n0, "absolute-value outer-pair recovery");
        recovery_col(st.pair_free_col[i], r.n0, "absolute-value outer-pair recovery");
      }
      for (int32_t i = 0; i < st.linked_count; ++i) {
        recovery_col(st.linked_slack_col[i], r.n0, "absolute-value linked-slack recovery");
        recovery_col(st.linked_t_col[i], r.n0, "absolute-value linked-slack recovery");
      }
      if (st.max_count > 0) {
        if (!st.max_starts || st.max_starts[0] != 0)
          bad_record("absolute-value max-slack recovery has mismatched payloads");
        for (int32_t i = 0; i < st.max_count; ++i) {
          if (st.max_starts[i + 1] < st.max_starts[i])
            bad_record("absolute-value max-slack recovery has mismatched payloads");
          recovery_col(st.max_slack_col[i], r.n0, "absolute-value max-slack recovery");
          for (gp_offset q = st.max_starts[i]; q < st.max_starts[i + 1]; ++q)
            recovery_col(st.max_t_cols[q], r.n0, "absolute-value max-slack recovery");
        }
      }
    } else if (cp.kind == GP_RECOVERY_VAR_PAIR) {
      if (cp.payload_index < 0 || cp.payload_index >= nvp || seen_v[cp.payload_index])
        bad_record("invalid or duplicate variable-pair recovery checkpoint");
      seen_v[cp.payload_index] = 1;
      const gp_var_pair_step& st = r.var_pair[cp.payload_index];
      if (st.count < 0) bad_record("variable-pair recovery has mismatched payloads");
      for (int32_t i = 0; i < st.count; ++i) {
        recovery_col(st.plus_cols[i], r.n0, "variable-pair recovery");
        recovery_col(st.minus_cols[i], r.n0, "variable-pair recovery");
        recovery_col(st.root_plus_cols[i], r.n0, "variable-pair recovery");
        recovery_col(st.root_minus_cols[i], r.n0, "variable-pair recovery");
      }
    } else {
      bad_record("unknown primal recovery checkpoint kind");
    }
  }
  for (uint8_t v : seen_a)
    if (!v) bad_record("absolute-value recovery payload is missing a checkpoint");
  for (uint8_t v : seen_v)
    if (!v) bad_record("variable-pair recovery payload is missing a checkpoint");
}

void validate_host_record(const gp_host_record& r) {
  if (r.m0 < 0 || r.n0 < 0 || r.m1 < 0 || r.n1 < 0 || r.m1 > r.m0 || r.n1 > r.n0 || r.fixed_count < 0)
    throw std::invalid_argument("invalid record dimensions");
  if ((r.m1 && !r.row_red2org) || (r.n1 && !r.col_red2org) || (r.fixed_count && (!r.fixed_idx || !r.fixed_val)))
    throw std::invalid_argument("record maps missing");
  for (int32_t i = 0; i < r.m1; ++i)
    if (r.row_red2org[i] < 0 || r.row_red2org[i] >= r.m0) bad_record("record row map out of range");
  for (int32_t j = 0; j < r.n1; ++j)
    if (r.col_red2org[j] < 0 || r.col_red2org[j] >= r.n0) bad_record("record column map out of range");
  for (int32_t k = 0; k < r.fixed_count; ++k)
    if (r.fixed_idx[k] < 0 || r.fixed_idx[k] >= r.n0) bad_record("record fixed column out of range");
  validate_host_tape(r);
  validate_timeline(r);
}

}  // namespace

}  // namespace gp

using namespace gp;

extern "C" int gp_postprocess(const double* x, const double* y, const double* z, const gp_summary* summary,
                              const gp_device_lp* original, double* x_out, double* y_out, double* z_out) {
  return guarded([&] {
    if (!summary || !summary->record) throw std::invalid_argument("summary has no record");
    const RecordView r = record_view(summary->record);
    const FoldState* fold = static_cast<const FoldState*>(r.fold);
    if (fold) {
      // postsolve against the folded LP, then unfold
      // (postsolve_and_unfold_to_host, folding.cu:211-262)
      PostsolvedDevice folded;
      const double *fx = x, *fy = y, *fz = z;
      if (summary->presolve_enabled) {
        const DeviceTape& t = *r.tape;
        PostsolveInputs in{};
        in.m0 = r.m0;
        in.n0 = r.n0;
        in.m1 = r.m1;
        in.n1 = r.n1;
        in.rows = r.rows;
        in.cols = r.cols;
        in.fixed_idx = r.fixed_idx;
        in.fixed_val = r.fixed_val;
        in.fixed_count = r.fixed_count;
        in.tape = PsTape{t.types.data(), t.istarts.data(), t.vstarts.data(), t.idx.data(), t.vals.data()};
        in.count = t.count;
        in.timeline = nullptr;
        postsolve_run(in, x, y, z, &fold->lp, nullptr, nullptr, nullptr, nullptr, &folded);
        fx = folded.x.data();
        fy = folded.y.data();
        fz = folded.z.data();
      }
      DVec<double> ux(static_cast<size_t>(fold->n ? fold->n : 1)), uy(static_cast<size_t>(fold->m ? fold->m : 1)),
          uz(static_cast<size_t>(fold->n ? fold->n : 1));
      unfold(*fold, fx, fy, fz, ux.data(), uy.data(), uz.data());
      if (x_out && fx) d2h(x_out, ux.data(), static_cast<size_t>(fold->n));
      if (y_out && fy) d2h(y_out, uy.data(), static_cast<size_t>(fold->m));
      if (z_out && fz) d2h(z_out, uz.data(), static_cast<size_t>(fold->n));
      return;
    }
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
    PostsolveInputs in{};
    in.m0 = r.m0;
    in.n0 = r.n0;
    in.m1 = r.m1;
    in.n1 = r.n1;
    in.rows = r.rows;
    in.cols = r.cols;
    in.fixed_idx = r.fixed_idx;
    in.fixed_val = r.fixed_val;
    in.fixed_count = r.fixed_count;
    in.tape = PsTape{t.types.data(), t.istarts.data(), t.vstarts.data(), t.idx.data(), t.vals.data()};
    in.count = t.count;
    in.timeline = nullptr;
    postsolve_run(in, x, y, z, original, nullptr, x_out, y_out, z_out);
  });
}

extern "C" int gp_postsolve_record(const double* x_red, const double* y_red, const double* z_red,
                                   const gp_host_record* record, const gp_device_lp* original,
                                   const double* original_l, double* x_out, double* y_out, double* z_out) {
  return guarded([&] {
    if (!record) throw std::invalid_argument("null record");
    const gp_host_record& r = *record;
    validate_host_record(r);
    if (original) {
      if (original->AT.row_ptr && original->AT.rows != r.n0) throw std::invalid_argument("original AT shape");
      if (original->A.row_ptr && original->A.rows != r.m0) throw std::invalid_argument("original A shape");
    }
    auto rows = upload(r.row_red2org, r.m1), cols = upload(r.col_red2org, r.n1);
    auto fidx = upload(r.fixed_idx, r.fixed_count);
    auto fval = upload(r.fixed_val, r.fixed_count);
    const int64_t count = r.record_count;
    const gp_offset zero = 0;
    auto types = upload(r.types, count);
    auto is = upload(count > 0 ? r.index_starts : &zero, count + 1);
    auto vs = upload(count > 0 ? r.value_starts : &zero, count + 1);
    auto idx = upload(r.indices, r.index_count);
    auto vals = upload(r.vals, r.value_count);
    PostsolveInputs in{};
    in.m0 = r.m0;
    in.n0 = r.n0;
    in.m1 = r.m1;
    in.n1 = r.n1;
    in.rows = rows.data();
    in.cols = cols.data();
    in.fixed_idx = fidx.data();
    in.fixed_val = fval.data();
    in.fixed_count = r.fixed_count;
    in.tape = PsTape{types.data(), is.data(), vs.data(), idx.data(), vals.data()};
    in.count = count;
    in.timeline = &r;
    postsolve_run(in, x_red, y_red, z_red, original, original_l, x_out, y_out, z_out);
  });
}
