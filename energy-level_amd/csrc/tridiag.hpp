// Host-side post-processing of many Lanczos runs: the two lowest genuine
// Ritz pairs of each FP64 tridiagonal T (alpha on the diagonal, the Lanczos
// norms beta off it), found without a full eigensolve.
//
// * eigenvalues by Sturm-sequence bisection, lowest first, only as many as
//   the selection below examines (usually two or three);
// * a candidate within tol (merge_rel·|T|) of the previous kept value is a
//   ghost copy and merges; a simple candidate (no other value within tol) that is also an
//   eigenvalue of T-hat (T without its first row and column; two Sturm
//   counts of T-hat around it decide) is spurious (Cullum-Willoughby) unless
//   its residual bound |beta_k · y[k-1]| shows it converged; the lowest
//   value is always kept (a spurious value never lies below the lowest
//   genuine one);
// * eigenvectors by inverse iteration on (T - theta) with a partially
//   pivoted tridiagonal LU, two sweeps from a fixed start, re-orthogonalised
//   against earlier vectors of the same instance whose values lie within
//   1e-3·|T| (as LAPACK's dstein does for clusters).
#pragma once
#include <cfloat>
#include <algorithm>
#include <cmath>
#include <vector>

namespace ryd {
namespace tri {

// eigenvalues of rows r0..k-1 of T below x
inline int sturm_count(const double* a, const double* b, int r0, int k, double x, double pivmin) {
  int c = 0;
  double q = 1.0;
  for (int i = r0; i < k; ++i) {
    q = a[i] - x - (i > r0 ? b[i - 1] * b[i - 1] / q : 0.0);
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++c;
  }
  return c;
}

// values j0 .. j0+n-1 (n <= 4) bisected together: the four Sturm recurrences
// of a pass are independent, so their divisions overlap
inline void bisect4(const double* a, const double* b, int k, int j0, int n, double lo, double hi,
                    double pivmin, double* out) {
  double l[4], h[4];
  for (int q = 0; q < 4; ++q) { l[q] = lo; h[q] = hi; }
  for (int it = 0; it < 256; ++it) {
    double x[4], qv[4];
    int c[4] = {0, 0, 0, 0};
    bool live = false;
    for (int q = 0; q < 4; ++q) {
      x[q] = 0.5 * (l[q] + h[q]);
      if (q < n && x[q] > l[q] && x[q] < h[q]) live = true;
    }
    if (!live) break;
    for (int q = 0; q < 4; ++q) {
      qv[q] = a[0] - x[q];
      if (std::fabs(qv[q]) < pivmin) qv[q] = -pivmin;
      c[q] += qv[q] < 0.0;
    }
    for (int i = 1; i < k; ++i) {
      const double bb = b[i - 1] * b[i - 1];
      for (int q = 0; q < 4; ++q) {
        double v = a[i] - x[q] - bb / qv[q];
        if (std::fabs(v) < pivmin) v = -pivmin;
        qv[q] = v;
        c[q] += v < 0.0;
      }
    }
    for (int q = 0; q < n; ++q) {
      if (!(x[q] > l[q] && x[q] < h[q])) continue;
      if (c[q] > j0 + q) h[q] = x[q];
      else l[q] = x[q];
    }
  }
  for (int q = 0; q < n; ++q) out[q] = 0.5 * (l[q] + h[q]);
}

// one solve (T - theta) x = rhs (in place) with partial pivoting
inline void shifted_solve(const double* a, const double* b, int k, double theta, double pivtiny,
                          std::vector<double>& x, std::vector<double>& d, std::vector<double>& du,
                          std::vector<double>& du2, std::vector<double>& dl, std::vector<int>& piv) {
  for (int i = 0; i < k; ++i) d[i] = a[i] - theta;
  for (int i = 0; i + 1 < k; ++i) { du[i] = b[i]; dl[i] = b[i]; }
  for (int i = 0; i + 2 < k; ++i) du2[i] = 0.0;
  for (int i = 0; i + 1 < k; ++i) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {          // no row interchange
      piv[i] = i;
      if (std::fabs(d[i]) < pivtiny) d[i] = pivtiny;
      const double f = dl[i] / d[i];
      dl[i] = f;
      d[i + 1] -= f * du[i];
    } else {                                            // swap rows i and i+1
      piv[i] = i + 1;
      const double f = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = d[i + 1];
      d[i + 1] = t - f * d[i + 1];
      if (i + 2 < k) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
    }
  }
  if (std::fabs(d[k - 1]) < pivtiny) d[k - 1] = pivtiny;
  for (int i = 0; i + 1 < k; ++i) {                     // L solve with the interchanges
    if (piv[i] != i) { const double t = x[i]; x[i] = x[i + 1]; x[i + 1] = t; }
    x[i + 1] -= dl[i] * x[i];
  }
  x[k - 1] /= d[k - 1];                                 // U solve
  if (k > 1) x[k - 2] = (x[k - 2] - du[k - 2] * x[k - 1]) / d[k - 2];
  for (int i = k - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
}

struct Work {
  std::vector<double> d, du, du2, dl, x;
  std::vector<int> piv;
  void resize(int k) {
    d.resize(k); du.resize(k); du2.resize(k); dl.resize(k); x.resize(k); piv.resize(k);
  }
};

// unit eigenvector of T for theta into y (length k), orthogonalised against
// the vectors in `prev` whose values lie within ortol of theta
inline void eigvec(const double* a, const double* b, int k, double theta, double scale,
                   const std::vector<const double*>& prev, const std::vector<double>& prev_theta,
                   double* y, Work& w) {
  w.resize(k);
  const double pivtiny = DBL_EPSILON * scale;
  const double ortol = 1e-3 * scale;
  uint64_t s = 0x9E3779B97F4A7C15ull;
  for (int i = 0; i < k; ++i) {                          // fixed, generic start vector
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    w.x[i] = 1.0 + 0.5 * ((double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5);
  }
  for (int sweep = 0; sweep < 3; ++sweep) {
    shifted_solve(a, b, k, theta, pivtiny, w.x, w.d, w.du, w.du2, w.dl, w.piv);
    for (size_t p = 0; p < prev.size(); ++p) {
      if (std::fabs(prev_theta[p] - theta) > ortol) continue;
      double dot = 0.0;
      for (int i = 0; i < k; ++i) dot += w.x[i] * prev[p][i];
      for (int i = 0; i < k; ++i) w.x[i] -= dot * prev[p][i];
    }
    double nrm = 0.0;
    for (int i = 0; i < k; ++i) nrm += w.x[i] * w.x[i];
    nrm = std::sqrt(nrm);
    if (!(nrm > 0.0)) nrm = 1.0;
    for (int i = 0; i < k; ++i) w.x[i] /= nrm;
  }
  for (int i = 0; i < k; ++i) y[i] = w.x[i];
}

// The two lowest genuine Ritz pairs of one k-step tridiagonal.  E[0..1],
// bound[0..1] (|beta_last · y[k-1]|; E1 = NaN, bound = inf when there is
// none), y0, y1 (length k, may be null).  Values within merge_rel·|T| of
// each other are one (the ghost copies of a recurrence agree to its own
// rounding: 1e-10 suits FP64 vectors, 1e-6 FP32 ones).
inline void lowest_pairs(const double* a, const double* b, int k, double beta_last,
                         double conv_tol, double merge_rel, double* E, double* bound, double* y0,
                         double* y1, Work& w) {
  E[0] = E[1] = NAN;
  bound[0] = bound[1] = INFINITY;
  if (k < 1) return;
  double lo = INFINITY, hi = -INFINITY, amax = 0.0, bmax = 0.0;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
    lo = std::fmin(lo, a[i] - r);
    hi = std::fmax(hi, a[i] + r);
    amax = std::fmax(amax, std::fabs(a[i]));
    if (i + 1 < k) bmax = std::fmax(bmax, std::fabs(b[i]));
  }
  const double scale = std::fmax(1.0, amax + 2.0 * bmax);
  const double span = hi - lo;
  lo -= 1e-12 * scale + 2.0 * DBL_EPSILON * span;
  hi += 1e-12 * scale + 2.0 * DBL_EPSILON * span;
  const double pivmin = DBL_MIN * std::fmax(1.0, bmax * bmax);
  const double tol = std::fmax(merge_rel * scale, 100.0 * k * DBL_EPSILON * scale);
  std::vector<double> th;                  // eigenvalues found so far, ascending
  std::vector<int> have;                   // index into vec of th[j], -1 if none
  std::vector<std::vector<double>> vec;    // eigenvectors computed so far
  std::vector<double> vth;                 // their eigenvalues
  auto value = [&](int j) {
    while ((int)th.size() <= j) {
      const int jj = (int)th.size();
      const int n = std::min(4, k - jj);
      const double l = jj > 0 ? th[jj - 1] - 2.0 * DBL_EPSILON * scale : lo;
      double out[4];
      bisect4(a, b, k, jj, n, std::fmin(l, hi), hi, pivmin, out);
      for (int q = 0; q < n; ++q) {
        th.push_back(out[q]);
        have.push_back(-1);
      }
    }
    return th[j];
  };
  auto vector_of = [&](int j) -> const std::vector<double>& {
    if (have[j] < 0) {
      std::vector<const double*> pv;
      for (const auto& v : vec) pv.push_back(v.data());
      std::vector<double> y(k);
      eigvec(a, b, k, th[j], scale, pv, vth, y.data(), w);
      vec.push_back(std::move(y));
      vth.push_back(th[j]);
      have[j] = (int)vec.size() - 1;
    }
    return vec[have[j]];
  };
  auto bound_of = [&](int j) { return std::fabs(beta_last * vector_of(j)[k - 1]); };
  // A kept value and its ghost copies (within tol) are one eigenvalue of H;
  // the copy with the smallest bound certifies it and gives its vector (the
  // newest copy of a converged value carries a large last component, so the
  // lowest copy alone can hold a bound far above the converged one).
  int kept[2] = {-1, -1};
  int best[2] = {-1, -1};
  double bb[2] = {INFINITY, INFINITY};
  int nk = 0;
  for (int j = 0; j < k; ++j) {
    const double t = value(j);
    if (nk > 0 && std::fabs(t - th[kept[nk - 1]]) <= tol) {   // ghost copy
      const double bj = bound_of(j);
      if (bj < bb[nk - 1]) {
        bb[nk - 1] = bj;
        best[nk - 1] = j;
      }
      continue;
    }
    if (nk == 2) break;
    if (nk > 0) {
      // simple: no other eigenvalue of T within tol (the values are sorted,
      // so the neighbours decide)
      const bool simple = !(j > 0 && std::fabs(th[j - 1] - t) <= tol) &&
                          !(j + 1 < k && std::fabs(value(j + 1) - t) <= tol);
      if (simple && k > 1) {
        const int above = sturm_count(a, b, 1, k, t + tol, pivmin);
        const int below = sturm_count(a, b, 1, k, t - tol, pivmin);
        if (above > below && !(bound_of(j) <= conv_tol * std::fmax(1.0, std::fabs(t)))) continue;
      }
    }
    kept[nk] = j;
    best[nk] = j;
    bb[nk] = bound_of(j);
    ++nk;
  }
  for (int q = 0; q < nk; ++q) {
    E[q] = th[kept[q]];
    bound[q] = bb[q];
    double* out = q == 0 ? y0 : y1;
    if (out) {
      const std::vector<double>& v = vector_of(best[q]);
      for (int i = 0; i < k; ++i) out[i] = v[i];
    }
  }
}

}  // namespace tri
}  // namespace ryd
