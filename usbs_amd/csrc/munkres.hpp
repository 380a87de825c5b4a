// Minimum-cost assignment (Kuhn-Munkres) for the QAP rounding, host code.
//
// The steps and their scan orders are those of the reference's munkres
// (utils/munkres.py:14-120, after Bob Pilgrim's description of the
// algorithm), so that ties between equal-cost assignments break the same
// way: zeros are starred in row-major order; each search takes the first
// uncovered zero in row-major order; an augmenting path alternates primes and
// stars from it; when no uncovered zero is left the smallest uncovered value
// is subtracted from every uncovered column and added to every covered row.
// The arithmetic on the working matrix is the reference's, operation for
// operation, so equality tests against zero see the same values.
//
// Header-only and free of HIP types so that a host-only build can check it.
#pragma once
#include <vector>

namespace usbs {

// cost: n x n row-major.  mask (n x n row-major) receives the reference's
// result matrix: 1 at the assigned entries.  The loop stops after n^2 steps,
// as the reference's does; if that leaves fewer than n stars the mask is
// returned as it stands (it may then also hold -1 where an augmenting path
// crossed an entry that was neither starred nor primed), exactly as the
// reference returns it.  Returns whether the assignment is complete.
inline bool munkres_mask(int n, const double* cost, std::vector<int>& mask) {
  std::vector<double> M(cost, cost + (size_t)n * n);
  // Step 1: subtract each row's minimum
  for (int i = 0; i < n; ++i) {
    double mn = M[(size_t)i * n];
    for (int j = 1; j < n; ++j) mn = M[(size_t)i * n + j] < mn ? M[(size_t)i * n + j] : mn;
    for (int j = 0; j < n; ++j) M[(size_t)i * n + j] = M[(size_t)i * n + j] - mn;
  }
  // mask: 0 none, 1 starred zero, 2 primed zero
  mask.assign((size_t)n * n, 0);
  std::vector<char> rc(n, 0), cc(n, 0);
  // Step 2: star zeros in row-major order, one per row and column
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (M[(size_t)i * n + j] == 0.0 && !rc[i] && !cc[j]) {
        mask[(size_t)i * n + j] = 1;
        rc[i] = 1;
        cc[j] = 1;
      }
  // covers for the main loop: rows none, columns holding a star
  auto cover_star_columns = [&]() {
    for (int i = 0; i < n; ++i) rc[i] = 0;
    for (int j = 0; j < n; ++j) {
      cc[j] = 0;
      for (int i = 0; i < n; ++i)
        if (mask[(size_t)i * n + j] == 1) cc[j] = 1;
    }
  };
  cover_star_columns();
  auto stars = [&]() {
    int s = 0;
    for (int x : mask) s += (x == 1);
    return s;
  };
  const long long limit = (long long)n * n;  // the reference's bound on the loop
  for (long long it = 0; it < limit && stars() < n; ++it) {
    // first uncovered zero in row-major order
    int zr = -1, zc = -1;
    for (int i = 0; i < n && zr < 0; ++i) {
      if (rc[i]) continue;
      for (int j = 0; j < n; ++j)
        if (!cc[j] && M[(size_t)i * n + j] == 0.0) { zr = i; zc = j; break; }
    }
    if (zr < 0) {
      // Step 6: smallest uncovered value (covered entries count as the maximum)
      double mx = M[0];
      for (double v : M) mx = v > mx ? v : mx;
      double mn = mx;
      bool first = true;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          const double v = (rc[i] || cc[j]) ? mx : M[(size_t)i * n + j];
          if (first || v < mn) { mn = v; first = false; }
        }
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          double v = M[(size_t)i * n + j];
          if (!cc[j]) v = v - mn;
          if (rc[i]) v = v + mn;
          M[(size_t)i * n + j] = v;
        }
      continue;
    }
    mask[(size_t)zr * n + zc] = 2;
    int star_col = -1;
    for (int j = 0; j < n; ++j)
      if (mask[(size_t)zr * n + j] == 1) { star_col = j; break; }
    if (star_col >= 0) {
      // Step 4: cover the row, uncover the star's column
      rc[zr] = 1;
      cc[star_col] = 0;
      continue;
    }
    // Step 5: augmenting path from the primed zero
    std::vector<char> path((size_t)n * n, 0);
    path[(size_t)zr * n + zc] = 1;
    int col = zc;
    for (int step = 0; step < n; ++step) {
      int row = -1;
      for (int i = 0; i < n; ++i)
        if (mask[(size_t)i * n + col] == 1) { row = i; break; }
      if (row < 0) break;
      int pcol = 0;
      for (int j = 0; j < n; ++j)
        if (mask[(size_t)row * n + j] == 2) { pcol = j; break; }
      path[(size_t)row * n + col] = 1;
      path[(size_t)row * n + pcol] = 1;
      col = pcol;
    }
    for (size_t e = 0; e < mask.size(); ++e) {
      int v = mask[e] - path[e];
      mask[e] = (v == 2) ? 0 : v;
    }
    cover_star_columns();
  }
  return stars() == n;
}

// assign[row] = the column of row's 1 in the mask, or -1 if it has none.
inline void mask_to_assign(int n, const std::vector<int>& mask, int* assign) {
  for (int i = 0; i < n; ++i) {
    assign[i] = -1;
    for (int j = 0; j < n; ++j)
      if (mask[(size_t)i * n + j] == 1) { assign[i] = j; break; }
  }
}

// trace(W M D M^T) for the assignment mask M (l x l row-major, integer
// entries), W, D: l x l row-major.  With a permutation mask this is
// sum_{a,b} W[a][b] D[pi(b)][pi(a)]; it is evaluated as written so that a
// partial mask is priced as the reference prices it.
inline double qap_objective_mask(int l, const double* W, const double* D,
                                 const std::vector<int>& M) {
  std::vector<double> T((size_t)l * l, 0.0);  // T = M D
  for (int b = 0; b < l; ++b)
    for (int i = 0; i < l; ++i) {
      const int mbi = M[(size_t)b * l + i];
      if (!mbi) continue;
      for (int j = 0; j < l; ++j) T[(size_t)b * l + j] += (double)mbi * D[(size_t)i * l + j];
    }
  double s = 0.0;
  for (int a = 0; a < l; ++a)
    for (int b = 0; b < l; ++b) {
      double tba = 0.0;  // (T M^T)[b][a]
      for (int j = 0; j < l; ++j) tba += T[(size_t)b * l + j] * (double)M[(size_t)a * l + j];
      s += W[(size_t)a * l + b] * tba;
    }
  return s;
}

}  // namespace usbs
