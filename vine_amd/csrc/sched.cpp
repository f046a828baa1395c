// DOP / AR scheduling rule (host).  See sched.hpp.
#include "sched.hpp"

#include <algorithm>
#include <cmath>

#include "vine.h"

namespace vine {

// Burg's recursion.  Forward/backward prediction errors f, b start as the
// series itself; each order k picks the reflection coefficient minimising the
// summed forward+backward error power and updates the lower-order
// coefficients Levinson-style.  The operation order follows the reference
// (include/AR.cpp:105-147) so both produce identical doubles.
bool ar_burg(const double* x, int n, int degree, double* coef) {
  if (n <= degree || degree < 1) return false;
  double mean = 0.0;
  for (int t = 0; t < n; t++) mean += x[t];
  mean /= (double)n;
  std::vector<double> w(n), fe(n + 1, 0.0), be(n + 1, 0.0), g(degree + 2, 0.0), h(degree + 2, 0.0);
  for (int t = 0; t < n; t++) w[t] = x[t] - mean;
  for (int k = 1; k <= degree; k++) {
    const int lag = k - 1;
    int span = n - lag - 1;
    double num = 0.0, den = 0.0;
    for (int j = 1; j <= span; j++) {
      const double f = w[j + lag] + fe[j];
      const double b = w[j - 1] + be[j];
      num -= 2.0 * f * b;
      den += (f * f) + (b * b);
    }
    const double rc = num / den;
    g[k + 1] = rc;
    if (lag != 0) {
      for (int j = 2; j <= k; j++) h[j] = g[j] + (rc * g[lag - j + 3]);
      for (int j = 2; j <= k; j++) g[j] = h[j];
      span--;
    }
    for (int j = 1; j <= span; j++) {
      be[j] += (rc * fe[j]) + (rc * w[j + k - 1]);
      fe[j] = fe[j + 1] + (rc * be[j + 1]) + (rc * w[j]);
    }
  }
  for (int j = 1; j <= degree; j++) coef[j - 1] = -g[j + 1];
  return true;
}

void DopRule::reset(double max_labels) {
  max_ = max_labels;
  dop_ = (int64_t)(max_labels / kIter);
  if (dop_ < 1) dop_ = 1;
  hist_.clear();
}

void DopRule::update(int64_t new_size) {
  const double lb = std::log2((double)new_size);
  const size_t k = hist_.size();
  // ending rule: three consecutive decreases -> widen by half
  if (k >= 3 && lb < hist_[k - 1] && hist_[k - 1] < hist_[k - 2] && hist_[k - 2] < hist_[k - 3]) {
    dop_ = (int64_t)std::min<double>(max_, (double)(dop_ + dop_ / 2));
    hist_.push_back(lb);
    return;
  }
  hist_.push_back(lb);
  if ((int)hist_.size() <= kDegree) return;
  double c[kDegree];
  if (!ar_burg(hist_.data(), (int)hist_.size(), kDegree, c)) return;
  // prediction on the raw (not mean-removed) series, as the reference does
  double est = 0.0;
  for (int j = 0; j < kDegree; j++) est += c[j] * hist_[hist_.size() - j - 1];
  est = std::exp2(est);
  if (est > (double)new_size) {
    double p = max_ / kLabelBytes / (est / (double)new_size - kTheta);
    if (p < 0) p = max_;
    p = std::min(p, max_);
    dop_ = (int64_t)std::llround(p);
    if (dop_ < 1) dop_ = 1;
  }
}

}  // namespace vine

extern "C" int vine_ar_burg(const double* x, int n, int degree, double* coef) {
  if (!x || !coef) return -1;
  return vine::ar_burg(x, n, degree, coef) ? 0 : -1;
}

extern "C" int64_t vine_dop_sequence(const int64_t* sizes, int n, double max_labels, int64_t* out_dop) {
  if (!sizes || !out_dop || n < 0 || !(max_labels > 0)) return -1;
  vine::DopRule r;
  r.reset(max_labels);
  for (int i = 0; i < n; i++) {
    r.update(sizes[i]);
    out_dop[i] = r.dop();
  }
  return n;
}
