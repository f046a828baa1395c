// Host-side bucket-width scheduling: the reference's DOP rule driven by an
// autoregressive model of the log2 frontier size (algorithm/cspp.cuh:33-47,
// 127-138, 254-284; include/AR.cpp:6-147).
#pragma once

#include <cstdint>
#include <vector>

namespace vine {

// Burg (maximum-entropy) AR fit of `degree` coefficients to the mean-removed
// series x[0..n); coef[j] multiplies x[t-1-j].  Returns false if n <= degree.
bool ar_burg(const double* x, int n, int degree, double* coef);

// The DOP ("degree of parallelism") rule.  dop() is the number of frontier
// labels to expand next round; update() is fed the new frontier size.
class DopRule {
 public:
  void reset(double max_labels);
  int64_t dop() const { return dop_; }
  void update(int64_t new_size);
  const std::vector<double>& history() const { return hist_; }

  static constexpr int kDegree = 3;       // AR_DEGREE, algorithm/cspp.cuh:26
  static constexpr double kTheta = 0.95;  // THETA, algorithm/cspp.cuh:27
  static constexpr int kIter = 1000;      // ITER, algorithm/cspp.cuh:24
  static constexpr int kLabelBytes = 12;  // sizeof(c_label)

 private:
  double max_ = 0;
  int64_t dop_ = 0;
  std::vector<double> hist_;
};

}  // namespace vine
