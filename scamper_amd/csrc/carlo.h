// Binning of scalar samples: the Carlo.rs Accumulator and Estimate
// (Carlo.rs/src/measurements.rs:60-280, Carlo.rs/src/estimate.rs:30-58).
#pragma once
#include <cmath>
#include <vector>

#include "host.h"

namespace scamper {

// Fixed-capacity binning of scalar samples (Accumulator::add).
class Accumulator {
 public:
  explicit Accumulator(uint64_t capacity) : cap_(capacity) {}
  void add(double v) {
    current_ += v;
    ++filling_;
    if (filling_ >= cap_) {
      bins_.push_back(current_ / (double)cap_);
      current_ = 0.0;
      filling_ = 0;
    }
  }
  // Accumulator::finalize (measurements.rs:162-182)
  scamper_estimate finalize() const {
    std::vector<double> all(bins_);
    if (filling_ > 0) all.push_back(current_ / ((double)filling_ * 1.0));
    return from_bins(all, autocorr_time());
  }

 private:
  // Accumulator::autocorr_time (measurements.rs:244-280)
  double autocorr_time() const {
    const size_t nb = bins_.size();
    if (nb < 2) return 1.0;
    const double n = (double)nb;
    double mu = -0.0;
    for (double b : bins_) mu += b;
    mu /= n;
    double var = -0.0;
    for (double b : bins_) var += (b - mu) * (b - mu);
    var /= (n - 1.0);
    const size_t rn = nb / 2;
    if (rn < 2) return 1.0;
    std::vector<double> r(rn);
    double rmu = -0.0;
    for (size_t i = 0; i < rn; ++i) {
      r[i] = (bins_[2 * i] + bins_[2 * i + 1]) / 2.0;
      rmu += r[i];
    }
    rmu /= (double)rn;
    double rvar = -0.0;
    for (double x : r) rvar += (x - rmu) * (x - rmu);
    rvar /= ((double)rn - 1.0);
    const double expected = var / 2.0;
    if (expected <= 0.0) return 1.0;
    const double t = rvar / expected - 1.0;
    return 0.5 * (t > 0.0 ? t : 0.0);
  }
  // Estimate::from_bins_with_autocorr (estimate.rs:30-58)
  static scamper_estimate from_bins(const std::vector<double> &bins, double tau) {
    scamper_estimate e{0.0, 0.0, 1.0, 0};
    if (bins.empty()) return e;
    const double n = (double)bins.size();
    double sum = -0.0;
    for (double b : bins) sum += b;
    const double mean = sum / n;
    double var = 0.0;
    if (n > 1.0) {
      double ss = -0.0;
      for (double b : bins) ss += (b - mean) * (b - mean);
      var = ss / (n - 1.0);
    }
    e.mean = mean;
    e.stderr_ = std::sqrt(var) / std::sqrt(n);
    e.autocorr_time = tau > 1.0 ? tau : 1.0;
    e.n_bins = bins.size();
    return e;
  }

  uint64_t cap_;
  std::vector<double> bins_;
  double current_ = 0.0;
  uint64_t filling_ = 0;
};

// DefaultObservableSet (CMC.rs/src/observables/common.rs:81-96): Energy, E2,
// Magnetization, M2, M4 of one measured sweep, moments by powi.
class DefaultObservables {
 public:
  explicit DefaultObservables(uint64_t binsize) : e_(binsize), e2_(binsize), m_(binsize), m2_(binsize), m4_(binsize) {}
  void add(double e, double m) {
    e_.add(e);
    e2_.add(e * e);
    m_.add(m);
    const double m2 = m * m;
    m2_.add(m2);
    m4_.add(m2 * m2);
  }
  void finalize(scamper_estimate out[5]) const {
    out[0] = e_.finalize();
    out[1] = e2_.finalize();
    out[2] = m_.finalize();
    out[3] = m2_.finalize();
    out[4] = m4_.finalize();
  }

 private:
  Accumulator e_, e2_, m_, m2_, m4_;
};

scamper_estimate binned_estimate(const double *samples, uint64_t n, uint64_t binsize);
void default_observables(const double *e, const double *m, uint64_t n, uint64_t binsize, scamper_estimate out[5]);

}  // namespace scamper
