// Host-side statistics and the run driver: the Carlo.rs layer above the
// sweep.
//
//   Accumulator / Estimate  -- Carlo.rs/src/measurements.rs:60-280,
//                              Carlo.rs/src/estimate.rs:30-58
//   run_one                 -- Carlo.rs/src/scheduler.rs:94-155 driving
//                              ClassicalMC (CMC.rs/src/classical_mc.rs:168-229,
//                              427-490) with DefaultObservableSet
//                              (observables/common.rs:81-96)
//
// The per-sweep energy and magnetisation come from the device series; the
// binning is FP64 on the host with the reference's fold order.
#include <cmath>
#include <cstring>
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

scamper_estimate binned_estimate(const double *samples, uint64_t n, uint64_t binsize) {
  Accumulator acc(binsize);
  for (uint64_t i = 0; i < n; ++i) acc.add(samples[i]);
  return acc.finalize();
}

// The five DefaultObservableSet estimates (Energy, E2, Magnetization, M2,
// M4; moments by powi, observables/common.rs:88-92) of a measured series.
void default_observables(const double *e, const double *m, uint64_t n, uint64_t binsize,
                         scamper_estimate out[5]) {
  Accumulator ae(binsize), ae2(binsize), am(binsize), am2(binsize), am4(binsize);
  for (uint64_t t = 0; t < n; ++t) {
    ae.add(e[t]);
    ae2.add(e[t] * e[t]);
    am.add(m[t]);
    const double m2 = m[t] * m[t];
    am2.add(m2);
    am4.add(m2 * m2);
  }
  out[0] = ae.finalize();
  out[1] = ae2.finalize();
  out[2] = am.finalize();
  out[3] = am2.finalize();
  out[4] = am4.finalize();
}

}  // namespace scamper

extern "C" int scamper_binned_estimate(const double *samples, uint64_t n, uint64_t binsize,
                                       scamper_estimate *out) {
  if (!out || (n && !samples)) return scamper::fail(SCAMPER_EINVAL, "null argument");
  if (binsize == 0) return scamper::fail(SCAMPER_EINVAL, "binsize must be positive");
  *out = scamper::binned_estimate(samples, n, binsize);
  return SCAMPER_OK;
}

extern "C" int scamper_run_one(const scamper_lattice *lat, const scamper_model *model, int algorithm,
                               double beta, int random_start, const scamper_run_config *cfg,
                               scamper_estimate out[5], double *energy, double *magnetization) {
  if (!lat || !model || !cfg || !out) return scamper::fail(SCAMPER_EINVAL, "null argument");
  if (cfg->binsize == 0) return scamper::fail(SCAMPER_EINVAL, "binsize must be positive");
  scamper_state *st = nullptr;
  int rc = scamper_state_create(lat, model, 1, cfg->base_seed, &st);
  if (rc) return rc;
  struct Guard {
    scamper_state *s;
    ~Guard() { scamper_state_free(s); }
  } guard{st};
  if ((rc = scamper_state_set_beta(st, 0, beta))) return rc;
  if (random_start && (rc = scamper_state_init(st, 0, 1))) return rc;
  // thermalisation (scheduler.rs:109-117), then measured sweeps (121-133)
  if (cfg->thermalization_sweeps && (rc = scamper_sweep(st, algorithm, cfg->thermalization_sweeps, 0))) return rc;
  if (cfg->measurement_sweeps && (rc = scamper_sweep(st, algorithm, cfg->measurement_sweeps, 1))) return rc;
  const uint64_t n = cfg->measurement_sweeps;
  std::vector<double> e(n), m(n);
  if (n && (rc = scamper_series_get(st, 0, e.data(), m.data(), n))) return rc;
  scamper::default_observables(e.data(), m.data(), n, cfg->binsize, out);
  if (energy) std::memcpy(energy, e.data(), n * sizeof(double));
  if (magnetization) std::memcpy(magnetization, m.data(), n * sizeof(double));
  return SCAMPER_OK;
}
