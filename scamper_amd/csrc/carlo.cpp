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

#include "carlo.h"

namespace scamper {

scamper_estimate binned_estimate(const double *samples, uint64_t n, uint64_t binsize) {
  Accumulator acc(binsize);
  for (uint64_t i = 0; i < n; ++i) acc.add(samples[i]);
  return acc.finalize();
}

// The five DefaultObservableSet estimates (Energy, E2, Magnetization, M2,
// M4; moments by powi, observables/common.rs:88-92) of a measured series.
void default_observables(const double *e, const double *m, uint64_t n, uint64_t binsize,
                         scamper_estimate out[5]) {
  DefaultObservables obs(binsize);
  for (uint64_t t = 0; t < n; ++t) obs.add(e[t], m[t]);
  obs.finalize(out);
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
