// Integer decision thresholds for the discrete kernels.
//
// The reference decides Metropolis moves as ln(max(u, MIN_POSITIVE)) < -βΔE
// (Carlo.rs/src/acceptance.rs:11-22, core/ensemble.rs:196-204) with f64 and
// the platform libm.  For u = k * 2^-32 the left side is monotone in k, so
// each decision is "k < K" for an integer K found here by bisection with the
// same glibc log; kernels then need no transcendental at all and agree with
// the reference rule bit for bit.  Heat-bath decisions "u < p" are handled
// the same way with p computed by the reference's formulas.
//
// This file is compiled with -ffp-contract=off so that every f64 expression
// rounds exactly like the reference's source.
#include <cfloat>
#include <cmath>

#include "host.h"

namespace scamper {

namespace {
constexpr uint64_t kTwo32 = 1ull << 32;
constexpr double kTwoM32 = 2.3283064365386963e-10;

bool accepts(double a, uint64_t k) {
  if (std::isnan(a)) return false;
  if (a >= 0.0) return true;
  if (a == -INFINITY) return false;
  return std::log(std::fmax((double)k * kTwoM32, DBL_MIN)) < a;
}
}  // namespace

// canonical_energy_term plus the zero log-Jacobian and proposal ratio
// (core/ensemble.rs:90-95,196-204; core/acceptance.rs:25-35).
double log_acceptance(double beta, double de) {
  double t;
  if (de == INFINITY) t = -INFINITY;
  else if (de == -INFINITY) t = INFINITY;
  else t = -beta * de;
  return (t + 0.0) + 0.0;
}

uint64_t accept_threshold(double a) {
  if (!accepts(a, 0)) return 0;
  if (accepts(a, kTwo32 - 1)) return kTwo32;
  uint64_t lo = 0, hi = kTwo32 - 1;  // accepts(lo), !accepts(hi)
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (accepts(a, mid)) lo = mid; else hi = mid;
  }
  return hi;
}

uint64_t cluster_bond_threshold(int kind, double beta, double j, double weight) {
  // cluster_bond_probability: coupling = J w; Ising 1 - exp(-2 beta c)
  // (models.rs:130-147), Potts 1 - exp(-beta c) (models.rs:312-329), then
  // clamp_probability (models.rs:16-24)
  const double coupling = j * weight;
  if (coupling == 0.0) return 0;
  const double arg = kind == kIsing ? -2.0 * beta * coupling : -beta * coupling;
  double p = 1.0 - std::exp(arg);
  if (std::isfinite(p)) p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);
  else p = p > 0.0 ? 1.0 : 0.0;
  return below_threshold(p);
}

uint64_t below_threshold(double p) {
  if (!(p > 0.0)) return 0;
  if ((double)(kTwo32 - 1) * kTwoM32 < p) return kTwo32;
  uint64_t lo = 0, hi = kTwo32 - 1;
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if ((double)mid * kTwoM32 < p) lo = mid; else hi = mid;
  }
  return hi;
}

// ΔE of an Ising flip with `anti` of z neighbours anti-aligned, summed in the
// order of PairInteraction::local_energy (interaction.rs:149-181): the
// anti-aligned incidences first.  Every partial sum is a multiple of J, so
// for |J| a power of two (J = 1 in every configured workload) the value does
// not depend on the order at all.
void ising_metropolis_table(double beta, double j, int z, Threshold *thr) {
  for (int anti = 0; anti <= z; ++anti) {
    double e_new = 0.0, e_old = 0.0;  // s = +1 without loss of generality
    for (int k = 0; k < z; ++k) {
      const double nb = k < anti ? -1.0 : 1.0;
      e_new += -j * 1.0 * (-1.0) * nb;
      e_old += -j * 1.0 * 1.0 * nb;
    }
    thr[anti] = make_threshold(accept_threshold(log_acceptance(beta, e_new - e_old)));
  }
}

// IsingModel::heat_bath_sample_site (models.rs:158-186): +1 iff u < p_plus,
// indexed by the number of -1 neighbours.
void ising_heat_bath_table(double beta, double j, int z, Threshold *thr) {
  for (int minus = 0; minus <= z; ++minus) {
    double field = 0.0;
    for (int k = 0; k < z; ++k) field += j * 1.0 * (k < minus ? -1.0 : 1.0);
    const double x = 2.0 * beta * field;
    const double p = x >= 0.0 ? 1.0 / (1.0 + std::exp(-x)) : std::exp(x) / (1.0 + std::exp(x));
    thr[minus] = make_threshold(below_threshold(p));
  }
}

// Graph lattices with integer bond weights: a site's weighted field
// h = sum w s_nb lies in [-W, W].  Metropolis is indexed by x = s h + W
// (ΔE = 2 J s h: every partial sum of ±J w is an integer multiple of J),
// heat bath by h + W.
void csr_ising_metropolis_table(double beta, double j, int w_max, Threshold *thr) {
  for (int x = -w_max; x <= w_max; ++x) {
    const double e_old = -j * (double)x, e_new = j * (double)x;  // local energies of s and -s
    thr[x + w_max] = make_threshold(accept_threshold(log_acceptance(beta, e_new - e_old)));
  }
}

void csr_ising_heat_bath_table(double beta, double j, int w_max, Threshold *thr) {
  for (int h = -w_max; h <= w_max; ++h) {
    const double field = j * (double)h;
    const double x = 2.0 * beta * field;
    const double p = x >= 0.0 ? 1.0 / (1.0 + std::exp(-x)) : std::exp(x) / (1.0 + std::exp(x));
    thr[h + w_max] = make_threshold(below_threshold(p));
  }
}

// Potts Metropolis (models.rs:230-236): ΔE = -J (n_new - n_old), indexed by
// d = n_new - n_old + z.
void potts_metropolis_table(double beta, double j, int z, Threshold *thr) {
  for (int d = -z; d <= z; ++d) {
    double e_new = 0.0, e_old = 0.0;
    for (int k = 0; k < (d > 0 ? d : 0); ++k) e_new += -j * 1.0;
    for (int k = 0; k < (d < 0 ? -d : 0); ++k) e_old += -j * 1.0;
    thr[d + z] = make_threshold(accept_threshold(log_acceptance(beta, e_new - e_old)));
  }
}

// Potts heat bath (models.rs:340-380): log weights are sums of n copies of
// beta*J*w; the sampler subtracts exp(lw(n) - lw(nmax)) in state order.
void potts_heat_bath_exp(double beta, double j, int z, double *table) {
  for (int n = 0; n <= z; ++n)
    for (int m = 0; m <= z; ++m) {
      double ln = 0.0, lm = 0.0;
      for (int k = 0; k < n; ++k) ln += beta * j * 1.0;
      for (int k = 0; k < m; ++k) lm += beta * j * 1.0;
      table[n * (z + 1) + m] = std::exp(ln - lm);
    }
}

}  // namespace scamper
