// Ising and Potts single-site updates on graph lattices with real bond
// weights (any finite CsrLattice weight, graph.rs:112-124), in f64 in the
// reference's operation order: this file is compiled without contraction.
//
// Per site, in the class order of the lattice (one launch per colour class,
// draws keyed by colour and rank as on the integer path):
// * Metropolis: ΔE = local(new) - local(old), each local energy summed from
//   0.0 over the site's incidences in CSR order, a self-loop adding half its
//   bond energy (interaction.rs:73-84, 149-181; models.rs:58-60, 230-236);
//   log A = (-β ΔE + 0) + 0 (ensemble.rs:196-204) and accept_log_probability
//   (acceptance.rs:11-22) with u = word 2^-32 of the acceptance stream; the
//   Potts proposal is floor(u' (q-1)), skipping the current state.
// * Heat bath: the Ising logistic p+ (models.rs:158-186) and the Potts
//   softmax with max-shift and sequential subtraction (models.rs:340-380),
//   self-loops left out of the fields.
// Spins are int8 (Ising +-1, Potts 0..q-1), site-major per replica.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kBlock = 256;
constexpr double kTwoM32 = 2.3283064365386963e-10;

__device__ __forceinline__ double bond(int kind, double j, double w, int l, int r) {
  return kind == kIsing ? -j * w * (double)l * (double)r : (l == r ? -j * w : 0.0);
}

// local(a) and local(b) in one pass over the incidences: each sum is still
// taken from 0.0 in CSR order, so both equal separate passes bit for bit.
__device__ void local_energies(const RealCsrGeom &g, int kind, double j, const int8_t *s, uint32_t site, int a, int b,
                               double *ea, double *eb) {
  double sa = 0.0, sb = 0.0;
  for (uint32_t i = g.offsets[site]; i < g.offsets[site + 1]; ++i) {
    const uint32_t t = g.neighbors[i];
    const double w = g.weights[i];
    if (t == site) {
      sa += 0.5 * bond(kind, j, w, a, a);
      sb += 0.5 * bond(kind, j, w, b, b);
    } else {
      const int n = s[t];
      sa += bond(kind, j, w, a, n);
      sb += bond(kind, j, w, b, n);
    }
  }
  *ea = sa;
  *eb = sb;
}

__device__ __forceinline__ bool accept_log(double log_a, double u) {
  if (isnan(log_a)) return false;
  if (log_a >= 0.0) return true;
  if (log_a == -INFINITY) return false;
  return log(fmax(u, 2.2250738585072014e-308)) < log_a;
}

template <int KIND, int ALG>
__global__ __launch_bounds__(kBlock) void k_csr_real(int8_t *__restrict__ spins, const ReplicaParams *__restrict__ reps,
                                                     RealCsrGeom g, uint32_t colour, uint64_t begin, uint64_t count,
                                                     uint64_t sweep, int q, double j) {
  const uint32_t rep = blockIdx.y;
  const uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= count) return;
  const ReplicaParams &P = reps[rep];
  int8_t *__restrict__ s = spins + (uint64_t)rep * g.n_sites;
  const uint32_t site = g.class_sites[begin + r];
  const int old = s[site];
  const double u = (double)site_word(P.key, r, sweep, colour, kStreamAccept) * kTwoM32;
  int value = old;
  if (ALG == kMetropolis) {
    int proposed;
    if (KIND == kIsing) {
      proposed = -old;
    } else {
      const uint32_t pw = site_word(P.key, r, sweep, colour, kStreamPropose);
      proposed = (int)(((uint64_t)pw * (uint64_t)(q - 1)) >> 32);
      if (proposed >= old) proposed += 1;
    }
    double e_new, e_old;
    local_energies(g, KIND, j, s, site, proposed, old, &e_new, &e_old);
    const double de = e_new - e_old;
    const double term = de == INFINITY ? -INFINITY : (de == -INFINITY ? INFINITY : -P.beta * de);
    if (accept_log((term + 0.0) + 0.0, u)) value = proposed;
  } else if (KIND == kIsing) {
    double field = -0.0;
    for (uint32_t i = g.offsets[site]; i < g.offsets[site + 1]; ++i) {
      const uint32_t t = g.neighbors[i];
      if (t != site) field += j * g.weights[i] * (double)s[t];
    }
    const double x = 2.0 * P.beta * field;
    const double p_plus = x >= 0.0 ? 1.0 / (1.0 + exp(-x)) : exp(x) / (1.0 + exp(x));
    value = u < p_plus ? 1 : -1;
  } else {
    double lw[kMaxQ];
    for (int k = 0; k < q; ++k) lw[k] = 0.0;
    for (uint32_t i = g.offsets[site]; i < g.offsets[site + 1]; ++i) {
      const uint32_t t = g.neighbors[i];
      if (t == site) continue;
      const int st = s[t];
      if (st < q) lw[st] += P.beta * j * g.weights[i];
    }
    double max_log = -INFINITY;
    for (int k = 0; k < q; ++k) max_log = fmax(max_log, lw[k]);
    double total = 0.0;
    for (int k = 0; k < q; ++k) {
      lw[k] = exp(lw[k] - max_log);
      total += lw[k];
    }
    double threshold = u * total;
    value = q - 1;
    for (int k = 0; k < q; ++k) {
      threshold -= lw[k];
      if (threshold <= 0.0) {
        value = k;
        break;
      }
    }
  }
  if (value != old) s[site] = (int8_t)value;
}

// em[0] (as a double) = sum over sites and non-loop incidences of w s_i s_j
// (Ising) or w delta(s_i, s_j) (Potts), every bond twice; em[1] = sum of
// spins (Ising) or em[2 + v] = sites in state v (Potts).
__global__ __launch_bounds__(kBlock) void k_csr_real_measure(const int8_t *__restrict__ spins, RealCsrGeom g, int kind,
                                                             int q, long long *__restrict__ em) {
  const uint32_t rep = blockIdx.y;
  const uint64_t site = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  double bonds = 0.0;
  int value = 0;
  const bool live = site < g.n_sites;
  if (live) {
    const int8_t *s = spins + (uint64_t)rep * g.n_sites;
    value = s[site];
    for (uint32_t i = g.offsets[site]; i < g.offsets[site + 1]; ++i) {
      const uint32_t t = g.neighbors[i];
      if (t == site) continue;
      bonds += kind == kIsing ? g.weights[i] * (double)(value * (int)s[t]) : ((int)s[t] == value ? g.weights[i] : 0.0);
    }
  }
  for (int off = 32; off > 0; off >>= 1) bonds += __shfl_xor(bonds, off);
  long long *rec = em + (uint64_t)rep * kEmStride;
  if ((threadIdx.x & 63) == 0) atomicAdd(reinterpret_cast<double *>(rec), bonds);
  if (!live) return;
  if (kind == kIsing) atomicAdd(reinterpret_cast<unsigned long long *>(rec + 1), (unsigned long long)(long long)value);
  else atomicAdd(reinterpret_cast<unsigned long long *>(rec + 2 + value), 1ull);
}

}  // namespace

hipError_t launch_csr_real(const RealCsrGeom &g, const uint64_t *class_start_host, uint32_t n_colours, int kind, int q,
                           int algorithm, int8_t *spins, const ReplicaParams *reps, uint32_t n_rep, uint64_t sweep,
                           double j, hipStream_t st) {
  for (uint32_t c = 0; c < n_colours; ++c) {
    const uint64_t begin = class_start_host[c], count = class_start_host[c + 1] - begin;
    if (count == 0) continue;
    const dim3 grid((unsigned)((count + kBlock - 1) / kBlock), n_rep);
#define SC_REAL(K, A) \
  hipLaunchKernelGGL((k_csr_real<K, A>), grid, dim3(kBlock), 0, st, spins, reps, g, c, begin, count, sweep, q, j)
    if (kind == kIsing && algorithm == kMetropolis) SC_REAL(kIsing, kMetropolis);
    else if (kind == kIsing) SC_REAL(kIsing, kHeatBath);
    else if (algorithm == kMetropolis) SC_REAL(kPotts, kMetropolis);
    else SC_REAL(kPotts, kHeatBath);
#undef SC_REAL
  }
  return hipGetLastError();
}

hipError_t launch_csr_real_measure(const RealCsrGeom &g, int kind, int q, const int8_t *spins, uint32_t n_rep,
                                   long long *em, hipStream_t st) {
  const dim3 grid((unsigned)((g.n_sites + kBlock - 1) / kBlock), n_rep);
  hipLaunchKernelGGL(k_csr_real_measure, grid, dim3(kBlock), 0, st, spins, g, kind, q, em);
  return hipGetLastError();
}

}  // namespace scamper
