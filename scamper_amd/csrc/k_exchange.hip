// Replica exchange for batches of independent replicas (scamper_state):
// parallel tempering over a temperature ladder held on the device.
//
// PtExchange::try_exchange (Carlo.rs/src/parallel_tempering.rs:362-419)
// pairs chains by temperature label with offset (step - 1) & 1; each side
// contributes w = (beta_self - beta_partner) * E_self, the log_weight_ratio
// of ClassicalMC (CMC.rs/src/classical_mc.rs:209-214), and the lower label
// decides with accept_log_probability(w_lower + w_upper, u)
// (Carlo.rs/src/acceptance.rs:11-22).  An accepted pair swaps labels;
// change_parameter (classical_mc.rs:216-228) then sets the new beta, which
// here is a copy of the label's decision tables into the replica's
// parameters (k_apply_labels).  Energies are exact recounts of the current
// spins, which is what change_parameter's recompute_energy makes the cached
// energy equal to.
//
// Built without floating-point contraction, so the energies are rounded as
// the host's observables are.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

namespace scamper {
namespace {

// Energy of one replica from its recount counters, with the host's own
// expressions (observables_from_em in capi.cpp).
__device__ double energy_of(const long long *em, const EnergyForm &f) {
  switch (f.kind) {
    case kEnergyStencil:
      return -f.j * (double)em[0];
    case kEnergyGraph:
      return -f.j * ((double)(em[0] / 2) + f.self_loops);
    case kEnergyGraphReal: {
      const double bonds = __longlong_as_double(em[0]);
      return -f.j * (0.5 * bonds + f.self_loops);
    }
    case kEnergyVecStencil:
      return -__longlong_as_double(em[0]);
    default: {  // kEnergyVecGraph
      const double d0 = __longlong_as_double(em[0]);
      return -0.5 * d0 - f.j * f.self_loops;
    }
  }
}

__global__ __launch_bounds__(1024) void k_state_exchange(int32_t *__restrict__ labels, const long long *__restrict__ em,
                                                         uint32_t n, EnergyForm form, const double *__restrict__ ladder,
                                                         uint64_t ex_key, uint64_t step, double *__restrict__ energy_out,
                                                         unsigned long long *__restrict__ stats,
                                                         const StepClock *__restrict__ clk) {
  if (clk) step = clk->exchange;
  __shared__ int32_t replica_of[1024];
  __shared__ double energy[1024];
  const uint32_t t = threadIdx.x;
  if (t < n) {
    energy[t] = energy_of(em + (uint64_t)t * kEmStride, form);
    energy_out[t] = energy[t];
    replica_of[labels[t]] = (int32_t)t;
  }
  __syncthreads();
  if (t == 0) atomicAdd(&stats[2], 1ull);
  const uint32_t offset = (uint32_t)((step > 0 ? step - 1 : 0) & 1u);
  const uint32_t lower = offset + 2 * t;
  if (n < 2 || lower + 1 >= n) return;
  const int32_t a = replica_of[lower], b = replica_of[lower + 1];
  const double bl = ladder[lower], bu = ladder[lower + 1];
  const double wa = (bl - bu) * energy[a];
  const double wb = (bu - bl) * energy[b];
  const double d = wa + wb;
  const U4 blk = philox4x32_10(lower, (uint32_t)step, ctr2(step, 0, kStreamExchange), 0u, (uint32_t)ex_key,
                               (uint32_t)(ex_key >> 32));
  const double u = (double)blk.x * 2.3283064365386963e-10;
  bool accept;
  if (isnan(d))
    accept = false;
  else if (d >= 0.0)
    accept = true;
  else if (d == -INFINITY)
    accept = false;
  else {
    const double lu = log(fmax(u, 2.2250738585072014e-308));
    accept = lu < d;
    // decisions closer than a few ulps are counted: the host libm could
    // round ln(u) differently there
    if (fabs(lu - d) <= 1e-13 * fmax(1.0, fabs(d))) atomicAdd(&stats[1], 1ull);
  }
  if (accept) {
    labels[a] = (int32_t)(lower + 1);
    labels[b] = (int32_t)lower;
    atomicAdd(&stats[0], 1ull);
  }
}

// Replica r takes the decision tables of the label it holds: beta and
// thresholds of its parameters (its key stays), and where present its tiled
// row's thresholds and byte tables and its Potts heat-bath weights.
__global__ __launch_bounds__(256) void k_apply_labels(const int32_t *__restrict__ labels, uint32_t n,
                                                      ReplicaParams *__restrict__ params,
                                                      const ReplicaParams *__restrict__ label_params, TileJob *rows,
                                                      const TileJob *__restrict__ label_rows, double *hbexp,
                                                      const double *__restrict__ label_hbexp, uint32_t hb_per) {
  const uint32_t r = blockIdx.x;
  if (r >= n) return;
  const int32_t l = labels[r];
  const uint32_t t = threadIdx.x;
  if (t == 0) params[r].beta = label_params[l].beta;
  if (t < (uint32_t)kMaxTable) params[r].thr[t] = label_params[l].thr[t];
  if (rows) {
    if (t < (uint32_t)kStencilTable) rows[r].thr[t] = label_rows[l].thr[t];
    if (t < 4) {
      rows[r].lut_k8[t] = label_rows[l].lut_k8[t];
      rows[r].lut_m8[t] = label_rows[l].lut_m8[t];
    }
  }
  if (hbexp)
    for (uint32_t k = t; k < hb_per; k += blockDim.x) hbexp[(uint64_t)r * hb_per + k] = label_hbexp[(uint64_t)l * hb_per + k];
}

__global__ void k_clock_set(StepClock *clk, uint64_t sweep, uint64_t exchange, uint64_t slot, void *series) {
  if (threadIdx.x == 0) *clk = StepClock{sweep, exchange, slot, series};
}

__global__ void k_clock_tick(StepClock *clk, uint64_t d_exchange, uint64_t d_slot) {
  if (threadIdx.x == 0) {
    clk->sweep += 1;
    clk->exchange += d_exchange;
    clk->slot += d_slot;
  }
}

__global__ __launch_bounds__(256) void k_series_from_em(const long long *__restrict__ em, const StepClock *clk,
                                                        uint64_t count) {
  long long *out = static_cast<long long *>(clk->series) + clk->slot * count;
  for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < count; k += (uint64_t)gridDim.x * 256) out[k] = em[k];
}

}  // namespace

hipError_t launch_clock_set(StepClock *clk, uint64_t sweep, uint64_t exchange, uint64_t slot, void *series,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(64), 0, st, clk, sweep, exchange, slot, series);
  return hipGetLastError();
}

hipError_t launch_clock_tick(StepClock *clk, uint64_t d_exchange, uint64_t d_slot, hipStream_t st) {
  hipLaunchKernelGGL(k_clock_tick, dim3(1), dim3(64), 0, st, clk, d_exchange, d_slot);
  return hipGetLastError();
}

hipError_t launch_series_from_em(const long long *em, const StepClock *clk, uint64_t count, hipStream_t st) {
  const unsigned blocks = (unsigned)std::min<uint64_t>((count + 255) / 256, 64);
  hipLaunchKernelGGL(k_series_from_em, dim3(blocks), dim3(256), 0, st, em, clk, count);
  return hipGetLastError();
}

hipError_t launch_state_exchange(int32_t *labels, const long long *em, uint32_t n, const EnergyForm &form,
                                 const double *ladder, uint64_t ex_key, uint64_t step, double *energy_out,
                                 unsigned long long *stats, hipStream_t st, const StepClock *clk) {
  hipLaunchKernelGGL(k_state_exchange, dim3(1), dim3(1024), 0, st, labels, em, n, form, ladder, ex_key, step,
                     energy_out, stats, clk);
  return hipGetLastError();
}

hipError_t launch_apply_labels(const int32_t *labels, uint32_t n, ReplicaParams *params,
                               const ReplicaParams *label_params, TileJob *rows, const TileJob *label_rows,
                               double *hbexp, const double *label_hbexp, uint32_t hb_per, hipStream_t st) {
  hipLaunchKernelGGL(k_apply_labels, dim3(n), dim3(256), 0, st, labels, n, params, label_params, rows, label_rows,
                     hbexp, label_hbexp, hb_per);
  return hipGetLastError();
}

}  // namespace scamper
