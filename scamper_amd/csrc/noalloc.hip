// scamper_ising_sweep_noalloc: the register-blocked Ising Metropolis sweep of
// replicas held in a caller-owned device buffer, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// scamper_sweep keeps each replica's decision tables and measurement
// counters in buffers its state allocates.  Here they live in one fixed,
// library-owned device array, `scratch`, of kScratchBytes: up to
// kMaxReplicas replicas' ReplicaParams, then their counters.  Each GPU holds
// its own copy of the array until the process ends; its device address is
// looked up once per device and used everywhere host code needs the array.
// Every call on one GPU shares it and leaves its contents behind: the
// counters are cleared by a kernel before each measured sweep, and calls
// that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "host.h"

namespace scamper {
namespace {

constexpr uint32_t kMaxReplicas = 64;
constexpr size_t kParamsBytes = (size_t)kMaxReplicas * sizeof(ReplicaParams);
constexpr size_t kEmOffset = (kParamsBytes + 255) & ~(size_t)255;
constexpr size_t kScratchBytes = kEmOffset + (size_t)kMaxReplicas * kEmStride * sizeof(long long);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on
// that device.
int scratch_address(void **out) {
  int dev = 0;
  SC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SCAMPER_EUNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(SCAMPER_EDEVICE, "scratch address: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SCAMPER_OK;
}

}  // namespace
}  // namespace scamper

using namespace scamper;

extern "C" int scamper_ising_sweep_noalloc(int8_t *d_spins, const uint64_t *dims, int n_dims, double j,
                                           const double *beta, uint32_t n_rep, uint64_t seed, uint32_t replica0,
                                           uint64_t sweep0, uint64_t n_sweeps, double *energy, double *magnetization,
                                           void *stream) {
  if (!d_spins || !dims || !beta) return fail(SCAMPER_EINVAL, "null argument");
  if (n_rep == 0 || n_rep > kMaxReplicas)
    return fail(SCAMPER_EINVAL, "n_rep must be in 1..%u (the fixed scratch array holds %u replicas)", kMaxReplicas,
                kMaxReplicas);
  if (n_dims != 2 && n_dims != 3) return fail(SCAMPER_EUNSUPPORTED, "2-D or 3-D lattices only");
  if (!std::isfinite(j)) return fail(SCAMPER_EINVAL, "coupling must be finite");
  StencilGeom g{};
  g.nd = n_dims;
  for (int d = 0; d < n_dims; ++d)
    if (dims[d] < 2 || dims[d] % 2 || dims[d] > (1ull << 31))
      return fail(SCAMPER_EINVAL, "sides must be even and in 2..2^31 (side %d is %llu)", d,
                  (unsigned long long)dims[d]);
  g.lx2 = (uint32_t)(dims[0] / 2);
  g.ly = (uint32_t)dims[1];
  g.lz = n_dims == 3 ? (uint32_t)dims[2] : 1u;
  g.nh = (uint64_t)g.lx2 * g.ly * g.lz;
  g.cs = g.nh;  // the caller's colour arrays are contiguous
  set_divisors(g);
  if (!ising_fast_path_ok(g))
    return fail(SCAMPER_EUNSUPPORTED, "the no-allocation sweep needs Lx a multiple of 32 (got %llu)",
                (unsigned long long)dims[0]);
  for (uint32_t r = 0; r < n_rep; ++r)
    if (!std::isfinite(beta[r]) || beta[r] < 0.0) return fail(SCAMPER_EINVAL, "beta must be finite and non-negative");
  void *base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  ReplicaParams *d_params = static_cast<ReplicaParams *>(base);
  long long *d_em = reinterpret_cast<long long *>(static_cast<unsigned char *>(base) + kEmOffset);
  hipStream_t st = static_cast<hipStream_t>(stream);

  // decision tables on the host (no allocation), then into the array
  ReplicaParams p[kMaxReplicas];
  for (uint32_t r = 0; r < n_rep; ++r) {
    memset(&p[r], 0, sizeof p[r]);
    p[r].key = stream_seed(seed, (uint64_t)replica0 + r, 0);
    p[r].beta = beta[r];
    ising_metropolis_table(beta[r], j, 2 * n_dims, p[r].thr);
  }
  // the tables live on this stack frame: the copy completes before return
  SC_HIP(hipMemcpyAsync(d_params, p, n_rep * sizeof(ReplicaParams), hipMemcpyHostToDevice, st));
  const bool measure = energy || magnetization;
  hipError_t e = hipSuccess;
  if (measure && n_sweeps > 0) e = launch_em_zero(d_em, (uint64_t)n_rep * kEmStride, st);
  for (uint64_t t = 0; t < n_sweeps && e == hipSuccess; ++t)
    for (uint32_t c = 0; c < 2 && e == hipSuccess; ++c)
      e = launch_ising_fast(g, d_spins, d_params, n_rep, c, sweep0 + t, (measure && c == 1 && t + 1 == n_sweeps) ? d_em : nullptr,
                            st);
  if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "sweep launch: %s", hipGetErrorString(e));
  long long em[kMaxReplicas * kEmStride];
  if (measure && n_sweeps > 0)
    SC_HIP(hipMemcpyAsync(em, d_em, (size_t)n_rep * kEmStride * sizeof(long long), hipMemcpyDeviceToHost, st));
  SC_HIP(hipStreamSynchronize(st));
  if (measure && n_sweeps > 0) {
    // IsingModel: E = -J sum_<ij> s_i s_j, |m| = |sum s| / N
    const double n = 2.0 * (double)g.nh;
    for (uint32_t r = 0; r < n_rep; ++r) {
      if (energy) energy[r] = -j * (double)em[(size_t)r * kEmStride];
      if (magnetization) magnetization[r] = std::fabs((double)em[(size_t)r * kEmStride + 1]) / n;
    }
  }
  return SCAMPER_OK;
}
