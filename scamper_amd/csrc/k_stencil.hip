// Checkerboard single-site updates of discrete spins (Ising +-1 and Potts
// 0..q-1 as int8) on periodic hypercubic lattices with even sides.
//
// Per-site rules follow the reference's MetropolisCore / HeatBathCore
// (CMC.rs/src/algorithms/metropolis.rs:67-106, heat_bath.rs:40-72) with the
// visit order "colour 0 then colour 1"; decisions come from per-replica
// threshold tables built on the host (tables.cpp), so a kernel only compares
// Philox draws with integers.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kBlock = 256;

// Blocks per replica of a grid-stride launch over nh sites.
inline unsigned capped_blocks(uint64_t nh) {
  const uint64_t b = (nh + kBlock - 1) / kBlock;
  return (unsigned)(b < 2048 ? b : 2048);
}

__device__ __forceinline__ uint32_t wrap_dec(uint32_t v, uint32_t n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ uint32_t wrap_inc(uint32_t v, uint32_t n) { return v + 1 == n ? 0 : v + 1; }

// Resolve "word < K" from the top byte; the lower 24 bits (mid byte and low
// half) are drawn only on a tie.  K = 2^32 (k8 = 256) accepts every word.
struct LazyRest {
  uint64_t key, j, sweep;
  uint32_t colour, stream;
  __device__ uint32_t operator()() const {
    return (site_byte(key, j, sweep, colour, stream, 1) << 16) | site_low(key, j, sweep, colour, stream);
  }
};

__device__ __forceinline__ bool decide(uint32_t top, Threshold t, const LazyRest &rest) {
  if (top != t.k8) return top < t.k8;
  return rest() < rest_of(t);
}

// Proposal index floor(word * m / 2^32) for m = q-1 (models.rs:212-218);
// the top byte decides unless its 2^24-wide bucket straddles a boundary.
__device__ __forceinline__ uint32_t propose_index(uint32_t top, uint32_t m, const LazyRest &rest) {
  const uint64_t base = (uint64_t)top << 24;
  const uint32_t a = (uint32_t)((base * m) >> 32);
  const uint32_t b = (uint32_t)(((base + 0xFFFFFFu) * m) >> 32);
  if (a == b) return a;
  return (uint32_t)(((base | rest()) * m) >> 32);
}

template <int ND>
struct Neighbours {
  int8_t v[2 * ND];
};

template <int ND>
__device__ __forceinline__ Neighbours<ND> gather(const int8_t *__restrict__ oth, const StencilGeom &g,
                                                 uint32_t row, uint32_t i, uint32_t y, uint32_t z,
                                                 uint32_t p) {
  Neighbours<ND> n;
  const uint64_t rb = (uint64_t)row * g.lx2;
  const uint32_t il = p ? i : wrap_dec(i, g.lx2);
  const uint32_t ir = p ? wrap_inc(i, g.lx2) : i;
  n.v[0] = oth[rb + il];
  n.v[1] = oth[rb + ir];
  if (ND >= 2) {
    const uint32_t ym = wrap_dec(y, g.ly), yp = wrap_inc(y, g.ly);
    n.v[2] = oth[((uint64_t)(ym + g.ly * z)) * g.lx2 + i];
    n.v[3] = oth[((uint64_t)(yp + g.ly * z)) * g.lx2 + i];
  }
  if (ND >= 3) {
    const uint32_t zm = wrap_dec(z, g.lz), zp = wrap_inc(z, g.lz);
    n.v[4] = oth[((uint64_t)(y + g.ly * zm)) * g.lx2 + i];
    n.v[5] = oth[((uint64_t)(y + g.ly * zp)) * g.lx2 + i];
  }
  return n;
}

// Block-wide sum of up to 1+1+q counters, then one atomic per counter.
template <int NC>
__device__ void block_accumulate(long long (&acc)[NC], int nc, long long *em) {
  __shared__ long long red[NC];
  if (threadIdx.x < NC) red[threadIdx.x] = 0;
  __syncthreads();
  for (int k = 0; k < nc; ++k) {
    long long v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0 && v != 0) atomicAdd((unsigned long long *)&red[k], (unsigned long long)v);
  }
  __syncthreads();
  if (threadIdx.x < nc && red[threadIdx.x] != 0)
    atomicAdd((unsigned long long *)&em[threadIdx.x], (unsigned long long)red[threadIdx.x]);
}

// ------------------------------------------------------------------------
// One thread per site of the colour being updated.  Handles every model,
// algorithm and size; the fast Ising kernel below replaces it where it can.
template <int ND, int KIND, int ALG, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_site(int8_t *__restrict__ spins,
                                                 const ReplicaParams *__restrict__ reps,
                                                 StencilGeom g, uint32_t colour, uint64_t sweep,
                                                 int q, const double *__restrict__ hb_exp, int hb_min,
                                                 long long *__restrict__ em) {
  constexpr int z = 2 * ND;
  const uint32_t rep = blockIdx.y;
  const ReplicaParams &P = reps[rep];
  __shared__ Threshold sthr[kMaxTable];
  if (threadIdx.x < kMaxTable) sthr[threadIdx.x] = P.thr[threadIdx.x];
  __syncthreads();
  long long acc[2 + kMaxQ];
#pragma unroll
  for (int k = 0; k < 2 + kMaxQ; ++k) acc[k] = 0;
  // grid-stride over sites: the grid is capped so that a measured launch
  // issues a bounded number of same-address atomics
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < g.nh; j += (uint64_t)gridDim.x * kBlock) {
    int8_t *__restrict__ own = spins + (2ull * rep + colour) * g.cs;
    const int8_t *__restrict__ oth = spins + (2ull * rep + 1 - colour) * g.cs;
    const SiteCoords sc_ = site_coords(j, g);
    const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
    const uint32_t p = (colour + y + zz) & 1u;
    const Neighbours<ND> n = gather<ND>(oth, g, row, i, y, zz, p);
    const int8_t s = own[j];
    int8_t ns = s;
    const uint64_t key = P.key;
    if (KIND == kIsing) {
      if (ALG == kMetropolis) {
        int anti = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) anti += (n.v[k] != s);
        const Threshold t = sthr[anti];
        if (t.k8 > 0xFFu) {
          ns = -s;
        } else if (t.k8 != 0 || rest_of(t) != 0) {
          const uint32_t top = site_byte(key, j, sweep, colour, kStreamAccept, 0);
          if (decide(top, t, LazyRest{key, j, sweep, colour, kStreamAccept})) ns = -s;
        }
      } else {  // heat bath: +1 iff word < K[#minus neighbours]
        int minus = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) minus += (n.v[k] < 0);
        const uint32_t top = site_byte(key, j, sweep, colour, kStreamAccept, 0);
        ns = decide(top, sthr[minus], LazyRest{key, j, sweep, colour, kStreamAccept}) ? 1 : -1;
      }
    } else {
      if (ALG == kMetropolis) {
        const uint32_t ptop = site_byte(key, j, sweep, colour, kStreamPropose, 0);
        uint32_t r = propose_index(ptop, (uint32_t)(q - 1), LazyRest{key, j, sweep, colour, kStreamPropose});
        if (r >= (uint32_t)s) r += 1;
        int d = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) d += ((uint32_t)n.v[k] == r) - (n.v[k] == s);
        const Threshold t = sthr[d + z];
        bool accept;
        if (t.k8 > 0xFFu)
          accept = true;
        else if (t.k8 == 0 && rest_of(t) == 0)
          accept = false;
        else
          accept = decide(site_byte(key, j, sweep, colour, kStreamAccept, 0), t,
                          LazyRest{key, j, sweep, colour, kStreamAccept});
        if (accept) ns = (int8_t)r;
      } else {
#pragma clang fp contract(off)
        // heat bath (models.rs:340-380) in f64: weights exp(lw(n) - max_log)
        // from the host's exp table, then the reference's sequential
        // threshold subtraction.  max_log is the log weight of the largest
        // count (beta J > 0) or of the smallest (hb_min: J < 0).
        int cnt[kMaxQ];
        for (int v = 0; v < q; ++v) cnt[v] = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) cnt[(uint8_t)n.v[k]] += 1;
        int nmax = hb_min ? z : 0;
        for (int v = 0; v < q; ++v) nmax = (hb_min ? cnt[v] < nmax : cnt[v] > nmax) ? cnt[v] : nmax;
        double total = 0.0;
        for (int v = 0; v < q; ++v) total = __dadd_rn(total, hb_exp[rep * (kMaxZ + 1) * (kMaxZ + 1) + cnt[v] * (z + 1) + nmax]);
        const uint32_t w = site_word(key, j, sweep, colour, kStreamAccept);
        double thr = __dmul_rn((double)w * 2.3283064365386963e-10, total);
        int pick = q - 1;
        for (int v = 0; v < q; ++v) {
          thr = __dsub_rn(thr, hb_exp[rep * (kMaxZ + 1) * (kMaxZ + 1) + cnt[v] * (z + 1) + nmax]);
          if (thr <= 0.0) {
            pick = v;
            break;
          }
        }
        ns = (int8_t)pick;
      }
    }
    own[j] = ns;
    if (MEAS) {
      const int8_t other_site = oth[j];
      if (KIND == kIsing) {
        int sum = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) sum += n.v[k];
        acc[0] += (long long)ns * sum;
        acc[1] += (long long)ns + other_site;
      } else {
        int same = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) same += (n.v[k] == ns);
        acc[0] += same;
        for (int v = 0; v < q; ++v) acc[2 + v] += (ns == v) + (other_site == v);
      }
    }
  }
  if (MEAS) block_accumulate<2 + kMaxQ>(acc, KIND == kIsing ? 2 : 2 + q, em + (uint64_t)rep * kEmStride);
}

// ------------------------------------------------------------------------
// Exact recomputation of the observables (the audit of audit.rs:41-52 and
// the measurement after uploads): over colour-1 sites, bond sums with their
// colour-0 neighbours (every bond has exactly one colour-1 end).
template <int ND, int KIND>
__global__ __launch_bounds__(kBlock) void k_measure(const int8_t *__restrict__ spins, StencilGeom g,
                                                    int q, long long *__restrict__ em) {
  constexpr int z = 2 * ND;
  const uint32_t rep = blockIdx.y;
  long long acc[2 + kMaxQ];
#pragma unroll
  for (int k = 0; k < 2 + kMaxQ; ++k) acc[k] = 0;
  for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < g.nh; j += (uint64_t)gridDim.x * kBlock) {
    const int8_t *own = spins + (2ull * rep + 1) * g.cs;
    const int8_t *oth = spins + (2ull * rep) * g.cs;
    const SiteCoords sc_ = site_coords(j, g);
    const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
    const uint32_t p = (1 + y + zz) & 1u;
    const Neighbours<ND> n = gather<ND>(oth, g, row, i, y, zz, p);
    const int8_t s = own[j], o = oth[j];
    if (KIND == kIsing) {
      int sum = 0;
#pragma unroll
      for (int k = 0; k < z; ++k) sum += n.v[k];
      acc[0] += (long long)s * sum;
      acc[1] += (long long)s + o;
    } else {
      int same = 0;
#pragma unroll
      for (int k = 0; k < z; ++k) same += (n.v[k] == s);
      acc[0] += same;
      for (int v = 0; v < q; ++v) acc[2 + v] += (s == v) + (o == v);
    }
  }
  block_accumulate<2 + kMaxQ>(acc, KIND == kIsing ? 2 : 2 + q, em + (uint64_t)rep * kEmStride);
}

// Cold (+1 / state 0) or random start keyed by (seed, replica, site).
__global__ __launch_bounds__(kBlock) void k_init(int8_t *__restrict__ spins, StencilGeom g, int kind,
                                                 int q, uint64_t key, uint32_t rep, int random) {
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint32_t colour = blockIdx.y;
  if (j >= g.nh) return;
  const SiteCoords sc_ = site_coords(j, g);
  const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
  const uint64_t site = 2ull * i + ((colour + y + zz) & 1u) + 2ull * g.lx2 * row;
  int8_t v;
  if (!random)
    v = kind == kIsing ? 1 : 0;
  else {
    v = kind == kIsing ? ((site_byte(key, site, 0, 0, kStreamInit, 0) & 1u) ? 1 : -1)
                       : (int8_t)(((uint64_t)site_word(key, site, 0, 0, kStreamInit) * (uint64_t)q) >> 32);
  }
  spins[(2ull * rep + colour) * g.cs + j] = v;
}

template <int ND>
hipError_t dispatch_site(const StencilGeom &g, int kind, int q, int alg, int8_t *spins,
                         const ReplicaParams *reps, uint32_t n_rep, uint32_t colour, uint64_t sweep,
                         const double *hb_exp, int hb_min, long long *em, hipStream_t st) {
  dim3 grid(capped_blocks(g.nh), n_rep);
  const bool meas = em != nullptr;
#define SC_LAUNCH(K, A, M) \
  hipLaunchKernelGGL((k_site<ND, K, A, M>), grid, dim3(kBlock), 0, st, spins, reps, g, colour, sweep, q, hb_exp, hb_min, em)
  if (kind == kIsing) {
    if (alg == kMetropolis) {
      if (meas) SC_LAUNCH(kIsing, kMetropolis, true); else SC_LAUNCH(kIsing, kMetropolis, false);
    } else {
      if (meas) SC_LAUNCH(kIsing, kHeatBath, true); else SC_LAUNCH(kIsing, kHeatBath, false);
    }
  } else {
    if (alg == kMetropolis) {
      if (meas) SC_LAUNCH(kPotts, kMetropolis, true); else SC_LAUNCH(kPotts, kMetropolis, false);
    } else {
      if (meas) SC_LAUNCH(kPotts, kHeatBath, true); else SC_LAUNCH(kPotts, kHeatBath, false);
    }
  }
#undef SC_LAUNCH
  return hipGetLastError();
}

}  // namespace

hipError_t launch_stencil_discrete(const StencilGeom &g, int kind, int q, int alg, int8_t *spins,
                                   const ReplicaParams *reps, uint32_t n_rep, uint32_t colour,
                                   uint64_t sweep, const double *hb_exp, int hb_min, long long *em,
                                   hipStream_t st) {
  if (g.cs < g.nh) return hipErrorInvalidValue;  // the colour stride is unset
  if (kind == kIsing && alg == kMetropolis && ising_fast_path_ok(g))
    return launch_ising_fast(g, spins, reps, n_rep, colour, sweep, em, st);
  switch (g.nd) {
    case 1: return dispatch_site<1>(g, kind, q, alg, spins, reps, n_rep, colour, sweep, hb_exp, hb_min, em, st);
    case 2: return dispatch_site<2>(g, kind, q, alg, spins, reps, n_rep, colour, sweep, hb_exp, hb_min, em, st);
    case 3: return dispatch_site<3>(g, kind, q, alg, spins, reps, n_rep, colour, sweep, hb_exp, hb_min, em, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stencil_measure(const StencilGeom &g, int kind, int q, const int8_t *spins,
                                  uint32_t n_rep, long long *em, hipStream_t st) {
  if (g.cs < g.nh) return hipErrorInvalidValue;
  dim3 grid(capped_blocks(g.nh), n_rep);
#define SC_M(ND)                                                                                   \
  if (kind == kIsing)                                                                              \
    hipLaunchKernelGGL((k_measure<ND, kIsing>), grid, dim3(kBlock), 0, st, spins, g, q, em);       \
  else                                                                                             \
    hipLaunchKernelGGL((k_measure<ND, kPotts>), grid, dim3(kBlock), 0, st, spins, g, q, em);
  switch (g.nd) {
    case 1: SC_M(1) break;
    case 2: SC_M(2) break;
    case 3: SC_M(3) break;
    default: return hipErrorInvalidValue;
  }
#undef SC_M
  return hipGetLastError();
}

hipError_t launch_stencil_init(const StencilGeom &g, int kind, int q, int8_t *spins, uint64_t key,
                               uint32_t rep, int random, hipStream_t st) {
  if (g.cs < g.nh) return hipErrorInvalidValue;
  dim3 grid((unsigned)((g.nh + kBlock - 1) / kBlock), 2);
  hipLaunchKernelGGL(k_init, grid, dim3(kBlock), 0, st, spins, g, kind, q, key, rep, random);
  return hipGetLastError();
}

}  // namespace scamper
