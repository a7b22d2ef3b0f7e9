// Graph-coloured single-site updates on arbitrary CSR lattices
// (CsrLattice, CMC.rs/src/lattice/graph.rs:112-124): one thread per site of
// the colour class being updated, neighbours read through the CSR arrays.
//
// Bond weights must be integers (|w| <= 127), so every local field is an
// integer h in [-W, W] and the host tables (tables.cpp) turn each decision
// into integer compares, exactly as on the stencil path.  Self-loops never
// change ΔE or the heat-bath conditionals for Ising and Potts spins
// (interaction.rs:149-181, models.rs:166-170, 355-358) and are not stored.
//
// Spins are int8, site-major per replica; the draw of a site is keyed by
// its colour and its rank inside the colour class.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kBlock = 256;

struct LazyRestCsr {
  uint64_t key, j, sweep;
  uint32_t colour, stream;
  __device__ uint32_t operator()() const {
    return (site_byte(key, j, sweep, colour, stream, 1) << 16) | site_low(key, j, sweep, colour, stream);
  }
};

__device__ __forceinline__ bool decide_csr(uint32_t top, Threshold t, const LazyRestCsr &rest) {
  if (top != t.k8) return top < t.k8;
  return rest() < rest_of(t);
}

template <int KIND, int ALG>
__global__ __launch_bounds__(kBlock) void k_csr_site(int8_t *__restrict__ spins, const ReplicaParams *__restrict__ reps,
                                                     CsrGeom g, uint32_t colour, uint64_t class_start,
                                                     uint64_t class_size, uint64_t sweep, int q,
                                                     const double *__restrict__ hb_exp, int hb_min) {
  const uint32_t rep = blockIdx.y;
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const ReplicaParams &P = reps[rep];
  __shared__ Threshold sthr[kMaxTable];
  if (threadIdx.x < kMaxTable) sthr[threadIdx.x] = P.thr[threadIdx.x];
  __syncthreads();
  if (j >= class_size) return;
  int8_t *__restrict__ s_all = spins + (uint64_t)rep * g.n_sites;
  const uint32_t site = g.class_sites[class_start + j];
  const int8_t s = s_all[site];
  const uint32_t b0 = g.offsets[site], b1 = g.offsets[site + 1];
  const int W = g.weight_max;
  const uint64_t key = P.key;
  int8_t ns = s;
  if (KIND == kIsing) {
    int h = 0;
    for (uint32_t b = b0; b < b1; ++b) h += (int)g.weights[b] * (int)s_all[g.neighbors[b]];
    if (ALG == kMetropolis) {
      const Threshold t = sthr[s * h + W];  // ΔE = 2 J s h
      bool flip;
      if (t.k8 > 0xFFu)
        flip = true;
      else if (t.k8 == 0 && rest_of(t) == 0)
        flip = false;
      else
        flip = decide_csr(site_byte(key, j, sweep, colour, kStreamAccept, 0), t,
                          LazyRestCsr{key, j, sweep, colour, kStreamAccept});
      if (flip) ns = -s;
    } else {
      ns = decide_csr(site_byte(key, j, sweep, colour, kStreamAccept, 0), sthr[h + W],
                      LazyRestCsr{key, j, sweep, colour, kStreamAccept}) ? 1 : -1;
    }
  } else {
    if (ALG == kMetropolis) {
      const uint32_t ptop = site_byte(key, j, sweep, colour, kStreamPropose, 0);
      const uint32_t m = (uint32_t)(q - 1);
      const uint64_t base = (uint64_t)ptop << 24;
      uint32_t r = (uint32_t)((base * m) >> 32);
      if (r != (uint32_t)(((base + 0xFFFFFFu) * m) >> 32))
        r = (uint32_t)(((base | LazyRestCsr{key, j, sweep, colour, kStreamPropose}()) * m) >> 32);
      if (r >= (uint32_t)s) r += 1;
      int d = 0;
      for (uint32_t b = b0; b < b1; ++b) {
        const int8_t v = s_all[g.neighbors[b]];
        d += (int)g.weights[b] * (((uint32_t)v == r) - (v == s));
      }
      const Threshold t = sthr[d + W];
      bool accept;
      if (t.k8 > 0xFFu)
        accept = true;
      else if (t.k8 == 0 && rest_of(t) == 0)
        accept = false;
      else
        accept = decide_csr(site_byte(key, j, sweep, colour, kStreamAccept, 0), t,
                            LazyRestCsr{key, j, sweep, colour, kStreamAccept});
      if (accept) ns = (int8_t)r;
    } else {
#pragma clang fp contract(off)
      int cnt[kMaxQ];
      for (int v = 0; v < q; ++v) cnt[v] = 0;
      for (uint32_t b = b0; b < b1; ++b) cnt[(uint8_t)s_all[g.neighbors[b]]] += g.weights[b];
      // the log weight of the largest count is max_log when beta J > 0, of
      // the smallest when J < 0 (hb_min); unit weights only on this path
      int nmax = hb_min ? (int)(b1 - b0) : 0;
      for (int v = 0; v < q; ++v) nmax = (hb_min ? cnt[v] < nmax : cnt[v] > nmax) ? cnt[v] : nmax;
      const double *E = hb_exp + (uint64_t)rep * (kMaxZ + 1) * (kMaxZ + 1);
      double total = 0.0;
      for (int v = 0; v < q; ++v) total = __dadd_rn(total, E[cnt[v] * (W + 1) + nmax]);
      const uint32_t w = site_word(key, j, sweep, colour, kStreamAccept);
      double thr = __dmul_rn((double)w * 2.3283064365386963e-10, total);
      int pick = q - 1;
      for (int v = 0; v < q; ++v) {
        thr = __dsub_rn(thr, E[cnt[v] * (W + 1) + nmax]);
        if (thr <= 0.0) {
          pick = v;
          break;
        }
      }
      ns = (int8_t)pick;
    }
  }
  s_all[site] = ns;
}

// Sum over sites of w-weighted neighbour agreement (each non-loop bond is
// seen from both ends) and the spin sum / state counts of one replica.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_csr_measure(const int8_t *__restrict__ spins, CsrGeom g, int q,
                                                        long long *__restrict__ em) {
  const uint32_t rep = blockIdx.y;
  const uint64_t site = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  long long acc[2 + kMaxQ];
#pragma unroll
  for (int k = 0; k < 2 + kMaxQ; ++k) acc[k] = 0;
  if (site < g.n_sites) {
    const int8_t *s_all = spins + (uint64_t)rep * g.n_sites;
    const int8_t s = s_all[site];
    long long bond = 0;
    for (uint32_t b = g.offsets[site]; b < g.offsets[site + 1]; ++b) {
      const int8_t v = s_all[g.neighbors[b]];
      bond += KIND == kIsing ? (long long)g.weights[b] * s * v : (long long)g.weights[b] * (v == s);
    }
    acc[0] = bond;
    if (KIND == kIsing)
      acc[1] = s;
    else
      for (int v = 0; v < q; ++v) acc[2 + v] = (s == v);
  }
  __shared__ long long red[2 + kMaxQ];
  if (threadIdx.x < 2 + kMaxQ) red[threadIdx.x] = 0;
  __syncthreads();
  const int nc = KIND == kIsing ? 2 : 2 + q;
  for (int k = 0; k < nc; ++k) {
    long long v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0 && v != 0) atomicAdd((unsigned long long *)&red[k], (unsigned long long)v);
  }
  __syncthreads();
  if (threadIdx.x < nc && red[threadIdx.x] != 0)
    atomicAdd((unsigned long long *)&em[(uint64_t)rep * kEmStride + threadIdx.x], (unsigned long long)red[threadIdx.x]);
}

__global__ __launch_bounds__(kBlock) void k_csr_init(int8_t *__restrict__ spins, uint64_t n, int kind, int q,
                                                     uint64_t key, uint32_t rep, int random) {
  const uint64_t site = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (site >= n) return;
  int8_t v;
  if (!random)
    v = kind == kIsing ? 1 : 0;
  else
    v = kind == kIsing ? ((site_byte(key, site, 0, 0, kStreamInit, 0) & 1u) ? 1 : -1)
                       : (int8_t)(((uint64_t)site_word(key, site, 0, 0, kStreamInit) * (uint64_t)q) >> 32);
  spins[(uint64_t)rep * n + site] = v;
}

}  // namespace

hipError_t launch_csr_discrete(const CsrGeom &g, const uint64_t *class_start_host, int kind, int q,
                               int alg, int8_t *spins, const ReplicaParams *reps, uint32_t n_rep,
                               uint64_t sweep, const double *hb_exp, int hb_min, hipStream_t st) {
  for (uint32_t c = 0; c < g.n_colours; ++c) {
    const uint64_t start = class_start_host[c], size = class_start_host[c + 1] - start;
    if (size == 0) continue;
    dim3 grid((unsigned)((size + kBlock - 1) / kBlock), n_rep);
#define SC_CSR(K, A) \
  hipLaunchKernelGGL((k_csr_site<K, A>), grid, dim3(kBlock), 0, st, spins, reps, g, c, start, size, sweep, q, hb_exp, hb_min)
    if (kind == kIsing) {
      if (alg == kMetropolis) SC_CSR(kIsing, kMetropolis); else SC_CSR(kIsing, kHeatBath);
    } else {
      if (alg == kMetropolis) SC_CSR(kPotts, kMetropolis); else SC_CSR(kPotts, kHeatBath);
    }
#undef SC_CSR
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_csr_measure(const CsrGeom &g, int kind, int q, const int8_t *spins, uint32_t n_rep,
                              long long *em, hipStream_t st) {
  dim3 grid((unsigned)((g.n_sites + kBlock - 1) / kBlock), n_rep);
  if (kind == kIsing)
    hipLaunchKernelGGL((k_csr_measure<kIsing>), grid, dim3(kBlock), 0, st, spins, g, q, em);
  else
    hipLaunchKernelGGL((k_csr_measure<kPotts>), grid, dim3(kBlock), 0, st, spins, g, q, em);
  return hipGetLastError();
}

hipError_t launch_csr_init(uint64_t n_sites, int kind, int q, int8_t *spins, uint64_t key, uint32_t rep,
                           int random, hipStream_t st) {
  hipLaunchKernelGGL(k_csr_init, dim3((unsigned)((n_sites + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                     spins, n_sites, kind, q, key, rep, random);
  return hipGetLastError();
}

}  // namespace scamper
