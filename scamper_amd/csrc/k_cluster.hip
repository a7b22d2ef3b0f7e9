// Swendsen-Wang and Wolff cluster updates for Ising and Potts spins
// (SWCore::sweep_with_phase, CMC.rs/src/algorithms/swendsen_wang.rs:73-135;
// WolffCore::sweep_with_phase, algorithms/wolff.rs:72-134; ClusterModel
// rules of IsingModel and PottsModel, lattice/models.rs:113-156, 295-338).
//
// One update of one replica:
//   1. parent[s] = s for every site;
//   2. every physical edge whose ends hold the same state is active when its
//      draw (stream 11, by edge index) is below the bond threshold of its
//      weight; active edges are united in a lock-free union-find that links
//      the larger root under the smaller, so each cluster's root is its
//      smallest site whatever order the unions run in;
//   3. Swendsen-Wang: every site takes the target its root draws (stream 12):
//      Ising +1 when the word is below 2^31, else -1; Potts floor(u q).
//      Wolff: the seed floor(u N) (stream 13, index 0) takes its cluster to
//      -seed (Ising) or another state floor(u' (q-1)) skipping the seed's
//      (Potts, u' from index 1).
// A Wolff cluster is the seed's component in the graph of active bonds, so
// every bond next to it is tested exactly once, as the reference's stack
// growth tests it.  The draws are the oracle's (orc_cluster_steps).
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ uint32_t load_parent(const uint32_t *p) {
  // bypass the non-coherent L1: other blocks link roots concurrently
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving (parents only ever move to smaller sites, and
// a non-root's parent is never the target of a link, so the plain halving
// store races only with other halving stores of valid ancestors).
__device__ uint32_t find_root(uint32_t *parent, uint32_t x) {
  while (true) {
    const uint32_t p = load_parent(parent + x);
    if (p == x) return x;
    const uint32_t gp = load_parent(parent + p);
    if (gp == p) return p;
    __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = gp;
  }
}

__device__ void unite(uint32_t *parent, uint32_t a, uint32_t b) {
  while (true) {
    a = find_root(parent, a);
    b = find_root(parent, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    // link the larger root a under b, unless a was linked meanwhile
    const uint32_t old = atomicCAS(parent + a, a, b);
    if (old == a) return;
    a = old;
  }
}

__device__ __forceinline__ uint64_t grid_threads() { return (uint64_t)gridDim.x * blockDim.x; }
__device__ __forceinline__ uint64_t thread_id() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }

__global__ void k_uf_init(uint32_t *__restrict__ parent, uint64_t n) {
  for (uint64_t i = thread_id(); i < n; i += grid_threads()) parent[i] = (uint32_t)i;
}

// Site-major index s of a periodic hypercube with even sides -> its byte in
// the colour-split replica at `spins`.
struct StencilSite {
  uint32_t x, y, z;
};
__device__ __forceinline__ StencilSite stencil_site(uint64_t s, uint32_t lx, uint32_t ly) {
  const uint64_t r = s / lx;
  return StencilSite{(uint32_t)(s - r * lx), (uint32_t)(r % ly), (uint32_t)(r / ly)};
}
__device__ __forceinline__ uint64_t stencil_byte(uint64_t s, const StencilSite &c, uint64_t nh) {
  return (uint64_t)((c.x + c.y + c.z) & 1u) * nh + (s >> 1);
}

// Edges of site s in build_hypercubic order: e = nd s + a, to +1 along axis a.
template <int ND>
__global__ void k_bonds_stencil(const int8_t *__restrict__ spins, uint32_t *parent, uint32_t lx, uint32_t ly,
                                uint32_t lz, uint64_t nh, uint64_t key, uint64_t step, uint64_t thr) {
  const uint64_t n = 2 * nh;
  for (uint64_t s = thread_id(); s < n; s += grid_threads()) {
    const StencilSite c = stencil_site(s, lx, ly);
    const int8_t v = spins[stencil_byte(s, c, nh)];
#pragma unroll
    for (int a = 0; a < ND; ++a) {
      StencilSite t = c;
      uint64_t ts;
      if (a == 0) {
        t.x = c.x + 1 == lx ? 0u : c.x + 1;
        ts = s - c.x + t.x;
      } else if (a == 1) {
        t.y = c.y + 1 == ly ? 0u : c.y + 1;
        ts = s + ((uint64_t)t.y - c.y) * lx;
      } else {
        t.z = c.z + 1 == lz ? 0u : c.z + 1;
        ts = s + ((uint64_t)t.z - c.z) * (uint64_t)lx * ly;
      }
      if (spins[stencil_byte(ts, t, nh)] != v) continue;
      if ((uint64_t)cluster_word(key, (uint64_t)ND * s + a, step, kStreamClusterBond) < thr)
        unite(parent, (uint32_t)s, (uint32_t)ts);
    }
  }
}

__global__ void k_bonds_csr(uint64_t n_edges, const ClusterEdge *__restrict__ edges, const int8_t *__restrict__ w,
                            const int8_t *__restrict__ spins, uint32_t *parent, uint64_t key, uint64_t step,
                            const uint64_t *__restrict__ thr) {
  for (uint64_t e = thread_id(); e < n_edges; e += grid_threads()) {
    const ClusterEdge b = edges[e];
    if (b.source == b.target || spins[b.source] != spins[b.target]) continue;  // a self-loop joins nothing
    if ((uint64_t)cluster_word(key, e, step, kStreamClusterBond) < thr[w[e]]) unite(parent, b.source, b.target);
  }
}

__device__ __forceinline__ int8_t sw_target(uint32_t root, int kind, int q, uint64_t key, uint64_t step) {
  const uint32_t w = cluster_word(key, root, step, kStreamClusterTarget);
  if (kind == kIsing) return w < 0x80000000u ? (int8_t)1 : (int8_t)-1;
  return (int8_t)(((uint64_t)w * (uint64_t)q) >> 32);
}

// The Wolff seed's root and target, from the spins before the update.
__global__ void k_wolff_seed(const int8_t *__restrict__ spins, uint32_t *parent, int stencil, uint32_t lx,
                             uint32_t ly, uint64_t nh, uint64_t n_sites, int kind, int q, uint64_t key,
                             uint64_t step, uint32_t *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t seed = ((uint64_t)cluster_word(key, 0, step, kStreamClusterSeed) * n_sites) >> 32;
  const int8_t s0 = stencil ? spins[stencil_byte(seed, stencil_site(seed, lx, ly), nh)] : spins[seed];
  int8_t target;
  if (kind == kIsing) {
    target = (int8_t)-s0;
  } else {
    uint32_t r = (uint32_t)(((uint64_t)cluster_word(key, 1, step, kStreamClusterSeed) * (uint64_t)(q - 1)) >> 32);
    r += r >= (uint32_t)s0;
    target = (int8_t)r;
  }
  out[0] = find_root(parent, (uint32_t)seed);
  out[1] = (uint32_t)(uint8_t)target;
}

__global__ void k_transform(int8_t *__restrict__ spins, uint32_t *parent, int stencil, uint32_t lx, uint32_t ly,
                            uint64_t nh, uint64_t n_sites, int kind, int q, int wolff, uint64_t key, uint64_t step,
                            const uint32_t *__restrict__ seed) {
  const uint32_t seed_root = wolff ? seed[0] : 0u;
  const int8_t wolff_target = wolff ? (int8_t)(uint8_t)seed[1] : (int8_t)0;
  for (uint64_t s = thread_id(); s < n_sites; s += grid_threads()) {
    const uint32_t root = find_root(parent, (uint32_t)s);
    int8_t v;
    if (wolff) {
      if (root != seed_root) continue;
      v = wolff_target;
    } else {
      v = sw_target(root, kind, q, key, step);
    }
    const uint64_t at = stencil ? stencil_byte(s, stencil_site(s, lx, ly), nh) : s;
    spins[at] = v;
  }
}

uint32_t blocks_for(uint64_t n) {
  const uint64_t b = (n + kThreads - 1) / kThreads;
  return (uint32_t)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

}  // namespace

hipError_t launch_cluster_bonds_stencil(const StencilGeom &g, const int8_t *spins, uint32_t *parent, uint64_t key,
                                        uint64_t step, uint64_t thr, hipStream_t stream) {
  const uint64_t n = 2 * g.nh;
  hipLaunchKernelGGL(k_uf_init, dim3(blocks_for(n)), dim3(kThreads), 0, stream, parent, n);
  const uint32_t lx = 2 * g.lx2;
  if (g.nd == 1)
    hipLaunchKernelGGL(k_bonds_stencil<1>, dim3(blocks_for(n)), dim3(kThreads), 0, stream, spins, parent, lx,
                       g.ly, g.lz, g.nh, key, step, thr);
  else if (g.nd == 2)
    hipLaunchKernelGGL(k_bonds_stencil<2>, dim3(blocks_for(n)), dim3(kThreads), 0, stream, spins, parent, lx,
                       g.ly, g.lz, g.nh, key, step, thr);
  else
    hipLaunchKernelGGL(k_bonds_stencil<3>, dim3(blocks_for(n)), dim3(kThreads), 0, stream, spins, parent, lx,
                       g.ly, g.lz, g.nh, key, step, thr);
  return hipGetLastError();
}

hipError_t launch_cluster_bonds_csr(uint64_t n_sites, uint64_t n_edges, const ClusterEdge *edges,
                                    const int8_t *weights, const int8_t *spins, uint32_t *parent, uint64_t key,
                                    uint64_t step, const uint64_t *thr, hipStream_t stream) {
  hipLaunchKernelGGL(k_uf_init, dim3(blocks_for(n_sites)), dim3(kThreads), 0, stream, parent, n_sites);
  if (n_edges)
    hipLaunchKernelGGL(k_bonds_csr, dim3(blocks_for(n_edges)), dim3(kThreads), 0, stream, n_edges, edges, weights,
                       spins, parent, key, step, thr);
  return hipGetLastError();
}

hipError_t launch_cluster_transform(const StencilGeom *g, uint64_t n_sites, int8_t *spins, uint32_t *parent,
                                    int kind, int q, int wolff, uint64_t key, uint64_t step, uint32_t *scratch,
                                    hipStream_t stream) {
  const int stencil = g != nullptr;
  const uint32_t lx = stencil ? 2 * g->lx2 : 0u, ly = stencil ? g->ly : 0u;
  const uint64_t nh = stencil ? g->nh : 0u;
  if (wolff)
    hipLaunchKernelGGL(k_wolff_seed, dim3(1), dim3(64), 0, stream, spins, parent, stencil, lx, ly, nh, n_sites, kind,
                       q, key, step, scratch);
  hipLaunchKernelGGL(k_transform, dim3(blocks_for(n_sites)), dim3(kThreads), 0, stream, spins, parent, stencil, lx, ly,
                     nh, n_sites, kind, q, wolff, key, step, scratch);
  return hipGetLastError();
}

}  // namespace scamper
