// Swendsen-Wang and Wolff cluster updates for Ising, Potts and O(3) spins
// (SWCore::sweep_with_phase, CMC.rs/src/algorithms/swendsen_wang.rs:73-135;
// WolffCore::sweep_with_phase, algorithms/wolff.rs:72-134; ClusterModel
// rules of IsingModel, PottsModel and ONModel, lattice/models.rs:113-156,
// 295-338, 516-577).  The O(3) embedding clusters are at the end of the file.
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
__device__ __forceinline__ uint64_t stencil_byte(uint64_t s, const StencilSite &c, uint64_t cs) {
  return (uint64_t)((c.x + c.y + c.z) & 1u) * cs + (s >> 1);
}

// Edges of site s in build_hypercubic order: e = nd s + a, to +1 along axis a.
template <int ND>
__global__ void k_bonds_stencil(const int8_t *__restrict__ spins, uint32_t *parent, uint32_t lx, uint32_t ly,
                                uint32_t lz, uint64_t nh, uint64_t cs, uint64_t key, uint64_t step, uint64_t thr) {
  const uint64_t n = 2 * nh;
  for (uint64_t s = thread_id(); s < n; s += grid_threads()) {
    const StencilSite c = stencil_site(s, lx, ly);
    const int8_t v = spins[stencil_byte(s, c, cs)];
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
      if (spins[stencil_byte(ts, t, cs)] != v) continue;
      if ((uint64_t)cluster_word(key, (uint64_t)ND * s + a, step, kStreamClusterBond) < thr)
        unite(parent, (uint32_t)s, (uint32_t)ts);
    }
  }
}

__global__ void k_bonds_csr(uint64_t n_edges, const ClusterEdge *__restrict__ edges, const int8_t *__restrict__ w,
                            const int8_t *__restrict__ spins, uint32_t *parent, uint64_t key, uint64_t step,
                            const uint64_t *__restrict__ thr, const uint64_t *__restrict__ edge_thr) {
  for (uint64_t e = thread_id(); e < n_edges; e += grid_threads()) {
    const ClusterEdge b = edges[e];
    if (b.source == b.target || spins[b.source] != spins[b.target]) continue;  // a self-loop joins nothing
    // J w >= 0 on every edge, so the coupling is |J| |w|: thr is indexed by
    // |w| for integer weights, edge_thr holds one threshold per edge otherwise
    const int wi = w[e];
    const uint64_t k = edge_thr ? edge_thr[e] : thr[wi < 0 ? -wi : wi];
    if ((uint64_t)cluster_word(key, e, step, kStreamClusterBond) < k) unite(parent, b.source, b.target);
  }
}

// ------------------------------------------------ tile-local labelling
// Stencils whose sides are multiples of the tile (2-D 32 x 32, 3-D 16 x 8 x 8
// sites, 1024 per block): each block unites the bonds inside its tile in LDS
// and writes every site's tile-local root (the tile's smallest member of its
// local cluster, which in site-major order is also the smallest global
// index); a second kernel unites the bonds that cross tile faces in global
// memory.  The draws and the final partition are those of the global path.
template <int ND>
struct Tile {
  static constexpr uint32_t X = ND == 2 ? 32 : 16, Y = ND == 2 ? 32 : 8, Z = ND == 2 ? 1 : 8;
  static constexpr uint32_t N = X * Y * Z;  // 1024
};

__device__ __forceinline__ uint32_t load_lds(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ uint32_t find_lds(uint32_t *lp, uint32_t x) {
  while (true) {
    const uint32_t p = load_lds(lp + x);
    if (p == x) return x;
    const uint32_t gp = load_lds(lp + p);
    if (gp == p) return p;
    __hip_atomic_store(lp + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = gp;
  }
}

__device__ void unite_lds(uint32_t *lp, uint32_t a, uint32_t b) {
  while (true) {
    a = find_lds(lp, a);
    b = find_lds(lp, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicCAS(lp + a, a, b);
    if (old == a) return;
    a = old;
  }
}

// One thread per quad of four consecutive sites along x (a tile row is X / 4
// quads).  The quad's 4 ND bonds (edge e = ND s + a, 4 | s) are the words of
// ND whole Philox blocks, and its spins are two 16-bit loads: sites s0 and
// s0 + 2 sit at bytes s0/2 and s0/2 + 1 of one colour, s0 + 1 and s0 + 3 at the
// same bytes of the other.  Bonds along x need no union-find: each site's
// parent is the first site of its run of active x bonds, found from the row's
// bond mask with one count of leading zeros, which is already a forest of
// depth 1 rooted at each run's smallest site.  Bonds along y (and z) then
// unite runs in LDS.
template <int ND>
__global__ __launch_bounds__(kThreads) void k_bonds_tile(const int8_t *__restrict__ spins, uint32_t *__restrict__ parent,
                                                         uint32_t lx, uint32_t ly, uint64_t cs, uint64_t key,
                                                         uint64_t step, uint64_t thr) {
  using T = Tile<ND>;
  static_assert(T::N == 4 * kThreads, "one quad per thread");
  constexpr uint32_t QR = T::X / 4;  // quads per tile row
  __shared__ uint32_t lp[T::N];
  __shared__ uint32_t sp4[kThreads];  // the quads' spins, one byte per site
  __shared__ uint8_t xb[kThreads];    // the quads' x-bond nibbles
  const uint32_t tiles_x = lx / T::X, tiles_y = ly / T::Y;
  const uint32_t tile = blockIdx.x;
  const uint32_t ox = (tile % tiles_x) * T::X, oy = ((tile / tiles_x) % tiles_y) * T::Y,
                 oz = (tile / (tiles_x * tiles_y)) * T::Z;
  const uint32_t q = threadIdx.x, i0 = 4 * q;
  const uint32_t x0 = i0 % T::X, y = (i0 / T::X) % T::Y, z = i0 / (T::X * T::Y);
  const uint32_t s0 = ox + x0 + lx * (oy + y + ly * (oz + z));
  const uint32_t c = (ox + x0 + oy + y + oz + z) & 1u;
  const uint32_t own = (uint16_t)*reinterpret_cast<const int16_t *>(spins + (uint64_t)c * cs + s0 / 2);
  const uint32_t oth = (uint16_t)*reinterpret_cast<const int16_t *>(spins + (uint64_t)(c ^ 1u) * cs + s0 / 2);
  uint32_t w[4 * ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const U4 blk = cluster_block(key, (uint64_t)ND * s0 / 4 + k, step, kStreamClusterBond);
    w[4 * k] = blk.x, w[4 * k + 1] = blk.y, w[4 * k + 2] = blk.z, w[4 * k + 3] = blk.w;
  }
  // bytes of sites s0 .. s0 + 3
  sp4[q] = (own & 0xFFu) | (oth & 0xFFu) << 8 | (own >> 8) << 16 | (oth >> 8) << 24;
  __syncthreads();
  const uint8_t *sp = reinterpret_cast<const uint8_t *>(sp4);
  auto bond = [&](uint32_t u, int a) { return (uint64_t)w[u * ND + a] < thr; };
  uint32_t xm = 0;  // bit u: the bond from site u to u + 1 is active and inside the tile
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    if (x0 + u + 1 < T::X && sp[i0 + u] == sp[i0 + u + 1] && bond(u, 0)) xm |= 1u << u;
  }
  xb[q] = (uint8_t)xm;
  __syncthreads();
  // the row's x-bond mask, bit x for the bond from x to x + 1
  const uint32_t r0 = q - q % QR;
  uint32_t row = 0;
#pragma unroll
  for (uint32_t k = 0; k < QR; ++k) row |= (uint32_t)xb[r0 + k] << (4 * k);
  const uint32_t starts = ~(row << 1);  // bit x: no active bond from x - 1 to x
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    const uint32_t x = x0 + u;
    const uint32_t first = 31u - (uint32_t)__builtin_clz(starts & (0xFFFFFFFFu >> (31u - x)));
    lp[i0 + u] = i0 - x0 + first;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    const uint32_t i = i0 + u;
    if (y + 1 < T::Y && sp[i] == sp[i + T::X] && bond(u, 1)) unite_lds(lp, i, i + T::X);
    if (ND == 3 && z + 1 < T::Z && sp[i] == sp[i + T::X * T::Y] && bond(u, ND - 1))
      unite_lds(lp, i, i + T::X * T::Y);
  }
  __syncthreads();
  uint32_t r[4];
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    const uint32_t l = find_lds(lp, i0 + u);
    r[u] = ox + l % T::X + lx * (oy + (l / T::X) % T::Y + ly * (oz + l / (T::X * T::Y)));
  }
  *reinterpret_cast<uint4 *>(parent + s0) = make_uint4(r[0], r[1], r[2], r[3]);
}

// Bonds from each tile's last layer along every axis to the next tile
// (periodic): X-face, Y-face (and Z-face) per tile, one thread each.
template <int ND>
__global__ void k_bonds_faces(const int8_t *__restrict__ spins, uint32_t *parent, uint32_t lx, uint32_t ly, uint32_t lz,
                              uint64_t cs, uint64_t n_tiles, uint64_t key, uint64_t step, uint64_t thr) {
  using T = Tile<ND>;
  constexpr uint32_t fx = T::Y * T::Z, fy = T::X * T::Z, fz = ND == 3 ? T::X * T::Y : 0u;
  constexpr uint32_t per_tile = fx + fy + fz;
  const uint32_t tiles_x = lx / T::X, tiles_y = ly / T::Y;
  for (uint64_t k = thread_id(); k < n_tiles * per_tile; k += grid_threads()) {
    const uint32_t tile = (uint32_t)(k / per_tile), f = (uint32_t)(k - (uint64_t)tile * per_tile);
    const uint32_t tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, tz = tile / (tiles_x * tiles_y);
    uint32_t x, y, z, a;
    if (f < fx) {  // +x bond of the tile's last column
      a = 0, x = tx * T::X + T::X - 1, y = ty * T::Y + f % T::Y, z = tz * T::Z + f / T::Y;
    } else if (f < fx + fy) {  // +y bond of its last row
      const uint32_t g = f - fx;
      a = 1, x = tx * T::X + g % T::X, y = ty * T::Y + T::Y - 1, z = tz * T::Z + g / T::X;
    } else {  // +z bond of its last plane
      const uint32_t g = f - fx - fy;
      a = 2, x = tx * T::X + g % T::X, y = ty * T::Y + g / T::X, z = tz * T::Z + T::Z - 1;
    }
    const uint32_t x2 = a == 0 ? (x + 1 == lx ? 0u : x + 1) : x;
    const uint32_t y2 = a == 1 ? (y + 1 == ly ? 0u : y + 1) : y;
    const uint32_t z2 = a == 2 ? (z + 1 == lz ? 0u : z + 1) : z;
    const uint64_t s = (uint64_t)x + (uint64_t)lx * ((uint64_t)y + (uint64_t)ly * z);
    const uint64_t t = (uint64_t)x2 + (uint64_t)lx * ((uint64_t)y2 + (uint64_t)ly * z2);
    if (spins[stencil_byte(s, StencilSite{x, y, z}, cs)] != spins[stencil_byte(t, StencilSite{x2, y2, z2}, cs)]) continue;
    // start from the tile-local roots, so that the entry of a site that is
    // not a local root keeps pointing at its own tile's local root
    if ((uint64_t)cluster_word(key, (uint64_t)ND * s + a, step, kStreamClusterBond) < thr)
      unite(parent, load_parent(parent + s), load_parent(parent + t));
  }
}

__device__ __forceinline__ int8_t sw_target(uint32_t root, int kind, int q, uint64_t key, uint64_t step) {
  const uint32_t w = cluster_word(key, root, step, kStreamClusterTarget);
  if (kind == kIsing) return w < 0x80000000u ? (int8_t)1 : (int8_t)-1;
  return (int8_t)(((uint64_t)w * (uint64_t)q) >> 32);
}

// The Wolff seed's root and target, from the spins before the update.
__global__ void k_wolff_seed(const int8_t *__restrict__ spins, uint32_t *parent, int stencil, uint32_t lx,
                             uint32_t ly, uint64_t cs, uint64_t n_sites, int kind, int q, uint64_t key,
                             uint64_t step, uint32_t *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t seed = ((uint64_t)cluster_word(key, 0, step, kStreamClusterSeed) * n_sites) >> 32;
  const int8_t s0 = stencil ? spins[stencil_byte(seed, stencil_site(seed, lx, ly), cs)] : spins[seed];
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
                            uint64_t cs, uint64_t n_sites, int kind, int q, int wolff, uint64_t key, uint64_t step,
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
    const uint64_t at = stencil ? stencil_byte(s, stencil_site(s, lx, ly), cs) : s;
    spins[at] = v;
  }
}

// Stencils with 4 | lx: one thread per four consecutive sites along x.  Sites
// s0 and s0 + 2 sit at bytes s0/2 and s0/2 + 1 of colour c = (y + z) & 1,
// s0 + 1 and s0 + 3 at the same bytes of the other colour, so each colour
// takes one 16-bit store.
__global__ void k_transform_stencil(int8_t *__restrict__ spins, uint32_t *parent, uint32_t lx, uint32_t ly,
                                    uint64_t cs, uint32_t n_sites, int kind, int q, int wolff, uint64_t key,
                                    uint64_t step, const uint32_t *__restrict__ seed) {
  const uint32_t seed_root = wolff ? seed[0] : 0u;
  const int8_t wolff_target = wolff ? (int8_t)(uint8_t)seed[1] : (int8_t)0;
  for (uint64_t g = thread_id(); g < n_sites / 4; g += grid_threads()) {
    const uint32_t s0 = (uint32_t)g * 4, r = s0 / lx, c = (r % ly + r / ly) & 1u;
    const uint4 p4 = *reinterpret_cast<const uint4 *>(parent + s0);
    const uint32_t first[4] = {p4.x, p4.y, p4.z, p4.w};
    int16_t *at_c = reinterpret_cast<int16_t *>(spins + (uint64_t)c * cs + s0 / 2);
    int16_t *at_o = reinterpret_cast<int16_t *>(spins + (uint64_t)(c ^ 1u) * cs + s0 / 2);
    uint32_t v[4];
    bool any = !wolff;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t root = first[k] == s0 + k ? first[k] : find_root(parent, first[k]);
      if (wolff) {
        v[k] = root == seed_root ? (uint32_t)(uint8_t)wolff_target : 0x100u;  // 0x100: keep
        any |= root == seed_root;
      } else {
        v[k] = (uint8_t)sw_target(root, kind, q, key, step);
      }
    }
    if (!any) continue;
    if (wolff) {
      const uint32_t oc = (uint16_t)*at_c, oo = (uint16_t)*at_o;
      v[0] = v[0] & 0x100u ? oc & 0xFFu : v[0];
      v[2] = v[2] & 0x100u ? oc >> 8 : v[2];
      v[1] = v[1] & 0x100u ? oo & 0xFFu : v[1];
      v[3] = v[3] & 0x100u ? oo >> 8 : v[3];
    }
    *at_c = (int16_t)(v[0] | v[2] << 8);
    *at_o = (int16_t)(v[1] | v[3] << 8);
  }
}

// The transform after tile labelling, one block per tile.  A site that is not
// a local root points at its tile's local root (the face merges start from
// local roots and never rewrite such entries), so the block resolves each
// local root once -- one global root walk and one target -- and every site
// reads its local root's result from LDS.  A local root linked to another
// tile's root walks on its own.
template <int ND>
__global__ __launch_bounds__(kThreads) void k_transform_tile(int8_t *__restrict__ spins, uint32_t *parent,
                                                             uint32_t lx, uint32_t ly, uint64_t cs, uint64_t div_lx2,
                                                             uint64_t div_ly, int kind, int q, int wolff, uint64_t key,
                                                             uint64_t step, const uint32_t *__restrict__ seed) {
  using T = Tile<ND>;
  constexpr uint32_t kKeep = 0x100u;  // Wolff: the site keeps its state
  __shared__ uint16_t res[T::N];
  __shared__ uint8_t need[T::N];
  const uint32_t seed_root = wolff ? seed[0] : 0u;
  const uint32_t wolff_target = wolff ? seed[1] & 0xFFu : 0u;
  const uint32_t tiles_x = lx / T::X, tiles_y = ly / T::Y;
  const uint32_t tile = blockIdx.x;
  const uint32_t ox = (tile % tiles_x) * T::X, oy = ((tile / tiles_x) % tiles_y) * T::Y,
                 oz = (tile / (tiles_x * tiles_y)) * T::Z;
  auto global_of = [&](uint32_t i) -> uint32_t {
    const uint32_t x = i % T::X, y = (i / T::X) % T::Y, z = i / (T::X * T::Y);
    return ox + x + lx * (oy + y + ly * (oz + z));
  };
  // tile-local index of global site g, or T::N when g lies in another tile
  auto local_of = [&](uint32_t g) -> uint32_t {
    const uint32_t row = magic_div(g >> 1, div_lx2), x = g - row * lx;
    const uint32_t z = magic_div(row, div_ly), y = row - z * ly;
    const uint32_t dx = x - ox, dy = y - oy, dz = z - oz;
    return dx < T::X && dy < T::Y && dz < T::Z ? dx + T::X * (dy + T::Y * dz) : T::N;
  };
  auto result_of = [&](uint32_t root) -> uint32_t {
    if (wolff) return root == seed_root ? wolff_target : kKeep;
    return (uint8_t)sw_target(root, kind, q, key, step);
  };
  for (uint32_t i = threadIdx.x; i < T::N; i += kThreads) need[i] = 0;
  __syncthreads();
  // four consecutive sites along x per thread (4 | X)
  const uint32_t i0 = 4 * threadIdx.x;
  const uint32_t s0 = global_of(i0);
  const uint4 p4 = *reinterpret_cast<const uint4 *>(parent + s0);
  const uint32_t p[4] = {p4.x, p4.y, p4.z, p4.w};
  uint32_t loc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    loc[k] = local_of(p[k]);
    if (loc[k] < T::N) need[loc[k]] = 1;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < T::N; i += kThreads)
    if (need[i]) res[i] = (uint16_t)result_of(find_root(parent, global_of(i)));
  __syncthreads();
  uint32_t v[4];
  bool any = !wolff;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = loc[k] < T::N ? res[loc[k]] : result_of(find_root(parent, p[k]));
    any |= v[k] != kKeep;
  }
  if (!any) return;
  // sites s0, s0 + 2 of colour c at bytes s0/2, s0/2 + 1; s0 + 1, s0 + 3 of the other
  const uint32_t x0 = i0 % T::X, y = (i0 / T::X) % T::Y, z = i0 / (T::X * T::Y);
  const uint32_t c = (ox + x0 + oy + y + oz + z) & 1u;
  int16_t *at_c = reinterpret_cast<int16_t *>(spins + (uint64_t)c * cs + s0 / 2);
  int16_t *at_o = reinterpret_cast<int16_t *>(spins + (uint64_t)(c ^ 1u) * cs + s0 / 2);
  if (wolff) {
    const uint32_t oc = (uint16_t)*at_c, oo = (uint16_t)*at_o;
    v[0] = v[0] == kKeep ? oc & 0xFFu : v[0];
    v[2] = v[2] == kKeep ? oc >> 8 : v[2];
    v[1] = v[1] == kKeep ? oo & 0xFFu : v[1];
    v[3] = v[3] == kKeep ? oo >> 8 : v[3];
  }
  *at_c = (int16_t)(v[0] | v[2] << 8);
  *at_o = (int16_t)(v[1] | v[3] << 8);
}

// ------------------------------------------------ O(N) embedding clusters
// (ClusterModel for ONModel<D>, lattice/models.rs:516-577), D = 2 or 3, on
// stencil lattices whose spins are FP32 component planes
// [colour][component][rank].  Arithmetic is f64 in the reference's operation
// order (sums left to right), and this file is compiled without contraction,
// so that the bond probabilities, the direction and the reflections round as
// the oracle's do.

template <int D>
__device__ __forceinline__ double dotn(const double (&a)[D], const double *b) {
  double acc = a[0] * b[0];
#pragma unroll
  for (int c = 1; c < D; ++c) acc = acc + a[c] * b[c];
  return acc;
}

// The update's direction: rejection from the square or cube, try k on words
// 4 + Dk .. 3 + D(k + 1) of the seed stream, accepted when
// 2^-40 < |x|^2 <= 1 and normalised as normalize (models.rs:432-439); 64
// rejections in a row (probability 0.215^64 or 0.476^64) leave the last
// unit axis.
template <int D>
__global__ void k_on_direction(uint64_t key, uint64_t step, double *r) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (uint64_t k = 0; k < 64; ++k) {
    double x[D];
#pragma unroll
    for (int c = 0; c < D; ++c)
      x[c] = 2.0 * ((double)cluster_word(key, 4 + D * k + c, step, kStreamClusterSeed) * 0x1p-32) - 1.0;
    const double n2 = dotn<D>(x, x);
    if (n2 > 0x1p-40 && n2 <= 1.0) {
      const double norm = sqrt(n2);
#pragma unroll
      for (int c = 0; c < D; ++c) r[c] = x[c] / norm;
      return;
    }
  }
#pragma unroll
  for (int c = 0; c < D; ++c) r[c] = c == D - 1 ? 1.0 : 0.0;
}

struct OnPlanes {
  float *v;     // one replica: planes (D colour + component) * nh
  uint64_t nh;  // ranks per colour
};

template <int D>
__device__ __forceinline__ void load_on(const OnPlanes &P, uint64_t s, const StencilSite &c, double (&out)[D]) {
  const uint64_t col = (c.x + c.y + c.z) & 1u, j = s >> 1;
#pragma unroll
  for (int k = 0; k < D; ++k) out[k] = (double)P.v[(D * col + k) * P.nh + j];
}

// Edges e = nd s + a to +1 along axis a (build_hypercubic order), weight 1:
// active when word(11, e) 2^-32 < p = 1 - exp(-2 beta J (s_i.r)(s_j.r)) for
// a positive product (cluster_bond_probability, models.rs:537-558).
template <int D, int ND>
__global__ void k_bonds_on(OnPlanes P, uint32_t *parent, uint32_t lx, uint32_t ly, uint32_t lz, uint64_t key,
                           uint64_t step, double beta, double coupling, const double *__restrict__ r_) {
  const uint64_t n = 2 * P.nh;
  double r[D];
#pragma unroll
  for (int c = 0; c < D; ++c) r[c] = r_[c];
  for (uint64_t s = thread_id(); s < n; s += grid_threads()) {
    const StencilSite c = stencil_site(s, lx, ly);
    double v[D];
    load_on<D>(P, s, c, v);
    const double ps = dotn<D>(v, r);
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
      double w[D];
      load_on<D>(P, ts, t, w);
      const double product = ps * dotn<D>(w, r);
      if (coupling == 0.0 || product <= 0.0) continue;
      double p = 1.0 - exp(-2.0 * beta * coupling * product);
      p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);  // clamp_probability; p is finite here
      if ((double)cluster_word(key, (uint64_t)ND * s + a, step, kStreamClusterBond) * 0x1p-32 < p)
        unite(parent, (uint32_t)s, (uint32_t)ts);
    }
  }
}

__global__ void k_on_seed(uint32_t *parent, uint64_t n_sites, uint64_t key, uint64_t step, uint32_t *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint64_t seed = ((uint64_t)cluster_word(key, 0, step, kStreamClusterSeed) * n_sites) >> 32;
  out[0] = find_root(parent, (uint32_t)seed);
}

// Reflection about r (transform_cluster_spin, models.rs:560-575), renormalised
// in f64 and stored as FP32: Swendsen-Wang reflects a cluster when its root's
// target word is below 2^31 (sw_cluster_auxiliary's coin), Wolff the seed's.
template <int D>
__global__ void k_transform_on(OnPlanes P, uint32_t *parent, uint32_t lx, uint32_t ly, int wolff, uint64_t key,
                               uint64_t step, const uint32_t *__restrict__ seed, const double *__restrict__ r_) {
  const uint64_t n = 2 * P.nh;
  const uint32_t seed_root = wolff ? seed[0] : 0u;
  double r[D];
#pragma unroll
  for (int c = 0; c < D; ++c) r[c] = r_[c];
  for (uint64_t s = thread_id(); s < n; s += grid_threads()) {
    const uint32_t root = find_root(parent, (uint32_t)s);
    const bool flip = wolff ? root == seed_root
                            : cluster_word(key, root, step, kStreamClusterTarget) < 0x80000000u;
    if (!flip) continue;
    const StencilSite c = stencil_site(s, lx, ly);
    double v[D];
    load_on<D>(P, s, c, v);
    const double proj = dotn<D>(v, r);
    double u[D];
#pragma unroll
    for (int k = 0; k < D; ++k) u[k] = v[k] - 2.0 * proj * r[k];
    const double norm = sqrt(dotn<D>(u, u));
    if (norm > 1e-15)
#pragma unroll
      for (int k = 0; k < D; ++k) u[k] /= norm;
    const uint64_t col = (c.x + c.y + c.z) & 1u, j = s >> 1;
#pragma unroll
    for (int k = 0; k < D; ++k) P.v[(D * col + k) * P.nh + j] = (float)u[k];
  }
}

// ------------------------------------------------ O(N) clusters, tile by tile
// Stencils whose sides are multiples of the tile: the embedding bonds inside
// a tile come from the sites' projections on r kept in LDS, x bonds as runs
// and y (and z) bonds through an LDS union-find, and every site's tile-local
// root is written; k_bonds_on_faces unites the bonds across tile faces in
// global memory; k_transform_on_tile resolves each local root once.  Draws,
// products, probabilities and reflections are the global kernels', in the
// same operation order, so the partition and the spins are theirs.

// Bond activity (cluster_bond_probability, models.rs:537-558) for the
// product (s_i.r)(s_j.r) of a bond's lower and upper site.
__device__ __forceinline__ bool on_bond_active(double product, double beta, double coupling, uint32_t word) {
  if (coupling == 0.0 || product <= 0.0) return false;
  double p = 1.0 - exp(-2.0 * beta * coupling * product);
  p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);
  return (double)word * 0x1p-32 < p;
}

// The quad of sites s0 .. s0 + 3 (4 | s0): s0 and s0 + 2 are ranks s0/2 and
// s0/2 + 1 of colour c, s0 + 1 and s0 + 3 the same ranks of the other.
template <int D>
__device__ __forceinline__ void load_quad_on(const OnPlanes &P, uint64_t s0, uint32_t c, double (&v)[4][D]) {
  const uint64_t j = s0 >> 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const float2 a = *reinterpret_cast<const float2 *>(P.v + (D * (uint64_t)c + k) * P.nh + j);
    const float2 b = *reinterpret_cast<const float2 *>(P.v + (D * (uint64_t)(c ^ 1u) + k) * P.nh + j);
    v[0][k] = (double)a.x;
    v[2][k] = (double)a.y;
    v[1][k] = (double)b.x;
    v[3][k] = (double)b.y;
  }
}

template <int D, int ND>
__global__ __launch_bounds__(kThreads) void k_bonds_on_tile(OnPlanes P, uint32_t *__restrict__ parent, uint32_t lx,
                                                            uint32_t ly, uint64_t key, uint64_t step, double beta,
                                                            double coupling, const double *__restrict__ r_) {
  using T = Tile<ND>;
  static_assert(T::N == 4 * kThreads, "one quad per thread");
  constexpr uint32_t QR = T::X / 4;  // quads per tile row
  __shared__ uint32_t lp[T::N];
  __shared__ double proj[T::N];    // s.r per site
  __shared__ uint8_t xb[kThreads];  // the quads' x-bond nibbles
  double r[D];
#pragma unroll
  for (int c = 0; c < D; ++c) r[c] = r_[c];
  const uint32_t tiles_x = lx / T::X, tiles_y = ly / T::Y;
  const uint32_t tile = blockIdx.x;
  const uint32_t ox = (tile % tiles_x) * T::X, oy = ((tile / tiles_x) % tiles_y) * T::Y,
                 oz = (tile / (tiles_x * tiles_y)) * T::Z;
  const uint32_t q = threadIdx.x, i0 = 4 * q;
  const uint32_t x0 = i0 % T::X, y = (i0 / T::X) % T::Y, z = i0 / (T::X * T::Y);
  const uint32_t s0 = ox + x0 + lx * (oy + y + ly * (oz + z));
  const uint32_t c = (ox + x0 + oy + y + oz + z) & 1u;
  double v[4][D];
  load_quad_on<D>(P, s0, c, v);
#pragma unroll
  for (int u = 0; u < 4; ++u) proj[i0 + u] = dotn<D>(v[u], r);
  uint32_t w[4 * ND];  // bond words of edges nd s + a for the quad's sites
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const U4 blk = cluster_block(key, (uint64_t)ND * s0 / 4 + k, step, kStreamClusterBond);
    w[4 * k] = blk.x, w[4 * k + 1] = blk.y, w[4 * k + 2] = blk.z, w[4 * k + 3] = blk.w;
  }
  __syncthreads();
  auto bond = [&](uint32_t u, int a, uint32_t i, uint32_t t) {
    return on_bond_active(proj[i] * proj[t], beta, coupling, w[u * ND + a]);
  };
  uint32_t xm = 0;  // bit u: the bond from site u to u + 1 is active and inside the tile
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u)
    if (x0 + u + 1 < T::X && bond(u, 0, i0 + u, i0 + u + 1)) xm |= 1u << u;
  xb[q] = (uint8_t)xm;
  __syncthreads();
  const uint32_t r0 = q - q % QR;
  uint32_t row = 0;
#pragma unroll
  for (uint32_t k = 0; k < QR; ++k) row |= (uint32_t)xb[r0 + k] << (4 * k);
  const uint32_t starts = ~(row << 1);  // bit x: no active bond from x - 1 to x
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    const uint32_t x = x0 + u;
    const uint32_t first = 31u - (uint32_t)__builtin_clz(starts & (0xFFFFFFFFu >> (31u - x)));
    lp[i0 + u] = i0 - x0 + first;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    const uint32_t i = i0 + u;
    if (y + 1 < T::Y && bond(u, 1, i, i + T::X)) unite_lds(lp, i, i + T::X);
    if (ND == 3 && z + 1 < T::Z && bond(u, ND - 1, i, i + T::X * T::Y)) unite_lds(lp, i, i + T::X * T::Y);
  }
  __syncthreads();
  uint32_t rt[4];
#pragma unroll
  for (uint32_t u = 0; u < 4; ++u) {
    const uint32_t l = find_lds(lp, i0 + u);
    rt[u] = ox + l % T::X + lx * (oy + (l / T::X) % T::Y + ly * (oz + l / (T::X * T::Y)));
  }
  *reinterpret_cast<uint4 *>(parent + s0) = make_uint4(rt[0], rt[1], rt[2], rt[3]);
}

// The embedding bonds from each tile's last layer along every axis to the
// next tile (periodic), one thread each.
template <int D, int ND>
__global__ void k_bonds_on_faces(OnPlanes P, uint32_t *parent, uint32_t lx, uint32_t ly, uint32_t lz, uint64_t n_tiles,
                                 uint64_t key, uint64_t step, double beta, double coupling,
                                 const double *__restrict__ r_) {
  using T = Tile<ND>;
  constexpr uint32_t fx = T::Y * T::Z, fy = T::X * T::Z, fz = ND == 3 ? T::X * T::Y : 0u;
  constexpr uint32_t per_tile = fx + fy + fz;
  double r[D];
#pragma unroll
  for (int c = 0; c < D; ++c) r[c] = r_[c];
  const uint32_t tiles_x = lx / T::X, tiles_y = ly / T::Y;
  for (uint64_t k = thread_id(); k < n_tiles * per_tile; k += grid_threads()) {
    const uint32_t tile = (uint32_t)(k / per_tile), f = (uint32_t)(k - (uint64_t)tile * per_tile);
    const uint32_t tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, tz = tile / (tiles_x * tiles_y);
    uint32_t x, y, z, a;
    if (f < fx) {  // +x bond of the tile's last column
      a = 0, x = tx * T::X + T::X - 1, y = ty * T::Y + f % T::Y, z = tz * T::Z + f / T::Y;
    } else if (f < fx + fy) {  // +y bond of its last row
      const uint32_t g = f - fx;
      a = 1, x = tx * T::X + g % T::X, y = ty * T::Y + T::Y - 1, z = tz * T::Z + g / T::X;
    } else {  // +z bond of its last plane
      const uint32_t g = f - fx - fy;
      a = 2, x = tx * T::X + g % T::X, y = ty * T::Y + g / T::X, z = tz * T::Z + T::Z - 1;
    }
    const uint32_t x2 = a == 0 ? (x + 1 == lx ? 0u : x + 1) : x;
    const uint32_t y2 = a == 1 ? (y + 1 == ly ? 0u : y + 1) : y;
    const uint32_t z2 = a == 2 ? (z + 1 == lz ? 0u : z + 1) : z;
    const uint64_t s = (uint64_t)x + (uint64_t)lx * ((uint64_t)y + (uint64_t)ly * z);
    const uint64_t t = (uint64_t)x2 + (uint64_t)lx * ((uint64_t)y2 + (uint64_t)ly * z2);
    double vs[D], vt[D];
    load_on<D>(P, s, StencilSite{x, y, z}, vs);
    load_on<D>(P, t, StencilSite{x2, y2, z2}, vt);
    const double product = dotn<D>(vs, r) * dotn<D>(vt, r);
    // start from the tile-local roots, so that the entry of a site that is
    // not a local root keeps pointing at its own tile's local root
    if (on_bond_active(product, beta, coupling, cluster_word(key, (uint64_t)ND * s + a, step, kStreamClusterBond)))
      unite(parent, load_parent(parent + s), load_parent(parent + t));
  }
}

// Reflection of the flipped clusters (transform_cluster_spin,
// models.rs:560-575), each tile-local root resolved once per tile.
template <int D, int ND>
__global__ __launch_bounds__(kThreads) void k_transform_on_tile(OnPlanes P, uint32_t *parent, uint32_t lx, uint32_t ly,
                                                                uint64_t div_lx2, uint64_t div_ly, int wolff,
                                                                uint64_t key, uint64_t step,
                                                                const uint32_t *__restrict__ seed,
                                                                const double *__restrict__ r_) {
  using T = Tile<ND>;
  __shared__ uint8_t res[T::N];
  __shared__ uint8_t need[T::N];
  const uint32_t seed_root = wolff ? seed[0] : 0u;
  const uint32_t tiles_x = lx / T::X, tiles_y = ly / T::Y;
  const uint32_t tile = blockIdx.x;
  const uint32_t ox = (tile % tiles_x) * T::X, oy = ((tile / tiles_x) % tiles_y) * T::Y,
                 oz = (tile / (tiles_x * tiles_y)) * T::Z;
  auto global_of = [&](uint32_t i) -> uint32_t {
    const uint32_t x = i % T::X, y = (i / T::X) % T::Y, z = i / (T::X * T::Y);
    return ox + x + lx * (oy + y + ly * (oz + z));
  };
  // tile-local index of global site g, or T::N when g lies in another tile
  auto local_of = [&](uint32_t g) -> uint32_t {
    const uint32_t row = magic_div(g >> 1, div_lx2), x = g - row * lx;
    const uint32_t z = magic_div(row, div_ly), y = row - z * ly;
    const uint32_t dx = x - ox, dy = y - oy, dz = z - oz;
    return dx < T::X && dy < T::Y && dz < T::Z ? dx + T::X * (dy + T::Y * dz) : T::N;
  };
  // Swendsen-Wang reflects a cluster when its root's target word is below
  // 2^31 (sw_cluster_auxiliary's coin), Wolff the seed's cluster
  auto flips = [&](uint32_t root) -> uint32_t {
    if (wolff) return root == seed_root ? 1u : 0u;
    return cluster_word(key, root, step, kStreamClusterTarget) < 0x80000000u ? 1u : 0u;
  };
  for (uint32_t i = threadIdx.x; i < T::N; i += kThreads) need[i] = 0;
  __syncthreads();
  const uint32_t i0 = 4 * threadIdx.x;
  const uint32_t s0 = global_of(i0);
  const uint4 p4 = *reinterpret_cast<const uint4 *>(parent + s0);
  const uint32_t p[4] = {p4.x, p4.y, p4.z, p4.w};
  uint32_t loc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    loc[k] = local_of(p[k]);
    if (loc[k] < T::N) need[loc[k]] = 1;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < T::N; i += kThreads)
    if (need[i]) res[i] = (uint8_t)flips(find_root(parent, global_of(i)));
  __syncthreads();
  uint32_t fm = 0;  // bit k: site s0 + k is reflected
#pragma unroll
  for (int k = 0; k < 4; ++k) fm |= (loc[k] < T::N ? (uint32_t)res[loc[k]] : flips(find_root(parent, p[k]))) << k;
  if (!fm) return;
  double r[D];
#pragma unroll
  for (int c = 0; c < D; ++c) r[c] = r_[c];
  const uint32_t x0 = i0 % T::X, y = (i0 / T::X) % T::Y, z = i0 / (T::X * T::Y);
  const uint32_t c = (ox + x0 + oy + y + oz + z) & 1u;
  double v[4][D];
  load_quad_on<D>(P, s0, c, v);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!((fm >> k) & 1u)) continue;
    const double pr = dotn<D>(v[k], r);
    double u[D];
#pragma unroll
    for (int m = 0; m < D; ++m) u[m] = v[k][m] - 2.0 * pr * r[m];
    const double norm = sqrt(dotn<D>(u, u));
    if (norm > 1e-15)
#pragma unroll
      for (int m = 0; m < D; ++m) u[m] /= norm;
    const uint64_t col = (k & 1) ? (c ^ 1u) : c, j = (s0 >> 1) + (uint64_t)(k >> 1);
#pragma unroll
    for (int m = 0; m < D; ++m) P.v[(D * col + m) * P.nh + j] = (float)u[m];
  }
}

uint32_t blocks_for(uint64_t n) {
  const uint64_t b = (n + kThreads - 1) / kThreads;
  return (uint32_t)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

// Graph lattices: edge e = (source, target, weight) in edge order,
// self-loops included (their union is a no-op, their word index counts);
// spins are one replica's planes [component][site].
template <int D>
__global__ void k_bonds_on_csr(const float *__restrict__ v, uint64_t n, const ClusterEdge *__restrict__ edges,
                               const double *__restrict__ weights, uint64_t n_edges, uint32_t *parent, uint64_t key,
                               uint64_t step, double beta, double j, const double *__restrict__ r_) {
  double r[D];
#pragma unroll
  for (int c = 0; c < D; ++c) r[c] = r_[c];
  for (uint64_t e = thread_id(); e < n_edges; e += grid_threads()) {
    const ClusterEdge b = edges[e];
    double a[D], t[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      a[k] = (double)v[(uint64_t)k * n + b.source];
      t[k] = (double)v[(uint64_t)k * n + b.target];
    }
    const double coupling = j * weights[e];
    const double product = dotn<D>(a, r) * dotn<D>(t, r);
    if (coupling == 0.0 || product <= 0.0) continue;
    double p = 1.0 - exp(-2.0 * beta * coupling * product);
    p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);  // clamp_probability; p is finite here
    if ((double)cluster_word(key, e, step, kStreamClusterBond) * 0x1p-32 < p) unite(parent, b.source, b.target);
  }
}

template <int D>
__global__ void k_transform_on_csr(float *__restrict__ v, uint64_t n, uint32_t *parent, int wolff, uint64_t key,
                                   uint64_t step, const uint32_t *__restrict__ seed, const double *__restrict__ r_) {
  const uint32_t seed_root = wolff ? seed[0] : 0u;
  double r[D];
#pragma unroll
  for (int c = 0; c < D; ++c) r[c] = r_[c];
  for (uint64_t s = thread_id(); s < n; s += grid_threads()) {
    const uint32_t root = find_root(parent, (uint32_t)s);
    const bool flip = wolff ? root == seed_root
                            : cluster_word(key, root, step, kStreamClusterTarget) < 0x80000000u;
    if (!flip) continue;
    double x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = (double)v[(uint64_t)k * n + s];
    const double proj = dotn<D>(x, r);
    double u[D];
#pragma unroll
    for (int k = 0; k < D; ++k) u[k] = x[k] - 2.0 * proj * r[k];
    const double norm = sqrt(dotn<D>(u, u));
    if (norm > 1e-15)
#pragma unroll
      for (int k = 0; k < D; ++k) u[k] /= norm;
#pragma unroll
    for (int k = 0; k < D; ++k) v[(uint64_t)k * n + s] = (float)u[k];
  }
}

// Stencils labelled tile by tile: 2-D or 3-D with every side a multiple of
// the tile's.
bool tiled(const StencilGeom &g) {
  const uint32_t lx = 2 * g.lx2;
  if (g.nd == 2) return lx % Tile<2>::X == 0 && g.ly % Tile<2>::Y == 0;
  if (g.nd == 3) return lx % Tile<3>::X == 0 && g.ly % Tile<3>::Y == 0 && g.lz % Tile<3>::Z == 0;
  return false;
}

template <int D>
void launch_on_update(const StencilGeom &g, float *planes, uint32_t *parent, uint64_t key, uint64_t step,
                      double beta, double j, int wolff, uint32_t *seed, double *r, hipStream_t stream) {
  const uint64_t n = 2 * g.nh;
  const uint32_t lx = 2 * g.lx2;
  const OnPlanes P{planes, g.nh};
  hipLaunchKernelGGL(k_on_direction<D>, dim3(1), dim3(64), 0, stream, key, step, r);
  if ((g.nd == 2 || g.nd == 3) && tiled(g)) {
#define SC_ON_TILED(ND)                                                                                              \
  {                                                                                                                  \
    using T = Tile<ND>;                                                                                              \
    constexpr uint64_t per_tile = T::Y * T::Z + T::X * T::Z + (ND == 3 ? T::X * T::Y : 0);                           \
    const uint64_t tiles = n / T::N;                                                                                 \
    hipLaunchKernelGGL((k_bonds_on_tile<D, ND>), dim3((uint32_t)tiles), dim3(kThreads), 0, stream, P, parent, lx,    \
                       g.ly, key, step, beta, j, r);                                                                 \
    hipLaunchKernelGGL((k_bonds_on_faces<D, ND>), dim3(blocks_for(tiles * per_tile)), dim3(kThreads), 0, stream, P,  \
                       parent, lx, g.ly, g.lz, tiles, key, step, beta, j, r);                                       \
    if (wolff) hipLaunchKernelGGL(k_on_seed, dim3(1), dim3(64), 0, stream, parent, n, key, step, seed);             \
    hipLaunchKernelGGL((k_transform_on_tile<D, ND>), dim3((uint32_t)tiles), dim3(kThreads), 0, stream, P, parent,   \
                       lx, g.ly, g.div_lx2, g.div_ly, wolff, key, step, seed, r);                                   \
  }
    if (g.nd == 2) SC_ON_TILED(2) else SC_ON_TILED(3)
#undef SC_ON_TILED
    return;
  }
  hipLaunchKernelGGL(k_uf_init, dim3(blocks_for(n)), dim3(kThreads), 0, stream, parent, n);
  if (g.nd == 1)
    hipLaunchKernelGGL((k_bonds_on<D, 1>), dim3(blocks_for(n)), dim3(kThreads), 0, stream, P, parent, lx, g.ly, g.lz,
                       key, step, beta, j, r);
  else if (g.nd == 2)
    hipLaunchKernelGGL((k_bonds_on<D, 2>), dim3(blocks_for(n)), dim3(kThreads), 0, stream, P, parent, lx, g.ly, g.lz,
                       key, step, beta, j, r);
  else
    hipLaunchKernelGGL((k_bonds_on<D, 3>), dim3(blocks_for(n)), dim3(kThreads), 0, stream, P, parent, lx, g.ly, g.lz,
                       key, step, beta, j, r);
  if (wolff) hipLaunchKernelGGL(k_on_seed, dim3(1), dim3(64), 0, stream, parent, n, key, step, seed);
  hipLaunchKernelGGL(k_transform_on<D>, dim3(blocks_for(n)), dim3(kThreads), 0, stream, P, parent, lx, g.ly, wolff,
                     key, step, seed, r);
}


}  // namespace

hipError_t launch_cluster_bonds_stencil(const StencilGeom &g, const int8_t *spins, uint32_t *parent, uint64_t key,
                                        uint64_t step, uint64_t thr, hipStream_t stream) {
  if (g.cs < g.nh) return hipErrorInvalidValue;  // the colour stride is unset
  const uint64_t n = 2 * g.nh;
  const uint32_t lxs = 2 * g.lx2;
  if (g.nd == 2 && tiled(g)) {
    const uint64_t tiles = n / Tile<2>::N;
    hipLaunchKernelGGL(k_bonds_tile<2>, dim3((uint32_t)tiles), dim3(kThreads), 0, stream, spins, parent, lxs, g.ly,
                       g.cs, key, step, thr);
    hipLaunchKernelGGL(k_bonds_faces<2>, dim3(blocks_for(tiles * 64)), dim3(kThreads), 0, stream, spins, parent, lxs,
                       g.ly, g.lz, g.cs, tiles, key, step, thr);
    return hipGetLastError();
  }
  if (g.nd == 3 && tiled(g)) {
    const uint64_t tiles = n / Tile<3>::N;
    hipLaunchKernelGGL(k_bonds_tile<3>, dim3((uint32_t)tiles), dim3(kThreads), 0, stream, spins, parent, lxs, g.ly,
                       g.cs, key, step, thr);
    hipLaunchKernelGGL(k_bonds_faces<3>, dim3(blocks_for(tiles * 320)), dim3(kThreads), 0, stream, spins, parent, lxs,
                       g.ly, g.lz, g.cs, tiles, key, step, thr);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_uf_init, dim3(blocks_for(n)), dim3(kThreads), 0, stream, parent, n);
  const uint32_t lx = 2 * g.lx2;
  if (g.nd == 1)
    hipLaunchKernelGGL(k_bonds_stencil<1>, dim3(blocks_for(n)), dim3(kThreads), 0, stream, spins, parent, lx,
                       g.ly, g.lz, g.nh, g.cs, key, step, thr);
  else if (g.nd == 2)
    hipLaunchKernelGGL(k_bonds_stencil<2>, dim3(blocks_for(n)), dim3(kThreads), 0, stream, spins, parent, lx,
                       g.ly, g.lz, g.nh, g.cs, key, step, thr);
  else
    hipLaunchKernelGGL(k_bonds_stencil<3>, dim3(blocks_for(n)), dim3(kThreads), 0, stream, spins, parent, lx,
                       g.ly, g.lz, g.nh, g.cs, key, step, thr);
  return hipGetLastError();
}

hipError_t launch_cluster_on(const StencilGeom &g, int dim, float *planes, uint32_t *parent, uint64_t key,
                             uint64_t step, double beta, double j, int wolff, void *scratch, hipStream_t stream) {
  uint32_t *seed = static_cast<uint32_t *>(scratch);
  double *r = reinterpret_cast<double *>(static_cast<char *>(scratch) + 8);
  if (dim == 2) launch_on_update<2>(g, planes, parent, key, step, beta, j, wolff, seed, r, stream);
  else if (dim == 3) launch_on_update<3>(g, planes, parent, key, step, beta, j, wolff, seed, r, stream);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_cluster_on_csr(uint64_t n_sites, uint64_t n_edges, const ClusterEdge *edges, const double *weights,
                                 int dim, float *planes, uint32_t *parent, uint64_t key, uint64_t step, double beta,
                                 double j, int wolff, void *scratch, hipStream_t stream) {
  uint32_t *seed = static_cast<uint32_t *>(scratch);
  double *r = reinterpret_cast<double *>(static_cast<char *>(scratch) + 8);
  if (dim != 2 && dim != 3) return hipErrorInvalidValue;
  if (dim == 2) hipLaunchKernelGGL(k_on_direction<2>, dim3(1), dim3(64), 0, stream, key, step, r);
  else hipLaunchKernelGGL(k_on_direction<3>, dim3(1), dim3(64), 0, stream, key, step, r);
  hipLaunchKernelGGL(k_uf_init, dim3(blocks_for(n_sites)), dim3(kThreads), 0, stream, parent, n_sites);
  if (n_edges) {
    if (dim == 2)
      hipLaunchKernelGGL(k_bonds_on_csr<2>, dim3(blocks_for(n_edges)), dim3(kThreads), 0, stream, planes, n_sites,
                         edges, weights, n_edges, parent, key, step, beta, j, r);
    else
      hipLaunchKernelGGL(k_bonds_on_csr<3>, dim3(blocks_for(n_edges)), dim3(kThreads), 0, stream, planes, n_sites,
                         edges, weights, n_edges, parent, key, step, beta, j, r);
  }
  if (wolff) hipLaunchKernelGGL(k_on_seed, dim3(1), dim3(64), 0, stream, parent, n_sites, key, step, seed);
  if (dim == 2)
    hipLaunchKernelGGL(k_transform_on_csr<2>, dim3(blocks_for(n_sites)), dim3(kThreads), 0, stream, planes, n_sites,
                       parent, wolff, key, step, seed, r);
  else
    hipLaunchKernelGGL(k_transform_on_csr<3>, dim3(blocks_for(n_sites)), dim3(kThreads), 0, stream, planes, n_sites,
                       parent, wolff, key, step, seed, r);
  return hipGetLastError();
}

hipError_t launch_cluster_bonds_csr(uint64_t n_sites, uint64_t n_edges, const ClusterEdge *edges,
                                    const int8_t *weights, const int8_t *spins, uint32_t *parent, uint64_t key,
                                    uint64_t step, const uint64_t *thr, const uint64_t *edge_thr,
                                    hipStream_t stream) {
  hipLaunchKernelGGL(k_uf_init, dim3(blocks_for(n_sites)), dim3(kThreads), 0, stream, parent, n_sites);
  if (n_edges)
    hipLaunchKernelGGL(k_bonds_csr, dim3(blocks_for(n_edges)), dim3(kThreads), 0, stream, n_edges, edges, weights,
                       spins, parent, key, step, thr, edge_thr);
  return hipGetLastError();
}

hipError_t launch_cluster_transform(const StencilGeom *g, uint64_t n_sites, int8_t *spins, uint32_t *parent,
                                    int kind, int q, int wolff, uint64_t key, uint64_t step, uint32_t *scratch,
                                    hipStream_t stream) {
  const int stencil = g != nullptr;
  if (stencil && g->cs < g->nh) return hipErrorInvalidValue;
  const uint32_t lx = stencil ? 2 * g->lx2 : 0u, ly = stencil ? g->ly : 0u;
  const uint64_t cs = stencil ? g->cs : 0u;  // bytes from colour 0 to colour 1
  if (wolff)
    hipLaunchKernelGGL(k_wolff_seed, dim3(1), dim3(64), 0, stream, spins, parent, stencil, lx, ly, cs, n_sites, kind,
                       q, key, step, scratch);
  if (stencil && tiled(*g)) {
    const uint32_t tiles = (uint32_t)(n_sites / Tile<2>::N);
    if (g->nd == 2)
      hipLaunchKernelGGL(k_transform_tile<2>, dim3(tiles), dim3(kThreads), 0, stream, spins, parent, lx, ly, cs,
                         g->div_lx2, g->div_ly, kind, q, wolff, key, step, scratch);
    else
      hipLaunchKernelGGL(k_transform_tile<3>, dim3(tiles), dim3(kThreads), 0, stream, spins, parent, lx, ly, cs,
                         g->div_lx2, g->div_ly, kind, q, wolff, key, step, scratch);
    return hipGetLastError();
  }
  if (stencil && g->nd >= 2 && lx % 4 == 0) {
    hipLaunchKernelGGL(k_transform_stencil, dim3(blocks_for(n_sites / 4)), dim3(kThreads), 0, stream, spins, parent,
                       lx, ly, cs, (uint32_t)n_sites, kind, q, wolff, key, step, scratch);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_transform, dim3(blocks_for(n_sites)), dim3(kThreads), 0, stream, spins, parent, stencil, lx, ly,
                     cs, n_sites, kind, q, wolff, key, step, scratch);
  return hipGetLastError();
}

}  // namespace scamper
