// Checkerboard single-site updates of discrete spins (Ising +-1 and Potts
// 0..q-1 as int8) on periodic hypercubic lattices with even sides.
//
// Per-site rules follow the reference's MetropolisCore / HeatBathCore
// (CMC.rs/src/algorithms/metropolis.rs:67-106, heat_bath.rs:40-72) with the
// visit order "colour 0 then colour 1"; decisions come from per-replica
// threshold tables built on the host (host_tables.cpp), so a kernel only
// compares 16-bit Philox halves with integers.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t wrap_dec(uint32_t v, uint32_t n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ uint32_t wrap_inc(uint32_t v, uint32_t n) { return v + 1 == n ? 0 : v + 1; }

// Resolve "word < K" from the high half; `lo` is drawn only on a tie.
struct LazyLo {
  uint64_t key, j, sweep;
  uint32_t colour, stream;
  __device__ uint32_t operator()() const { return site_half(key, j, sweep, colour, stream + 1); }
};

__device__ __forceinline__ bool decide(uint32_t hi, Threshold t, const LazyLo &lo) {
  if (hi != t.khi) return hi < t.khi;
  return lo() < t.klo;
}

// Proposal index floor(word * m / 2^32) for m = q-1 (models.rs:212-218);
// the high half decides unless its 2^16-wide bucket straddles a boundary.
__device__ __forceinline__ uint32_t propose_index(uint32_t hi, uint32_t m, const LazyLo &lo) {
  const uint64_t base = (uint64_t)hi << 16;
  const uint32_t a = (uint32_t)((base * m) >> 32);
  const uint32_t b = (uint32_t)(((base + 0xFFFFu) * m) >> 32);
  if (a == b) return a;
  return (uint32_t)(((base | lo()) * m) >> 32);
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
                                                 int q, const double *__restrict__ hb_exp,
                                                 long long *__restrict__ em) {
  constexpr int z = 2 * ND;
  const uint32_t rep = blockIdx.y;
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const ReplicaParams &P = reps[rep];
  __shared__ Threshold sthr[kMaxTable];
  if (threadIdx.x < kMaxTable) sthr[threadIdx.x] = P.thr[threadIdx.x];
  __syncthreads();
  long long acc[2 + kMaxQ];
#pragma unroll
  for (int k = 0; k < 2 + kMaxQ; ++k) acc[k] = 0;
  if (j < g.nh) {
    int8_t *__restrict__ own = spins + (2ull * rep + colour) * g.nh;
    const int8_t *__restrict__ oth = spins + (2ull * rep + 1 - colour) * g.nh;
    const uint32_t row = (uint32_t)(j / g.lx2), i = (uint32_t)(j - (uint64_t)row * g.lx2);
    const uint32_t y = row % g.ly, zz = row / g.ly;
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
        if (t.khi > 0xFFFFu) {
          ns = -s;
        } else if (t.khi != 0 || t.klo != 0) {
          const uint32_t hi = site_half(key, j, sweep, colour, kStreamAccept);
          if (decide(hi, t, LazyLo{key, j, sweep, colour, kStreamAccept})) ns = -s;
        }
      } else {  // heat bath: +1 iff word < K[#minus neighbours]
        int minus = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) minus += (n.v[k] < 0);
        const uint32_t hi = site_half(key, j, sweep, colour, kStreamAccept);
        ns = decide(hi, sthr[minus], LazyLo{key, j, sweep, colour, kStreamAccept}) ? 1 : -1;
      }
    } else {
      if (ALG == kMetropolis) {
        const uint32_t phi = site_half(key, j, sweep, colour, kStreamPropose);
        uint32_t r = propose_index(phi, (uint32_t)(q - 1), LazyLo{key, j, sweep, colour, kStreamPropose});
        if (r >= (uint32_t)s) r += 1;
        int d = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) d += ((uint32_t)n.v[k] == r) - (n.v[k] == s);
        const Threshold t = sthr[d + z];
        bool accept;
        if (t.khi > 0xFFFFu)
          accept = true;
        else if (t.khi == 0 && t.klo == 0)
          accept = false;
        else
          accept = decide(site_half(key, j, sweep, colour, kStreamAccept), t,
                          LazyLo{key, j, sweep, colour, kStreamAccept});
        if (accept) ns = (int8_t)r;
      } else {
#pragma clang fp contract(off)
        // heat bath (models.rs:340-380) in f64: weights from the host's exp
        // table, then the reference's sequential threshold subtraction.
        int cnt[kMaxQ];
        for (int v = 0; v < q; ++v) cnt[v] = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) cnt[(uint8_t)n.v[k]] += 1;
        int nmax = 0;
        for (int v = 0; v < q; ++v) nmax = cnt[v] > nmax ? cnt[v] : nmax;
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
        acc[0] = (long long)ns * sum;
        acc[1] = (long long)ns + other_site;
      } else {
        int same = 0;
#pragma unroll
        for (int k = 0; k < z; ++k) same += (n.v[k] == ns);
        acc[0] = same;
        for (int v = 0; v < q; ++v) acc[2 + v] = (ns == v) + (other_site == v);
      }
    }
  }
  if (MEAS) block_accumulate<2 + kMaxQ>(acc, KIND == kIsing ? 2 : 2 + q, em + (uint64_t)rep * kEmStride);
}

// ------------------------------------------------------------------------
// Fast Ising Metropolis (2-D and 3-D): one thread owns 16 consecutive bytes
// of a row and walks R rows of one plane, keeping the other colour's rows
// y-1, y, y+1 in registers.  Anti-aligned neighbour counts are formed four
// sites per 32-bit word (spins are 0x01 / 0xFF, so (s ^ n) has bit 1 set
// exactly when the two differ).  Requires lx2 % 16 == 0 and ly % R == 0.
struct Chunk {
  uint32_t w[4];
};

__device__ __forceinline__ Chunk load16(const int8_t *p) {
  const uint4 v = *reinterpret_cast<const uint4 *>(p);
  return Chunk{{v.x, v.y, v.z, v.w}};
}
__device__ __forceinline__ void store16(int8_t *p, const Chunk &c) {
  *reinterpret_cast<uint4 *>(p) = make_uint4(c.w[0], c.w[1], c.w[2], c.w[3]);
}
__device__ __forceinline__ uint32_t diff_bits(uint32_t a, uint32_t b) { return ((a ^ b) >> 1) & 0x01010101u; }

template <int ND, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_ising_fast(int8_t *__restrict__ spins,
                                                       const ReplicaParams *__restrict__ reps,
                                                       StencilGeom g, uint32_t colour,
                                                       uint64_t sweep, uint32_t R,
                                                       long long *__restrict__ em) {
  const uint32_t rep = blockIdx.y;
  // XCD-aware order: the hardware deals consecutive blocks round-robin over
  // 8 XCDs; remap so each XCD walks one contiguous band of rows and the
  // halo rows it shares stay in that XCD's L2.
  uint32_t b = blockIdx.x;
  const uint32_t nb = gridDim.x;
  if ((nb & 7u) == 0) b = (b & 7u) * (nb >> 3) + (b >> 3);
  const ReplicaParams &P = reps[rep];
  __shared__ Threshold sthr[2 * ND + 1];
  if (threadIdx.x < 2 * ND + 1) sthr[threadIdx.x] = P.thr[threadIdx.x];
  __syncthreads();

  const uint32_t cpr = g.lx2 >> 4;
  const uint64_t tid = (uint64_t)b * kBlock + threadIdx.x;
  const uint32_t groups_per_plane = g.ly / R;
  const uint64_t total = (uint64_t)cpr * groups_per_plane * g.lz;
  long long acc0 = 0, acc1 = 0;
  if (tid < total) {
    const uint32_t chunk = (uint32_t)(tid % cpr);
    const uint64_t rg = tid / cpr;
    const uint32_t zz = (uint32_t)(rg / groups_per_plane);
    const uint32_t y0 = (uint32_t)(rg % groups_per_plane) * R;
    const uint32_t i0 = chunk * 16;
    int8_t *__restrict__ own = spins + (2ull * rep + colour) * g.nh;
    const int8_t *__restrict__ oth = spins + (2ull * rep + 1 - colour) * g.nh;
    const uint64_t plane = (uint64_t)g.ly * zz;
    const uint64_t key = P.key;
    const uint32_t left_i = chunk == 0 ? g.lx2 - 4 : i0 - 4;      // dword holding byte i0-1
    const uint32_t right_i = i0 + 16 == g.lx2 ? 0 : i0 + 16;       // dword holding byte i0+16

    Chunk up = load16(oth + (plane + wrap_dec(y0, g.ly)) * g.lx2 + i0);
    Chunk mid = load16(oth + (plane + y0) * g.lx2 + i0);
    for (uint32_t r = 0; r < R; ++r) {
      const uint32_t y = y0 + r;
      const uint64_t row = plane + y;
      const Chunk down = load16(oth + (plane + wrap_inc(y, g.ly)) * g.lx2 + i0);
      Chunk o = load16(own + row * g.lx2 + i0);
      const uint32_t p = (colour + y + zz) & 1u;
      // horizontal partner: bytes i0-1..i0+14 (p = 0) or i0+1..i0+16 (p = 1)
      Chunk sh;
      if (p == 0) {
        const uint32_t edge = *reinterpret_cast<const uint32_t *>(oth + row * g.lx2 + left_i);
        sh.w[0] = (mid.w[0] << 8) | (edge >> 24);
        sh.w[1] = (mid.w[1] << 8) | (mid.w[0] >> 24);
        sh.w[2] = (mid.w[2] << 8) | (mid.w[1] >> 24);
        sh.w[3] = (mid.w[3] << 8) | (mid.w[2] >> 24);
      } else {
        const uint32_t edge = *reinterpret_cast<const uint32_t *>(oth + row * g.lx2 + right_i);
        sh.w[0] = (mid.w[0] >> 8) | (mid.w[1] << 24);
        sh.w[1] = (mid.w[1] >> 8) | (mid.w[2] << 24);
        sh.w[2] = (mid.w[2] >> 8) | (mid.w[3] << 24);
        sh.w[3] = (mid.w[3] >> 8) | (edge << 24);
      }
      Chunk zm, zp;
      if (ND == 3) {
        zm = load16(oth + ((uint64_t)g.ly * wrap_dec(zz, g.lz) + y) * g.lx2 + i0);
        zp = load16(oth + ((uint64_t)g.ly * wrap_inc(zz, g.lz) + y) * g.lx2 + i0);
      }
      uint32_t anti[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        uint32_t a = diff_bits(o.w[k], up.w[k]) + diff_bits(o.w[k], down.w[k]) +
                     diff_bits(o.w[k], mid.w[k]) + diff_bits(o.w[k], sh.w[k]);
        if (ND == 3) a += diff_bits(o.w[k], zm.w[k]) + diff_bits(o.w[k], zp.w[k]);
        anti[k] = a;
      }
      const uint64_t j0 = row * g.lx2 + i0;
      const U4 b0 = group_block(key, j0 >> 3, sweep, colour, kStreamAccept);
      const U4 b1 = group_block(key, (j0 >> 3) + 1, sweep, colour, kStreamAccept);
      uint32_t flip[4] = {0, 0, 0, 0};
      uint32_t ties = 0;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const uint32_t a = (anti[k >> 2] >> (8 * (k & 3))) & 0xFFu;
        const Threshold t = sthr[a];
        const uint32_t hi = half_of(k < 8 ? b0 : b1, k & 7);
        if (hi < t.khi) flip[k >> 2] |= 0xFEu << (8 * (k & 3));
        if (hi == t.khi) ties |= 1u << k;
      }
      if (ties) {
        const U4 l0 = (ties & 0xFFu) ? group_block(key, j0 >> 3, sweep, colour, kStreamAccept + 1) : U4{0, 0, 0, 0};
        const U4 l1 = (ties >> 8) ? group_block(key, (j0 >> 3) + 1, sweep, colour, kStreamAccept + 1) : U4{0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          if (ties & (1u << k)) {
            const uint32_t a = (anti[k >> 2] >> (8 * (k & 3))) & 0xFFu;
            if (half_of(k < 8 ? l0 : l1, k & 7) < sthr[a].klo) flip[k >> 2] |= 0xFEu << (8 * (k & 3));
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) o.w[k] ^= flip[k];
      store16(own + row * g.lx2 + i0, o);
      if (MEAS) {
        // bond sum of the new spins: sum over sites of (z - 2 * anti_new),
        // anti_new = z - anti where flipped; spin sums from sign bits.
        constexpr uint32_t z = 2 * ND;
        uint32_t anti_sum = 0, neg = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t fl = (flip[k] >> 1) & 0x01010101u;  // 1 per flipped byte
          // anti_new = anti + fl * (z - 2 anti), bytewise without borrow:
          const uint32_t an = anti[k] + fl * z - 2u * (anti[k] & (fl * 0xFFu));
          anti_sum += __builtin_amdgcn_sad_u8(an, 0u, 0u);
          neg += __builtin_popcount(o.w[k] & 0x80808080u) + __builtin_popcount(mid.w[k] & 0x80808080u);
        }
        acc0 += (long long)(16 * z) - 2ll * anti_sum;
        acc1 += 32ll - 2ll * neg;
      }
      up = mid;
      mid = down;
    }
  }
  if (MEAS) {
    long long a[2] = {acc0, acc1};
    block_accumulate<2>(a, 2, em + (uint64_t)rep * kEmStride);
  }
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
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  long long acc[2 + kMaxQ];
#pragma unroll
  for (int k = 0; k < 2 + kMaxQ; ++k) acc[k] = 0;
  if (j < g.nh) {
    const int8_t *own = spins + (2ull * rep + 1) * g.nh;
    const int8_t *oth = spins + (2ull * rep) * g.nh;
    const uint32_t row = (uint32_t)(j / g.lx2), i = (uint32_t)(j - (uint64_t)row * g.lx2);
    const uint32_t y = row % g.ly, zz = row / g.ly;
    const uint32_t p = (1 + y + zz) & 1u;
    const Neighbours<ND> n = gather<ND>(oth, g, row, i, y, zz, p);
    const int8_t s = own[j], o = oth[j];
    if (KIND == kIsing) {
      int sum = 0;
#pragma unroll
      for (int k = 0; k < z; ++k) sum += n.v[k];
      acc[0] = (long long)s * sum;
      acc[1] = (long long)s + o;
    } else {
      int same = 0;
#pragma unroll
      for (int k = 0; k < z; ++k) same += (n.v[k] == s);
      acc[0] = same;
      for (int v = 0; v < q; ++v) acc[2 + v] = (s == v) + (o == v);
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
  const uint32_t row = (uint32_t)(j / g.lx2), i = (uint32_t)(j - (uint64_t)row * g.lx2);
  const uint32_t y = row % g.ly, zz = row / g.ly;
  const uint64_t site = 2ull * i + ((colour + y + zz) & 1u) + 2ull * g.lx2 * row;
  int8_t v;
  if (!random)
    v = kind == kIsing ? 1 : 0;
  else {
    const uint32_t h = site_half(key, site, 0, 0, kStreamInit);
    v = kind == kIsing ? ((h & 1u) ? 1 : -1) : (int8_t)((h * (uint32_t)q) >> 16);
  }
  spins[(2ull * rep + colour) * g.nh + j] = v;
}

template <int ND>
hipError_t dispatch_site(const StencilGeom &g, int kind, int q, int alg, int8_t *spins,
                         const ReplicaParams *reps, uint32_t n_rep, uint32_t colour, uint64_t sweep,
                         const double *hb_exp, long long *em, hipStream_t st) {
  dim3 grid((unsigned)((g.nh + kBlock - 1) / kBlock), n_rep);
  const bool meas = em != nullptr;
#define SC_LAUNCH(K, A, M) \
  hipLaunchKernelGGL((k_site<ND, K, A, M>), grid, dim3(kBlock), 0, st, spins, reps, g, colour, sweep, q, hb_exp, em)
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

uint32_t fast_rows(const StencilGeom &g) {
  for (uint32_t r : {16u, 8u, 4u, 2u, 1u})
    if (g.ly % r == 0) return r;
  return 1;
}

}  // namespace

bool stencil_fast_path_ok(const StencilGeom &g) {
  return (g.nd == 2 || g.nd == 3) && g.lx2 % 16 == 0 && g.ly >= 2;
}

hipError_t launch_stencil_discrete(const StencilGeom &g, int kind, int q, int alg, int8_t *spins,
                                   const ReplicaParams *reps, uint32_t n_rep, uint32_t colour,
                                   uint64_t sweep, const double *hb_exp, long long *em,
                                   hipStream_t st) {
  if (kind == kIsing && alg == kMetropolis && stencil_fast_path_ok(g)) {
    const uint32_t R = fast_rows(g);
    const uint64_t threads = (uint64_t)(g.lx2 / 16) * (g.ly / R) * g.lz;
    dim3 grid((unsigned)((threads + kBlock - 1) / kBlock), n_rep);
    if (g.nd == 2) {
      if (em) hipLaunchKernelGGL((k_ising_fast<2, true>), grid, dim3(kBlock), 0, st, spins, reps, g, colour, sweep, R, em);
      else hipLaunchKernelGGL((k_ising_fast<2, false>), grid, dim3(kBlock), 0, st, spins, reps, g, colour, sweep, R, em);
    } else {
      if (em) hipLaunchKernelGGL((k_ising_fast<3, true>), grid, dim3(kBlock), 0, st, spins, reps, g, colour, sweep, R, em);
      else hipLaunchKernelGGL((k_ising_fast<3, false>), grid, dim3(kBlock), 0, st, spins, reps, g, colour, sweep, R, em);
    }
    return hipGetLastError();
  }
  switch (g.nd) {
    case 1: return dispatch_site<1>(g, kind, q, alg, spins, reps, n_rep, colour, sweep, hb_exp, em, st);
    case 2: return dispatch_site<2>(g, kind, q, alg, spins, reps, n_rep, colour, sweep, hb_exp, em, st);
    case 3: return dispatch_site<3>(g, kind, q, alg, spins, reps, n_rep, colour, sweep, hb_exp, em, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stencil_measure(const StencilGeom &g, int kind, int q, const int8_t *spins,
                                  uint32_t n_rep, long long *em, hipStream_t st) {
  dim3 grid((unsigned)((g.nh + kBlock - 1) / kBlock), n_rep);
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
  dim3 grid((unsigned)((g.nh + kBlock - 1) / kBlock), 2);
  hipLaunchKernelGGL(k_init, grid, dim3(kBlock), 0, st, spins, g, kind, q, key, rep, random);
  return hipGetLastError();
}

}  // namespace scamper
