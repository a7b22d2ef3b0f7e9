// Register-blocked checkerboard Metropolis for the Ising model on 2-D and
// 3-D periodic lattices (int8 spins, colour-split rows).
//
// One thread owns 16 consecutive bytes (16 sites) of a row and walks R rows
// of one plane, keeping the other colour's rows y-1, y, y+1 in registers and
// issuing row r+1's loads before row r is computed.  Per row it needs two
// Philox blocks: their 16 bytes are the top and second bytes of the 16
// sites' draws, laid out exactly like the spins, so every decision is made
// four sites at a time with byte-wise integer arithmetic (SWAR):
//
//   anti  = number of anti-aligned neighbours per byte (spins are 0x01/0xFF,
//           so (s ^ n) has bit 1 set exactly when the two differ)
//   t8,m8 = top and second byte of the threshold K[anti]   (v_perm lookups)
//   flip  = (top, mid) < (t8, m8), or K[anti] = 2^32 (every word accepts);
//           the pairs are interleaved into 16-bit lanes by v_perm and
//           compared two sites at a time with packed u16 max/min
//   tie   = (top, mid) == (t8, m8)   (one uphill site in ~65536)
//
// A tie needs the low 16 bits of the draw from a third block.  The row loop
// only records where ties are (one LDS word per row, no branch); after the R
// rows each tied site is settled from its neighbours.
//
// The per-site rule is MetropolisCore's (CMC.rs/src/algorithms/metropolis.rs:
// 67-106) with accept_log_probability (Carlo.rs/src/acceptance.rs:11-22)
// turned into integer thresholds on the host (tables.cpp).
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMsb = 0x80808080u;
constexpr uint32_t kMaxRows = 16;

__device__ __forceinline__ uint32_t wrap_dec(uint32_t v, uint32_t n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ uint32_t wrap_inc(uint32_t v, uint32_t n) { return v + 1 == n ? 0 : v + 1; }

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

// 0x02 in every byte where a and b differ (spins 0x01 / 0xFF).
__device__ __forceinline__ uint32_t differ2(uint32_t a, uint32_t b) { return (a ^ b) & 0x02020202u; }

// Packed unsigned 16-bit max / min of two lanes (v_pk_max_u16, v_pk_min_u16).
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_max_u16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t,
                            __builtin_elementwise_max(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_min_u16(uint32_t a, uint32_t b) {
  return __builtin_bit_cast(uint32_t,
                            __builtin_elementwise_min(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}

// v_perm selectors: the 16-bit values (top << 8 | mid) of sites 0, 1 (kPair01)
// or 2, 3 (kPair23) of a word, from perm(top word, mid word); and the bytes 0
// and 2 of two such lane words gathered back into site order (kGather).
constexpr uint32_t kPair01 = 0x05010400u, kPair23 = 0x07030602u, kGather = 0x06040200u;

// Settle one tied site: its spin is still the pre-sweep value; recount its
// anti-aligned neighbours and compare the low 24 bits of its draw.
template <int ND, bool MEAS>
__device__ void settle_tie(int8_t *__restrict__ own, const int8_t *__restrict__ oth, const StencilGeom &g,
                           uint32_t y, uint32_t zz, uint32_t i, uint32_t colour, uint64_t sweep,
                           const ReplicaParams &P, long long &acc0, long long &acc1) {
  constexpr int z = 2 * ND;
  const uint64_t plane = (uint64_t)g.ly * zz;
  const uint64_t j = (plane + y) * g.lx2 + i;
  const int8_t s = own[j];
  const uint32_t p = (colour + y + zz) & 1u;
  const uint64_t rb = (plane + y) * g.lx2;
  int anti = (oth[rb + (p ? i : wrap_dec(i, g.lx2))] != s) + (oth[rb + (p ? wrap_inc(i, g.lx2) : i)] != s) +
             (oth[(plane + wrap_dec(y, g.ly)) * g.lx2 + i] != s) + (oth[(plane + wrap_inc(y, g.ly)) * g.lx2 + i] != s);
  if (ND == 3)
    anti += (oth[((uint64_t)g.ly * wrap_dec(zz, g.lz) + y) * g.lx2 + i] != s) +
            (oth[((uint64_t)g.ly * wrap_inc(zz, g.lz) + y) * g.lx2 + i] != s);
  if (site_low(P.key, j, sweep, colour, kStreamAccept) < P.thr[anti].k16) {
    own[j] = -s;
    if (MEAS) {
      acc0 += 4ll * anti - 2ll * z;  // bond sum of the site changes sign
      acc1 += s < 0 ? 2 : -2;
    }
  }
}

template <int ND, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_ising_fast(int8_t *__restrict__ spins,
                                                       const ReplicaParams *__restrict__ reps,
                                                       StencilGeom g, uint32_t colour,
                                                       uint64_t sweep, uint32_t R,
                                                       long long *__restrict__ em) {
  constexpr uint32_t z = 2 * ND;
  __shared__ uint32_t tie_rows[kMaxRows * kBlock];
  const uint32_t rep = blockIdx.y;
  // XCD-aware order: the hardware deals consecutive blocks round-robin over
  // 8 XCDs; remap so each XCD walks one contiguous band of rows and the
  // halo rows it shares stay in that XCD's L2.
  uint32_t b = blockIdx.x;
  const uint32_t nb = gridDim.x;
  if ((nb & 7u) == 0) b = (b & 7u) * (nb >> 3) + (b >> 3);
  const ReplicaParams &P = reps[rep];

  // byte LUTs indexed by the anti-aligned count (<= 6 < 8): top and second
  // byte of K, and 1 where K = 2^32
  uint32_t t_lo = 0, t_hi = 0, m_lo = 0, m_hi = 0, al_lo = 0, al_hi = 0;
#pragma unroll
  for (uint32_t a = 0; a <= z; ++a) {
    const uint32_t k8 = P.thr[a].k8;
    const uint32_t tb = k8 > 0xFFu ? 0xFFu : k8, ab = k8 > 0xFFu ? 1u : 0u;
    const uint32_t mb = P.thr[a].m8;
    if (a < 4) {
      t_lo |= tb << (8 * a);
      m_lo |= mb << (8 * a);
      al_lo |= ab << (8 * a);
    } else {
      t_hi |= tb << (8 * (a - 4));
      m_hi |= mb << (8 * (a - 4));
      al_hi |= ab << (8 * (a - 4));
    }
  }

  // 1 in both 16-bit lanes, opaque to the optimiser, which would otherwise
  // turn min(x, 1) into per-lane compares and selects
  uint32_t lanes1 = 0x00010001u;
  asm volatile("" : "+v"(lanes1));

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
    const uint32_t left_i = chunk == 0 ? g.lx2 - 4 : i0 - 4;  // dword holding byte i0-1
    const uint32_t right_i = i0 + 16 == g.lx2 ? 0 : i0 + 16;   // dword holding byte i0+16
    uint32_t tied_rows = 0;

    // Row r+1's loads are issued before row r is computed (the rows are
    // distinct and the other colour is read-only during a half-sweep), so
    // memory latency overlaps the draws and decisions.
    auto edge_of = [&](uint32_t y) {
      const uint32_t pe = (colour + y + zz) & 1u;
      return *reinterpret_cast<const uint32_t *>(oth + (plane + y) * g.lx2 + (pe ? right_i : left_i));
    };
    const uint64_t zm_plane = (uint64_t)g.ly * wrap_dec(zz, g.lz), zp_plane = (uint64_t)g.ly * wrap_inc(zz, g.lz);
    Chunk up = load16(oth + (plane + wrap_dec(y0, g.ly)) * g.lx2 + i0);
    Chunk mid = load16(oth + (plane + y0) * g.lx2 + i0);
    Chunk down = load16(oth + (plane + wrap_inc(y0, g.ly)) * g.lx2 + i0);
    Chunk o = load16(own + (plane + y0) * g.lx2 + i0);
    uint32_t edge = edge_of(y0);
    Chunk zm, zp;
    if (ND == 3) {
      zm = load16(oth + (zm_plane + y0) * g.lx2 + i0);
      zp = load16(oth + (zp_plane + y0) * g.lx2 + i0);
    }
    for (uint32_t r = 0; r < R; ++r) {
      const uint32_t y = y0 + r;
      const uint64_t row = plane + y;
      Chunk down_n, o_n, zm_n, zp_n;
      uint32_t edge_n = 0;
      if (r + 1 < R) {
        down_n = load16(oth + (plane + wrap_inc(y + 1, g.ly)) * g.lx2 + i0);
        o_n = load16(own + (row + 1) * g.lx2 + i0);
        edge_n = edge_of(y + 1);
        if (ND == 3) {
          zm_n = load16(oth + (zm_plane + y + 1) * g.lx2 + i0);
          zp_n = load16(oth + (zp_plane + y + 1) * g.lx2 + i0);
        }
      }
      const uint32_t p = (colour + y + zz) & 1u;
      // horizontal partner: bytes i0-1..i0+14 (p = 0) or i0+1..i0+16 (p = 1)
      Chunk sh;
      if (p == 0) {
        sh.w[0] = __builtin_amdgcn_alignbyte(mid.w[0], edge, 3);
        sh.w[1] = __builtin_amdgcn_alignbyte(mid.w[1], mid.w[0], 3);
        sh.w[2] = __builtin_amdgcn_alignbyte(mid.w[2], mid.w[1], 3);
        sh.w[3] = __builtin_amdgcn_alignbyte(mid.w[3], mid.w[2], 3);
      } else {
        sh.w[0] = __builtin_amdgcn_alignbyte(mid.w[1], mid.w[0], 1);
        sh.w[1] = __builtin_amdgcn_alignbyte(mid.w[2], mid.w[1], 1);
        sh.w[2] = __builtin_amdgcn_alignbyte(mid.w[3], mid.w[2], 1);
        sh.w[3] = __builtin_amdgcn_alignbyte(edge, mid.w[3], 1);
      }
      const uint64_t j0 = row * g.lx2 + i0;
      const U4 T = byte_block(key, j0 >> 4, sweep, colour, kStreamAccept, 0);
      const U4 M = byte_block(key, j0 >> 4, sweep, colour, kStreamAccept, 1);
      uint32_t tpos = 0;  // bit 8b+k: site 4k+b is tied
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        uint32_t a2 = differ2(o.w[k], up.w[k]) + differ2(o.w[k], down.w[k]) +
                      differ2(o.w[k], mid.w[k]) + differ2(o.w[k], sh.w[k]);
        if (ND == 3) a2 += differ2(o.w[k], zm.w[k]) + differ2(o.w[k], zp.w[k]);
        const uint32_t a = a2 >> 1;  // anti-aligned count per byte, 0..z
        const uint32_t t8 = __builtin_amdgcn_perm(t_hi, t_lo, a);
        const uint32_t m8 = __builtin_amdgcn_perm(m_hi, m_lo, a);
        const uint32_t always = __builtin_amdgcn_perm(al_hi, al_lo, a);  // 1 where K = 2^32
        const uint32_t ht = word_of(T, k), hm = word_of(M, k);
        // word < K from its top two bytes: (top, mid) < (t8, m8) as 16-bit
        // lanes, two sites per packed compare; u < k is max(u, k) != u
        const uint32_t u01 = __builtin_amdgcn_perm(ht, hm, kPair01), u23 = __builtin_amdgcn_perm(ht, hm, kPair23);
        const uint32_t k01 = __builtin_amdgcn_perm(t8, m8, kPair01), k23 = __builtin_amdgcn_perm(t8, m8, kPair23);
        const uint32_t lt01 = pk_min_u16(pk_max_u16(u01, k01) ^ u01, lanes1);
        const uint32_t lt23 = pk_min_u16(pk_max_u16(u23, k23) ^ u23, lanes1);
        const uint32_t m = __builtin_amdgcn_perm(lt23, lt01, kGather) | always;  // 1 per accepted site
        // (top, mid) == (t8, m8): a tie, settled from the low bytes unless K = 2^32
        const uint32_t ne01 = pk_min_u16(u01 ^ k01, lanes1), ne23 = pk_min_u16(u23 ^ k23, lanes1);
        const uint32_t tie = ~__builtin_amdgcn_perm(ne23, ne01, kGather) & ~always & 0x01010101u;
        tpos |= tie << k;
        // bytes 0x00 / 0xFF from selectors 0 / 1 into 0x0000FF00
        const uint32_t flip = __builtin_amdgcn_perm(0u, 0x0000FF00u, m) & 0xFEFEFEFEu;
        if (MEAS) {
          // bond sum over the chunk's sites with their new spins:
          // sum (z - 2 * anti_new), anti_new = z - anti where flipped
          const uint32_t an = a + m * z - 2u * (a & (m * 0xFFu));
          acc0 += 4ll * z - 2ll * __builtin_amdgcn_sad_u8(an, 0u, 0u);
          acc1 += 8ll - 2ll * (__builtin_popcount((o.w[k] ^ flip) & kMsb) + __builtin_popcount(mid.w[k] & kMsb));
        }
        o.w[k] ^= flip;
      }
      store16(own + row * g.lx2 + i0, o);
      tie_rows[r * kBlock + threadIdx.x] = tpos;
      tied_rows |= (tpos != 0 ? 1u : 0u) << r;
      up = mid;
      mid = down;
      down = down_n;
      o = o_n;
      edge = edge_n;
      if (ND == 3) {
        zm = zm_n;
        zp = zp_n;
      }
    }
    // settle the tied sites (tied sites were left unflipped above)
    while (tied_rows) {
      const uint32_t r = __builtin_ctz(tied_rows);
      tied_rows &= tied_rows - 1;
      uint32_t tp = tie_rows[r * kBlock + threadIdx.x];
      while (tp) {
        const uint32_t bit = __builtin_ctz(tp);
        tp &= tp - 1;
        const uint32_t s16 = 4 * (bit & 7u) + (bit >> 3);
        settle_tie<ND, MEAS>(own, oth, g, y0 + r, zz, i0 + s16, colour, sweep, P, acc0, acc1);
      }
    }
  }
  if (MEAS) {
    __shared__ long long red[2];
    if (threadIdx.x < 2) red[threadIdx.x] = 0;
    __syncthreads();
    long long v0 = acc0, v1 = acc1;
    for (int off = 32; off > 0; off >>= 1) {
      v0 += __shfl_xor(v0, off);
      v1 += __shfl_xor(v1, off);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd((unsigned long long *)&red[0], (unsigned long long)v0);
      atomicAdd((unsigned long long *)&red[1], (unsigned long long)v1);
    }
    __syncthreads();
    if (threadIdx.x < 2) atomicAdd((unsigned long long *)&em[(uint64_t)rep * kEmStride + threadIdx.x],
                                   (unsigned long long)red[threadIdx.x]);
  }
}

uint32_t fast_rows(const StencilGeom &g) {
  for (uint32_t r : {16u, 8u, 4u, 2u, 1u})
    if (r <= kMaxRows && g.ly % r == 0) return r;
  return 1;
}

}  // namespace

bool ising_fast_path_ok(const StencilGeom &g) {
  return (g.nd == 2 || g.nd == 3) && g.lx2 % 16 == 0 && g.ly >= 2;
}

hipError_t launch_ising_fast(const StencilGeom &g, int8_t *spins, const ReplicaParams *reps,
                             uint32_t n_rep, uint32_t colour, uint64_t sweep, long long *em,
                             hipStream_t st) {
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

}  // namespace scamper
