// Register-blocked checkerboard Metropolis for the Ising model on 2-D and
// 3-D periodic lattices (int8 spins, colour-split rows).
//
// One thread owns 16 consecutive bytes (16 sites) of a row and walks R rows
// of one plane, keeping the other colour's rows y-1, y, y+1 in registers and
// issuing row r+1's loads before row r is computed.  Per row it needs one
// Philox block: its 16 bytes are the top bytes of the 16 sites' draws, laid
// out exactly like the spins, so every decision is made four sites at a time
// with byte-wise integer arithmetic (SWAR):
//
//   anti  = number of anti-aligned neighbours per byte (spins are 0x01/0xFF,
//           so (s ^ n) has bit 1 set exactly when the two differ)
//   t8    = top byte of the threshold K[anti]            (v_perm lookup)
//   flip  = top < t8, or K[anti] = 2^32 (every word accepts)
//   tie   = top == t8 and K[anti] < 2^32  (one uphill site in 256)
//
// A tie is decided by the lower 24 bits of its draw against those of K: the
// mid byte (a second block, shared by the row's 16 sites) and, when that
// ties too, the low half (a third block).  The row loop parks a tied row in
// a small per-thread LDS list together with the anti-aligned counts and spins
// of its 16 sites, and the ties are settled after the R rows, so the second
// block is drawn for ~2% of rows instead of all of them.  A thread whose list
// is full settles the row at once.
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
constexpr uint32_t kSlots = 4;
constexpr uint32_t kPrefetch = 1;  // rows whose loads are in flight ahead of the row computed
// DPP controls: every lane takes the value of lane - 1 (shr) or lane + 1 (shl)
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;  // tied rows a thread can park before settling at once

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

// Bit 7 of each byte set where byte x < byte y (unsigned).  d's bytes are
// (x | 0x80) - (y & 0x7f) >= 1, so no borrow crosses a byte, and bit 7 of d
// is set exactly when x mod 128 >= y mod 128; x < y when the top bits say so,
// or they agree and the low seven bits are smaller.
__device__ __forceinline__ uint32_t bytes_lt(uint32_t x, uint32_t y) {
  const uint32_t d = (x | kMsb) - (y & ~kMsb);
  return ((~x & y) | (~(x ^ y) & ~d)) & kMsb;
}
// Bit 7 of each byte set where byte x == byte y: (e & 0x7f) + 0x7f reaches
// bit 7 exactly when the low seven bits of e = x ^ y are not all zero.
__device__ __forceinline__ uint32_t bytes_eq(uint32_t x, uint32_t y) {
  const uint32_t e = x ^ y;
  return ~(((e & ~kMsb) + ~kMsb) | e) & kMsb;
}

// One row's loads beyond the rows above it.
struct RowIn {
  Chunk down, o, zm, zp;
  uint32_t edge;
};

// A parked row: which of its sites tie on the top byte, and per site the
// anti-aligned count (3 bits) and whether the spin is -1 (1 bit).
struct TiedRow {
  uint32_t meta;  // bit 8b + 7 - k: site 4k + b ties; bits 0..3: row r
  uint32_t a01;   // byte b: site b in bits 0..3, site 4 + b in bits 4..7
  uint32_t a23;   // byte b: site 8 + b in bits 0..3, site 12 + b in bits 4..7
};

// Settle the ties of one row.  Their spins still hold the pre-sweep values;
// a tied site flips when the lower 24 bits of its draw, mid byte then low
// half, are below those of K.
template <int ND, bool MEAS>
__device__ __forceinline__ void settle_row(int8_t *__restrict__ own, uint64_t j0, const TiedRow &t, uint64_t sweep,
                                           uint32_t colour, const ReplicaParams &P, long long &acc0,
                                           long long &acc1) {
  constexpr int z = 2 * ND;
  const U4 M = byte_block(P.key, j0 >> 4, sweep, colour, kStreamAccept, 1);
  uint32_t tp = t.meta & 0xF0F0F0F0u;
  while (tp) {
    const uint32_t bit = __builtin_ctz(tp);
    tp &= tp - 1;
    const uint32_t b = bit >> 3, k = 7u - (bit & 7u);
    const uint32_t s16 = 4 * k + b;
    const uint32_t nib = ((k < 2 ? t.a01 : t.a23) >> (8 * b + 4 * (k & 1u))) & 0xFu;
    const uint32_t anti = nib & 7u;
    const Threshold th = P.thr[anti];
    const uint32_t mid = (word_of(M, k) >> (8 * b)) & 0xFFu;
    bool flip = mid < th.m8;
    if (mid == th.m8) flip = site_low(P.key, j0 + s16, sweep, colour, kStreamAccept) < th.k16;
    if (flip) {
      const bool neg = (nib >> 3) != 0;
      own[j0 + s16] = neg ? 1 : -1;
      if (MEAS) {
        acc0 += 4ll * anti - 2ll * z;  // bond sum of the site changes sign
        acc1 += neg ? 2 : -2;
      }
    }
  }
}

// R rows per thread, a compile-time constant so the row loop unrolls fully
// and the prefetched rows stay in their own registers (a rolled loop copies
// them into the current row's registers, which waits for the loads).
template <int ND, bool MEAS, uint32_t R, typename Off, uint32_t PF, bool NBR>
__global__ __launch_bounds__(kBlock) void k_ising_fast(int8_t *__restrict__ spins,
                                                       const ReplicaParams *__restrict__ reps,
                                                       StencilGeom g, uint32_t colour,
                                                       uint64_t sweep, long long *__restrict__ em,
                                                       const StepClock *__restrict__ clk) {
  constexpr uint32_t z = 2 * ND;
  if (clk) sweep = clk->sweep;
  __shared__ TiedRow parked[kSlots][kBlock];
  const uint32_t rep = blockIdx.y;
  // XCD-aware order: the hardware deals consecutive blocks round-robin over
  // 8 XCDs; remap so each XCD walks one contiguous band of rows and the
  // halo rows it shares stay in that XCD's L2.
  uint32_t b = blockIdx.x;
  const uint32_t nb = gridDim.x;
  if ((nb & 7u) == 0) b = (b & 7u) * (nb >> 3) + (b >> 3);
  const ReplicaParams &P = reps[rep];

  // byte LUTs indexed by the anti-aligned count (<= 6 < 8): the top byte of
  // K, and 0x80 where K = 2^32
  uint32_t t_lo = 0, t_hi = 0, al_lo = 0, al_hi = 0;
#pragma unroll
  for (uint32_t a = 0; a <= z; ++a) {
    const uint32_t k8 = P.thr[a].k8;
    const uint32_t tb = k8 > 0xFFu ? 0xFFu : k8, ab = k8 > 0xFFu ? 0x80u : 0u;
    if (a < 4) {
      t_lo |= tb << (8 * a);
      al_lo |= ab << (8 * a);
    } else {
      t_hi |= tb << (8 * (a - 4));
      al_hi |= ab << (8 * (a - 4));
    }
  }

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
    int8_t *__restrict__ own = spins + (2ull * rep + colour) * g.cs;
    const int8_t *__restrict__ oth = spins + (2ull * rep + 1 - colour) * g.cs;
    // offsets inside one colour array: 32-bit when it has < 2^32 sites, so
    // every load takes one offset register beside the array's scalar base
    const Off lx2 = g.lx2;
    const Off plane = (Off)g.ly * zz;
    const uint64_t key = P.key;
    const uint32_t left_i = chunk == 0 ? g.lx2 - 4 : i0 - 4;  // dword holding byte i0-1
    const uint32_t right_i = i0 + 16 == g.lx2 ? 0 : i0 + 16;   // dword holding byte i0+16
    uint32_t n_parked = 0;

    // Row r+PF's loads are issued before row r is computed (the rows are
    // distinct and the other colour is read-only during a half-sweep), so
    // memory latency overlaps the draws and decisions.
    // The horizontal partner of the chunk's first (p = 0) or last (p = 1)
    // site is a byte of the next chunk to the left or right.  With NBR a
    // wave's lanes hold 64 consecutive chunks of one row, so that byte is in
    // the neighbouring lane's registers (a DPP wave shift); only the lane at
    // the wave's edge loads it.
    const uint32_t lane = threadIdx.x & 63u;
    auto edge_of = [&](uint32_t y) {
      const uint32_t pe = (colour + y + zz) & 1u;
      const uint32_t *src = reinterpret_cast<const uint32_t *>(oth + ((plane + y) * lx2 + (pe ? right_i : left_i)));
      if (!NBR) return *src;
      uint32_t v = 0;
      if (lane == (pe ? 63u : 0u)) v = *src;
      return v;
    };
    const Off zm_plane = (Off)g.ly * wrap_dec(zz, g.lz), zp_plane = (Off)g.ly * wrap_inc(zz, g.lz);
    // what row y0 + rr needs beyond the rows above it: the other colour's row
    // below, its own spins, the horizontal edge and (3-D) the planes beside
    auto fetch = [&](uint32_t rr) {
      RowIn v;
      const uint32_t yy = y0 + rr;
      v.down = load16(oth + ((plane + wrap_inc(yy, g.ly)) * lx2 + i0));
      v.o = load16(own + ((plane + yy) * lx2 + i0));
      v.edge = edge_of(yy);
      if (ND == 3) {
        v.zm = load16(oth + ((zm_plane + yy) * lx2 + i0));
        v.zp = load16(oth + ((zp_plane + yy) * lx2 + i0));
      }
      return v;
    };
    Chunk up = load16(oth + ((plane + wrap_dec(y0, g.ly)) * lx2 + i0));
    Chunk mid = load16(oth + ((plane + y0) * lx2 + i0));
    RowIn cur = fetch(0), n1, n2;
    if (PF == 2 && R > 1) n1 = fetch(1);
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
      const uint32_t y = y0 + r;
      const Off row = plane + y;
      if (PF == 1 && r + 1 < R) n1 = fetch(r + 1);
      if (PF == 2 && r + 2 < R) n2 = fetch(r + 2);
      const Chunk &down = cur.down, &zm = cur.zm, &zp = cur.zp;
      const uint32_t p = (colour + y + zz) & 1u;
      uint32_t edge = cur.edge;
      if (NBR) {
        // lane - 1's last word (p = 0) or lane + 1's first word (p = 1)
        const uint32_t from_nbr = p ? (uint32_t)__builtin_amdgcn_mov_dpp((int)mid.w[0], kDppWaveShl1, 0xF, 0xF, false)
                                    : (uint32_t)__builtin_amdgcn_mov_dpp((int)mid.w[3], kDppWaveShr1, 0xF, 0xF, false);
        if (lane != (p ? 63u : 0u)) edge = from_nbr;
      }
      Chunk &o = cur.o;
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
      const Off j0 = row * lx2 + i0;
      const U4 T = byte_block(key, j0 >> 4, sweep, colour, kStreamAccept, 0);
      uint32_t tpos = 0;        // bit 8b + 7 - k: site 4k + b ties
      uint32_t anti[2] = {0, 0};  // packed counts, as TiedRow::a01 / a23
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        uint32_t a2 = differ2(o.w[k], up.w[k]) + differ2(o.w[k], down.w[k]) +
                      differ2(o.w[k], mid.w[k]) + differ2(o.w[k], sh.w[k]);
        if (ND == 3) a2 += differ2(o.w[k], zm.w[k]) + differ2(o.w[k], zp.w[k]);
        const uint32_t a = a2 >> 1;  // anti-aligned count per byte, 0..z
        anti[k >> 1] |= a << (4 * (k & 1));
        const uint32_t t8 = __builtin_amdgcn_perm(t_hi, t_lo, a);
        const uint32_t always = __builtin_amdgcn_perm(al_hi, al_lo, a);  // 0x80 where K = 2^32
        const uint32_t top = word_of(T, k);
        const uint32_t acc = bytes_lt(top, t8) | always;  // 0x80 per accepted site
        tpos |= (bytes_eq(top, t8) & ~always) >> k;
        const uint32_t m = acc >> 7;  // 1 per accepted site
        // bytes 0x00 / 0xFE from selectors 0 / 1 into 0x0000FE00
        const uint32_t flip = __builtin_amdgcn_perm(0u, 0x0000FE00u, m);
        if (MEAS) {
          // bond sum over the chunk's sites with their new spins:
          // sum (z - 2 * anti_new), anti_new = z - anti where flipped
          const uint32_t an = a + m * z - 2u * (a & (m * 0xFFu));
          acc0 += 4ll * z - 2ll * __builtin_amdgcn_sad_u8(an, 0u, 0u);
          acc1 += 8ll - 2ll * (__builtin_popcount((o.w[k] ^ flip) & kMsb) + __builtin_popcount(mid.w[k] & kMsb));
        }
        o.w[k] ^= flip;
      }
      store16(own + (row * lx2 + i0), o);
      if (tpos) {
        // tied sites kept their spins: record the spins' signs beside the counts
        TiedRow t;
        t.meta = tpos | r;
        t.a01 = anti[0] | ((o.w[0] & kMsb) >> 4) | (o.w[1] & kMsb);
        t.a23 = anti[1] | ((o.w[2] & kMsb) >> 4) | (o.w[3] & kMsb);
        if (n_parked < kSlots)
          parked[n_parked++][threadIdx.x] = t;
        else
          settle_row<ND, MEAS>(own, j0, t, sweep, colour, P, acc0, acc1);
      }
      up = mid;
      mid = down;
      cur = n1;
      if (PF == 2) n1 = n2;
    }
    for (uint32_t s = 0; s < n_parked; ++s) {
      const TiedRow t = parked[s][threadIdx.x];
      const Off j0 = (plane + y0 + (t.meta & 0xFu)) * lx2 + i0;
      settle_row<ND, MEAS>(own, j0, t, sweep, colour, P, acc0, acc1);
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
  // ly is even (the checkerboard needs even sides), so 2 always divides it
  for (uint32_t r : {16u, 8u, 4u})
    if (r <= kMaxRows && g.ly % r == 0) return r;
  return 2;
}

template <int ND, bool MEAS>
void launch_rows(uint32_t R, dim3 grid, hipStream_t st, int8_t *spins, const ReplicaParams *reps,
                 const StencilGeom &g, uint32_t colour, uint64_t sweep, long long *em, const StepClock *clk) {
  // lanes hold consecutive chunks of one row when a row is whole waves wide
  const bool nbr = (g.lx2 / 16) % 64 == 0;
#define SC_ROWS(RR)                                                                                        \
  if (g.nh >= (1ull << 32))                                                                                \
    hipLaunchKernelGGL((k_ising_fast<ND, MEAS, RR, uint64_t, kPrefetch, false>), grid, dim3(kBlock), 0, st, \
                       spins, reps, g, colour, sweep, em, clk);                                            \
  else if (nbr)                                                                                            \
    hipLaunchKernelGGL((k_ising_fast<ND, MEAS, RR, uint32_t, kPrefetch, true>), grid, dim3(kBlock), 0, st,  \
                       spins, reps, g, colour, sweep, em, clk);                                            \
  else                                                                                                     \
    hipLaunchKernelGGL((k_ising_fast<ND, MEAS, RR, uint32_t, kPrefetch, false>), grid, dim3(kBlock), 0, st, \
                       spins, reps, g, colour, sweep, em, clk)
  switch (R) {
    case 16: SC_ROWS(16); break;
    case 8: SC_ROWS(8); break;
    case 4: SC_ROWS(4); break;
    default: SC_ROWS(2); break;
  }
#undef SC_ROWS
}

}  // namespace

bool ising_fast_path_ok(const StencilGeom &g) {
  return (g.nd == 2 || g.nd == 3) && g.lx2 % 16 == 0 && g.ly >= 2;
}

hipError_t launch_ising_fast(const StencilGeom &g, int8_t *spins, const ReplicaParams *reps,
                             uint32_t n_rep, uint32_t colour, uint64_t sweep, long long *em,
                             hipStream_t st, const StepClock *clk) {
  if (g.cs < g.nh) return hipErrorInvalidValue;  // the colour stride is unset
  const uint32_t R = fast_rows(g);
  const uint64_t threads = (uint64_t)(g.lx2 / 16) * (g.ly / R) * g.lz;
  dim3 grid((unsigned)((threads + kBlock - 1) / kBlock), n_rep);
  if (g.nd == 2) {
    if (em) launch_rows<2, true>(R, grid, st, spins, reps, g, colour, sweep, em, clk);
    else launch_rows<2, false>(R, grid, st, spins, reps, g, colour, sweep, em, clk);
  } else {
    if (em) launch_rows<3, true>(R, grid, st, spins, reps, g, colour, sweep, em, clk);
    else launch_rows<3, false>(R, grid, st, spins, reps, g, colour, sweep, em, clk);
  }
  return hipGetLastError();
}

}  // namespace scamper
