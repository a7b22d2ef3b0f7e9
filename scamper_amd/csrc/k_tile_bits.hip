// Checkerboard Metropolis for Ising and Potts (q <= 4) spins held as bit
// planes, over the same table of independent periodic lattices (jobs) as the
// int8 tiled kernel, with the same draws and the same decisions.
//
// Per-site rule: MetropolisCore::sweep_with_phase
// (CMC.rs/src/algorithms/metropolis.rs:67-106) with the proposal of
// IsingModel::propose_from (-s, lattice/models.rs:98-100) or
// PottsModel::random_other_state (models.rs:212-218), ΔE from the
// nearest-neighbour bonds (interaction.rs:73-84, 149-181) and
// accept_log_probability (Carlo.rs/src/acceptance.rs:11-22) folded into the
// host's integer thresholds.  Visit order is colour 0 then colour 1.
//
// Layout: colour class c of a job is ceil(nh / 32) words per plane, bit k of
// word w holding rank 32w + k; a Potts state v is planes (v & 1, v >> 1), an
// Ising spin one plane (1 for -1).  Each plane has one spare word before and
// after it, so that 32 bits starting anywhere in [-32, 32 words) are two
// aligned loads and a funnel shift.  A job is admitted with int8 arrays;
// they are packed into its planes at its first chunk, the planes are its
// state until its last chunk unpacks them.
//
// A lane updates one word (32 ranks, two Philox groups).  Its neighbours in
// the other colour are that colour's bits at rank offsets 0, -1 / +1 (by row
// parity), -+lx2 and -+lx2 ly, plus the periodic wraps where a row, plane or
// the lattice ends, chosen bit by bit with masks built from the rows the word
// covers.  Equality with the own and the proposed state is two bit operations
// per neighbour, and the counts are bit-sliced adders.  The acceptance index
// t = #(n == r) + 7 - #(n == s) (Ising: 7 - #(n == s)) is expanded to one byte
// per site, four sites at a time, for the byte look-up of the threshold's top
// byte and a byte-wise compare with the draws' top bytes, as in the int8
// kernel.  Sites whose top byte equals the threshold's (and, for q = 4,
// whose proposal byte straddles a bucket boundary) are settled one by one
// with the exact rule.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kTileWords = 256;  // words per tile: one per lane of a block
constexpr uint32_t kMsb = 0x80808080u;

// Plane words through global-memory pointers (a generic pointer read from
// the table would make every access a flat one).
typedef __attribute__((address_space(1))) uint32_t g32;
typedef __attribute__((address_space(1))) const char gbyte;

// Word i of a plane, addressed by a 32-bit byte offset from its uniform base
// (planes are far below 4 GiB), so that the load takes a scalar base and one
// offset register.
__device__ __forceinline__ uint32_t ld(const g32 *p, uint32_t i) {
  return *reinterpret_cast<const g32 *>(reinterpret_cast<gbyte *>(p) + (i << 2));
}

__device__ __forceinline__ uint32_t div_by(uint32_t n, uint64_t magic) {
  return magic ? (uint32_t)__umul64hi((uint64_t)n, magic) : n;
}

__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c) { return (a & b) | (c & (a | b)); }

// Byte-wise unsigned x < y and x == y, reported in bit 7 of each byte.
__device__ __forceinline__ uint32_t lt_bytes(uint32_t x, uint32_t y) {
  const uint32_t same = ~(x ^ y);
  const uint32_t d = ((x | kMsb) - (y & ~kMsb)) ^ (same & kMsb);
  return ((~x & y) | (same & d)) & kMsb;
}
__device__ __forceinline__ uint32_t eq_bytes(uint32_t x, uint32_t y) {
  const uint32_t t = x ^ y;
  return ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & kMsb;
}
__device__ __forceinline__ uint32_t spread_msb(uint32_t x) { return __builtin_amdgcn_perm(x << 8, x, 0x090B080Au); }

// Bit 0 of each byte of x (other bits zero) gathered into bits 0..3.
__device__ __forceinline__ uint32_t gather4(uint32_t x) { return (x * 0x01020408u) >> 24; }
// Bits 0..3 of n spread to bit p of bytes 0..3.
template <int P>
__device__ __forceinline__ uint32_t spread4(uint32_t n) {
  return (n * (0x00204081u << P)) & (0x01010101u << P);
}

__device__ __forceinline__ uint32_t bit_of(uint32_t w, uint32_t k) { return (w >> k) & 1u; }

// x where c is 1, y where it is 0 (bit masks, so that a runtime choice
// among registers stays in registers).
__device__ __forceinline__ uint32_t msel(uint32_t c, uint32_t x, uint32_t y) {
  const uint32_t m = 0u - c;
  return (x & m) | (y & ~m);
}

// Byte k (0..31) of the 32 bytes of blocks a and b.
__device__ __forceinline__ uint32_t byte32(const U4 &a, const U4 &b, uint32_t k) {
  const uint32_t odd = (k >> 2) & 1u, high = (k >> 3) & 1u;
  const uint32_t wa = msel(high, msel(odd, a.w, a.z), msel(odd, a.y, a.x));
  const uint32_t wb = msel(high, msel(odd, b.w, b.z), msel(odd, b.y, b.x));
  return (msel(k >> 4, wb, wa) >> (8u * (k & 3u))) & 0xFFu;
}

// Exact proposal floor(word * m / 2^32) from its top byte, reading the finer
// draws only when the byte's bucket straddles a boundary.
__device__ uint32_t propose_exact(uint32_t top, uint32_t m, uint64_t key, uint32_t j, uint64_t sweep,
                                  uint32_t colour) {
  const uint64_t b1 = (uint64_t)top << 24;
  uint32_t a = (uint32_t)((b1 * m) >> 32);
  if (a == (uint32_t)(((b1 + 0xFFFFFFu) * m) >> 32)) return a;
  const uint64_t b2 = b1 | ((uint64_t)site_byte(key, j, sweep, colour, kStreamPropose, 1) << 16);
  a = (uint32_t)((b2 * m) >> 32);
  if (a == (uint32_t)(((b2 + 0xFFFFu) * m) >> 32)) return a;
  return (uint32_t)(((b2 | site_low(key, j, sweep, colour, kStreamPropose)) * m) >> 32);
}

// Byte t (0..15) of a 16-entry byte table held in four uniform words.
__device__ __forceinline__ uint32_t lut_byte(const uint32_t (&L)[4], uint32_t t) {
  const uint32_t lo = __builtin_amdgcn_perm(L[1], L[0], t & 7u), hi = __builtin_amdgcn_perm(L[3], L[2], t & 7u);
  return ((t & 8u) ? hi : lo) & 0xFFu;
}

// Exact "word < K" for look-up index t (K = 2^32 where bit t of tal is set),
// from the top byte, reading the finer draws on a tie.
__device__ bool decide_exact(uint32_t top, const TileJob &J, uint32_t t, uint64_t key, uint32_t j, uint64_t sweep,
                             uint32_t colour) {
  if ((J.tal >> t) & 1u) return true;
  const uint32_t k8 = lut_byte(J.lut_t8, t);
  if (top != k8) return top < k8;
  const uint32_t mid = site_byte(key, j, sweep, colour, kStreamAccept, 1), m8 = lut_byte(J.lut_tm8, t);
  if (mid != m8) return mid < m8;
  return site_low(key, j, sweep, colour, kStreamAccept) < (lut_byte(J.lut_tk16lo, t) | lut_byte(J.lut_tk16hi, t) << 8);
}

// Bit-sliced count of Z one-bit inputs: c0 + 2 c1 + 4 c2.
template <int Z>
__device__ __forceinline__ void count_bits(const uint32_t (&e)[Z], uint32_t &c0, uint32_t &c1, uint32_t &c2) {
  if (Z == 2) {
    c0 = e[0] ^ e[1];
    c1 = e[0] & e[1];
    c2 = 0;
  } else if (Z == 4) {
    const uint32_t s1 = e[0] ^ e[1], k1 = e[0] & e[1], s2 = e[2] ^ e[3], k2 = e[2] & e[3];
    c0 = s1 ^ s2;
    const uint32_t k3 = s1 & s2;
    c1 = k1 ^ k2 ^ k3;
    c2 = maj3(k1, k2, k3);
  } else {
    const uint32_t s1 = e[0] ^ e[1] ^ e[2], k1 = maj3(e[0], e[1], e[2]);
    const uint32_t s2 = e[3] ^ e[4] ^ e[5], k2 = maj3(e[3], e[4], e[5]);
    c0 = s1 ^ s2;
    const uint32_t k3 = s1 & s2;
    c1 = k1 ^ k2 ^ k3;
    c2 = maj3(k1, k2, k3);
  }
}

__device__ __forceinline__ int weighted_popc(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t vm) {
  return __builtin_popcount(c0 & vm) + 2 * __builtin_popcount(c1 & vm) + 4 * __builtin_popcount(c2 & vm);
}

// Per-lane measured counters of one job (summed over the lane's words).
struct BMeas {
  int e = 0, m = 0;
  int cnt[4] = {0, 0, 0, 0};
};

// Sum of v over the 64 lanes of a wave (every lane active).
__device__ __forceinline__ int wave_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm [1, 0, 3, 2]
  v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm [2, 3, 0, 1]
  v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
  v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);  // row_mirror
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1 and 3
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2 and 3
  return __builtin_amdgcn_readlane(v, 63);
}

// Block sum of the lanes' counters, added to the record at `rec` (nc
// values).  Flush k writes partial buffer k & 1: a wave can rewrite a buffer
// only after passing the next flush's barrier, which every wave reaches
// after reading this one.
struct BLds {
  int part[2][kThreads / 64][6];
};
__device__ __forceinline__ void flush_block(BLds &L, BMeas &M, int nc, bool ising, long long *rec, uint32_t &flushes) {
  const uint32_t wave = threadIdx.x >> 6, buf = flushes & 1u;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    if (c >= nc) break;
    const int v = c == 0 ? M.e : c == 1 ? (ising ? M.m : 0) : M.cnt[c - 2];
    const int sum = wave_sum(v);
    if ((threadIdx.x & 63u) == 0) L.part[buf][wave][c] = sum;
  }
  __syncthreads();
  if ((int)threadIdx.x < nc) {
    long long t = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) t += L.part[buf][w][threadIdx.x];
    if (t != 0) atomicAdd(reinterpret_cast<unsigned long long *>(rec) + threadIdx.x, (unsigned long long)t);
  }
  M = BMeas{};
  ++flushes;
}

// The words of one plane a lane needs: its own word and those beside it (the
// same rank, and the row neighbours at rank offsets -1 / +1), and word pairs
// for the 32 bits at each other neighbour offset, with the periodic-wrap
// pairs loaded only where the lane has sites that wrap.
struct Pair {
  uint32_t lo = 0, hi = 0;
};
__device__ __forceinline__ Pair pair_at(const g32 *__restrict__ p, int32_t o, int32_t hi) {
  // storage index of word o >> 5, never negative after the clamp: an
  // unsigned 32-bit offset from the plane's uniform base
  const uint32_t i = (uint32_t)((min(max(o, -32), hi) >> 5) + 1);
  return Pair{ld(p, i), ld(p, i + 1u)};
}
__device__ __forceinline__ uint32_t bits_of(const Pair &q, int32_t o) {
  return __builtin_amdgcn_alignbit(q.hi, q.lo, (uint32_t)o & 31u);
}

template <int ND>
struct Raw {
  uint32_t before, at, after;  // words w - 1, w, w + 1
  Pair wl, wr, up, dn, y0, yl, bk, fw, z0, zl;

  __device__ __forceinline__ void load(const g32 *__restrict__ p, uint32_t w, int32_t o, int32_t hi, int32_t x2,
                                       int32_t pl, int32_t all, uint32_t need_wl, uint32_t need_wr, uint32_t m_y0,
                                       uint32_t m_yl, uint32_t m_z0, uint32_t m_zl) {
    before = ld(p, w);
    at = ld(p, w + 1u);
    after = ld(p, w + 2u);
    if (need_wl) wl = pair_at(p, o - 1 + x2, hi);
    if (need_wr) wr = pair_at(p, o + 1 - x2, hi);
    if (ND >= 2) {
      up = pair_at(p, o - x2, hi);
      dn = pair_at(p, o + x2, hi);
      if (m_y0) y0 = pair_at(p, o - x2 + pl, hi);
      if (m_yl) yl = pair_at(p, o + x2 - pl, hi);
    }
    if (ND == 3) {
      bk = pair_at(p, o - pl, hi);
      fw = pair_at(p, o + pl, hi);
      if (m_z0) z0 = pair_at(p, o - pl + all, hi);
      if (m_zl) zl = pair_at(p, o + pl - all, hi);
    }
  }

  // neighbour words: same rank, row neighbour, then y -/+ and z -/+
  __device__ __forceinline__ void combine(uint32_t (&N)[2 * ND], int32_t o, int32_t x2, int32_t pl, int32_t all,
                                          uint32_t m_p1, uint32_t m_first, uint32_t m_last, uint32_t m_y0,
                                          uint32_t m_yl, uint32_t m_z0, uint32_t m_zl) const {
    N[0] = at;
    const uint32_t lm = __builtin_amdgcn_alignbit(at, before, 31u), rp = __builtin_amdgcn_alignbit(after, at, 1u);
    const uint32_t left = (lm & ~m_first) | (bits_of(wl, o - 1 + x2) & m_first);
    const uint32_t right = (rp & ~m_last) | (bits_of(wr, o + 1 - x2) & m_last);
    N[1] = (right & m_p1) | (left & ~m_p1);
    if (ND >= 2) {
      N[2] = (bits_of(up, o - x2) & ~m_y0) | (bits_of(y0, o - x2 + pl) & m_y0);
      N[3] = (bits_of(dn, o + x2) & ~m_yl) | (bits_of(yl, o + x2 - pl) & m_yl);
    }
    if (ND == 3) {
      N[4] = (bits_of(bk, o - pl) & ~m_z0) | (bits_of(z0, o - pl + all) & m_z0);
      N[5] = (bits_of(fw, o + pl) & ~m_zl) | (bits_of(zl, o + pl - all) & m_zl);
    }
  }
};

// One word (32 ranks) of colour `colour` of job J.  QK: 0 for one plane
// (Ising, or Potts q = 2), 3 or 4 for Potts q = 3 or 4.
template <int ND, bool POTTS, int QK>
__device__ __forceinline__ void run_word(const TileJob &J, uint32_t w, uint64_t sweep, uint32_t colour, bool meas,
                                         BMeas &M) {
  constexpr int Z = 2 * ND;
  constexpr bool TWO = QK != 0;  // two planes
  const uint32_t nh = J.nh, lx2 = J.lx2, ly = J.ly, lz = J.lz;
  const uint32_t n_words = (nh + 31u) >> 5;
  if (w >= n_words) return;
  const uint32_t j0 = w << 5;
  const uint32_t nv = min(nh - j0, 32u);
  const uint32_t vm = nv == 32u ? ~0u : (1u << nv) - 1u;
  const uint64_t key = J.key;
  const uint32_t nwp = J.nwp;
  g32 *const base = (g32 *)J.bits;
  g32 *__restrict__ own0 = base + (colour * 2u) * nwp;
  g32 *__restrict__ own1 = own0 + nwp;
  const g32 *__restrict__ oth0 = base + ((1u - colour) * 2u) * nwp;
  const g32 *__restrict__ oth1 = oth0 + nwp;
  const int32_t hi = (int32_t)(32u * n_words) - 1;

  // rows the word covers: masks of the sites with row parity 1, at a row's
  // first or last rank, and in the first or last row of a plane (y) or of
  // the lattice (z)
  uint32_t m_p1 = 0, m_first = 0, m_last = 0, m_y0 = 0, m_yl = 0, m_z0 = 0, m_zl = 0;
  {
    uint32_t row = div_by(j0, J.div_lx2);
    uint32_t i = j0 - row * lx2;
    uint32_t zz = ND == 3 ? div_by(row, J.div_ly) : 0u;
    uint32_t y = ND >= 2 ? row - zz * ly : 0u;
    uint32_t pos = 0;
    while (pos < nv) {
      const uint32_t len = min(lx2 - i, nv - pos);
      const uint32_t rng = (len == 32u ? ~0u : (1u << len) - 1u) << pos;
      if ((colour + y + zz) & 1u) m_p1 |= rng;
      if (i == 0) m_first |= 1u << pos;
      if (i + len == lx2) m_last |= 1u << (pos + len - 1u);
      if (ND >= 2) {
        if (y == 0) m_y0 |= rng;
        if (y + 1u == ly) m_yl |= rng;
      }
      if (ND == 3) {
        if (zz == 0) m_z0 |= rng;
        if (zz + 1u == lz) m_zl |= rng;
      }
      pos += len;
      i = 0;
      if (ND >= 2 && ++y == ly) {
        y = 0;
        ++zz;
      }
    }
  }

  // neighbour planes: the same rank, the row neighbour on the other side, and
  // the rows and planes beside
  // every load first (the wrap words only where a lane has such sites, but
  // none of them used before the Philox rounds), so that one memory round
  // trip covers them all
  const int32_t o = (int32_t)j0;
  const int32_t pl = (int32_t)(lx2 * ly), all = (int32_t)nh, x2 = (int32_t)lx2;
  Raw<ND> A0, A1;
  A0.load(oth0, w, o, hi, x2, pl, all, m_first & ~m_p1, m_last & m_p1, m_y0, m_yl, m_z0, m_zl);
  if (TWO) A1.load(oth1, w, o, hi, x2, pl, all, m_first & ~m_p1, m_last & m_p1, m_y0, m_yl, m_z0, m_zl);
  const uint32_t S0 = ld(own0, w + 1u);
  const uint32_t S1 = TWO ? ld(own1, w + 1u) : 0u;

  const uint32_t g0 = j0 >> 4;
  const U4 T0 = byte_block(key, g0, sweep, colour, kStreamAccept, 0);
  const U4 T1 = byte_block(key, g0 + 1u, sweep, colour, kStreamAccept, 0);
  U4 P0{0, 0, 0, 0}, P1{0, 0, 0, 0};
  if (POTTS && QK != 0) {
    P0 = byte_block(key, g0, sweep, colour, kStreamPropose, 0);
    P1 = byte_block(key, g0 + 1u, sweep, colour, kStreamPropose, 0);
  }
  uint32_t N0[Z], N1[Z];
  A0.combine(N0, o, x2, pl, all, m_p1, m_first, m_last, m_y0, m_yl, m_z0, m_zl);
  if (TWO) A1.combine(N1, o, x2, pl, all, m_p1, m_first, m_last, m_y0, m_yl, m_z0, m_zl);

  // proposed states (bit planes R0, R1); q = 4: sites whose proposal byte
  // straddles a bucket boundary are settled exactly below
  uint32_t R0, R1 = 0, strad = 0;
  if (!POTTS) {
    R0 = ~S0;
  } else if (QK == 0) {
    R0 = ~S0;
  } else {
    uint32_t a0 = 0, a1 = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8u; ++i) {
      const uint32_t pt = word_of(i < 4u ? P0 : P1, i & 3u);
      if (QK == 3) {
        a0 |= gather4((pt >> 7) & 0x01010101u) << (4u * i);  // floor(u 2): the top bit
      } else {
        // floor(u 3) from the top byte: 3 top in 16-bit lanes, bits 8..9
        const uint32_t pe = (pt & 0x00FF00FFu) * 3u, po = ((pt >> 8) & 0x00FF00FFu) * 3u;
        const uint32_t a = __builtin_amdgcn_perm(po, pe, 0x07030501u);   // 3 top >> 8, per byte
        const uint32_t lo = __builtin_amdgcn_perm(po, pe, 0x06020400u);  // 3 top & 0xff, per byte
        a0 |= gather4(a & 0x01010101u) << (4u * i);
        a1 |= gather4((a >> 1) & 0x01010101u) << (4u * i);
        // the bucket [top, top + 1) 2^24 straddles when (3 top & 0xff) >= 254
        strad |= gather4(eq_bytes(lo | 0x01010101u, 0xFFFFFFFFu) >> 7) << (4u * i);
      }
    }
    // r = a + (a >= s)
    const uint32_t lt = (~a1 & S1) | (~(a1 ^ S1) & ~a0 & S0);  // a < s
    const uint32_t ge = ~lt;
    R0 = a0 ^ ge;
    R1 = a1 ^ (a0 & ge);
  }

  // equality counts and the look-up index t (bit planes t0..t3)
  uint32_t es[Z], er[Z];
#pragma unroll
  for (int c = 0; c < Z; ++c) {
    if (TWO) {
      es[c] = ~((N0[c] ^ S0) | (N1[c] ^ S1));
      er[c] = ~((N0[c] ^ R0) | (N1[c] ^ R1));
    } else {
      es[c] = ~(N0[c] ^ S0);
      er[c] = ~es[c];
    }
  }
  uint32_t cs0, cs1, cs2;
  count_bits<Z>(es, cs0, cs1, cs2);
  uint32_t t0, t1, t2, t3 = 0;
  const uint32_t u0 = ~cs0, u1 = ~cs1, u2 = ~cs2;  // 7 - #(n == s)
  if (!POTTS) {
    t0 = u0, t1 = u1, t2 = u2;  // anti-aligned count + 7 - z
  } else {
    uint32_t cr0, cr1, cr2;
    count_bits<Z>(er, cr0, cr1, cr2);
    t0 = cr0 ^ u0;
    const uint32_t k0 = cr0 & u0;
    t1 = cr1 ^ u1 ^ k0;
    const uint32_t k1 = maj3(cr1, u1, k0);
    t2 = cr2 ^ u2 ^ k1;
    t3 = maj3(cr2, u2, k1);
  }

  // decisions on the top bytes, four sites at a time
  const uint32_t L0 = J.lut_t8[0], L1 = J.lut_t8[1], L2 = J.lut_t8[2], L3 = J.lut_t8[3];
  uint32_t acc = 0, tie = 0;
#pragma unroll
  for (uint32_t i = 0; i < 8u; ++i) {
    const uint32_t sh = 4u * i;
    uint32_t tb = spread4<0>((t0 >> sh) & 0xFu) | spread4<1>((t1 >> sh) & 0xFu) | spread4<2>((t2 >> sh) & 0xFu);
    uint32_t k8;
    if (POTTS) {
      tb |= spread4<3>((t3 >> sh) & 0xFu);
      const uint32_t lo = __builtin_amdgcn_perm(L1, L0, tb & 0x07070707u);
      const uint32_t hib = __builtin_amdgcn_perm(L3, L2, tb & 0x07070707u);
      const uint32_t mk = spread_msb(tb << 4);
      k8 = (hib & mk) | (lo & ~mk);
    } else {
      k8 = __builtin_amdgcn_perm(L1, L0, tb);
    }
    const uint32_t top = word_of(i < 4u ? T0 : T1, i & 3u);
    acc |= gather4(lt_bytes(top, k8) >> 7) << sh;
    tie |= gather4(eq_bytes(top, k8) >> 7) << sh;
  }
  const uint32_t spec = (tie | strad) & vm;
  acc &= vm & ~spec;
  uint32_t NS0 = (R0 & acc) | (S0 & ~acc);
  uint32_t NS1 = TWO ? (R1 & acc) | (S1 & ~acc) : 0u;

  // the open sites, one at a time, by the exact rule
  uint32_t open = spec;
  while (open) {
    const uint32_t k = (uint32_t)__builtin_ctz(open);
    open &= open - 1u;
    const uint32_t j = j0 + k;
    const uint32_t s = bit_of(S0, k) | (TWO ? bit_of(S1, k) << 1 : 0u);
    uint32_t r, tt;
    if (QK != 4) {
      // a tie (q <= 3 proposals never straddle): the planes' proposal and
      // look-up index are exact
      r = bit_of(R0, k) | (TWO ? bit_of(R1, k) << 1 : 0u);
      tt = bit_of(t0, k) | bit_of(t1, k) << 1 | bit_of(t2, k) << 2 | bit_of(t3, k) << 3;
    } else {
      r = propose_exact(byte32(P0, P1, k), (uint32_t)QK - 1u, key, j, sweep, colour);
      r += r >= s;
      uint32_t ns = 0, nr = 0;
#pragma unroll
      for (int c = 0; c < Z; ++c) {
        const uint32_t n = bit_of(N0[c], k) | (TWO ? bit_of(N1[c], k) << 1 : 0u);
        ns += (uint32_t)(n == s);
        nr += (uint32_t)(n == r);
      }
      tt = nr + 7u - ns;  // index + 7 - z
    }
    const uint32_t top = byte32(T0, T1, k);
    const uint32_t v = decide_exact(top, J, tt, key, j, sweep, colour) ? r : s;
    const uint32_t b = 1u << k;
    NS0 = (v & 1u) ? NS0 | b : NS0 & ~b;
    if (TWO) NS1 = (v & 2u) ? NS1 | b : NS1 & ~b;
  }

  if (NS0 != S0) own0[w + 1] = NS0;
  if (TWO && NS1 != S1) own1[w + 1] = NS1;

  if (meas) {
    // bonds of the updated sites with their new states (each bond has one
    // end in colour 1), and the states of both colours at these ranks
    uint32_t en[Z];
#pragma unroll
    for (int c = 0; c < Z; ++c) en[c] = TWO ? ~((N0[c] ^ NS0) | (N1[c] ^ NS1)) : ~(N0[c] ^ NS0);
    uint32_t c0, c1, c2;
    count_bits<Z>(en, c0, c1, c2);
    const int like = weighted_popc(c0, c1, c2, vm);
    const int n = (int)nv;
    if (!POTTS) {
      M.e += 2 * like - Z * n;  // sum of s_i s_j over the bonds
      M.m += 2 * n - 2 * (__builtin_popcount(NS0 & vm) + __builtin_popcount(N0[0] & vm));
    } else {
      M.e += like;
      const uint32_t C0 = N0[0], C1 = TWO ? N1[0] : 0u;
      if (!TWO) {
        M.cnt[1] += __builtin_popcount(NS0 & vm) + __builtin_popcount(C0 & vm);
        M.cnt[0] += 2 * n - __builtin_popcount(NS0 & vm) - __builtin_popcount(C0 & vm);
      } else {
        M.cnt[0] += __builtin_popcount(~NS1 & ~NS0 & vm) + __builtin_popcount(~C1 & ~C0 & vm);
        M.cnt[1] += __builtin_popcount(~NS1 & NS0 & vm) + __builtin_popcount(~C1 & C0 & vm);
        M.cnt[2] += __builtin_popcount(NS1 & ~NS0 & vm) + __builtin_popcount(C1 & ~C0 & vm);
        if (QK == 4) M.cnt[3] += __builtin_popcount(NS1 & NS0 & vm) + __builtin_popcount(C1 & C0 & vm);
      }
    }
  }
}

// Blocks take tiles of 256 words of one job; a block adds up its measured
// counters over consecutive tiles of one job and flushes them when the job
// changes and at the end.
__global__ __launch_bounds__(kThreads) void k_tile_bits(const TileHeader *__restrict__ hdr,
                                                        const TileJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ job_of_tile, uint64_t step,
                                                        uint32_t colour, int measure, EmLayout em) {
  __shared__ BLds L;
  // One job: block b walks a contiguous share of its XCD's eighth of the
  // tiles (blocks b, b + 8, ... run on one XCD), so halo words stay in that
  // XCD's L2.  Several jobs: block b takes tiles b, b + grid, ..., which
  // spreads every job over all XCDs (their costs differ by q and
  // dimension), and its first tile's row is read together with the header
  // (the map holds at least one entry per block).
  const uint32_t first = job_of_tile[blockIdx.x];
  const uint32_t n_tiles = hdr->pad[0];
  const bool spread = hdr->n_jobs > 1;
  uint32_t t0 = blockIdx.x, end = n_tiles, stride = gridDim.x;
  if (!spread) {
    const uint32_t xcd = blockIdx.x & 7u;
    t0 = (uint32_t)(((uint64_t)n_tiles * xcd) >> 3) + (blockIdx.x >> 3);
    end = (uint32_t)(((uint64_t)n_tiles * (xcd + 1)) >> 3);
    stride = gridDim.x >> 3;
  }
  BMeas M;
  long long *rec = nullptr;
  int nc = 0;
  bool ising = false, pending = false;
  uint32_t cur = ~0u, flushes = 0;
  for (uint32_t t = t0; t < end; t += stride) {
    const uint32_t a = !spread ? 0u : t == blockIdx.x ? first : job_of_tile[t];
    const TileJob &J = jobs[a];
    const uint64_t sweep = J.sweep_base + step;
    if (sweep >= J.sweep_end) continue;
    const bool meas = measure && colour == 1 && sweep >= J.therm;
    if (pending && a != cur) {
      flush_block(L, M, nc, ising, rec, flushes);
      pending = false;
    }
    if (meas) {
      cur = a;
      rec = em.base + J.em_row * em.row_stride + step * em.step_stride +
            (uint64_t)(blockIdx.x & em.sub_mask) * em.sub_stride;
      ising = J.kind == kIsing;
      nc = ising ? 2 : 2 + (int)J.q;
      pending = true;
    }
    const uint32_t w = (t - J.bit_tile_start) * kTileWords + threadIdx.x;
    const uint32_t sel = (J.nd - 1u) * 4u + (J.kind == kIsing ? 0u : J.q <= 2u ? 1u : J.q - 1u);
    switch (sel) {
      case 0: run_word<1, false, 0>(J, w, sweep, colour, meas, M); break;
      case 1: run_word<1, true, 0>(J, w, sweep, colour, meas, M); break;
      case 2: run_word<1, true, 3>(J, w, sweep, colour, meas, M); break;
      case 3: run_word<1, true, 4>(J, w, sweep, colour, meas, M); break;
      case 4: run_word<2, false, 0>(J, w, sweep, colour, meas, M); break;
      case 5: run_word<2, true, 0>(J, w, sweep, colour, meas, M); break;
      case 6: run_word<2, true, 3>(J, w, sweep, colour, meas, M); break;
      case 7: run_word<2, true, 4>(J, w, sweep, colour, meas, M); break;
      case 8: run_word<3, false, 0>(J, w, sweep, colour, meas, M); break;
      case 9: run_word<3, true, 0>(J, w, sweep, colour, meas, M); break;
      case 10: run_word<3, true, 3>(J, w, sweep, colour, meas, M); break;
      default: run_word<3, true, 4>(J, w, sweep, colour, meas, M); break;
    }
  }
  if (pending) flush_block(L, M, nc, ising, rec, flushes);
}

// int8 colour arrays <-> bit planes, one thread per (word, colour) of the
// bit-plane jobs in the table that start (pack) or end (unpack) in this chunk.
__device__ __forceinline__ uint32_t planes_of4(uint32_t x, bool ising, uint32_t &b1) {
  if (ising) {
    b1 = 0;
    return gather4((x >> 7) & 0x01010101u);
  }
  b1 = gather4((x >> 1) & 0x01010101u);
  return gather4(x & 0x01010101u);
}

__global__ void k_bits_pack(const TileHeader *__restrict__ hdr, const TileJob *__restrict__ jobs,
                              const uint32_t *__restrict__ job_of_tile) {
  const uint32_t n_tiles = hdr->pad[0];
  const uint64_t n = (uint64_t)n_tiles * kTileWords * 2u;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t colour = (uint32_t)(i & 1u), word = (uint32_t)(i >> 1);
    const uint32_t t = word / kTileWords;
    const TileJob &J = jobs[job_of_tile[t]];
    const uint32_t w = (t - J.bit_tile_start) * kTileWords + word % kTileWords;
    const uint32_t nh = J.nh;
    if (!(J.bit_io & 1u) || w >= ((nh + 31u) >> 5)) continue;
    const bool ising = J.kind == kIsing, two = !ising && J.q > 2u;
    const int8_t *src = J.spins + (uint64_t)colour * J.cs + 32ull * w;
    uint32_t p0 = 0, p1 = 0;
    if (32u * w + 32u <= nh) {
      const uint4 a = reinterpret_cast<const uint4 *>(src)[0], b = reinterpret_cast<const uint4 *>(src)[1];
      const uint32_t x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        uint32_t h;
        p0 |= planes_of4(x[k], ising, h) << (4 * k);
        p1 |= h << (4 * k);
      }
    } else {
      for (uint32_t k = 0; k < nh - 32u * w; ++k) {
        const uint32_t v = (uint8_t)src[k];
        p0 |= (ising ? v >> 7 : v & 1u) << k;
        p1 |= (ising ? 0u : (v >> 1) & 1u) << k;
      }
    }
    uint32_t *dst = J.bits + (colour * 2u) * J.nwp + 1u + w;
    dst[0] = p0;
    if (two) dst[J.nwp] = p1;
  }
}

__global__ void k_bits_unpack(const TileHeader *__restrict__ hdr, const TileJob *__restrict__ jobs,
                              const uint32_t *__restrict__ job_of_tile) {
  const uint32_t n_tiles = hdr->pad[0];
  const uint64_t n = (uint64_t)n_tiles * kTileWords * 2u;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t colour = (uint32_t)(i & 1u), word = (uint32_t)(i >> 1);
    const uint32_t t = word / kTileWords;
    const TileJob &J = jobs[job_of_tile[t]];
    const uint32_t w = (t - J.bit_tile_start) * kTileWords + word % kTileWords;
    const uint32_t nh = J.nh;
    if (!(J.bit_io & 2u) || w >= ((nh + 31u) >> 5)) continue;
    const bool ising = J.kind == kIsing, two = !ising && J.q > 2u;
    const uint32_t *src = J.bits + (colour * 2u) * J.nwp + 1u + w;
    const uint32_t p0 = src[0], p1 = two ? src[J.nwp] : 0u;
    int8_t *dst = J.spins + (uint64_t)colour * J.cs + 32ull * w;
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t b0 = spread4<0>((p0 >> (4 * k)) & 0xFu);
      x[k] = ising ? b0 * 0xFEu + 0x01010101u : b0 | spread4<1>((p1 >> (4 * k)) & 0xFu);
    }
    if (32u * w + 32u <= nh) {
      reinterpret_cast<uint4 *>(dst)[0] = make_uint4(x[0], x[1], x[2], x[3]);
      reinterpret_cast<uint4 *>(dst)[1] = make_uint4(x[4], x[5], x[6], x[7]);
    } else {
      for (uint32_t k = 0; k < nh - 32u * w; ++k) dst[k] = (int8_t)((x[k >> 2] >> (8u * (k & 3u))) & 0xFFu);
    }
  }
}

}  // namespace

// Byte look-up table over t = index + 7 - z (the bit kernel's index).
void fill_bit_lut(TileJob &t) {
  const int z = 2 * (int)t.nd;
  for (int w = 0; w < 4; ++w) t.lut_t8[w] = t.lut_tm8[w] = t.lut_tk16lo[w] = t.lut_tk16hi[w] = 0;
  t.tal = 0;
  for (int v = 0; v < 16; ++v) {
    const int idx = v - 7 + z;
    if (idx < 0 || idx >= kStencilTable) continue;
    const Threshold &th = t.thr[idx];
    const uint32_t sh = 8u * (uint32_t)(v & 3);
    t.lut_t8[v >> 2] |= (th.k8 > 0xFFu ? 0xFFu : th.k8) << sh;
    t.lut_tm8[v >> 2] |= (uint32_t)(th.m8 & 0xFFu) << sh;
    t.lut_tk16lo[v >> 2] |= (uint32_t)(th.k16 & 0xFFu) << sh;
    t.lut_tk16hi[v >> 2] |= (uint32_t)(th.k16 >> 8) << sh;
    if (th.k8 > 0xFFu) t.tal |= 1u << v;
  }
}

hipError_t launch_tile_bits(const TileHeader *hdr, const TileJob *jobs, const uint32_t *job_of_tile, uint32_t grid,
                            uint64_t step, uint32_t colour, int measure, EmLayout em, hipStream_t stream) {
  if (grid == 0 || (grid & 7u)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tile_bits, dim3(grid), dim3(kThreads), 0, stream, hdr, jobs, job_of_tile, step, colour, measure,
                     em);
  return hipGetLastError();
}

hipError_t launch_bits_pack(const TileHeader *hdr, const TileJob *jobs, const uint32_t *job_of_tile, int unpack,
                            hipStream_t stream) {
  if (unpack)
    hipLaunchKernelGGL(k_bits_unpack, dim3(1024), dim3(kThreads), 0, stream, hdr, jobs, job_of_tile);
  else
    hipLaunchKernelGGL(k_bits_pack, dim3(1024), dim3(kThreads), 0, stream, hdr, jobs, job_of_tile);
  return hipGetLastError();
}

}  // namespace scamper
