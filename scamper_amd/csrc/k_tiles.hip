// Tiled checkerboard Metropolis for Ising and Potts spins over a table of
// independent periodic lattices (jobs).  Serves both the single-lattice
// state path (one row per replica) and the batched job pool (one row per
// admitted job, one launch per colour per step for all of them).
//
// Per-site rule: MetropolisCore::sweep_with_phase
// (CMC.rs/src/algorithms/metropolis.rs:67-106) with the proposal of
// IsingModel::propose_from (-s, lattice/models.rs:98-100) or
// PottsModel::random_other_state (models.rs:212-218), ΔE from the
// nearest-neighbour bonds (interaction.rs:73-84, 149-181) and
// accept_log_probability (Carlo.rs/src/acceptance.rs:11-22) folded into the
// host's integer thresholds.  Visit order is colour 0 then colour 1.
//
// Layout of work: a tile is 4096 consecutive ranks j of one colour class of
// one job, 16 per lane: one Philox group, so a lane's draws are exactly the
// bytes of its own blocks (acceptance top and mid bytes, and the proposal's
// for Potts).  Four sites share a 32-bit word and are updated together with
// byte-wise integer arithmetic (SWAR).  When a lane's 16 sites lie in at most
// two rows, its own spins and its neighbours in the other colour are
// 16-byte windows (two or four aligned 8-byte loads and a funnel shift); the
// left/right windows are the same-row window shifted by one byte with one
// edge byte, plus the periodic wrap where the lane crosses a row end.  Lanes
// in rows shorter than 16 sites, or whose windows would leave the colour
// array, take a per-site path with byte loads.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ uint32_t wrap_dec(uint32_t v, uint32_t n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ uint32_t wrap_inc(uint32_t v, uint32_t n) { return v + 1 == n ? 0 : v + 1; }

// floor(n / d) for 32-bit n and d >= 2 from magic = ceil(2^64 / d): the
// error term n * (magic * d - 2^64) / 2^64 < n * d / 2^64 < 1 / d.
__device__ __forceinline__ uint32_t div_by(uint32_t n, uint64_t magic) {
  return magic ? (uint32_t)__umul64hi((uint64_t)n, magic) : n;
}

// "word < K" from (top, mid) bytes, the low half only on a 16-bit tie.
__device__ __forceinline__ bool decide2(uint32_t top, uint32_t mid, const Threshold &t, uint64_t key, uint32_t j,
                                        uint64_t sweep, uint32_t colour) {
  if (top != t.k8) return top < t.k8;
  if (mid != t.m8) return mid < t.m8;
  return site_low(key, j, sweep, colour, kStreamAccept) < t.k16;
}

// floor(word * m / 2^32): the top byte decides unless its bucket straddles a
// boundary, then the top 16 bits, then the whole word.
__device__ __forceinline__ uint32_t propose2(uint32_t top, uint32_t mid, uint32_t m, uint64_t key, uint32_t j,
                                             uint64_t sweep, uint32_t colour) {
  const uint64_t b1 = (uint64_t)top << 24;
  uint32_t a = (uint32_t)((b1 * m) >> 32);
  if (a == (uint32_t)(((b1 + 0xFFFFFFu) * m) >> 32)) return a;
  const uint64_t b2 = b1 | ((uint64_t)mid << 16);
  a = (uint32_t)((b2 * m) >> 32);
  if (a == (uint32_t)(((b2 + 0xFFFFu) * m) >> 32)) return a;
  return (uint32_t)(((b2 | site_low(key, j, sweep, colour, kStreamPropose)) * m) >> 32);
}

// Per-wave partial sums of a flush, double-buffered so that consecutive
// flushes need one barrier each.
struct TileLds {
  int part[2][kThreads / 64][2 + kMaxQ];
};

__device__ __forceinline__ uint32_t byte4(uint32_t w, int k) { return (w >> (8 * k)) & 0xFFu; }

// A 16-byte window as four dwords: byte k is byte (k & 3) of d[k >> 2].
struct W16 {
  uint32_t d[4];
};

// Unaligned loads and stores through global-memory pointers: gfx950 global
// memory instructions take any byte address, so a window is one dwordx4 (or
// dword) access, addressed as a 32-bit offset from a uniform base.  Callers
// keep every window inside the spin allocation (its slack included).
typedef __attribute__((address_space(1))) int8_t g8;
struct __attribute__((packed)) Bytes16 {
  uint32_t d[4];
};
struct __attribute__((packed)) Bytes4 {
  uint32_t d;
};
typedef __attribute__((address_space(1))) Bytes16 gBytes16;
typedef __attribute__((address_space(1))) Bytes4 gBytes4;
__device__ __forceinline__ W16 load16(const g8 *p) {
  const gBytes16 *q = reinterpret_cast<const gBytes16 *>(p);
  return W16{{q->d[0], q->d[1], q->d[2], q->d[3]}};
}
__device__ __forceinline__ void store16(g8 *p, const W16 &w) {
  gBytes16 *q = reinterpret_cast<gBytes16 *>(p);
#pragma unroll
  for (int k = 0; k < 4; ++k) q->d[k] = w.d[k];
}
__device__ __forceinline__ uint32_t load4(const g8 *p) { return reinterpret_cast<const gBytes4 *>(p)->d; }
__device__ __forceinline__ void store4(g8 *p, uint32_t w) { reinterpret_cast<gBytes4 *>(p)->d = w; }

// a where mask, b elsewhere (bitwise).
__device__ __forceinline__ uint32_t pick(uint32_t a, uint32_t b, uint32_t mask) { return (a & mask) | (b & ~mask); }
__device__ __forceinline__ W16 pick16(const W16 &a, const W16 &b, const W16 &mask) {
  W16 r;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.d[k] = pick(a.d[k], b.d[k], mask.d[k]);
  return r;
}

// 0xff in bytes 0 .. n-1 of a window (1 <= n <= 16).
__device__ __forceinline__ W16 head_mask16(uint32_t n) {
  W16 r;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = min(max((int)n - 4 * k, 0), 4);
    r.d[k] = t == 0 ? 0u : ~0u >> (32 - 8 * t);
  }
  return r;
}

// 0xff in byte `pos` of a window; no byte when pos > 15.
__device__ __forceinline__ W16 byte_mask16(uint32_t pos) {
  const uint32_t b = 0xFFu << (8u * (pos & 3u)), dw = pos >> 2;
  W16 r;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.d[k] = dw == (uint32_t)k ? b : 0u;
  return r;
}

// Accumulators of a measured half-sweep (per lane).
struct Meas {
  int e = 0, m = 0;
  uint64_t cnt_lo = 0, cnt_hi = 0;  // byte v of (lo, hi): sites in Potts state v (< 16)
};

// A tile's counts of Potts states 0 .. 3 (q <= 4), per byte position: byte b
// of n01 is (#state 0) + 16 (#state 1), of n23 (#state 2) + 16 (#state 3);
// a word adds at most 2 per byte, a tile at most 8.
struct Nibbles {
  uint32_t n01 = 0, n23 = 0;
};

// Add a tile's nibble counts to the lane's byte counters of states 0 .. 3.
__device__ __forceinline__ void fold_nibbles(Meas &M, const Nibbles &B) {
  const uint32_t lo = 0x0F0F0F0Fu;
  const uint32_t packed = __builtin_amdgcn_sad_u8(B.n01 & lo, 0u, 0u) |
                          (__builtin_amdgcn_sad_u8((B.n01 >> 4) & lo, 0u, 0u) << 8) |
                          (__builtin_amdgcn_sad_u8(B.n23 & lo, 0u, 0u) << 16) |
                          (__builtin_amdgcn_sad_u8((B.n23 >> 4) & lo, 0u, 0u) << 24);
  M.cnt_lo += packed;  // every byte stays below 256 between flushes: no carries
}

// One site: new spin from its own spin, its z neighbours and its draws.
template <int ND, int KIND>
__device__ __forceinline__ int8_t update_site(int8_t s, const int8_t (&n)[2 * ND], uint32_t top, uint32_t mid,
                                              uint32_t ptop, uint32_t pmid, const Threshold *thr, uint32_t q,
                                              uint64_t key, uint32_t j, uint64_t sweep, uint32_t colour) {
  constexpr int z = 2 * ND;
  if (KIND == kIsing) {
    int anti = 0;
#pragma unroll
    for (int c = 0; c < z; ++c) anti += (n[c] != s);
    return decide2(top, mid, thr[anti], key, j, sweep, colour) ? (int8_t)-s : s;
  }
  uint32_t r;
  if (q == 2) {
    r = 1u - (uint32_t)s;
  } else {
    r = propose2(ptop, pmid, q - 1, key, j, sweep, colour);
    r += (r >= (uint32_t)s);
  }
  int d = 0;
#pragma unroll
  for (int c = 0; c < z; ++c) d += ((uint32_t)n[c] == r) - (n[c] == s);
  return decide2(top, mid, thr[d + z], key, j, sweep, colour) ? (int8_t)r : s;
}

template <int ND, int KIND>
__device__ __forceinline__ void measure_site(Meas &M, int8_t ns, int8_t o, const int8_t (&n)[2 * ND]) {
  constexpr int z = 2 * ND;
  if (KIND == kIsing) {
    int sum = 0;
#pragma unroll
    for (int c = 0; c < z; ++c) sum += n[c];
    M.e += ns * sum;
    M.m += ns + o;
  } else {
    int same = 0;
#pragma unroll
    for (int c = 0; c < z; ++c) same += (n[c] == ns);
    M.e += same;
    const uint64_t one_ns = 1ull << (8 * (ns & 7)), one_o = 1ull << (8 * (o & 7));
    if (ns < 8) M.cnt_lo += one_ns; else M.cnt_hi += one_ns;
    if (o < 8) M.cnt_lo += one_o; else M.cnt_hi += one_o;
  }
}

// Draw bytes of four sites (one word of each block).
struct LaneDraws {
  uint32_t top, mid, ptop, pmid;
};

__device__ __forceinline__ LaneDraws quad_draws(uint64_t key, uint32_t g, uint64_t sweep, uint32_t colour,
                                                bool need_prop, bool need_pmid) {
  const uint32_t sub = threadIdx.x & 3u;
  U4 B{0, 0, 0, 0};
  if (sub < 2u || (sub == 2u && need_prop) || (sub == 3u && need_pmid)) {
    const uint32_t stream = sub < 2u ? kStreamAccept : kStreamPropose;
    B = byte_block(key, g, sweep, colour, stream, sub & 1u);
  }
  // 4x4 transpose across the quad in two butterfly stages (lane bit 0, then
  // bit 1); every exchange picks its operands with lane-constant selects.
  const bool b0 = sub & 1u, b1 = sub & 2u;
  const uint32_t r0 = __shfl_xor(b0 ? B.x : B.y, 1), r2 = __shfl_xor(b0 ? B.z : B.w, 1);
  const uint32_t a0 = b0 ? r0 : B.x, a1 = b0 ? B.y : r0, a2 = b0 ? r2 : B.z, a3 = b0 ? B.w : r2;
  const uint32_t s0 = __shfl_xor(b1 ? a0 : a2, 2), s1 = __shfl_xor(b1 ? a1 : a3, 2);
  return LaneDraws{b1 ? s0 : a0, b1 ? s1 : a1, b1 ? a2 : s0, b1 ? a3 : s1};
}


// ---------------------------------------------------------------- SWAR
// Four sites per 32-bit word, one per byte.
constexpr uint32_t kMsb = 0x80808080u, kOnes = 0x01010101u;

// Byte-wise unsigned x < y, reported in bit 7 of each byte.  The subtraction
// is done on 7-bit fields (no borrow crosses a byte), bit 7 is repaired, and
// the borrow out of bit 7 is the comparison.
__device__ __forceinline__ uint32_t lt_bytes(uint32_t x, uint32_t y) {
  const uint32_t same = ~(x ^ y);
  const uint32_t d = ((x | kMsb) - (y & ~kMsb)) ^ (same & kMsb);
  return ((~x & y) | (same & d)) & kMsb;
}

// Byte-wise x == y, reported in bit 7 of each byte.
__device__ __forceinline__ uint32_t eq_bytes(uint32_t x, uint32_t y) {
  const uint32_t t = x ^ y;
  return ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & kMsb;
}

// 0xff in every byte of x whose bit 7 is set, 0 in the others: v_perm's
// selectors 8 .. 11 replicate bits 15, 31, 47 and 63 of {hi, lo}, which with
// lo = x and hi = x << 8 are bit 7 of bytes 1, 3, 0 and 2 of x.
__device__ __forceinline__ uint32_t spread_msb(uint32_t x) { return __builtin_amdgcn_perm(x << 8, x, 0x090B080Au); }

__device__ __forceinline__ uint32_t put_byte(uint32_t w, uint32_t k, uint32_t v) {
  return (w & ~(0xFFu << (8u * k))) | (v << (8u * k));
}

// Per-byte look-up in a 16-entry byte table (4 words), index < 16 per byte.
__device__ __forceinline__ uint32_t lut16(const uint32_t (&t)[4], uint32_t idx) {
  const uint32_t sel = idx & 0x07070707u;
  const uint32_t lo = __builtin_amdgcn_perm(t[1], t[0], sel);
  const uint32_t hi = __builtin_amdgcn_perm(t[3], t[2], sel);
  const uint32_t mk = spread_msb(idx << 4);  // 0xff where index bit 3 is set (indices < 16)
  return (hi & mk) | (lo & ~mk);
}

// Accept mask (bit 7 per byte) of "word < K[idx]" for the four sites: the
// top and second bytes decide; 16-bit ties (one decision in 65536, also
// every draw with top bytes 0xffff where K = 2^32) are settled exactly.
template <bool SMALL_INDEX>
__device__ __forceinline__ uint32_t accept_bytes(uint32_t idx, const LaneDraws &D, const TileJob &J,
                                                 uint64_t key, uint32_t j0, uint64_t sweep,
                                                 uint32_t colour) {
  uint32_t k8, m8;
  if (SMALL_INDEX) {  // every index < 8
    k8 = __builtin_amdgcn_perm(J.lut_k8[1], J.lut_k8[0], idx);
    m8 = __builtin_amdgcn_perm(J.lut_m8[1], J.lut_m8[0], idx);
  } else {
    k8 = lut16(J.lut_k8, idx);
    m8 = lut16(J.lut_m8, idx);
  }
  const uint32_t eq_t = eq_bytes(D.top, k8);
  uint32_t acc = lt_bytes(D.top, k8) | (eq_t & lt_bytes(D.mid, m8));
  uint32_t tie = eq_t & eq_bytes(D.mid, m8);
  while (tie) {
    const uint32_t k = (uint32_t)__builtin_ctz(tie) >> 3;
    tie &= tie - 1;
    const Threshold &t = J.thr[byte4(idx, (int)k)];
    if (t.k8 > 0xFFu || site_low(key, j0 + k, sweep, colour, kStreamAccept) < t.k16) acc |= 0x80u << (8u * k);
  }
  return acc;
}

// 16 in every byte where a and b differ, for byte values below 16 (Potts
// states): (a ^ b) + 15 reaches 16 exactly when a ^ b != 0, and no byte
// carries into the next.
__device__ __forceinline__ uint32_t ne16(uint32_t a, uint32_t b) { return ((a ^ b) + 0x0F0F0F0Fu) & 0x10101010u; }

template <int ND>
__device__ __forceinline__ uint32_t ising_word(uint32_t S, const uint32_t (&N)[2 * ND], uint32_t C,
                                               const LaneDraws &D, const TileJob &J, uint64_t key,
                                               uint32_t j0, uint64_t sweep, uint32_t colour, bool meas, Meas &M,
                                               uint32_t vm) {
  constexpr int z = 2 * ND;
  uint32_t a2 = 0;  // spins are 0x01 / 0xff: bit 1 of s ^ n is set exactly when they differ
#pragma unroll
  for (int c = 0; c < z; ++c) a2 += (S ^ N[c]) & 0x02020202u;
  const uint32_t anti = a2 >> 1;
  const uint32_t flip = spread_msb(accept_bytes<true>(anti, D, J, key, j0, sweep, colour));
  const uint32_t NS = S ^ (flip & 0xFEFEFEFEu);
  if (meas) {
    // a flipped site has z - anti anti-aligned neighbours; a missing site
    // counts z / 2 of them, which adds z - 2 (z / 2) = 0
    const uint32_t b2 = pick(pick((uint32_t)(2 * z) * kOnes - a2, a2, flip), (uint32_t)z * kOnes, vm);
    M.e += 4 * z - 2 * (int)__builtin_amdgcn_sad_u8(b2 >> 1, 0u, 0u);
    M.m += 2 * __builtin_popcount(vm & kMsb) -
           2 * (__builtin_popcount(NS & vm & kMsb) + __builtin_popcount(C & vm & kMsb));
  }
  return NS;
}

template <int ND>
__device__ __forceinline__ uint32_t potts_word(uint32_t S, const uint32_t (&N)[2 * ND], uint32_t C,
                                               const LaneDraws &D, const TileJob &J, uint32_t q,
                                               uint64_t key, uint32_t j0, uint64_t sweep, uint32_t colour, bool meas,
                                               Meas &M, Nibbles &B, uint32_t vm) {
  constexpr int z = 2 * ND;
  uint32_t R, idx;  // proposed states; table index d + z = z + #(n != s) - #(n != r) per byte, 0 .. 2z
  uint32_t ds = 0, dr;  // 16 #(n != s), 16 #(n != r) per byte
  if (q == 2) {
    // r is the other state: #(n != r) = z - #(n != s), index 2 #(n != s)
    R = S ^ kOnes;
#pragma unroll
    for (int c = 0; c < z; ++c) ds += ne16(N[c], S);
    dr = (uint32_t)(16 * z) * kOnes - ds;
    idx = ds >> 3;  // bytes of ds are multiples of 16: nothing crosses a byte
  } else {
    const uint32_t m = q - 1;
    uint32_t a;  // floor(u * m) per byte, u the proposal draw
    if ((m & (m - 1)) == 0) {
      // m = 2^k: the top k bits of the top byte, never a straddled bucket
      const uint32_t k = 31u - (uint32_t)__builtin_clz(m);
      a = (D.ptop >> (8u - k)) & ((m - 1u) * kOnes);
    } else {
      // floor(top * m / 256) per byte from 16-bit lane products (top * m <
      // 4096); a byte whose bucket straddles an index boundary (low byte of
      // the product + m > 256) is settled with the finer bytes.
      const uint32_t m8 = m & 0xFFu;  // m < 16: both factors below 2^24 (v_mul_u32_u24)
      const uint32_t pe = (D.ptop & 0x00FF00FFu) * m8, po = ((D.ptop >> 8) & 0x00FF00FFu) * m8;
      a = ((pe >> 8) & 0x000F000Fu) | (po & 0x0F000F00u);
      const uint32_t bias = (m + 255u) * 0x00010001u;
      uint32_t strad = (((((pe & 0x00FF00FFu) + bias) >> 9) & 0x00010001u) |
                        (((((po & 0x00FF00FFu) + bias) >> 9) & 0x00010001u) << 8));
      while (strad) {
        const uint32_t k = (uint32_t)__builtin_ctz(strad) >> 3;
        strad &= strad - 1;
        a = put_byte(a, k, propose2(byte4(D.ptop, (int)k), byte4(D.pmid, (int)k), m, key, j0 + k, sweep, colour));
      }
    }
    R = a + ((((a | kMsb) - S) & kMsb) >> 7);  // skip the current state
    dr = 0;
#pragma unroll
    for (int c = 0; c < z; ++c) {
      ds += ne16(N[c], S);
      dr += ne16(N[c], R);
    }
    // 16 z + ds - dr lies in [0, 32 z] per byte, a multiple of 16
    idx = ((uint32_t)(16 * z) * kOnes + ds - dr) >> 4;
  }
  const uint32_t am = spread_msb(accept_bytes<ND == 1>(idx, D, J, key, j0, sweep, colour));  // 0xff where accepted
  const uint32_t NS = pick(R, S, am);
  if (meas) {
    // bonds to unlike neighbours after the update: those of r where accepted
    // (a missing site: z unlike bonds, so it adds nothing)
    M.e += 4 * z - (int)(__builtin_amdgcn_sad_u8(pick(pick(dr, ds, am), (uint32_t)(16 * z) * kOnes, vm), 0u, 0u) >> 4);
    // sites in each state among the four updated and the four other-colour
    // sites: v_perm look-ups give the nibble counts of states 0 .. 3 (q <= 4)
    // or 1 in every byte in state v (q <= 8); a missing site's byte becomes
    // a selector of zero (v_perm) or a value of no state (byte compares)
    const uint32_t none = q <= 8u ? 0x0C0C0C0Cu : 0xFFFFFFFFu;
    const uint32_t NSm = pick(NS, none, vm), Cm = pick(C, none, vm);
    if (q <= 4u) {
      constexpr uint32_t k01 = 0x00001001u, k23 = 0x10010000u;  // state -> nibble
      B.n01 += __builtin_amdgcn_perm(0u, k01, NSm) + __builtin_amdgcn_perm(0u, k01, Cm);
      if (q > 2u) B.n23 += __builtin_amdgcn_perm(0u, k23, NSm) + __builtin_amdgcn_perm(0u, k23, Cm);
      return NS;
    }
#pragma unroll
    for (uint32_t v = 0; v < (uint32_t)kMaxQ; ++v) {
      if (v >= q) break;
      uint32_t n;
      if (q <= 8u) {
        const uint32_t lo = v < 4u ? 1u << (8u * v) : 0u, hi = v >= 4u ? 1u << (8u * (v - 4u)) : 0u;
        n = __builtin_amdgcn_sad_u8(__builtin_amdgcn_perm(hi, lo, NSm), 0u,
                                    __builtin_amdgcn_sad_u8(__builtin_amdgcn_perm(hi, lo, Cm), 0u, 0u));
      } else {
        const uint32_t vv = v * kOnes;
        n = __builtin_amdgcn_sad_u8(eq_bytes(NSm, vv) >> 7, 0u, __builtin_amdgcn_sad_u8(eq_bytes(Cm, vv) >> 7, 0u, 0u));
      }
      if (v < 8u) M.cnt_lo += (uint64_t)n << (8u * v); else M.cnt_hi += (uint64_t)n << (8u * (v - 8u));
    }
  }
  return NS;
}

// Where a lane's W consecutive sites (W = 4 or 16, rows of >= W sites) and
// their neighbours lie.  Sites 0 .. ks-1 are in row a = (y, zz) from column
// i0, sites ks .. W-1 (when two_rows) at the start of row b, the next row.
// Row parities pa, pb = (colour + y + z) & 1 differ unless row b starts a
// new plane (y wraps: the parity changes by ly - 2).  Neighbour windows are loaded
// at positions measured from obase = other colour - kSpinSlack in 32-bit
// arithmetic (colour classes have < 2^31 sites, every window starts at
// rank >= -16).
template <int ND, uint32_t W>
struct LaneSpan {
  uint32_t i0, ks, pa, pb, left;
  bool two_rows;
  uint32_t off_a[2 * ND - 2], off_b[2 * ND - 2];  // U, D (, B, F) offsets of rows a and b, modulo 2^32
  uint32_t pos_l, pos_r;                          // byte whose left / right neighbour wraps (W = none)

  __device__ __forceinline__ LaneSpan(const TileJob &J, uint32_t j0, uint32_t colour) {
    const uint32_t lx2 = J.lx2, ly = J.ly, lz = J.lz;
    const uint32_t row = div_by(j0, J.div_lx2);
    i0 = j0 - row * lx2;
    const uint32_t zz = ND == 3 ? div_by(row, J.div_ly) : 0u;
    const uint32_t y = ND == 1 ? 0u : row - zz * ly;
    left = lx2 - i0;
    // a lane past the end of the colour class has its sites in the last
    // row: row b would lie outside the lattice
    two_rows = left < W && j0 + W <= J.nh;
    ks = left < W ? left : W;
    pa = (colour + y + zz) & 1u;
    const uint32_t yb = ND == 1 ? 0u : (y + 1u == ly ? 0u : y + 1u);
    const uint32_t zb = ND == 3 ? (y + 1u == ly ? zz + 1u : zz) : 0u;
    pb = (colour + yb + zb) & 1u;
    if (ND >= 2) {
      const uint32_t wrap = (ly - 1u) * lx2;
      off_a[0] = y == 0u ? wrap : 0u - lx2;
      off_a[1] = y + 1u == ly ? 0u - wrap : lx2;
      off_b[0] = yb == 0u ? wrap : 0u - lx2;
      off_b[1] = yb + 1u == ly ? 0u - wrap : lx2;
      if (ND == 3) {
        const uint32_t plane = lx2 * ly, wrapz = (lz - 1u) * plane;
        off_a[2] = zz == 0u ? wrapz : 0u - plane;
        off_a[3] = zz + 1u == lz ? 0u - wrapz : plane;
        off_b[2] = zb == 0u ? wrapz : 0u - plane;
        off_b[3] = zb + 1u == lz ? 0u - wrapz : plane;
      }
    }
    // the left neighbour of a parity-0 site at i = 0 is the row's last site
    // (byte 0 when the lane starts a row, byte ks when row b starts inside);
    // the right neighbour of a parity-1 site at i = lx2-1 is the row's first
    pos_l = (i0 == 0u && !pa) ? 0u : (two_rows && !pb) ? ks : W;
    pos_r = (left <= W && pa) ? ks - 1u : W;
  }
  // window start (relative to obase) of neighbour c for rows a and b
  __device__ __forceinline__ uint32_t at_a(uint32_t js, int c) const { return js + off_a[c]; }
  __device__ __forceinline__ uint32_t at_b(uint32_t js, int c) const { return js + (two_rows ? off_b[c] : off_a[c]); }
};

// One tile of 1024 sites, 4 per lane (quad-shared Philox group); measured
// counters are added to the lane's M.
template <int ND, int KIND>
__device__ __forceinline__ void run_tile4(const TileJob &J, uint32_t tile, uint64_t sweep, uint32_t colour,
                                         bool meas, Meas &M) {
  constexpr int z = 2 * ND;
  const uint32_t nh = J.nh;
  const uint32_t j0 = tile * (4u * kThreads) + threadIdx.x * 4u;  // first of the lane's 4 sites
  const uint64_t key = J.key;
  const uint32_t q = J.q;
  const uint32_t m = q - 1;
  const bool need_prop = KIND == kPotts && q > 2;
  const bool need_pmid = need_prop && (m & (m - 1)) != 0;

  // the quad's group must be drawn by all four lanes even where some of
  // them are past the end of the colour class
  const uint32_t g = j0 >> 4;
  const bool group_active = (g << 4) < nh;
  if (group_active) {
    int8_t *__restrict__ own = J.spins + (uint64_t)colour * J.cs;
    const int8_t *__restrict__ oth = J.spins + (uint64_t)(1 - colour) * J.cs;
    const uint32_t lx2 = J.lx2, ly = J.ly, lz = J.lz;
    const bool vec = lx2 >= 4u && j0 + 4u <= nh;
    g8 *const gown = (g8 *)own;
    const g8 *const obase = (const g8 *)oth - kSpinSlack;
    const uint32_t js = j0 + kSpinSlack;
    const LaneSpan<ND, 4> sp(J, vec ? j0 : 0u, colour);

    // windows first, so that the loads are in flight during the Philox rounds
    uint32_t O = 0, C = 0, Lm = 0, Rp = 0, Lx = 0, Rx = 0, Na[z - 2 > 0 ? z - 2 : 1], Nb[z - 2 > 0 ? z - 2 : 1];
    if (vec) {
      O = load4(gown + j0);
      C = load4(obase + js);
      Lm = load4(obase + (js - 1u));
      Rp = load4(obase + (js + 1u));
      Lx = load4(obase + (sp.pos_l < 4u ? js - 1u + lx2 : js));
      Rx = load4(obase + (sp.pos_r < 4u ? js + 1u - lx2 : js));
#pragma unroll
      for (int c = 0; c < z - 2; ++c) {
        Na[c] = load4(obase + sp.at_a(js, c));
        Nb[c] = load4(obase + sp.at_b(js, c));
      }
    }
    const LaneDraws D = quad_draws(key, g, sweep, colour, need_prop, need_pmid);

    if (vec) {
      const uint32_t amask = sp.two_rows ? (1u << (8u * sp.ks)) - 1u : ~0u;  // bytes of row a
      const uint32_t pmask = (sp.pa ? amask : 0u) | (sp.pb ? ~amask : 0u);  // bytes of parity 1
      uint32_t N[z];
      N[0] = pick(Lx, pick(C, Lm, pmask), sp.pos_l < 4u ? 0xFFu << (8u * sp.pos_l) : 0u);
      N[1] = pick(Rx, pick(Rp, C, pmask), sp.pos_r < 4u ? 0xFFu << (8u * sp.pos_r) : 0u);
#pragma unroll
      for (int c = 0; c < z - 2; ++c) N[2 + c] = pick(Na[c], Nb[c], amask);
      Nibbles B;
      const uint32_t NS = KIND == kIsing ? ising_word<ND>(O, N, C, D, J, key, j0, sweep, colour, meas, M, ~0u)
                                         : potts_word<ND>(O, N, C, D, J, q, key, j0, sweep, colour, meas, M, B, ~0u);
      if (KIND == kPotts && meas && q <= 4u) fold_nibbles(M, B);
      if (NS != O) store4(gown + j0, NS);
    } else {
#pragma unroll
      for (uint32_t k = 0; k < 4u; ++k) {  // unrolled: draw bytes are picked by constant k
        const uint32_t j = j0 + k;
        if (j >= nh) continue;
        const uint32_t r2 = div_by(j, J.div_lx2);
        const uint32_t i = j - r2 * lx2;
        const uint32_t z2 = ND == 3 ? div_by(r2, J.div_ly) : 0u;
        const uint32_t y2 = ND == 1 ? 0u : r2 - z2 * ly;
        const uint32_t p2 = (colour + y2 + z2) & 1u;
        const uint32_t rb2 = r2 * lx2;
        int8_t n[z];
        n[0] = oth[rb2 + (p2 ? i : wrap_dec(i, lx2))];
        n[1] = oth[rb2 + (p2 ? wrap_inc(i, lx2) : i)];
        if (ND >= 2) {
          n[2] = oth[(wrap_dec(y2, ly) + ly * z2) * lx2 + i];
          n[3] = oth[(wrap_inc(y2, ly) + ly * z2) * lx2 + i];
        }
        if (ND >= 3) {
          n[4] = oth[(y2 + ly * wrap_dec(z2, lz)) * lx2 + i];
          n[5] = oth[(y2 + ly * wrap_inc(z2, lz)) * lx2 + i];
        }
        const int8_t s = own[j];
        const int8_t ns = update_site<ND, KIND>(s, n, byte4(D.top, k), byte4(D.mid, k), byte4(D.ptop, k),
                                                byte4(D.pmid, k), J.thr, q, key, j, sweep, colour);
        if (ns != s) own[j] = ns;
        if (meas) measure_site<ND, KIND>(M, ns, oth[j], n);
      }
    }
  }

}

// One tile of 4096 sites, 16 per lane (one Philox group each); measured
// counters are added to the lane's M.
template <int ND, int KIND>
__device__ __forceinline__ void run_tile16(const TileJob &J, uint32_t tile, uint64_t sweep, uint32_t colour,
                                         bool meas, Meas &M) {
  constexpr int z = 2 * ND;
  const uint32_t nh = J.nh;
  const uint32_t j0 = tile * (16u * kThreads) + threadIdx.x * 16u;  // first of the lane's 16 sites
  const uint64_t key = J.key;
  const uint32_t q = J.q;
  const uint32_t m = q - 1;
  const bool need_prop = KIND == kPotts && q > 2;
  const bool need_pmid = need_prop && (m & (m - 1)) != 0;

  if (j0 < nh) {
    int8_t *__restrict__ own = J.spins + (uint64_t)colour * J.cs;
    const int8_t *__restrict__ oth = J.spins + (uint64_t)(1 - colour) * J.cs;
    const uint32_t lx2 = J.lx2, ly = J.ly, lz = J.lz;
    // Rows of >= 16 sites: 16-byte windows.  A lane with fewer than 16 ranks
    // left lies inside the class's last row; its windows reach into the gap
    // of >= 16 bytes after the colour array, and its missing sites are
    // masked out (so no lane falls back to the site-by-site path there).
    // The rank test is always true (j0 < nh, a multiple of 16), written so
    // that the compiler keeps the register allocation of the full lanes.
    const bool vec = lx2 >= 16u && j0 + 16u <= ((nh + 15u) & ~15u);
    g8 *const gown = (g8 *)own;
    const g8 *const obase = (const g8 *)oth - kSpinSlack;
    const uint32_t js = j0 + kSpinSlack;
    const LaneSpan<ND, 16> sp(J, vec ? j0 : 0u, colour);

    // windows first, so that the loads are in flight during the Philox rounds
    W16 O{}, C{}, Lm{}, Rp{}, Lx{}, Rx{}, Na[z - 2 > 0 ? z - 2 : 1];
    if (vec) {
      O = load16(gown + j0);
      C = load16(obase + js);
      Lm = load16(obase + (js - 1u));
      Rp = load16(obase + (js + 1u));
      Lx = load16(obase + (sp.pos_l < 16u ? js - 1u + lx2 : js));
      Rx = load16(obase + (sp.pos_r < 16u ? js + 1u - lx2 : js));
#pragma unroll
      for (int c = 0; c < z - 2; ++c) Na[c] = load16(obase + sp.at_a(js, c));
    }
    const uint32_t g = j0 >> 4;
    const U4 T = byte_block(key, g, sweep, colour, kStreamAccept, 0);
    const U4 Mb = byte_block(key, g, sweep, colour, kStreamAccept, 1);
    const U4 PT = need_prop ? byte_block(key, g, sweep, colour, kStreamPropose, 0) : U4{0, 0, 0, 0};
    const U4 PM = need_pmid ? byte_block(key, g, sweep, colour, kStreamPropose, 1) : U4{0, 0, 0, 0};

    if (vec) {
      const W16 amask = head_mask16(sp.ks);  // bytes of row a
      W16 pmask;                             // bytes of parity 1
#pragma unroll
      for (int k = 0; k < 4; ++k) pmask.d[k] = (sp.pa ? amask.d[k] : 0u) | (sp.pb ? ~amask.d[k] : 0u);
      const W16 Lw = pick16(Lx, pick16(C, Lm, pmask), byte_mask16(sp.pos_l));
      const W16 Rw = pick16(Rx, pick16(Rp, C, pmask), byte_mask16(sp.pos_r));
      // rows a and b take different U/D/B/F offsets only next to a periodic
      // wrap: only those lanes load a second window
      W16 Nw[z - 2 > 0 ? z - 2 : 1];
      bool split = false;
#pragma unroll
      for (int c = 0; c < z - 2; ++c) {
        Nw[c] = Na[c];
        split |= sp.two_rows && sp.off_b[c] != sp.off_a[c];
      }
      if (split) {
#pragma unroll
        for (int c = 0; c < z - 2; ++c) Nw[c] = pick16(Na[c], load16(obase + sp.at_b(js, c)), amask);
      }
      W16 NS;
      Nibbles B;
      bool changed = false;
      // Where the colour class ends inside the lane, the bytes past its end
      // hold no site: they are updated like sites (every byte-wise carry
      // and borrow runs towards higher bytes, away from the real sites
      // below them), written back into the gap, and left out of the
      // measured counters (vm: 0xff in the bytes of real sites).
      const uint32_t nv = min(nh - j0, 16u);
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        uint32_t N[z];
        N[0] = Lw.d[w];
        N[1] = Rw.d[w];
#pragma unroll
        for (int c = 0; c < z - 2; ++c) N[2 + c] = Nw[c].d[w];
        const LaneDraws D{word_of(T, (uint32_t)w), word_of(Mb, (uint32_t)w), word_of(PT, (uint32_t)w),
                          word_of(PM, (uint32_t)w)};
        const uint32_t S = O.d[w], jw = j0 + 4u * (uint32_t)w;
        const uint32_t have = nv - min(nv, 4u * (uint32_t)w);  // real sites from byte 0 of this word on
        const uint32_t vm = have >= 4u ? ~0u : (1u << (8u * have)) - 1u;
        NS.d[w] = KIND == kIsing ? ising_word<ND>(S, N, C.d[w], D, J, key, jw, sweep, colour, meas, M, vm)
                                 : potts_word<ND>(S, N, C.d[w], D, J, q, key, jw, sweep, colour, meas, M, B, vm);
        changed |= NS.d[w] != S;
      }
      if (KIND == kPotts && meas && q <= 4u) fold_nibbles(M, B);
      if (changed) store16(gown + j0, NS);
    } else {
#pragma unroll
      for (uint32_t k = 0; k < 16u; ++k) {  // unrolled: draw bytes are picked by constant k
        const uint32_t j = j0 + k;
        if (j >= nh) continue;
        const uint32_t r2 = div_by(j, J.div_lx2);
        const uint32_t i = j - r2 * lx2;
        const uint32_t z2 = ND == 3 ? div_by(r2, J.div_ly) : 0u;
        const uint32_t y2 = ND == 1 ? 0u : r2 - z2 * ly;
        const uint32_t p2 = (colour + y2 + z2) & 1u;
        const uint32_t rb2 = r2 * lx2;
        int8_t n[z];
        n[0] = oth[rb2 + (p2 ? i : wrap_dec(i, lx2))];
        n[1] = oth[rb2 + (p2 ? wrap_inc(i, lx2) : i)];
        if (ND >= 2) {
          n[2] = oth[(wrap_dec(y2, ly) + ly * z2) * lx2 + i];
          n[3] = oth[(wrap_inc(y2, ly) + ly * z2) * lx2 + i];
        }
        if (ND >= 3) {
          n[4] = oth[(y2 + ly * wrap_dec(z2, lz)) * lx2 + i];
          n[5] = oth[(y2 + ly * wrap_inc(z2, lz)) * lx2 + i];
        }
        const int8_t s = own[j];
        const uint32_t wk = k >> 2;
        const int8_t ns = update_site<ND, KIND>(s, n, byte4(word_of(T, wk), (int)(k & 3)), byte4(word_of(Mb, wk), (int)(k & 3)),
                                                byte4(word_of(PT, wk), (int)(k & 3)), byte4(word_of(PM, wk), (int)(k & 3)),
                                                J.thr, q, key, j, sweep, colour);
        if (ns != s) own[j] = ns;
        if (meas) measure_site<ND, KIND>(M, ns, oth[j], n);
      }
    }
  }

}

// Sum of v over the 64 lanes of a wave (every lane active): quad, half-row
// and row sums by DPP, then rows 0 + 1 and 2 + 3 by row broadcasts, whose
// total lands in lane 63.
__device__ __forceinline__ int wave_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm [1, 0, 3, 2]
  v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm [2, 3, 0, 1]
  v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
  v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);  // row_mirror
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1 and 3
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2 and 3
  return __builtin_amdgcn_readlane(v, 63);
}

// Block sum of the lanes' counters, added to the record at `rec` (nc values).
// Every lane's counters stay far below 2^31 between flushes (at most 2 LANE
// sites and 2 LANE z bonds per tile, 255 / (2 LANE) tiles), and so do a
// wave's and the block's.  Flush k writes partial buffer k & 1: a wave can
// rewrite a buffer only after passing the next flush's barrier, which every
// wave reaches after reading this one.
__device__ __forceinline__ void flush(TileLds &L, Meas &M, int nc, bool ising, long long *rec, uint32_t &flushes) {
  const uint32_t wave = threadIdx.x >> 6, buf = flushes & 1u;
  for (int c = 0; c < nc; ++c) {
    int v;
    if (c == 0) v = M.e;
    else if (c == 1) v = ising ? M.m : 0;
    else v = (int)(((c - 2 < 8 ? M.cnt_lo : M.cnt_hi) >> (8 * ((c - 2) & 7))) & 0xFFu);
    const int sum = wave_sum(v);
    if ((threadIdx.x & 63) == 0) L.part[buf][wave][c] = sum;
  }
  __syncthreads();
  if ((int)threadIdx.x < nc) {
    long long t = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) t += L.part[buf][w][threadIdx.x];
    if (t != 0) atomicAdd(reinterpret_cast<unsigned long long *>(rec) + threadIdx.x, (unsigned long long)t);
  }
  M = Meas{};
  ++flushes;
}

// At most 96 VGPRs, so five waves share a SIMD: the 3-D Potts variant
// alone would take 113 (four waves); the compiler spills one value, saved
// at the start of the kernel and restored at its end.
template <int LANE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_tile_sweep(const TileHeader *__restrict__ hdr,
                                                         const TileJob *__restrict__ jobs, uint64_t step,
                                                         uint32_t colour, int measure, EmLayout em) {
  __shared__ TileLds L;
  const uint32_t n_tiles = hdr->n_tiles, n_jobs = hdr->n_jobs;
  // XCD-aware walk: blocks b, b+8, ... run on one XCD and take one
  // contiguous eighth of the tiles, so halo rows stay in that XCD's L2.  A
  // table of jobs of different costs (header flag pad[1] bit 0, set by the
  // pool for several jobs) is dealt tile by tile over all blocks instead.
  const bool spread = (hdr->pad[1] & 1u) != 0u;
  const uint32_t xcd = blockIdx.x & 7u;
  const uint32_t lo = spread ? blockIdx.x : (uint32_t)(((uint64_t)n_tiles * xcd) >> 3) + (blockIdx.x >> 3);
  const uint32_t hi = spread ? n_tiles : (uint32_t)(((uint64_t)n_tiles * (xcd + 1)) >> 3);
  const uint32_t per = spread ? gridDim.x : gridDim.x >> 3;
  // A block adds up its measured counters over consecutive tiles of one job
  // and flushes them when the job changes, before a lane's 8-bit state
  // counters can overflow (they gain at most 2 * LANE per tile) and at the end.
  constexpr uint32_t kFlushTiles = 255 / (2 * LANE);
  Meas M;
  long long *rec = nullptr;
  int nc = 0;
  bool ising = false;
  uint32_t cur = ~0u, pending = 0, flushes = 0;
  for (uint32_t t = lo; t < hi; t += per) {
    uint32_t a = 0, b = n_jobs;  // last job with tile_start <= t
    while (b - a > 1) {
      const uint32_t c = (a + b) >> 1;
      if (jobs[c].tile_start <= t) a = c; else b = c;
    }
    const TileJob &J = jobs[a];
    const uint64_t sweep = J.sweep_base + step;
    if (sweep >= J.sweep_end) continue;
    const bool meas = measure && colour == 1 && sweep >= J.therm;
    if (pending && (a != cur || pending == kFlushTiles)) {
      flush(L, M, nc, ising, rec, flushes);
      pending = 0;
    }
    if (meas) {
      cur = a;
      rec = em.base + J.em_row * em.row_stride + step * em.step_stride +
            (uint64_t)(blockIdx.x & em.sub_mask) * em.sub_stride;
      ising = J.kind == kIsing;
      nc = ising ? 2 : 2 + (int)J.q;
      ++pending;
    }
    const uint32_t tile = t - J.tile_start;
    const uint32_t sel = J.nd * 2 + (J.kind == kPotts ? 1 : 0);
    if (LANE == 16) {
      switch (sel) {
        case 2: run_tile16<1, kIsing>(J, tile, sweep, colour, meas, M); break;
        case 3: run_tile16<1, kPotts>(J, tile, sweep, colour, meas, M); break;
        case 4: run_tile16<2, kIsing>(J, tile, sweep, colour, meas, M); break;
        case 5: run_tile16<2, kPotts>(J, tile, sweep, colour, meas, M); break;
        case 6: run_tile16<3, kIsing>(J, tile, sweep, colour, meas, M); break;
        default: run_tile16<3, kPotts>(J, tile, sweep, colour, meas, M); break;
      }
    } else {
      switch (sel) {
        case 2: run_tile4<1, kIsing>(J, tile, sweep, colour, meas, M); break;
        case 3: run_tile4<1, kPotts>(J, tile, sweep, colour, meas, M); break;
        case 4: run_tile4<2, kIsing>(J, tile, sweep, colour, meas, M); break;
        case 5: run_tile4<2, kPotts>(J, tile, sweep, colour, meas, M); break;
        case 6: run_tile4<3, kIsing>(J, tile, sweep, colour, meas, M); break;
        default: run_tile4<3, kPotts>(J, tile, sweep, colour, meas, M); break;
      }
    }
  }
  if (pending) flush(L, M, nc, ising, rec, flushes);
}

__global__ void k_tile_zero(const TileHeader *__restrict__ hdr, long long *__restrict__ p, uint64_t per_row) {
  const uint64_t n = (uint64_t)hdr->n_jobs * per_row;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_em_zero(long long *__restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_em_fold(const long long *__restrict__ ring, long long *__restrict__ out, uint64_t n_rec,
                          uint32_t width) {
  const uint64_t n = n_rec * width;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t rec = i / width, c = i - rec * width;
    long long s = 0;
#pragma unroll
    for (int sub = 0; sub < 8; ++sub) s += ring[(rec * 8 + sub) * width + c];
    out[i] = s;
  }
}

__global__ void k_tile_fold(const TileHeader *__restrict__ hdr, const long long *__restrict__ ring,
                            long long *__restrict__ compact, uint32_t steps) {
  const uint64_t n = (uint64_t)hdr->n_jobs * steps * 8;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t rec = i >> 3, c = i & 7;
    long long s = 0;
#pragma unroll
    for (int sub = 0; sub < 8; ++sub) s += ring[(rec * 8 + sub) * 8 + c];
    compact[i] = s;
  }
}

}  // namespace

uint64_t div_magic(uint32_t d) {
  if (d <= 1) return 0;
  return ~0ull / d + 1;  // ceil(2^64 / d) for d >= 2 (exact when d divides 2^64)
}

// 16 sites per lane once a launch covers >= 2 Mi colour-class sites (>= 2048
// waves of 64 lanes); below that, 4 per lane keeps enough waves in flight
// (measured crossover on MI355X: 1.5 M sites favour 4, 2 M sites favour 16).
uint32_t tile_lane_sites(uint64_t total_sites) { return total_sites >= (2ull << 20) ? 16u : 4u; }

hipError_t launch_tile_sweep(const TileHeader *hdr, const TileJob *jobs, uint32_t grid, uint32_t lane_sites,
                             uint64_t step, uint32_t colour, int measure, EmLayout em, hipStream_t stream) {
  if (grid == 0 || (grid & 7u)) return hipErrorInvalidValue;
  if (lane_sites == 16)
    hipLaunchKernelGGL(k_tile_sweep<16>, dim3(grid), dim3(kThreads), 0, stream, hdr, jobs, step, colour, measure, em);
  else if (lane_sites == 4)
    hipLaunchKernelGGL(k_tile_sweep<4>, dim3(grid), dim3(kThreads), 0, stream, hdr, jobs, step, colour, measure, em);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_tile_zero(const TileHeader *hdr, long long *p, uint64_t per_row, hipStream_t stream) {
  hipLaunchKernelGGL(k_tile_zero, dim3(256), dim3(256), 0, stream, hdr, p, per_row);
  return hipGetLastError();
}

hipError_t launch_em_zero(long long *p, uint64_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_em_zero, dim3(64), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

hipError_t launch_em_fold(const long long *ring, long long *out, uint64_t n_rec, uint32_t width, hipStream_t stream) {
  hipLaunchKernelGGL(k_em_fold, dim3(64), dim3(256), 0, stream, ring, out, n_rec, width);
  return hipGetLastError();
}

hipError_t launch_tile_fold(const TileHeader *hdr, const long long *ring, long long *compact, uint32_t steps,
                            hipStream_t stream) {
  hipLaunchKernelGGL(k_tile_fold, dim3(64), dim3(256), 0, stream, hdr, ring, compact, steps);
  return hipGetLastError();
}

}  // namespace scamper
