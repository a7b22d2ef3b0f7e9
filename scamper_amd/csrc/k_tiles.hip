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

struct TileLds {
  Threshold thr[kStencilTable];
  unsigned long long red[2 + kMaxQ];
};

__device__ __forceinline__ uint32_t byte4(uint32_t w, int k) { return (w >> (8 * k)) & 0xFFu; }

// A 16-byte window: byte k is byte k of (lo, hi) in little-endian order.
struct W16 {
  uint64_t lo, hi;
};

__device__ __forceinline__ uint32_t word16(const W16 &w, int k) {
  return (uint32_t)((k < 2 ? w.lo : w.hi) >> (32 * (k & 1)));
}

// The 16 bytes at p, any alignment: the aligned 16 bytes at p, or the
// aligned 32 bytes around it (the caller has checked they are allocated).
__device__ __forceinline__ W16 load16(const int8_t *p) {
  const uintptr_t a = (uintptr_t)p;
  const uint64_t *A = reinterpret_cast<const uint64_t *>(a & ~(uintptr_t)15);
  const uint32_t s = (uint32_t)(a & 15);
  const uint64_t a0 = A[0], a1 = A[1];
  if (s == 0) return W16{a0, a1};
  const uint64_t a2 = A[2], a3 = A[3];
  const bool q = s >= 8;
  const uint32_t r = 8 * (s & 7);
  const uint64_t e0 = q ? a1 : a0, e1 = q ? a2 : a1, e2 = q ? a3 : a2;
  if (r == 0) return W16{e0, e1};
  return W16{(e0 >> r) | (e1 << (64 - r)), (e1 >> r) | (e2 << (64 - r))};
}

// w moved one byte up (byte 0 = e: the left neighbours) or down (byte 15 = e:
// the right neighbours).
__device__ __forceinline__ W16 shift_up(const W16 &w, uint32_t e) {
  return W16{(w.lo << 8) | e, (w.hi << 8) | (w.lo >> 56)};
}
__device__ __forceinline__ W16 shift_down(const W16 &w, uint32_t e) {
  return W16{(w.lo >> 8) | (w.hi << 56), (w.hi >> 8) | ((uint64_t)e << 56)};
}

__device__ __forceinline__ W16 put_byte16(const W16 &w, uint32_t k, uint32_t v) {
  const uint32_t sh = 8u * (k & 7u);
  const uint64_t m = 0xFFull << sh, b = (uint64_t)v << sh;
  return k < 8 ? W16{(w.lo & ~m) | b, w.hi} : W16{w.lo, (w.hi & ~m) | b};
}

__device__ __forceinline__ W16 select16(const W16 &a, const W16 &b, const W16 &mask) {  // a where mask
  return W16{(a.lo & mask.lo) | (b.lo & ~mask.lo), (a.hi & mask.hi) | (b.hi & ~mask.hi)};
}

__device__ __forceinline__ void store16(int8_t *p, const W16 &w) {
  const uintptr_t a = (uintptr_t)p;
  if ((a & 15) == 0) {
    *reinterpret_cast<uint4 *>(p) = make_uint4((uint32_t)w.lo, (uint32_t)(w.lo >> 32), (uint32_t)w.hi,
                                               (uint32_t)(w.hi >> 32));
  } else if ((a & 7) == 0) {
    reinterpret_cast<uint64_t *>(p)[0] = w.lo;
    reinterpret_cast<uint64_t *>(p)[1] = w.hi;
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) p[k] = (int8_t)((k < 8 ? w.lo : w.hi) >> (8 * (k & 7)));
  }
}

// p lies in [begin, end) and its aligned read span ends by `end` (the span
// may start before `begin`: the preceding array of the same allocation,
// never before the allocation).
__device__ __forceinline__ bool window_ok16(const int8_t *p, const int8_t *begin, const int8_t *end) {
  return p >= begin && (((uintptr_t)p + 16 + 15) & ~(uintptr_t)15) <= (uintptr_t)end;
}

// The 4 bytes at p, any alignment: the aligned dword at p, or the two aligned
// dwords around it (the caller has checked that they lie in the allocation).
__device__ __forceinline__ uint32_t load4(const int8_t *p) {
  const uintptr_t a = (uintptr_t)p;
  const uint32_t *A = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
  const uint32_t s = (uint32_t)(a & 3);
  const uint32_t lo = A[0];
  if (s == 0) return lo;
  return __builtin_amdgcn_alignbyte(A[1], lo, s);
}

// p lies in [begin, end) with its aligned read span ending by `end`; the
// span may start before `begin` (the preceding array of the same
// allocation), never before the allocation.
__device__ __forceinline__ bool window_ok4(const int8_t *p, const int8_t *begin, const int8_t *end) {
  return p >= begin && (((uintptr_t)p + 4 + 3) & ~(uintptr_t)3) <= (uintptr_t)end;
}

__device__ __forceinline__ void store4(int8_t *p, uint32_t w) {
  if (((uintptr_t)p & 3) == 0) {
    *reinterpret_cast<uint32_t *>(p) = w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) p[k] = (int8_t)byte4(w, k);
  }
}

// Accumulators of a measured half-sweep (per lane).
struct Meas {
  int e = 0, m = 0;
  uint64_t cnt_lo = 0, cnt_hi = 0;  // byte v of (lo, hi): sites in Potts state v (< 16)
};

// One site: new spin from its own spin, its z neighbours and its draws.
template <int ND, int KIND>
__device__ __forceinline__ int8_t update_site(int8_t s, const int8_t (&n)[2 * ND], uint32_t top, uint32_t mid,
                                              uint32_t ptop, uint32_t pmid, const TileLds &L, uint32_t q,
                                              uint64_t key, uint32_t j, uint64_t sweep, uint32_t colour) {
  constexpr int z = 2 * ND;
  if (KIND == kIsing) {
    int anti = 0;
#pragma unroll
    for (int c = 0; c < z; ++c) anti += (n[c] != s);
    return decide2(top, mid, L.thr[anti], key, j, sweep, colour) ? (int8_t)-s : s;
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
  return decide2(top, mid, L.thr[d + z], key, j, sweep, colour) ? (int8_t)r : s;
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

__device__ __forceinline__ uint32_t put_byte(uint32_t w, uint32_t k, uint32_t v) {
  return (w & ~(0xFFu << (8u * k))) | (v << (8u * k));
}

// Per-byte look-up in a 16-entry byte table (4 words), index < 16 per byte.
__device__ __forceinline__ uint32_t lut16(const uint32_t (&t)[4], uint32_t idx) {
  const uint32_t sel = idx & 0x07070707u;
  const uint32_t lo = __builtin_amdgcn_perm(t[1], t[0], sel);
  const uint32_t hi = __builtin_amdgcn_perm(t[3], t[2], sel);
  const uint32_t mk = ((idx >> 3) & kOnes) * 0xFFu;
  return (hi & mk) | (lo & ~mk);
}

// Accept mask (bit 7 per byte) of "word < K[idx]" for the four sites: the
// top and second bytes decide; 16-bit ties (one decision in 65536, also
// every draw with top bytes 0xffff where K = 2^32) are settled exactly.
template <bool SMALL_INDEX>
__device__ __forceinline__ uint32_t accept_bytes(uint32_t idx, const LaneDraws &D, const TileJob &J,
                                                 const TileLds &L, uint64_t key, uint32_t j0, uint64_t sweep,
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
    const Threshold &t = L.thr[byte4(idx, (int)k)];
    if (t.k8 > 0xFFu || site_low(key, j0 + k, sweep, colour, kStreamAccept) < t.k16) acc |= 0x80u << (8u * k);
  }
  return acc;
}

// 1 in every byte where a and b differ, for byte values below 16 (Potts
// states): (a ^ b) + 0x7f sets bit 7 of a byte exactly when a ^ b != 0.
__device__ __forceinline__ uint32_t ne01(uint32_t a, uint32_t b) { return (((a ^ b) + 0x7F7F7F7Fu) >> 7) & kOnes; }

template <int ND>
__device__ __forceinline__ uint32_t ising_word(uint32_t S, const uint32_t (&N)[2 * ND], uint32_t C,
                                               const LaneDraws &D, const TileJob &J, const TileLds &L, uint64_t key,
                                               uint32_t j0, uint64_t sweep, uint32_t colour, bool meas, Meas &M) {
  constexpr int z = 2 * ND;
  uint32_t a2 = 0;  // spins are 0x01 / 0xff: bit 1 of s ^ n is set exactly when they differ
#pragma unroll
  for (int c = 0; c < z; ++c) a2 += (S ^ N[c]) & 0x02020202u;
  const uint32_t anti = a2 >> 1;
  const uint32_t m = accept_bytes<true>(anti, D, J, L, key, j0, sweep, colour) >> 7;
  const uint32_t NS = S ^ (((m << 8) - m) & 0xFEFEFEFEu);
  if (meas) {
    uint32_t b2 = 0;
#pragma unroll
    for (int c = 0; c < z; ++c) b2 += (NS ^ N[c]) & 0x02020202u;
    M.e += 4 * z - 2 * (int)__builtin_amdgcn_sad_u8(b2 >> 1, 0u, 0u);
    M.m += 8 - 2 * (__builtin_popcount(NS & kMsb) + __builtin_popcount(C & kMsb));
  }
  return NS;
}

template <int ND>
__device__ __forceinline__ uint32_t potts_word(uint32_t S, const uint32_t (&N)[2 * ND], uint32_t C,
                                               const LaneDraws &D, const TileJob &J, const TileLds &L, uint32_t q,
                                               uint64_t key, uint32_t j0, uint64_t sweep, uint32_t colour, bool meas,
                                               Meas &M) {
  constexpr int z = 2 * ND;
  uint32_t R;
  if (q == 2) {
    R = S ^ kOnes;
  } else {
    // floor(top * m / 256) per byte from 16-bit lane products (top * m <
    // 4096); a byte whose bucket straddles an index boundary (low byte of
    // the product + m > 256) is settled with the finer bytes.
    const uint32_t m = q - 1;
    const uint32_t pe = (D.ptop & 0x00FF00FFu) * m, po = ((D.ptop >> 8) & 0x00FF00FFu) * m;
    uint32_t idx = ((pe >> 8) & 0x000F000Fu) | (po & 0x0F000F00u);
    const uint32_t bias = (m + 255u) * 0x00010001u;
    uint32_t strad = ((((pe & 0x00FF00FFu) + bias) >> 9) & 0x00010001u) |
                     (((((po & 0x00FF00FFu) + bias) >> 9) & 0x00010001u) << 8);
    while (strad) {
      const uint32_t k = (uint32_t)__builtin_ctz(strad) >> 3;
      strad &= strad - 1;
      idx = put_byte(idx, k, propose2(byte4(D.ptop, (int)k), byte4(D.pmid, (int)k), m, key, j0 + k, sweep, colour));
    }
    R = idx + ((((idx | kMsb) - S) & kMsb) >> 7);  // skip the current state
  }
  // d + z = z + #(n != s) - #(n != r) per byte, 0 .. 2z
  uint32_t ns_ = (uint32_t)z * kOnes, nr = 0;
#pragma unroll
  for (int c = 0; c < z; ++c) {
    ns_ += ne01(N[c], S);
    nr += ne01(N[c], R);
  }
  const uint32_t idx = ns_ - nr;
  const uint32_t am = (accept_bytes<false>(idx, D, J, L, key, j0, sweep, colour) >> 7) * 0xFFu;
  const uint32_t NS = (R & am) | (S & ~am);
  if (meas) {
    uint32_t differ = 0;
#pragma unroll
    for (int c = 0; c < z; ++c) differ += ne01(N[c], NS);
    M.e += 4 * z - (int)__builtin_amdgcn_sad_u8(differ, 0u, 0u);
    for (uint32_t v = 0; v < q; ++v) {
      const uint32_t vv = v * kOnes;
      const uint64_t n = __builtin_amdgcn_sad_u8(eq_bytes(NS, vv) >> 7, 0u, 0u) +
                         __builtin_amdgcn_sad_u8(eq_bytes(C, vv) >> 7, 0u, 0u);
      if (v < 8) M.cnt_lo += n << (8 * v); else M.cnt_hi += n << (8 * (v - 8));
    }
  }
  return NS;
}

// One tile of 1024 sites, 4 per lane (quad-shared Philox group); measured
// counters are added to the lane's M.
template <int ND, int KIND>
__device__ __forceinline__ void run_tile4(TileLds &L, const TileJob &J, uint32_t tile, uint64_t sweep, uint32_t colour,
                                         bool meas, Meas &M) {
  constexpr int z = 2 * ND;
  const uint32_t nh = J.nh;
  const uint32_t j0 = tile * (4u * kThreads) + threadIdx.x * 4u;  // first of the lane's 4 sites
  const uint64_t key = J.key;
  const uint32_t q = J.q;
  const uint32_t m = q - 1;
  const bool need_prop = KIND == kPotts && q > 2;
  const bool need_pmid = need_prop && (m & (m - 1)) != 0;

  if (threadIdx.x < kStencilTable) L.thr[threadIdx.x] = J.thr[threadIdx.x];
  __syncthreads();

  // the quad's group must be drawn by all four lanes even where some of
  // them are past the end of the colour class
  const uint32_t g = j0 >> 4;
  const bool group_active = (g << 4) < nh;
  if (group_active) {
    int8_t *__restrict__ own = J.spins + (uint64_t)colour * nh;
    const int8_t *__restrict__ oth = J.spins + (uint64_t)(1 - colour) * nh;
    const uint32_t lx2 = J.lx2, ly = J.ly, lz = J.lz;
    const uint32_t row = div_by(j0 < nh ? j0 : 0u, J.div_lx2);
    const uint32_t i0 = (j0 < nh ? j0 : 0u) - row * lx2;
    const uint32_t zz = ND == 3 ? div_by(row, J.div_ly) : 0u;
    const uint32_t y = ND == 1 ? 0u : row - zz * ly;
    // sites 0 .. ks-1 of the lane lie in row a = `row`, sites ks .. 3 in
    // row b = row + 1 (rows of >= 4 sites on the window path)
    const uint32_t ks = lx2 - i0 < 4u ? lx2 - i0 : 4u;
    const bool two_rows = ks < 4u;
    const uint32_t yb = ND == 1 ? 0u : (y + 1 == ly ? 0u : y + 1);
    const uint32_t zb = ND == 3 ? (y + 1 == ly ? zz + 1 : zz) : 0u;
    const uint32_t pa = (colour + y + zz) & 1u, pb = (colour + yb + zb) & 1u;
    const uint32_t rba = row * lx2, rbb = rba + lx2;
    const int64_t plane = (int64_t)lx2 * ly;
    const int64_t upa = ND >= 2 ? (y == 0 ? (int64_t)(ly - 1) * lx2 : -(int64_t)lx2) : 0;
    const int64_t dna = ND >= 2 ? (y + 1 == ly ? -(int64_t)(ly - 1) * lx2 : (int64_t)lx2) : 0;
    const int64_t bka = ND >= 3 ? (zz == 0 ? (int64_t)(lz - 1) * plane : -plane) : 0;
    const int64_t fra = ND >= 3 ? (zz + 1 == lz ? -(int64_t)(lz - 1) * plane : plane) : 0;
    const int64_t upb = ND >= 2 ? (yb == 0 ? (int64_t)(ly - 1) * lx2 : -(int64_t)lx2) : 0;
    const int64_t dnb = ND >= 2 ? (yb + 1 == ly ? -(int64_t)(ly - 1) * lx2 : (int64_t)lx2) : 0;
    const int64_t bkb = ND >= 3 ? (zb == 0 ? (int64_t)(lz - 1) * plane : -plane) : 0;
    const int64_t frb = ND >= 3 ? (zb + 1 == lz ? -(int64_t)(lz - 1) * plane : plane) : 0;
    const int8_t *own_end = own + nh, *oth_end = oth + nh;
    bool vec = lx2 >= 4u && j0 + 4u <= nh && window_ok4(own + j0, own, own_end) &&
               window_ok4(oth + j0, oth, oth_end);
    if (ND >= 2) {
      vec = vec && window_ok4(oth + j0 + upa, oth, oth_end) && window_ok4(oth + j0 + dna, oth, oth_end);
      if (two_rows) vec = vec && window_ok4(oth + j0 + upb, oth, oth_end) && window_ok4(oth + j0 + dnb, oth, oth_end);
    }
    if (ND >= 3) {
      vec = vec && window_ok4(oth + j0 + bka, oth, oth_end) && window_ok4(oth + j0 + fra, oth, oth_end);
      if (two_rows) vec = vec && window_ok4(oth + j0 + bkb, oth, oth_end) && window_ok4(oth + j0 + frb, oth, oth_end);
    }

    // windows first, so that the loads are in flight during the Philox rounds
    uint32_t O = 0, C = 0, Ua = 0, Da = 0, Ba = 0, Fa = 0, Ub = 0, Db = 0, Bb = 0, Fb = 0, el = 0, er = 0;
    int8_t wrap_r = 0, wrap_l = 0;
    if (vec) {
      O = load4(own + j0);
      C = load4(oth + j0);
      // left neighbour of site 0 when pa = 0; right neighbour of site 3 when
      // its row's parity is 1 (site 3 is never at the end of row b)
      el = (uint8_t)oth[pa ? j0 : (i0 == 0 ? rba + lx2 - 1u : j0 - 1u)];
      const uint32_t last_p = two_rows ? pb : pa;
      const uint32_t er_idx = (!two_rows && i0 + 4u == lx2) ? rba : j0 + 4u;
      er = (uint8_t)oth[last_p ? er_idx : j0];
      wrap_r = oth[rba];                             // right of site ks-1 (i = lx2-1) when pa = 1
      wrap_l = two_rows ? oth[rbb + lx2 - 1u] : 0;  // left of site ks (i = 0) when pb = 0
      if (ND >= 2) {
        Ua = load4(oth + j0 + upa);
        Da = load4(oth + j0 + dna);
        Ub = two_rows && upb != upa ? load4(oth + j0 + upb) : Ua;
        Db = two_rows && dnb != dna ? load4(oth + j0 + dnb) : Da;
      }
      if (ND >= 3) {
        Ba = load4(oth + j0 + bka);
        Fa = load4(oth + j0 + fra);
        Bb = two_rows && bkb != bka ? load4(oth + j0 + bkb) : Ba;
        Fb = two_rows && frb != fra ? load4(oth + j0 + frb) : Fa;
      }
    }
    const LaneDraws D = quad_draws(key, g, sweep, colour, need_prop, need_pmid);

    if (vec) {
      const uint32_t Su = (C << 8) | el;          // Su byte k = C byte k-1
      const uint32_t Sd = (C >> 8) | (er << 24);  // Sd byte k = C byte k+1
      // bytes of row a, and bytes whose row parity is 1
      const uint32_t amask = two_rows ? ((1u << (8u * ks)) - 1u) : 0xFFFFFFFFu;
      const uint32_t pmask = (pa ? amask : 0u) | (pb ? ~amask : 0u);
      uint32_t N[z];
      N[0] = (C & pmask) | (Su & ~pmask);
      N[1] = (Sd & pmask) | (C & ~pmask);
      if (two_rows && pa) N[1] = put_byte(N[1], ks - 1u, (uint8_t)wrap_r);
      if (two_rows && !pb) N[0] = put_byte(N[0], ks, (uint8_t)wrap_l);
      if (ND >= 2) {
        N[2] = (Ua & amask) | (Ub & ~amask);
        N[3] = (Da & amask) | (Db & ~amask);
      }
      if (ND >= 3) {
        N[4] = (Ba & amask) | (Bb & ~amask);
        N[5] = (Fa & amask) | (Fb & ~amask);
      }
      const uint32_t NS = KIND == kIsing ? ising_word<ND>(O, N, C, D, J, L, key, j0, sweep, colour, meas, M)
                                         : potts_word<ND>(O, N, C, D, J, L, q, key, j0, sweep, colour, meas, M);
      if (NS != O) store4(own + j0, NS);
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
                                                byte4(D.pmid, k), L, q, key, j, sweep, colour);
        if (ns != s) own[j] = ns;
        if (meas) measure_site<ND, KIND>(M, ns, oth[j], n);
      }
    }
  }

  __syncthreads();  // the block's next tile rewrites L.thr
}

// One tile of 4096 sites, 16 per lane (one Philox group each); measured
// counters are added to the lane's M.
template <int ND, int KIND>
__device__ __forceinline__ void run_tile16(TileLds &L, const TileJob &J, uint32_t tile, uint64_t sweep, uint32_t colour,
                                         bool meas, Meas &M) {
  constexpr int z = 2 * ND;
  const uint32_t nh = J.nh;
  const uint32_t j0 = tile * (16u * kThreads) + threadIdx.x * 16u;  // first of the lane's 16 sites
  const uint64_t key = J.key;
  const uint32_t q = J.q;
  const uint32_t m = q - 1;
  const bool need_prop = KIND == kPotts && q > 2;
  const bool need_pmid = need_prop && (m & (m - 1)) != 0;

  if (threadIdx.x < kStencilTable) L.thr[threadIdx.x] = J.thr[threadIdx.x];
  __syncthreads();

  if (j0 < nh) {
    int8_t *__restrict__ own = J.spins + (uint64_t)colour * nh;
    const int8_t *__restrict__ oth = J.spins + (uint64_t)(1 - colour) * nh;
    const uint32_t lx2 = J.lx2, ly = J.ly, lz = J.lz;
    const uint32_t row = div_by(j0, J.div_lx2);
    const uint32_t i0 = j0 - row * lx2;
    const uint32_t zz = ND == 3 ? div_by(row, J.div_ly) : 0u;
    const uint32_t y = ND == 1 ? 0u : row - zz * ly;
    // sites 0 .. ks-1 of the lane lie in row a = `row`, sites ks .. 15 in
    // row b = row + 1 (rows of >= 16 sites on the window path)
    const uint32_t ks = lx2 - i0 < 16u ? lx2 - i0 : 16u;
    const bool two_rows = ks < 16u;
    const uint32_t yb = ND == 1 ? 0u : (y + 1 == ly ? 0u : y + 1);
    const uint32_t zb = ND == 3 ? (y + 1 == ly ? zz + 1 : zz) : 0u;
    const uint32_t pa = (colour + y + zz) & 1u, pb = (colour + yb + zb) & 1u;
    const uint32_t rba = row * lx2, rbb = rba + lx2;
    const int64_t plane = (int64_t)lx2 * ly;
    const int64_t upa = ND >= 2 ? (y == 0 ? (int64_t)(ly - 1) * lx2 : -(int64_t)lx2) : 0;
    const int64_t dna = ND >= 2 ? (y + 1 == ly ? -(int64_t)(ly - 1) * lx2 : (int64_t)lx2) : 0;
    const int64_t bka = ND >= 3 ? (zz == 0 ? (int64_t)(lz - 1) * plane : -plane) : 0;
    const int64_t fra = ND >= 3 ? (zz + 1 == lz ? -(int64_t)(lz - 1) * plane : plane) : 0;
    const int64_t upb = ND >= 2 ? (yb == 0 ? (int64_t)(ly - 1) * lx2 : -(int64_t)lx2) : 0;
    const int64_t dnb = ND >= 2 ? (yb + 1 == ly ? -(int64_t)(ly - 1) * lx2 : (int64_t)lx2) : 0;
    const int64_t bkb = ND >= 3 ? (zb == 0 ? (int64_t)(lz - 1) * plane : -plane) : 0;
    const int64_t frb = ND >= 3 ? (zb + 1 == lz ? -(int64_t)(lz - 1) * plane : plane) : 0;
    const int8_t *own_end = own + nh, *oth_end = oth + nh;
    bool vec = lx2 >= 16u && j0 + 16u <= nh && window_ok16(own + j0, own, own_end) &&
               window_ok16(oth + j0, oth, oth_end);
    if (ND >= 2) {
      vec = vec && window_ok16(oth + j0 + upa, oth, oth_end) && window_ok16(oth + j0 + dna, oth, oth_end);
      if (two_rows) vec = vec && window_ok16(oth + j0 + upb, oth, oth_end) && window_ok16(oth + j0 + dnb, oth, oth_end);
    }
    if (ND >= 3) {
      vec = vec && window_ok16(oth + j0 + bka, oth, oth_end) && window_ok16(oth + j0 + fra, oth, oth_end);
      if (two_rows) vec = vec && window_ok16(oth + j0 + bkb, oth, oth_end) && window_ok16(oth + j0 + frb, oth, oth_end);
    }

    // windows first, so that the loads are in flight during the Philox rounds
    W16 O{0, 0}, C{0, 0}, Ua{0, 0}, Da{0, 0}, Ba{0, 0}, Fa{0, 0}, Ub{0, 0}, Db{0, 0}, Bb{0, 0}, Fb{0, 0};
    uint32_t el = 0, er = 0;
    int8_t wrap_r = 0, wrap_l = 0;
    if (vec) {
      O = load16(own + j0);
      C = load16(oth + j0);
      // left neighbour of site 0 when pa = 0; right neighbour of site 15
      // when its row's parity is 1 (site 15 is never at the end of row b)
      el = (uint8_t)oth[pa ? j0 : (i0 == 0 ? rba + lx2 - 1u : j0 - 1u)];
      const uint32_t last_p = two_rows ? pb : pa;
      const uint32_t er_idx = (!two_rows && i0 + 16u == lx2) ? rba : j0 + 16u;
      er = (uint8_t)oth[last_p ? er_idx : j0];
      wrap_r = oth[rba];                             // right of site ks-1 (i = lx2-1) when pa = 1
      wrap_l = two_rows ? oth[rbb + lx2 - 1u] : 0;  // left of site ks (i = 0) when pb = 0
      if (ND >= 2) {
        Ua = load16(oth + j0 + upa);
        Da = load16(oth + j0 + dna);
        Ub = two_rows && upb != upa ? load16(oth + j0 + upb) : Ua;
        Db = two_rows && dnb != dna ? load16(oth + j0 + dnb) : Da;
      }
      if (ND >= 3) {
        Ba = load16(oth + j0 + bka);
        Fa = load16(oth + j0 + fra);
        Bb = two_rows && bkb != bka ? load16(oth + j0 + bkb) : Ba;
        Fb = two_rows && frb != fra ? load16(oth + j0 + frb) : Fa;
      }
    }
    const uint32_t g = j0 >> 4;
    const U4 T = byte_block(key, g, sweep, colour, kStreamAccept, 0);
    const U4 Mb = byte_block(key, g, sweep, colour, kStreamAccept, 1);
    const U4 PT = need_prop ? byte_block(key, g, sweep, colour, kStreamPropose, 0) : U4{0, 0, 0, 0};
    const U4 PM = need_pmid ? byte_block(key, g, sweep, colour, kStreamPropose, 1) : U4{0, 0, 0, 0};

    if (vec) {
      // bytes of row a, and bytes whose row parity is 1
      const W16 amask = two_rows ? W16{ks >= 8 ? ~0ull : ((1ull << (8u * ks)) - 1u),
                                       ks > 8 ? ((1ull << (8u * (ks - 8))) - 1u) : 0ull}
                                 : W16{~0ull, ~0ull};
      const W16 pmask{(pa ? amask.lo : 0ull) | (pb ? ~amask.lo : 0ull), (pa ? amask.hi : 0ull) | (pb ? ~amask.hi : 0ull)};
      const W16 Su = shift_up(C, el), Sd = shift_down(C, er);
      W16 Lw = select16(C, Su, pmask), Rw = select16(Sd, C, pmask);
      if (two_rows && pa) Rw = put_byte16(Rw, ks - 1u, (uint8_t)wrap_r);
      if (two_rows && !pb) Lw = put_byte16(Lw, ks, (uint8_t)wrap_l);
      W16 Uw{0, 0}, Dw{0, 0}, Bw{0, 0}, Fw{0, 0};
      if (ND >= 2) {
        Uw = select16(Ua, Ub, amask);
        Dw = select16(Da, Db, amask);
      }
      if (ND >= 3) {
        Bw = select16(Ba, Bb, amask);
        Fw = select16(Fa, Fb, amask);
      }
      uint32_t NSw[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        uint32_t N[z];
        N[0] = word16(Lw, w);
        N[1] = word16(Rw, w);
        if (ND >= 2) {
          N[2] = word16(Uw, w);
          N[3] = word16(Dw, w);
        }
        if (ND >= 3) {
          N[4] = word16(Bw, w);
          N[5] = word16(Fw, w);
        }
        const LaneDraws D{word_of(T, (uint32_t)w), word_of(Mb, (uint32_t)w), word_of(PT, (uint32_t)w),
                          word_of(PM, (uint32_t)w)};
        const uint32_t S = word16(O, w), Cw = word16(C, w), jw = j0 + 4u * (uint32_t)w;
        NSw[w] = KIND == kIsing ? ising_word<ND>(S, N, Cw, D, J, L, key, jw, sweep, colour, meas, M)
                                : potts_word<ND>(S, N, Cw, D, J, L, q, key, jw, sweep, colour, meas, M);
      }
      const W16 NS{(uint64_t)NSw[0] | ((uint64_t)NSw[1] << 32), (uint64_t)NSw[2] | ((uint64_t)NSw[3] << 32)};
      if (NS.lo != O.lo || NS.hi != O.hi) store16(own + j0, NS);
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
                                                L, q, key, j, sweep, colour);
        if (ns != s) own[j] = ns;
        if (meas) measure_site<ND, KIND>(M, ns, oth[j], n);
      }
    }
  }

  __syncthreads();  // the block's next tile rewrites L.thr
}

// Block sum of the lanes' counters, added to the record at `rec` (nc values).
__device__ __forceinline__ void flush(TileLds &L, Meas &M, int nc, bool ising, long long *rec) {
  const uint32_t lane = threadIdx.x & 63;
  if (threadIdx.x < 2 + kMaxQ) L.red[threadIdx.x] = 0;
  __syncthreads();
  for (int c = 0; c < nc; ++c) {
    long long v;
    if (c == 0) v = M.e;
    else if (c == 1) v = ising ? M.m : 0;
    else v = (long long)(((c - 2 < 8 ? M.cnt_lo : M.cnt_hi) >> (8 * ((c - 2) & 7))) & 0xFFu);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0 && v != 0) atomicAdd(&L.red[c], (unsigned long long)v);
  }
  __syncthreads();
  if ((int)threadIdx.x < nc && L.red[threadIdx.x] != 0)
    atomicAdd(reinterpret_cast<unsigned long long *>(rec) + threadIdx.x, L.red[threadIdx.x]);
  M = Meas{};
  __syncthreads();
}

template <int LANE>
__global__ __launch_bounds__(kThreads) void k_tile_sweep(const TileHeader *__restrict__ hdr,
                                                         const TileJob *__restrict__ jobs, uint64_t step,
                                                         uint32_t colour, int measure, EmLayout em) {
  __shared__ TileLds L;
  const uint32_t n_tiles = hdr->n_tiles, n_jobs = hdr->n_jobs;
  // XCD-aware walk: blocks b, b+8, ... run on one XCD and take one
  // contiguous eighth of the tiles, so halo rows stay in that XCD's L2.
  const uint32_t xcd = blockIdx.x & 7u, local = blockIdx.x >> 3, per = gridDim.x >> 3;
  const uint32_t lo = (uint32_t)(((uint64_t)n_tiles * xcd) >> 3);
  const uint32_t hi = (uint32_t)(((uint64_t)n_tiles * (xcd + 1)) >> 3);
  // A block adds up its measured counters over consecutive tiles of one job
  // and flushes them when the job changes, before a lane's 8-bit state
  // counters can overflow (they gain at most 2 * LANE per tile) and at the end.
  constexpr uint32_t kFlushTiles = 255 / (2 * LANE);
  Meas M;
  long long *rec = nullptr;
  int nc = 0;
  bool ising = false;
  uint32_t cur = ~0u, pending = 0;
  for (uint32_t t = lo + local; t < hi; t += per) {
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
      flush(L, M, nc, ising, rec);
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
        case 2: run_tile16<1, kIsing>(L, J, tile, sweep, colour, meas, M); break;
        case 3: run_tile16<1, kPotts>(L, J, tile, sweep, colour, meas, M); break;
        case 4: run_tile16<2, kIsing>(L, J, tile, sweep, colour, meas, M); break;
        case 5: run_tile16<2, kPotts>(L, J, tile, sweep, colour, meas, M); break;
        case 6: run_tile16<3, kIsing>(L, J, tile, sweep, colour, meas, M); break;
        default: run_tile16<3, kPotts>(L, J, tile, sweep, colour, meas, M); break;
      }
    } else {
      switch (sel) {
        case 2: run_tile4<1, kIsing>(L, J, tile, sweep, colour, meas, M); break;
        case 3: run_tile4<1, kPotts>(L, J, tile, sweep, colour, meas, M); break;
        case 4: run_tile4<2, kIsing>(L, J, tile, sweep, colour, meas, M); break;
        case 5: run_tile4<2, kPotts>(L, J, tile, sweep, colour, meas, M); break;
        case 6: run_tile4<3, kIsing>(L, J, tile, sweep, colour, meas, M); break;
        default: run_tile4<3, kPotts>(L, J, tile, sweep, colour, meas, M); break;
      }
    }
  }
  if (pending) flush(L, M, nc, ising, rec);
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

// 16 sites per lane once a launch covers >= 4 Mi colour-class sites (>= 4096
// waves of 64 lanes, 4 per SIMD); below that, 4 per lane keeps enough waves
// in flight.
uint32_t tile_lane_sites(uint64_t total_sites) { return total_sites >= (4ull << 20) ? 16u : 4u; }

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
