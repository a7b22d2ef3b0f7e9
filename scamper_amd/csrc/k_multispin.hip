// Multi-spin coded Ising Metropolis with one temperature per bit, and
// replica exchange on the device.
//
// Site words are u64: bit r is replica 64b + r of bundle b (1 = spin +1,
// multi_spin.rs:28-33).  Replica r draws from its own Philox key, so every
// bit follows exactly the trajectory the single-replica kernels give that
// replica; the reference's MultiSpinIsing (multi_spin.rs:107-151) shares one
// β over the 64 bits, here each bit's β comes from its temperature label.
//
// Per thread: 16 consecutive sites of one colour class (one Philox group)
// and one quarter of the 64 bits.  The anti-aligned neighbour count of every
// bit is formed bit-sliced (three planes from a full-adder tree over the
// neighbour XORs); then for each of the thread's 16 bits the top bytes of
// its draws are compared with the bit's threshold table, four sites per
// instruction.  Measured sweeps add the new anti-aligned counts into
// bit-sliced vertical counters and popcount them per bit.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kH = 0x80808080u;  // the top bit of every byte
// v_perm selectors between the two layouts of four sites' 16-bit slices:
// pairs (sites 4w and 4w+2 in one register, 4w+1 and 4w+3 in the other,
// low half first) and bytes (byte b of word h = bits 8h .. 8h+7 of site
// 4w+b's slice).  The same two selectors convert either way.
constexpr uint32_t kLoBytes = 0x06020400u, kHiBytes = 0x07030501u;

__device__ __forceinline__ uint32_t wrap_dec(uint32_t v, uint32_t n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ uint32_t wrap_inc(uint32_t v, uint32_t n) { return v + 1 == n ? 0 : v + 1; }
__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) {
  return __builtin_amdgcn_perm(hi, lo, sel);
}
__device__ __forceinline__ uint32_t rotr(uint32_t x, uint32_t s) { return __builtin_amdgcn_alignbit(x, x, s); }
__device__ __forceinline__ uint32_t maj(uint32_t a, uint32_t b, uint32_t c) { return (a & b) | (c & (a ^ b)); }

struct Planes {
  uint32_t b0, b1, b2;  // count = b0 + 2 b1 + 4 b2
};

// Bit-sliced count of set bits across z words (z = 2, 4 or 6).
template <int Z>
__device__ __forceinline__ Planes count_bits(const uint32_t *x) {
  Planes p;
  if (Z == 2) {
    p.b0 = x[0] ^ x[1];
    p.b1 = x[0] & x[1];
    p.b2 = 0;
  } else if (Z == 4) {
    const uint32_t s1 = x[0] ^ x[1] ^ x[2], c1 = maj(x[0], x[1], x[2]);
    p.b0 = s1 ^ x[3];
    const uint32_t k = s1 & x[3];
    p.b1 = c1 ^ k;
    p.b2 = c1 & k;
  } else {
    const uint32_t s1 = x[0] ^ x[1] ^ x[2], c1 = maj(x[0], x[1], x[2]);
    const uint32_t s2 = x[3] ^ x[4] ^ x[5], c2 = maj(x[3], x[4], x[5]);
    p.b0 = s1 ^ s2;
    const uint32_t k = s1 & s2;
    p.b1 = c1 ^ c2 ^ k;
    p.b2 = maj(c1, c2, k);
  }
  return p;
}

// Anti-aligned neighbour count of bit `bit` at rank j, from the words in
// memory (the other colour is not written during the half-sweep, and bit
// `bit` of rank j only by its own thread, after the decisions are settled);
// `spin` gets the bit itself.
template <int ND>
__device__ __forceinline__ uint32_t anti_of(const uint64_t *__restrict__ own, const uint64_t *__restrict__ oth,
                                            const StencilGeom &g, uint32_t colour, uint64_t j, uint32_t bit,
                                            uint32_t &spin) {
  const SiteCoords sc_ = site_coords(j, g);
  const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
  const uint32_t p = (colour + y + zz) & 1u;
  const uint64_t rb = (uint64_t)row * g.lx2;
  const uint64_t w = own[j];
  spin = (uint32_t)(w >> bit) & 1u;
  uint64_t x = (w ^ oth[rb + (p ? i : wrap_dec(i, g.lx2))]) >> bit & 1u;
  x += (w ^ oth[rb + (p ? wrap_inc(i, g.lx2) : i)]) >> bit & 1u;
  if (ND >= 2) {
    x += (w ^ oth[((uint64_t)(wrap_dec(y, g.ly) + g.ly * zz)) * g.lx2 + i]) >> bit & 1u;
    x += (w ^ oth[((uint64_t)(wrap_inc(y, g.ly) + g.ly * zz)) * g.lx2 + i]) >> bit & 1u;
  }
  if (ND >= 3) {
    x += (w ^ oth[((uint64_t)(y + g.ly * wrap_dec(zz, g.lz))) * g.lx2 + i]) >> bit & 1u;
    x += (w ^ oth[((uint64_t)(y + g.ly * wrap_inc(zz, g.lz))) * g.lx2 + i]) >> bit & 1u;
  }
  return (uint32_t)x;
}

// The thread's dword (half `hsel` of the u64 words) of site j's own word, of
// its neighbours' words (left, right, y-, y+, z-, z+) and, for a measured
// sweep, of the other colour's word at the same rank.  The neighbour columns
// come from arithmetic on values, so every load is unconditional.
template <int ND, bool MEAS>
__device__ __forceinline__ void site_dwords(const uint32_t *__restrict__ own32, const uint32_t *__restrict__ oth32,
                                            const StencilGeom &g, uint32_t colour, uint64_t j, uint32_t hsel,
                                            uint32_t (&d)[8]) {
  const SiteCoords s = site_coords(j, g);
  const uint32_t p = (colour + s.y + s.z) & 1u;
  const uint64_t rb = (uint64_t)s.row * g.lx2;
  const uint32_t t = s.i + p;  // right neighbour's column before wrapping; the left one is t - 1
  const uint32_t xl = t == 0 ? g.lx2 - 1 : t - 1, xr = t == g.lx2 ? 0 : t;
  d[0] = own32[2 * j + hsel];
  d[1] = oth32[2 * (rb + xl) + hsel];
  d[2] = oth32[2 * (rb + xr) + hsel];
  if (ND >= 2) {
    d[3] = oth32[2 * ((uint64_t)(wrap_dec(s.y, g.ly) + g.ly * s.z) * g.lx2 + s.i) + hsel];
    d[4] = oth32[2 * ((uint64_t)(wrap_inc(s.y, g.ly) + g.ly * s.z) * g.lx2 + s.i) + hsel];
  }
  if (ND >= 3) {
    d[5] = oth32[2 * ((uint64_t)(s.y + g.ly * wrap_dec(s.z, g.lz)) * g.lx2 + s.i) + hsel];
    d[6] = oth32[2 * ((uint64_t)(s.y + g.ly * wrap_inc(s.z, g.lz)) * g.lx2 + s.i) + hsel];
  }
  if (MEAS) d[7] = oth32[2 * j + hsel];
}

// One thread per (group of 16 consecutive sites, quarter q of the 64 bits):
// the group's Philox block of replica 16q + r holds the 16 sites' top draw
// bytes, so no block is computed twice.  The thread loads only the dword of
// each site word that holds its quarter, forms the anti-aligned counts of
// its 16 bits bit-sliced (two sites per register), and turns the three
// count planes into the byte layout, where one rotate-and-insert per plane
// gives the counts of four sites for one bit as the bytes of a v_perm
// selector into that bit's table of threshold top bytes.  The four
// decisions "top draw byte < threshold byte" are then SWAR byte compares.
// Equal bytes (one decision in 256) are settled after the loop: the wave
// queues them in LDS with a ballot and settles them one per lane with the
// draw's second byte and, on a 16-bit tie, its low half, as the reference
// comparison of the full 32-bit draw with K requires.  The thread writes
// only its own 16-bit slice of every site word.
template <int ND, bool MEAS, bool ONE_ROW>
__global__ __launch_bounds__(kBlock) void k_ms_sweep(uint64_t *__restrict__ words, const Threshold *__restrict__ thr_by_label,
                                                     const int32_t *__restrict__ labels,
                                                     const uint64_t *__restrict__ keys, StencilGeom g,
                                                     uint32_t colour, uint64_t sweep,
                                                     unsigned long long *__restrict__ counts,
                                                     const StepClock *__restrict__ clk) {
  constexpr int z = 2 * ND;
  if (clk) sweep = clk->sweep;
  const uint32_t bundle = blockIdx.y;
  __shared__ Threshold sthr[64 * (z + 1)];
  __shared__ uint32_t sk8[64][2];  // per bit: top byte of K by anti-aligned count (0xff where K = 2^32), v_perm table
  __shared__ uint64_t skey[64];
  __shared__ uint32_t sred[64];  // measured sweeps, per bit: anti-aligned bonds | +1 spins << 16 (block sums < 2^16)
  __shared__ uint32_t sq[kWaves][64 * 16];  // per wave: queued ties (lane | bit r << 6 | tied sites << 10)
  __shared__ uint32_t sfix[2][4][kBlock];   // per thread: settled accepts, byte layout
  for (uint32_t e = threadIdx.x; e < 64 * (z + 1); e += kBlock) {
    const uint32_t r = e / (z + 1), a = e % (z + 1);
    sthr[e] = thr_by_label[(uint32_t)labels[bundle * 64 + r] * (z + 1) + a];
  }
  if (threadIdx.x < 64) {
    const uint32_t r = threadIdx.x;
    uint32_t lut[2] = {0u, 0u};
    for (int a = 0; a <= z; ++a) {
      const uint32_t k8 = thr_by_label[(uint32_t)labels[bundle * 64 + r] * (z + 1) + a].k8;
      lut[a >> 2] |= (k8 > 0xFFu ? 0xFFu : k8) << (8 * (a & 3));
    }
    sk8[r][0] = lut[0];
    sk8[r][1] = lut[1];
    skey[r] = keys[bundle * 64 + r];
  }
  if (MEAS && threadIdx.x < 64) sred[threadIdx.x] = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int w = 0; w < 4; ++w) sfix[h][w][threadIdx.x] = 0;
  __syncthreads();

  const uint64_t tid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t grp = tid >> 2;
  const uint32_t q = (uint32_t)(tid & 3u), sh0 = 16u * q, hsel = q >> 1;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t j0 = grp * 16;
  // sites of the group inside the colour class; a group past the end loads
  // from rank 0 and decides nothing
  const uint32_t nv = j0 < g.nh ? (uint32_t)(g.nh - j0 < 16 ? g.nh - j0 : 16) : 0u;
  const uint64_t jc0 = nv ? j0 : 0;
  uint64_t *__restrict__ own = words + (2ull * bundle + colour) * g.nh;
  const uint64_t *__restrict__ oth = words + (2ull * bundle + 1 - colour) * g.nh;
  const uint32_t *__restrict__ own32 = reinterpret_cast<const uint32_t *>(own);
  const uint32_t *__restrict__ oth32 = reinterpret_cast<const uint32_t *>(oth);
  const uint32_t selp = (q & 1u) ? 0x07060302u : 0x05040100u;  // the quarter's 16 bits of two dwords

  // ONE_ROW (rows of a multiple of 16 sites): the group lies in one row,
  // whose coordinates are found once; the left and right neighbours of its
  // 16 sites are the 17 consecutive words at columns t0 - 1 .. t0 + 15
  // (wrapping only at either end).
  uint32_t t0 = 0, xw0 = 0, xw16 = 0;
  const uint32_t *xrow = oth32, *rym = oth32, *ryp = oth32, *rzm = oth32, *rzp = oth32;
  uint32_t p0 = 0;
  if (ONE_ROW) {
    const SiteCoords s = site_coords(jc0, g);
    p0 = (colour + s.y + s.z) & 1u;
    t0 = s.i + p0;
    xrow = oth32 + 2 * (uint64_t)s.row * g.lx2 + hsel;
    xw0 = t0 == 0 ? g.lx2 - 1 : t0 - 1;
    xw16 = t0 + 15 == g.lx2 ? 0 : t0 + 15;
    if (ND >= 2) {
      rym = oth32 + 2 * ((uint64_t)(wrap_dec(s.y, g.ly) + g.ly * s.z) * g.lx2 + s.i) + hsel;
      ryp = oth32 + 2 * ((uint64_t)(wrap_inc(s.y, g.ly) + g.ly * s.z) * g.lx2 + s.i) + hsel;
    }
    if (ND >= 3) {
      rzm = oth32 + 2 * ((uint64_t)(s.y + g.ly * wrap_dec(s.z, g.lz)) * g.lx2 + s.i) + hsel;
      rzp = oth32 + 2 * ((uint64_t)(s.y + g.ly * wrap_inc(s.z, g.lz)) * g.lx2 + s.i) + hsel;
    }
  }

  uint32_t pw[3][4][2];  // count planes, byte layout: [plane][quad][bits 0-7 / 8-15]
  uint32_t wp[4][2];     // own slices, pair layout
  uint32_t cp[4][2];     // other colour at the same ranks (measured sweeps), pair layout
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    uint32_t d[4][8];
    if (ONE_ROW) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int k = 4 * w + b;
        d[b][0] = own32[2 * (jc0 + k) + hsel];
        d[b][1] = k == 0 ? xrow[2 * xw0] : xrow[2 * (t0 - 1 + k)];
        d[b][2] = k == 15 ? xrow[2 * xw16] : xrow[2 * (t0 + k)];
        if (ND >= 2) {
          d[b][3] = rym[2 * k];
          d[b][4] = ryp[2 * k];
        }
        if (ND >= 3) {
          d[b][5] = rzm[2 * k];
          d[b][6] = rzp[2 * k];
        }
        if (MEAS) d[b][7] = p0 ? d[b][1] : d[b][2];
      }
    } else {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const uint64_t j = j0 + 4 * w + b;
        site_dwords<ND, MEAS>(own32, oth32, g, colour, 4 * w + b < (int)nv ? j : 0, hsel, d[b]);
      }
    }
    const uint32_t oa = perm(d[2][0], d[0][0], selp), ob = perm(d[3][0], d[1][0], selp);
    uint32_t xa[z], xb[z];
#pragma unroll
    for (int e = 0; e < z; ++e) {
      xa[e] = oa ^ perm(d[2][e + 1], d[0][e + 1], selp);
      xb[e] = ob ^ perm(d[3][e + 1], d[1][e + 1], selp);
    }
    const Planes pa = count_bits<z>(xa), pb = count_bits<z>(xb);
    pw[0][w][0] = perm(pb.b0, pa.b0, kLoBytes);
    pw[0][w][1] = perm(pb.b0, pa.b0, kHiBytes);
    pw[1][w][0] = perm(pb.b1, pa.b1, kLoBytes);
    pw[1][w][1] = perm(pb.b1, pa.b1, kHiBytes);
    pw[2][w][0] = z >= 4 ? perm(pb.b2, pa.b2, kLoBytes) : 0u;
    pw[2][w][1] = z >= 4 ? perm(pb.b2, pa.b2, kHiBytes) : 0u;
    wp[w][0] = oa;
    wp[w][1] = ob;
    if (MEAS) {
      cp[w][0] = perm(d[2][7], d[0][7], selp);
      cp[w][1] = perm(d[3][7], d[1][7], selp);
    }
  }

  // valid sites: bytes of the byte layout, and bit 8b + w for site 4w + b
  uint32_t vb[4], vcode = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    vb[w] = nv >= 4u * w + 4 ? ~0u : nv <= 4u * w ? 0u : (1u << (8 * (nv - 4 * w))) - 1u;
    vcode |= (vb[w] & kH) >> (7 - w);
  }

  uint32_t acc[2][4];  // accepted flips, byte layout
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int w = 0; w < 4; ++w) acc[h][w] = 0;
  uint32_t n_items = 0;  // the wave's queued ties (the same in every lane)
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    for (uint32_t i = 0; i < 8; ++i) {
      const uint32_t r = 8u * h + i, bit = sh0 + r;
      const U4 T = byte_block(skey[bit], grp, sweep, colour, kStreamAccept, 0);
      const uint32_t lut0 = sk8[bit][0], lut1 = sk8[bit][1];
      const uint32_t s1 = (i - 1u) & 31u, s2 = (i - 2u) & 31u, mi = 0x01010101u << i;
      uint32_t code = 0;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        uint32_t sel = ((pw[0][w][h] >> i) & 0x01010101u) | (rotr(pw[1][w][h], s1) & 0x02020202u);
        if (z >= 4) sel |= rotr(pw[2][w][h], s2) & 0x04040404u;
        const uint32_t k = perm(lut1, lut0, sel);  // threshold top bytes of the four sites
        const uint32_t x = w == 0 ? T.x : w == 1 ? T.y : w == 2 ? T.z : T.w;  // their top draw bytes
        // per byte: x < k from the borrow-free difference of the low 7 bits and the top bits
        const uint32_t dl = (x | kH) - (k & ~kH);
        const uint32_t lt = (~x & k) | (~(x ^ k) & ~dl);
        acc[h][w] |= (lt >> (7u - i)) & mi;
        // per byte: x == k
        const uint32_t e = x ^ k;
        code |= (~(((e & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | e) & kH) >> (7 - w);
      }
      code &= vcode;
      const unsigned long long has = __ballot(code != 0);
      if (has) {
        if (code) {
          const uint32_t before =
              __builtin_amdgcn_mbcnt_hi((uint32_t)(has >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)has, 0u));
          const uint32_t c = code | (code >> 4);  // site 4w + b -> bit 4b + w
          sq[wave][n_items + before] = lane | (r << 6) | (((c & 0xFFu) | ((c >> 8) & 0xFF00u)) << 10);
        }
        n_items += (uint32_t)__popcll(has);
      }
    }
  }

  // Measured sweep: the new anti-aligned counts and +1 spins of the flips
  // decided so far; each settled tie below adds its own change.
  if (MEAS) {
    constexpr uint32_t z0 = (z & 1) ? ~0u : 0u, z1 = (z & 2) ? ~0u : 0u, z2 = (z & 4) ? ~0u : 0u;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t sel = h ? kHiBytes : kLoBytes;
      uint32_t v[6] = {0, 0, 0, 0, 0, 0};  // anti-aligned bond counts with the new spins, bit-sliced
      uint32_t up[5] = {0, 0, 0, 0, 0};    // +1 spins (own new and other colour), bit-sliced
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        // sites past the colour class contribute nothing
        const uint32_t valid = vb[w];
        const uint32_t f = acc[h][w], c0in = pw[0][w][h], c1in = pw[1][w][h], c2in = pw[2][w][h];
        // z - count where flipped (the anti-aligned count of the new spin)
        const uint32_t c0 = z0 ^ c0in;
        uint32_t borrow = ~z0 & c0in;
        const uint32_t c1 = z1 ^ c1in ^ borrow;
        borrow = (~z1 & c1in) | (~(z1 ^ c1in) & borrow);
        const uint32_t c2b = z2 ^ c2in ^ borrow;
        const uint32_t n0 = ((f & c0) | (~f & c0in)) & valid, n1 = ((f & c1) | (~f & c1in)) & valid,
                       n2 = ((f & c2b) | (~f & c2in)) & valid;
        uint32_t cy = v[0] & n0;
        v[0] ^= n0;
        uint32_t t1 = v[1] ^ n1 ^ cy;
        cy = (v[1] & n1) | (cy & (v[1] ^ n1));
        v[1] = t1;
        t1 = v[2] ^ n2 ^ cy;
        cy = (v[2] & n2) | (cy & (v[2] ^ n2));
        v[2] = t1;
        t1 = v[3] ^ cy;
        cy &= v[3];
        v[3] = t1;
        t1 = v[4] ^ cy;
        cy &= v[4];
        v[4] = t1;
        v[5] ^= cy;
        const uint32_t s0 = (perm(wp[w][1], wp[w][0], sel) ^ f) & valid, s1 = perm(cp[w][1], cp[w][0], sel) & valid;
        uint32_t d2 = up[0] & s0;
        up[0] ^= s0;
        uint32_t d3 = up[0] & s1;
        up[0] ^= s1;
        uint32_t t2 = up[1] ^ d2 ^ d3;
        uint32_t cc = (up[1] & d2) | (d3 & (up[1] ^ d2));
        up[1] = t2;
        t2 = up[2] ^ cc;
        cc &= up[2];
        up[2] = t2;
        t2 = up[3] ^ cc;
        cc &= up[3];
        up[3] = t2;
        up[4] ^= cc;
      }
      // per bit of the quarter: over the four sites of every word, then over
      // the 16 lanes of the row with the same quarter (two DPP rotates), then
      // over the rows and waves of the block (LDS atomics); a lane's counts
      // are below 2^8, so both halves of the packed word stay exact
      for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t m = 0x01010101u << i, r = 8u * h + i;
        uint32_t anti = __popc(v[5] & m), plus = __popc(up[4] & m);
#pragma unroll
        for (int b = 4; b >= 0; --b) anti = 2u * anti + __popc(v[b] & m);
#pragma unroll
        for (int b = 3; b >= 0; --b) plus = 2u * plus + __popc(up[b] & m);
        uint32_t x = anti | (plus << 16);
        x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xF, 0xF, false);  // row_ror:8
        x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x124, 0xF, 0xF, false);  // row_ror:4
        if ((lane & 12u) == 0) atomicAdd(&sred[16 * (lane & 3u) + r], x);
      }
    }
  }
  // settle the wave's ties, one queued (thread, bit) per lane
  const uint64_t tid_w = (uint64_t)blockIdx.x * kBlock + wave * 64;
  for (uint32_t it = lane; it < n_items; it += 64) {
    const uint32_t v = sq[wave][it];
    const uint32_t ol = v & 63u, r = (v >> 6) & 15u;
    uint32_t tied = v >> 10;
    const uint64_t ogrp = (tid_w + ol) >> 2;
    const uint32_t obit = 16u * (uint32_t)((tid_w + ol) & 3u) + r;
    const uint64_t key = skey[obit];
    const U4 T = byte_block(key, ogrp, sweep, colour, kStreamAccept, 0);
    const U4 M = byte_block(key, ogrp, sweep, colour, kStreamAccept, 1);
    const Threshold *tr = sthr + obit * (z + 1);
    while (tied) {
      const uint32_t n = (uint32_t)__builtin_ctz(tied);
      tied &= tied - 1;
      const uint32_t k = 4u * (n & 3u) + (n >> 2);
      const uint64_t j = ogrp * 16 + k;
      uint32_t spin;
      const uint32_t a = anti_of<ND>(own, oth, g, colour, j, obit, spin);
      const Threshold t = tr[a];
      const uint32_t top = (word_of(T, k >> 2) >> (8 * (k & 3))) & 0xFFu;
      const uint32_t mid = (word_of(M, k >> 2) >> (8 * (k & 3))) & 0xFFu;
      bool ok;
      if (top != t.k8)  // K = 2^32: the table's 0xff tied with a top byte of 0xff
        ok = top < t.k8;
      else if (mid != t.m8)
        ok = mid < t.m8;
      else
        ok = site_low(key, j, sweep, colour, kStreamAccept) < t.k16;
      if (ok) {
        atomicOr(&sfix[r >> 3][k >> 2][wave * 64 + ol], 1u << (8 * (k & 3) + (r & 7)));
        if (MEAS)  // the flip turns a anti-aligned bonds into z - a, and the spin's sign (mod 2^32)
          atomicAdd(&sred[obit], (uint32_t)(z - 2 * (int)a) + (spin ? 0xFFFF0000u : 0x00010000u));
      }
    }
  }
  if (n_items) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int w = 0; w < 4; ++w) acc[h][w] |= sfix[h][w][threadIdx.x];
  }

#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint32_t na = wp[w][0] ^ perm(acc[1][w], acc[0][w], kLoBytes);  // sites 4w, 4w+2
    const uint32_t nb = wp[w][1] ^ perm(acc[1][w], acc[0][w], kHiBytes);  // sites 4w+1, 4w+3
    if (4u * w + 0 < nv) reinterpret_cast<uint16_t *>(own + j0 + 4 * w + 0)[q] = (uint16_t)na;
    if (4u * w + 1 < nv) reinterpret_cast<uint16_t *>(own + j0 + 4 * w + 1)[q] = (uint16_t)nb;
    if (4u * w + 2 < nv) reinterpret_cast<uint16_t *>(own + j0 + 4 * w + 2)[q] = (uint16_t)(na >> 16);
    if (4u * w + 3 < nv) reinterpret_cast<uint16_t *>(own + j0 + 4 * w + 3)[q] = (uint16_t)(nb >> 16);
  }
  if (MEAS) {
    __syncthreads();
    if (threadIdx.x < 64) {
      const uint32_t x = sred[threadIdx.x];
      unsigned long long *c = counts + ((uint64_t)bundle * kMsCountSlots + blockIdx.x % kMsCountSlots) * 128;
      if (x & 0xFFFFu) atomicAdd(&c[threadIdx.x], (unsigned long long)(x & 0xFFFFu));
      if (x >> 16) atomicAdd(&c[64 + threadIdx.x], (unsigned long long)(x >> 16));
    }
  }
}

// Cold or random start of every bit of one bundle; bit r of site s follows
// the plain kernels' start of replica 64b + r (keyed by the replica).
__global__ __launch_bounds__(kBlock) void k_ms_init(uint64_t *__restrict__ words, const uint64_t *__restrict__ keys,
                                                    StencilGeom g, uint32_t bundle, int random) {
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint32_t colour = blockIdx.y;
  if (j >= g.nh) return;
  const SiteCoords sc_ = site_coords(j, g);
  const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
  const uint64_t site = 2ull * i + ((colour + y + zz) & 1u) + 2ull * g.lx2 * row;
  uint64_t v = ~0ull;
  if (random) {
    v = 0;
    for (uint32_t r = 0; r < 64; ++r)
      v |= (uint64_t)(site_byte(keys[bundle * 64 + r], site, 0, 0, kStreamInit, 0) & 1u) << r;
  }
  words[(2ull * bundle + colour) * g.nh + j] = v;
}

// One replica-exchange round over the ladder of all 64 * n_bundles replicas
// (PtExchange::try_exchange, Carlo.rs/src/parallel_tempering.rs:362-419):
// pairing offset (step-1)&1 by temperature label, w = (β_self − β_partner)
// E_self for each side, the lower label decides with
// accept_log_probability(w_lower + w_upper, u), accepted pairs swap labels.
// Also turns the bit counters of the sweep into energies and clears them.
__global__ __launch_bounds__(1024) void k_ms_exchange(int32_t *__restrict__ labels, unsigned long long *__restrict__ counts,
                                                      const double *__restrict__ beta_by_label, uint32_t n,
                                                      double n_bonds, double n_sites, double j, uint64_t ex_key,
                                                      uint64_t step, int do_exchange, double *__restrict__ energy_out,
                                                      double *__restrict__ mag_out,
                                                      unsigned long long *__restrict__ stats,
                                                      StepClock *__restrict__ clk, int measure_to_clock, int tick) {
  if (clk) {
    step = clk->exchange;
    if (measure_to_clock) {
      energy_out = static_cast<double *>(clk->series) + clk->slot * 2ull * n;
      mag_out = energy_out + n;
    }
  }
  __shared__ int32_t cfg_of[1024];
  __shared__ double energy[1024];
  const uint32_t t = threadIdx.x;
  if (t < n) {
    unsigned long long anti = 0, plus = 0;
    unsigned long long *c = counts + (uint64_t)(t / 64) * kMsCountSlots * 128 + (t % 64);
    for (uint32_t s = 0; s < kMsCountSlots; ++s) {
      anti += c[s * 128];
      plus += c[s * 128 + 64];
      c[s * 128] = 0;
      c[s * 128 + 64] = 0;
    }
    energy[t] = -j * (n_bonds - 2.0 * (double)anti);
    if (energy_out) energy_out[t] = energy[t];
    if (mag_out) mag_out[t] = fabs(2.0 * (double)plus - n_sites) / n_sites;
    cfg_of[labels[t]] = (int32_t)t;
  }
  __syncthreads();
  // every thread has read the clock: the step's tick (the next kernel on
  // the stream sees it)
  if (tick && t == 0) {
    clk->sweep += 1;
    clk->exchange += do_exchange ? 1u : 0u;
    clk->slot += measure_to_clock ? 1u : 0u;
  }
  if (!do_exchange || n < 2) return;
  const uint32_t offset = (uint32_t)((step > 0 ? step - 1 : 0) & 1u);
  const uint32_t lower = offset + 2 * t;
  if (lower + 1 >= n) return;
  const int32_t a = cfg_of[lower], b = cfg_of[lower + 1];
  const double bl = beta_by_label[lower], bu = beta_by_label[lower + 1];
  const double wa = (bl - bu) * energy[a];
  const double wb = (bu - bl) * energy[b];
  const double d = wa + wb;
  const U4 blk = philox4x32_10(lower, (uint32_t)step, ctr2(step, 0, kStreamExchange), 0u, (uint32_t)ex_key,
                               (uint32_t)(ex_key >> 32));
  const double u = (double)blk.x * 2.3283064365386963e-10;
  bool accept;
  if (isnan(d))
    accept = false;
  else if (d >= 0.0)
    accept = true;
  else if (d == -INFINITY)
    accept = false;
  else {
    const double lu = log(fmax(u, 2.2250738585072014e-308));
    accept = lu < d;
    // decisions closer than a few ulps are flagged: the host libm could round
    // ln(u) differently there
    if (fabs(lu - d) <= 1e-13 * fmax(1.0, fabs(d))) atomicAdd(&stats[1], 1ull);
  }
  if (accept) {
    labels[a] = (int32_t)(lower + 1);
    labels[b] = (int32_t)lower;
    atomicAdd(&stats[0], 1ull);
  }
}

}  // namespace

hipError_t launch_ms_init(uint64_t *words, const uint64_t *keys, const StencilGeom &g, uint32_t bundle, int random,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_ms_init, dim3((unsigned)((g.nh + kBlock - 1) / kBlock), 2), dim3(kBlock), 0, st, words, keys, g,
                     bundle, random);
  return hipGetLastError();
}

hipError_t launch_ms_sweep(uint64_t *words, const Threshold *thr_by_label, const int32_t *labels, const uint64_t *keys,
                           const StencilGeom &g, uint32_t n_bundles, uint32_t colour, uint64_t sweep,
                           unsigned long long *counts, hipStream_t st, const StepClock *clk) {
  const uint64_t threads = ((g.nh + 15) / 16) * 4;  // four quarters per group of 16 sites
  dim3 grid((unsigned)((threads + kBlock - 1) / kBlock), n_bundles);
  const bool one_row = (g.lx2 & 15u) == 0;  // a group of 16 ranks never crosses a row
#define SC_MS2(ND, MEAS)                                                                                          \
  if (one_row)                                                                                                    \
    hipLaunchKernelGGL((k_ms_sweep<ND, MEAS, true>), grid, dim3(kBlock), 0, st, words, thr_by_label, labels, keys, \
                       g, colour, sweep, counts, clk);                                                            \
  else                                                                                                            \
    hipLaunchKernelGGL((k_ms_sweep<ND, MEAS, false>), grid, dim3(kBlock), 0, st, words, thr_by_label, labels,      \
                       keys, g, colour, sweep, counts, clk);
#define SC_MS(ND)       \
  if (counts) {         \
    SC_MS2(ND, true)    \
  } else {              \
    SC_MS2(ND, false)   \
  }
  switch (g.nd) {
    case 1: SC_MS(1) break;
    case 2: SC_MS(2) break;
    case 3: SC_MS(3) break;
    default: return hipErrorInvalidValue;
  }
#undef SC_MS
#undef SC_MS2
  return hipGetLastError();
}

hipError_t launch_ms_exchange(int32_t *labels, unsigned long long *counts, const double *beta_by_label, uint32_t n,
                              double n_bonds, double n_sites, double j, uint64_t ex_key, uint64_t step, int do_exchange,
                              double *energy_out, double *mag_out, unsigned long long *stats, hipStream_t st,
                              StepClock *clk, int measure_to_clock, int tick) {
  if (tick && !clk) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ms_exchange, dim3(1), dim3(1024), 0, st, labels, counts, beta_by_label, n, n_bonds, n_sites, j,
                     ex_key, step, do_exchange, energy_out, mag_out, stats, clk, measure_to_clock, tick);
  return hipGetLastError();
}

}  // namespace scamper
