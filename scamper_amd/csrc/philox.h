// Counter-based random numbers shared by host and device code.
//
// Philox4x32-10 is written from Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC'11): ten rounds of two 32x32->64 products
// with the multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.
//
// Draw layout (the "injected uniforms" every kernel and the CPU checker use).
// A site is (colour c, rank j of the site inside colour class c); a draw is
// one 32-bit word per (key, sweep, c, j, stream s), built from three Philox
// blocks with x(t) = (sweep>>32 & 0xffff) | c<<16 | t<<24:
//
//   top(j)  = byte (j & 15) of Philox({ j>>4, sweep, x(s),   j>>36 }, key)
//   mid(j)  = byte (j & 15) of Philox({ j>>4, sweep, x(s+1), j>>36 }, key)
//   low(j)  = half (j & 7)  of Philox({ j>>3, sweep, x(s+2), j>>35 }, key)
//   word(j) = top << 24 | mid << 16 | low,      uniform u = word * 2^-32
//
// Byte k of a 32-bit word is bits 8k..8k+7, half k bits 16k..16k+15.  The
// top and mid blocks hold the bytes of 16 consecutive sites in the same order
// as 16 int8 spins, so kernels compare four draws at a time with byte-wise
// arithmetic; the low block is needed only when the top 16 bits equal those
// of the threshold (one decision in ~65536).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SC_HD __host__ __device__ __forceinline__
#else
#define SC_HD inline
#endif

namespace scamper {

enum Stream : uint32_t {
  kStreamAccept = 0,        // acceptance / heat-bath draws (uses 0, 1, 2)
  kStreamPropose = 3,       // Potts proposal draws (uses 3, 4, 5)
  kStreamInit = 6,          // discrete random start (j = site, colour 0, sweep 0)
  kStreamCont = 7,          // continuous spins: two 32-bit words per site
  kStreamContFallback = 8,  // continuous isotropic redraw (zero field)
  kStreamExchange = 9,      // replica exchange
  kStreamContInit = 10,     // continuous random start
  kStreamClusterBond = 11,   // cluster updates: bond activation by physical edge index
  kStreamClusterTarget = 12, // Swendsen-Wang cluster target by the cluster's smallest site
  kStreamClusterSeed = 13,   // Wolff seed site (index 0) and Potts target (index 1)
  kStreamVonMises = 14,      // O(2) heat bath: block (rank, try) of the von Mises rejection sampler
};

struct U4 {
  uint32_t x, y, z, w;
};

// a ^ b ^ c in one v_bitop3_b32 on the device (truth table 0x96).
SC_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

SC_HD U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                       uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32->64 product per lane pair (v_mad_u64_u32 on gfx950)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    c0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    c1 = (uint32_t)p1;
    c2 = xor3((uint32_t)(p0 >> 32), c3, k1);
    c3 = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}

// Third counter word: sweep bits 32..47, colour and stream.
SC_HD uint32_t ctr2(uint64_t sweep, uint32_t colour, uint32_t stream) {
  return ((uint32_t)(sweep >> 32) & 0xFFFFu) | ((colour & 0xFFu) << 16) | ((stream & 0xFFu) << 24);
}

// Block holding the top bytes (part = 0) or mid bytes (part = 1) of sites
// 16g .. 16g+15.
SC_HD U4 byte_block(uint64_t key, uint64_t g, uint64_t sweep, uint32_t colour, uint32_t stream,
                    uint32_t part) {
  return philox4x32_10((uint32_t)g, (uint32_t)sweep, ctr2(sweep, colour, stream + part),
                       (uint32_t)(g >> 32), (uint32_t)key, (uint32_t)(key >> 32));
}

// Block holding the low halves of sites 8h .. 8h+7.
SC_HD U4 low_block(uint64_t key, uint64_t h, uint64_t sweep, uint32_t colour, uint32_t stream) {
  return philox4x32_10((uint32_t)h, (uint32_t)sweep, ctr2(sweep, colour, stream + 2),
                       (uint32_t)(h >> 32), (uint32_t)key, (uint32_t)(key >> 32));
}

SC_HD uint32_t word_of(const U4 &b, uint32_t k) {
  // two selects on values (an indexed form gets lowered to a stack array)
  const uint32_t lo = (k & 1) ? b.y : b.x;
  const uint32_t hi = (k & 1) ? b.w : b.z;
  return (k & 2) ? hi : lo;
}

// Cluster-update draw for index idx of update `step` in stream s: word
// idx & 3 of the block {idx >> 2, step, x(s), idx >> 34}.
// The block holding the cluster words of indices 4 blk .. 4 blk + 3.
SC_HD U4 cluster_block(uint64_t key, uint64_t blk, uint64_t step, uint32_t stream) {
  return philox4x32_10((uint32_t)blk, (uint32_t)step, ctr2(step, 0, stream), (uint32_t)(blk >> 32), (uint32_t)key,
                       (uint32_t)(key >> 32));
}

SC_HD uint32_t cluster_word(uint64_t key, uint64_t idx, uint64_t step, uint32_t stream) {
  const U4 b = philox4x32_10((uint32_t)(idx >> 2), (uint32_t)step, ctr2(step, 0, stream), (uint32_t)(idx >> 34),
                             (uint32_t)key, (uint32_t)(key >> 32));
  return word_of(b, (uint32_t)idx & 3u);
}

SC_HD uint32_t site_byte(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t stream,
                         uint32_t part) {
  const U4 b = byte_block(key, j >> 4, sweep, colour, stream, part);
  return (word_of(b, (uint32_t)(j >> 2) & 3) >> (8 * (j & 3))) & 0xFFu;
}

SC_HD uint32_t site_low(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t stream) {
  const U4 b = low_block(key, j >> 3, sweep, colour, stream);
  return (word_of(b, (uint32_t)(j >> 1) & 3) >> (16 * (j & 1))) & 0xFFFFu;
}

SC_HD uint32_t site_word(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t stream) {
  return (site_byte(key, j, sweep, colour, stream, 0) << 24) |
         (site_byte(key, j, sweep, colour, stream, 1) << 16) | site_low(key, j, sweep, colour, stream);
}

// A decision "accept iff word < K" with K in [0, 2^32] stored as
// k8 = K >> 24 (0..256), m8 = (K >> 16) & 0xff, k16 = K & 0xffff.
struct Threshold {
  uint32_t k8;
  uint16_t m8;
  uint16_t k16;
};

SC_HD Threshold make_threshold(uint64_t k) {
  return Threshold{(uint32_t)(k >> 24), (uint16_t)((k >> 16) & 0xFFu), (uint16_t)(k & 0xFFFFu)};
}

// The low 16 bits of K as one integer (for the rare final comparison).
SC_HD uint32_t rest_of(const Threshold &t) { return ((uint32_t)t.m8 << 16) | t.k16; }

}  // namespace scamper
