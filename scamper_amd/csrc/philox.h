// Counter-based random numbers shared by host and device code.
//
// Philox4x32-10 is written from Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC'11): ten rounds of two 32x32->64 products
// with the multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.
//
// Draw layout (the "injected uniforms" every kernel and the CPU checker use):
//   key   = 64-bit stream seed of (job seed, replica)
//   site  = (colour c, rank j of the site inside colour class c)
//   ctr   = { j>>3, sweep lo32, (sweep>>32 & 0xffff) | c<<16 | stream<<24, j>>35 }
//   half  = 16-bit lane (j & 7) of Philox(ctr, key); lane k is word k>>1,
//           low half for even k
//   word  = half(stream) << 16 | half(stream + 1), uniform u = word * 2^-32
// Kernels always draw the high half; the low half is drawn only when the
// high half alone cannot settle a decision (probability ~2^-16 per site).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SC_HD __host__ __device__ __forceinline__
#else
#define SC_HD inline
#endif

namespace scamper {

enum Stream : uint32_t {
  kStreamAccept = 0,      // acceptance / heat-bath word (hi), 1 = lo
  kStreamPropose = 2,     // Potts proposal word (hi), 3 = lo
  kStreamInit = 4,        // discrete random start (j = site, colour 0, sweep 0)
  kStreamCont = 5,        // continuous spins: two 32-bit words per site
  kStreamContFallback = 6,// continuous isotropic redraw (zero field)
  kStreamExchange = 7,    // replica exchange
  kStreamContInit = 8,    // continuous random start
};

struct U4 {
  uint32_t x, y, z, w;
};

SC_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

SC_HD U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                       uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = mulhi32(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = mulhi32(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}

// Third counter word: sweep bits 32..47, colour and stream.
SC_HD uint32_t ctr2(uint64_t sweep, uint32_t colour, uint32_t stream) {
  return ((uint32_t)(sweep >> 32) & 0xFFFFu) | ((colour & 0xFFu) << 16) | ((stream & 0xFFu) << 24);
}

// Philox block holding the 16-bit halves of sites 8g .. 8g+7.
SC_HD U4 group_block(uint64_t key, uint64_t g, uint64_t sweep, uint32_t colour, uint32_t stream) {
  return philox4x32_10((uint32_t)g, (uint32_t)sweep, ctr2(sweep, colour, stream),
                       (uint32_t)(g >> 32), (uint32_t)key, (uint32_t)(key >> 32));
}

SC_HD uint32_t half_of(const U4 &b, uint32_t lane) {
  const uint32_t w = (lane & 4) ? ((lane & 2) ? b.w : b.z) : ((lane & 2) ? b.y : b.x);
  return (lane & 1) ? (w >> 16) : (w & 0xFFFFu);
}

SC_HD uint32_t site_half(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour,
                         uint32_t stream) {
  return half_of(group_block(key, j >> 3, sweep, colour, stream), (uint32_t)(j & 7));
}

SC_HD uint32_t site_word(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour,
                         uint32_t stream_hi) {
  return (site_half(key, j, sweep, colour, stream_hi) << 16) |
         site_half(key, j, sweep, colour, stream_hi + 1);
}

// A decision "accept iff word < K" with K in [0, 2^32] stored as
// khi = K >> 16 (0..65536) and klo = K & 0xffff.  The high half settles the
// decision unless it equals khi.
struct Threshold {
  uint32_t khi;
  uint32_t klo;
};

SC_HD Threshold make_threshold(uint64_t k) {
  return Threshold{(uint32_t)(k >> 16), (uint32_t)(k & 0xFFFFu)};
}

}  // namespace scamper
