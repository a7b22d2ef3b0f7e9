// Multi-spin coded Ising Metropolis with one temperature per bit, and
// replica exchange on the device.
//
// Site words are u64: bit r is replica 64b + r of bundle b (1 = spin +1,
// multi_spin.rs:28-33).  Replica r draws from its own Philox key, so every
// bit follows exactly the trajectory the single-replica kernels give that
// replica; the reference's MultiSpinIsing (multi_spin.rs:107-151) shares one
// β over the 64 bits, here each bit's β comes from its temperature label.
//
// Per thread: 8 consecutive sites of one colour class.  The anti-aligned
// neighbour count of every bit is formed bit-sliced (three u64 planes from a
// full-adder tree over the neighbour XORs); then for each bit the Philox
// bytes of its draw are compared with the bit's threshold table.  Measured
// sweeps add the new anti-aligned counts into bit-sliced vertical counters,
// transpose them across the wavefront (64 lanes x 64 bits) and popcount, so
// each lane ends with the anti-aligned bond count of one bit.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kBlock = 256;
constexpr int kSitesPerThread = 8;

__device__ __forceinline__ uint32_t wrap_dec(uint32_t v, uint32_t n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ uint32_t wrap_inc(uint32_t v, uint32_t n) { return v + 1 == n ? 0 : v + 1; }

__device__ __forceinline__ uint64_t maj(uint64_t a, uint64_t b, uint64_t c) { return (a & b) | (c & (a ^ b)); }

struct Planes {
  uint64_t b0, b1, b2;  // count = b0 + 2 b1 + 4 b2
};

// Bit-sliced count of set bits across z words (z = 2, 4 or 6).
template <int Z>
__device__ __forceinline__ Planes count_bits(const uint64_t *x) {
  Planes p;
  if (Z == 2) {
    p.b0 = x[0] ^ x[1];
    p.b1 = x[0] & x[1];
    p.b2 = 0;
  } else if (Z == 4) {
    const uint64_t s1 = x[0] ^ x[1] ^ x[2], c1 = maj(x[0], x[1], x[2]);
    p.b0 = s1 ^ x[3];
    const uint64_t k = s1 & x[3];
    p.b1 = c1 ^ k;
    p.b2 = c1 & k;
  } else {
    const uint64_t s1 = x[0] ^ x[1] ^ x[2], c1 = maj(x[0], x[1], x[2]);
    const uint64_t s2 = x[3] ^ x[4] ^ x[5], c2 = maj(x[3], x[4], x[5]);
    p.b0 = s1 ^ s2;
    const uint64_t k = s1 & s2;
    p.b1 = c1 ^ c2 ^ k;
    p.b2 = maj(c1, c2, k);
  }
  return p;
}

// z - count, bit-sliced (count <= z <= 6).
template <int Z>
__device__ __forceinline__ Planes complement(const Planes &c) {
  // subtract (b2 b1 b0) from the constant Z with borrow propagation
  constexpr uint64_t z0 = (Z & 1) ? ~0ull : 0ull, z1 = (Z & 2) ? ~0ull : 0ull, z2 = (Z & 4) ? ~0ull : 0ull;
  Planes r;
  r.b0 = z0 ^ c.b0;
  uint64_t borrow = ~z0 & c.b0;
  r.b1 = z1 ^ c.b1 ^ borrow;
  borrow = (~z1 & c.b1) | (~(z1 ^ c.b1) & borrow);
  r.b2 = z2 ^ c.b2 ^ borrow;
  return r;
}

// Wavefront transpose of a 64x64 bit matrix (lane l holds row l); after it
// lane l holds column l.  Six butterfly stages with lane exchanges.
__device__ __forceinline__ uint64_t transpose64(uint64_t x) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t masks[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                             0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int st = 0; st < 6; ++st) {
    const int s = 32 >> st;
    const uint64_t m = masks[st];
    const uint64_t y = __shfl_xor(x, s);
    if (lane & s)
      x = ((y >> s) & m) | (x & ~m);
    else
      x = (x & m) | ((y & m) << s);
  }
  return x;
}

template <int ND, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_ms_sweep(uint64_t *__restrict__ words, const Threshold *__restrict__ thr_by_label,
                                                     const int32_t *__restrict__ labels,
                                                     const uint64_t *__restrict__ keys, StencilGeom g,
                                                     uint32_t colour, uint64_t sweep,
                                                     unsigned long long *__restrict__ counts) {
  constexpr int z = 2 * ND;
  const uint32_t bundle = blockIdx.y;
  __shared__ Threshold sthr[64 * (z + 1)];
  __shared__ unsigned long long sred[2 * 64];
  for (uint32_t e = threadIdx.x; e < 64 * (z + 1); e += kBlock) {
    const uint32_t r = e / (z + 1), a = e % (z + 1);
    sthr[e] = thr_by_label[(uint32_t)labels[bundle * 64 + r] * (z + 1) + a];
  }
  if (MEAS && threadIdx.x < 128) sred[threadIdx.x] = 0;
  __syncthreads();

  const uint64_t j0 = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * kSitesPerThread;
  uint64_t *__restrict__ own = words + (2ull * bundle + colour) * g.nh;
  const uint64_t *__restrict__ oth = words + (2ull * bundle + 1 - colour) * g.nh;
  uint64_t w[kSitesPerThread], cen[kSitesPerThread];
  Planes pl[kSitesPerThread];
  bool valid[kSitesPerThread];
#pragma unroll
  for (int k = 0; k < kSitesPerThread; ++k) {
    const uint64_t j = j0 + k;
    valid[k] = j < g.nh;
    w[k] = 0;
    cen[k] = 0;
    pl[k] = Planes{0, 0, 0};
    if (!valid[k]) continue;
    const SiteCoords sc_ = site_coords(j, g);
    const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
    const uint32_t p = (colour + y + zz) & 1u;
    const uint64_t rb = (uint64_t)row * g.lx2;
    uint64_t x[z];
    w[k] = own[j];
    const uint64_t nl = oth[rb + (p ? i : wrap_dec(i, g.lx2))];
    const uint64_t nr = oth[rb + (p ? wrap_inc(i, g.lx2) : i)];
    cen[k] = oth[j];
    x[0] = w[k] ^ nl;
    x[1] = w[k] ^ nr;
    if (ND >= 2) {
      x[2] = w[k] ^ oth[((uint64_t)(wrap_dec(y, g.ly) + g.ly * zz)) * g.lx2 + i];
      x[3] = w[k] ^ oth[((uint64_t)(wrap_inc(y, g.ly) + g.ly * zz)) * g.lx2 + i];
    }
    if (ND >= 3) {
      x[4] = w[k] ^ oth[((uint64_t)(y + g.ly * wrap_dec(zz, g.lz))) * g.lx2 + i];
      x[5] = w[k] ^ oth[((uint64_t)(y + g.ly * wrap_inc(zz, g.lz))) * g.lx2 + i];
    }
    pl[k] = count_bits<z>(x);
  }

  uint64_t flip[kSitesPerThread];
#pragma unroll
  for (int k = 0; k < kSitesPerThread; ++k) flip[k] = 0;
  const uint32_t half = (uint32_t)(j0 >> 3) & 1u;  // which 8 of the block's 16 bytes
  for (uint32_t r = 0; r < 64; ++r) {
    const uint64_t key = keys[bundle * 64 + r];
    const U4 T = byte_block(key, j0 >> 4, sweep, colour, kStreamAccept, 0);
    const U4 M = byte_block(key, j0 >> 4, sweep, colour, kStreamAccept, 1);
    const uint64_t top8 = half ? ((uint64_t)T.w << 32 | T.z) : ((uint64_t)T.y << 32 | T.x);
    const uint64_t mid8 = half ? ((uint64_t)M.w << 32 | M.z) : ((uint64_t)M.y << 32 | M.x);
    const Threshold *tr = sthr + r * (z + 1);
#pragma unroll
    for (int k = 0; k < kSitesPerThread; ++k) {
      const uint32_t a = (uint32_t)((pl[k].b0 >> r) & 1u) | (uint32_t)((pl[k].b1 >> r) & 1u) << 1 |
                         (uint32_t)((pl[k].b2 >> r) & 1u) << 2;
      const Threshold t = tr[a];
      const uint32_t top = (uint32_t)(top8 >> (8 * k)) & 0xFFu, mid = (uint32_t)(mid8 >> (8 * k)) & 0xFFu;
      bool acc;
      if (top != t.k8)
        acc = top < t.k8;
      else if (mid != t.m8)
        acc = mid < t.m8;
      else
        acc = valid[k] && site_low(key, j0 + k, sweep, colour, kStreamAccept) < t.k16;
      flip[k] |= (uint64_t)acc << r;
    }
  }

  uint64_t v[6] = {0, 0, 0, 0, 0, 0};  // anti-aligned bond counts, bit-sliced
  uint64_t up[5] = {0, 0, 0, 0, 0};    // number of +1 spins, bit-sliced
#pragma unroll
  for (int k = 0; k < kSitesPerThread; ++k) {
    if (!valid[k]) continue;
    const uint64_t f = flip[k];
    own[j0 + k] = w[k] ^ f;
    if (MEAS) {
      const Planes c = complement<z>(pl[k]);
      const uint64_t n0 = (f & c.b0) | (~f & pl[k].b0), n1 = (f & c.b1) | (~f & pl[k].b1),
                     n2 = (f & c.b2) | (~f & pl[k].b2);
      uint64_t cy = v[0] & n0;
      v[0] ^= n0;
      uint64_t t1 = v[1] ^ n1 ^ cy;
      cy = maj(v[1], n1, cy);
      v[1] = t1;
      t1 = v[2] ^ n2 ^ cy;
      cy = maj(v[2], n2, cy);
      v[2] = t1;
      t1 = v[3] ^ cy;
      cy &= v[3];
      v[3] = t1;
      t1 = v[4] ^ cy;
      cy &= v[4];
      v[4] = t1;
      v[5] ^= cy;
      // +1 spins of this site (new) and of the other colour's site at j
      const uint64_t s0 = w[k] ^ f, s1 = cen[k];
      uint64_t c2 = up[0] & s0;
      up[0] ^= s0;
      uint64_t c3 = up[0] & s1;
      up[0] ^= s1;
      uint64_t t2 = up[1] ^ c2 ^ c3;
      uint64_t cc = maj(up[1], c2, c3);
      up[1] = t2;
      t2 = up[2] ^ cc;
      cc &= up[2];
      up[2] = t2;
      t2 = up[3] ^ cc;
      cc &= up[3];
      up[3] = t2;
      up[4] ^= cc;
    }
  }
  if (MEAS) {
    // lane l ends with bit l's totals over the wavefront
    unsigned long long anti = 0, plus = 0;
#pragma unroll
    for (int pidx = 0; pidx < 6; ++pidx) anti += (unsigned long long)__popcll(transpose64(v[pidx])) << pidx;
#pragma unroll
    for (int pidx = 0; pidx < 5; ++pidx) plus += (unsigned long long)__popcll(transpose64(up[pidx])) << pidx;
    const uint32_t lane = threadIdx.x & 63;
    atomicAdd(&sred[lane], anti);
    atomicAdd(&sred[64 + lane], plus);
    __syncthreads();
    if (threadIdx.x < 128) atomicAdd(&counts[(uint64_t)bundle * 128 + threadIdx.x], sred[threadIdx.x]);
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
                                                      unsigned long long *__restrict__ stats) {
  __shared__ int32_t cfg_of[1024];
  __shared__ double energy[1024];
  const uint32_t t = threadIdx.x;
  if (t < n) {
    const unsigned long long anti = counts[(uint64_t)(t / 64) * 128 + (t % 64)];
    const unsigned long long plus = counts[(uint64_t)(t / 64) * 128 + 64 + (t % 64)];
    energy[t] = -j * (n_bonds - 2.0 * (double)anti);
    if (energy_out) energy_out[t] = energy[t];
    if (mag_out) mag_out[t] = fabs(2.0 * (double)plus - n_sites) / n_sites;
    counts[(uint64_t)(t / 64) * 128 + (t % 64)] = 0;
    counts[(uint64_t)(t / 64) * 128 + 64 + (t % 64)] = 0;
    cfg_of[labels[t]] = (int32_t)t;
  }
  __syncthreads();
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
                           unsigned long long *counts, hipStream_t st) {
  const uint64_t threads = (g.nh + kSitesPerThread - 1) / kSitesPerThread;
  dim3 grid((unsigned)((threads + kBlock - 1) / kBlock), n_bundles);
#define SC_MS(ND)                                                                                                     \
  if (counts)                                                                                                         \
    hipLaunchKernelGGL((k_ms_sweep<ND, true>), grid, dim3(kBlock), 0, st, words, thr_by_label, labels, keys, g, colour, \
                       sweep, counts);                                                                                \
  else                                                                                                                \
    hipLaunchKernelGGL((k_ms_sweep<ND, false>), grid, dim3(kBlock), 0, st, words, thr_by_label, labels, keys, g,       \
                       colour, sweep, counts);
  switch (g.nd) {
    case 1: SC_MS(1) break;
    case 2: SC_MS(2) break;
    case 3: SC_MS(3) break;
    default: return hipErrorInvalidValue;
  }
#undef SC_MS
  return hipGetLastError();
}

hipError_t launch_ms_exchange(int32_t *labels, unsigned long long *counts, const double *beta_by_label, uint32_t n,
                              double n_bonds, double n_sites, double j, uint64_t ex_key, uint64_t step, int do_exchange,
                              double *energy_out, double *mag_out, unsigned long long *stats, hipStream_t st) {
  hipLaunchKernelGGL(k_ms_exchange, dim3(1), dim3(1024), 0, st, labels, counts, beta_by_label, n, n_bonds, n_sites, j,
                     ex_key, step, do_exchange, energy_out, mag_out, stats);
  return hipGetLastError();
}

}  // namespace scamper
