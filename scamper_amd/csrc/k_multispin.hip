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
// neighbour XORs); then for each of the thread's 16 bits the Philox bytes of
// its draw are compared with the bit's threshold table.  Measured sweeps add
// the new anti-aligned counts into bit-sliced vertical counters and popcount
// them per bit.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kBlock = 256;

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

// One thread per (group of 16 consecutive sites, quarter q of the 64 bits):
// the group's Philox blocks of replica 16q + r hold the 16 sites' draws, so
// no block is computed twice.  The thread keeps the 16-bit slices (bits
// 16q .. 16q+15) of the site words and of the bit-sliced anti-aligned
// counts, two sites per register, and writes only its own 16-bit slice of
// every site word.
template <int ND, bool MEAS>
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
  __shared__ uint64_t skey[64];
  __shared__ unsigned long long sred[2 * 64];
  for (uint32_t e = threadIdx.x; e < 64 * (z + 1); e += kBlock) {
    const uint32_t r = e / (z + 1), a = e % (z + 1);
    sthr[e] = thr_by_label[(uint32_t)labels[bundle * 64 + r] * (z + 1) + a];
  }
  if (threadIdx.x < 64) skey[threadIdx.x] = keys[bundle * 64 + threadIdx.x];
  if (MEAS && threadIdx.x < 128) sred[threadIdx.x] = 0;
  __syncthreads();

  const uint64_t tid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t grp = tid >> 2;
  const uint32_t q = (uint32_t)(tid & 3u), sh0 = 16u * q;
  const uint64_t j0 = grp * 16;
  uint64_t *__restrict__ own = words + (2ull * bundle + colour) * g.nh;
  const uint64_t *__restrict__ oth = words + (2ull * bundle + 1 - colour) * g.nh;

  // slices of sites 2i (low half) and 2i+1 (high half)
  uint32_t w2[8], c2[8], b0[8], b1[8], b2[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) w2[k] = c2[k] = b0[k] = b1[k] = b2[k] = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint64_t j = j0 + k;
    if (j >= g.nh) continue;
    const SiteCoords sc_ = site_coords(j, g);
    const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
    const uint32_t p = (colour + y + zz) & 1u;
    const uint64_t rb = (uint64_t)row * g.lx2;
    const uint64_t w = own[j];
    uint64_t x[z];
    x[0] = w ^ oth[rb + (p ? i : wrap_dec(i, g.lx2))];
    x[1] = w ^ oth[rb + (p ? wrap_inc(i, g.lx2) : i)];
    if (ND >= 2) {
      x[2] = w ^ oth[((uint64_t)(wrap_dec(y, g.ly) + g.ly * zz)) * g.lx2 + i];
      x[3] = w ^ oth[((uint64_t)(wrap_inc(y, g.ly) + g.ly * zz)) * g.lx2 + i];
    }
    if (ND >= 3) {
      x[4] = w ^ oth[((uint64_t)(y + g.ly * wrap_dec(zz, g.lz))) * g.lx2 + i];
      x[5] = w ^ oth[((uint64_t)(y + g.ly * wrap_inc(zz, g.lz))) * g.lx2 + i];
    }
    const Planes pl = count_bits<z>(x);
    const uint32_t hs = 16u * (uint32_t)(k & 1);
    w2[k >> 1] |= (uint32_t)((w >> sh0) & 0xFFFFu) << hs;
    b0[k >> 1] |= (uint32_t)((pl.b0 >> sh0) & 0xFFFFu) << hs;
    b1[k >> 1] |= (uint32_t)((pl.b1 >> sh0) & 0xFFFFu) << hs;
    b2[k >> 1] |= (uint32_t)((pl.b2 >> sh0) & 0xFFFFu) << hs;
    if (MEAS) c2[k >> 1] |= (uint32_t)((oth[j] >> sh0) & 0xFFFFu) << hs;
  }

  uint32_t f2[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) f2[k] = 0;
  if (j0 < g.nh) {
    for (uint32_t r = 0; r < 16; ++r) {
      const uint32_t bit = sh0 + r;
      const uint64_t key = skey[bit];
      const U4 T = byte_block(key, grp, sweep, colour, kStreamAccept, 0);
      const U4 M = byte_block(key, grp, sweep, colour, kStreamAccept, 1);
      const Threshold *tr = sthr + bit * (z + 1);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const uint32_t s = r + 16u * (uint32_t)(k & 1);
        const uint32_t a = ((b0[k >> 1] >> s) & 1u) | (((b1[k >> 1] >> s) & 1u) << 1) | (((b2[k >> 1] >> s) & 1u) << 2);
        const Threshold t = tr[a];
        const uint32_t top = (word_of(T, (uint32_t)(k >> 2)) >> (8 * (k & 3))) & 0xFFu;
        const uint32_t mid = (word_of(M, (uint32_t)(k >> 2)) >> (8 * (k & 3))) & 0xFFu;
        bool acc;
        if (top != t.k8)
          acc = top < t.k8;
        else if (mid != t.m8)
          acc = mid < t.m8;
        else
          acc = j0 + k < g.nh && site_low(key, j0 + k, sweep, colour, kStreamAccept) < t.k16;
        f2[k >> 1] |= (uint32_t)acc << s;
      }
    }
  }

  uint32_t v[6] = {0, 0, 0, 0, 0, 0};  // anti-aligned bond counts with the new spins, bit-sliced (two sites per lane)
  uint32_t up[5] = {0, 0, 0, 0, 0};    // +1 spins (own new and other colour), bit-sliced
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint64_t j = j0 + k;
    if (j >= g.nh) continue;
    const uint32_t hs = 16u * (uint32_t)(k & 1);
    const uint16_t ns = (uint16_t)(((w2[k >> 1] ^ f2[k >> 1]) >> hs) & 0xFFFFu);
    reinterpret_cast<uint16_t *>(own + j)[q] = ns;
  }
  if (MEAS) {
    constexpr uint32_t z0 = (z & 1) ? ~0u : 0u, z1 = (z & 2) ? ~0u : 0u, z2 = (z & 4) ? ~0u : 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      // sites that fall past the colour class contribute nothing: their
      // slices are zero and their planes are masked out
      const uint32_t valid = (j0 + 2 * k < g.nh ? 0xFFFFu : 0u) | (j0 + 2 * k + 1 < g.nh ? 0xFFFF0000u : 0u);
      const uint32_t f = f2[k];
      // z - count where flipped (the anti-aligned count of the new spin)
      const uint32_t c0 = z0 ^ b0[k];
      uint32_t borrow = ~z0 & b0[k];
      const uint32_t c1 = z1 ^ b1[k] ^ borrow;
      borrow = (~z1 & b1[k]) | (~(z1 ^ b1[k]) & borrow);
      const uint32_t c2b = z2 ^ b2[k] ^ borrow;
      const uint32_t n0 = ((f & c0) | (~f & b0[k])) & valid, n1 = ((f & c1) | (~f & b1[k])) & valid,
                     n2 = ((f & c2b) | (~f & b2[k])) & valid;
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
      const uint32_t s0 = (w2[k] ^ f) & valid, s1 = c2[k] & valid;
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
    // per replica of the quarter: sum of both halves, then over the lanes of
    // the wave with the same quarter, then over the block
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t r = 0; r < 16; ++r) {
      const uint32_t m = (1u << r) | (1u << (16 + r));
      unsigned long long anti = 0, plus = 0;
#pragma unroll
      for (int b = 0; b < 6; ++b) anti += (unsigned long long)__popc(v[b] & m) << b;
#pragma unroll
      for (int b = 0; b < 5; ++b) plus += (unsigned long long)__popc(up[b] & m) << b;
      for (int off = 4; off < 64; off <<= 1) {
        anti += __shfl_xor(anti, off);
        plus += __shfl_xor(plus, off);
      }
      if (lane < 4) {
        atomicAdd(&sred[16 * lane + r], anti);
        atomicAdd(&sred[64 + 16 * lane + r], plus);
      }
    }
    __syncthreads();
    if (threadIdx.x < 128 && sred[threadIdx.x] != 0)
      atomicAdd(&counts[(uint64_t)bundle * 128 + threadIdx.x], sred[threadIdx.x]);
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
                                                      const StepClock *__restrict__ clk, int measure_to_clock) {
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
                           unsigned long long *counts, hipStream_t st, const StepClock *clk) {
  const uint64_t threads = ((g.nh + 15) / 16) * 4;  // four quarters per group of 16 sites
  dim3 grid((unsigned)((threads + kBlock - 1) / kBlock), n_bundles);
#define SC_MS(ND)                                                                                                     \
  if (counts)                                                                                                         \
    hipLaunchKernelGGL((k_ms_sweep<ND, true>), grid, dim3(kBlock), 0, st, words, thr_by_label, labels, keys, g, colour, \
                       sweep, counts, clk);                                                                           \
  else                                                                                                                \
    hipLaunchKernelGGL((k_ms_sweep<ND, false>), grid, dim3(kBlock), 0, st, words, thr_by_label, labels, keys, g,       \
                       colour, sweep, counts, clk);
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
                              double *energy_out, double *mag_out, unsigned long long *stats, hipStream_t st,
                              const StepClock *clk, int measure_to_clock) {
  hipLaunchKernelGGL(k_ms_exchange, dim3(1), dim3(1024), 0, st, labels, counts, beta_by_label, n, n_bonds, n_sites, j,
                     ex_key, step, do_exchange, energy_out, mag_out, stats, clk, measure_to_clock);
  return hipGetLastError();
}

}  // namespace scamper
