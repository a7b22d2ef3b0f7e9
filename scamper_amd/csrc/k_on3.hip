// Classical Heisenberg (O(3)) spins on periodic hypercubic lattices with even
// sides: checkerboard heat bath (ContinuousHeatBathCore with
// ONModel<3>::heat_bath_sample_site, CMC.rs/src/algorithms/heat_bath.rs:
// 102-134, lattice/models.rs:617-662) and over-relaxation
// (MicrocanonicalCore, algorithms/microcanonical.rs:42-80).
//
// Spin arithmetic is FP32 with unit-norm renormalisation after every update;
// observables are accumulated in FP64 (products of FP32 components are exact
// in FP64).  Spins are stored as three component planes per colour class
// (x, y, z of rank j at [c][comp][j]) so every load is a coalesced dword.
//
// The inverse-CDF draw of cos(theta) is written in the forms
//   1 - cos = -log1p((1 - u) * expm1(-2 kappa)) / kappa,
//   1 + cos =  log1p(u * expm1(2 kappa)) / kappa,
// algebraically the reference's 1 + ln(u + (1 - u) e^{-2 kappa}) / kappa but
// without its FP32 cancellation at small kappa and in the tails, and
// sin(theta) = sqrt((1 - cos)(1 + cos)) without that of 1 - cos^2.  Each
// draw evaluates one logarithm, whichever form it takes, so a wave does not
// run both forms when its lanes differ.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t wrap_dec(uint32_t v, uint32_t n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ uint32_t wrap_inc(uint32_t v, uint32_t n) { return v + 1 == n ? 0 : v + 1; }

struct V3 {
  float x, y, z;
};

// sin and cos of 2 pi t for t in [0, 1): the hardware's v_sin_f32 /
// v_cos_f32 take their argument in revolutions, so no range reduction.
__device__ __forceinline__ void sincos_turns(float t, float *s, float *c) {
  *s = __builtin_amdgcn_sinf(t);
  *c = __builtin_amdgcn_cosf(t);
}

__device__ __forceinline__ V3 load3(const float *base, uint64_t nh, uint64_t j) {
  return V3{base[j], base[nh + j], base[2 * nh + j]};
}

__device__ __forceinline__ void store3(float *base, uint64_t nh, uint64_t j, V3 v) {
  base[j] = v.x;
  base[nh + j] = v.y;
  base[2 * nh + j] = v.z;
}

__device__ __forceinline__ V3 normalised(V3 v) {
  const float n2 = v.x * v.x + v.y * v.y + v.z * v.z;
  if (!(n2 > 1e-30f)) return v;
  const float inv = rsqrtf(n2);
  return V3{v.x * inv, v.y * inv, v.z * inv};
}

// Box-Muller pairs from four words, normalised (random_unit_vector,
// models.rs:405-422).
__device__ __forceinline__ V3 random_unit(const U4 &w) {
  const float two_m32 = 2.3283064365386963e-10f;
  const float u1 = fmaxf((float)w.x * two_m32, 1.17549435e-38f), u2 = (float)w.y * two_m32;
  const float u3 = fmaxf((float)w.z * two_m32, 1.17549435e-38f), u4 = (float)w.w * two_m32;
  const float r1 = sqrtf(-2.0f * logf(u1)), r2 = sqrtf(-2.0f * logf(u3));
  float s1, c1, s2, c2;
  sincos_turns(u2, &s1, &c1);
  sincos_turns(u4, &s2, &c2);
  return normalised(V3{r1 * c1, r1 * s1, r2 * c2});
}

template <int ND>
__device__ __forceinline__ V3 local_field(const float *__restrict__ oth, const StencilGeom &g, uint64_t j, uint32_t row,
                                          uint32_t i, uint32_t y, uint32_t zz, uint32_t p, float jc, V3 *centre) {
  const uint64_t rb = (uint64_t)row * g.lx2;
  const uint64_t il = rb + (p ? i : wrap_dec(i, g.lx2)), ir = rb + (p ? wrap_inc(i, g.lx2) : i);
  uint64_t nb[2 * ND];
  nb[0] = il;
  nb[1] = ir;
  if (ND >= 2) {
    nb[2] = ((uint64_t)(wrap_dec(y, g.ly) + g.ly * zz)) * g.lx2 + i;
    nb[3] = ((uint64_t)(wrap_inc(y, g.ly) + g.ly * zz)) * g.lx2 + i;
  }
  if (ND >= 3) {
    nb[4] = ((uint64_t)(y + g.ly * wrap_dec(zz, g.lz))) * g.lx2 + i;
    nb[5] = ((uint64_t)(y + g.ly * wrap_inc(zz, g.lz))) * g.lx2 + i;
  }
  V3 h{0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 2 * ND; ++k) {
    const V3 s = load3(oth, g.nh, nb[k]);
    h.x += s.x;
    h.y += s.y;
    h.z += s.z;
    if (nb[k] == j) *centre = s;
  }
  return V3{jc * h.x, jc * h.y, jc * h.z};
}

// ALG: kContinuousHeatBath or kOverRelaxation.  MEAS accumulates, over the
// colour-1 sites after their update, E = -J sum s.h and the magnetisation
// of both colours (each other-colour site is the centre neighbour of
// exactly one colour-1 site).
// One heat-bath draw (heat_bath_sample_site, models.rs:617-662): cos(theta)
// about the local field h from the word w0, the azimuth from w1, rotated
// into h's frame.  A vanishing field draws a uniform direction from a block
// of its own (stream kStreamContFallback, by rank).
//
// The arithmetic uses the hardware's exp2, log2, reciprocal and square-root
// instructions (about 1 ulp each): log1p and expm1 are formed from log and
// exp with the correction of Goldberg and Kahan (f(w) z / (w - 1), which keeps
// the relative accuracy of f at small arguments), divisions are products with
// a reciprocal, and square roots of sums of squares come from one reciprocal
// square root.  FP32 spins are renormalised after every update, and the
// error against the FP64 oracle stays at the level of the FP32 arithmetic.
__device__ __forceinline__ float fast_log1p(float z) {  // z > -1
  const float w = 1.0f + z;
  if (w == 1.0f) return z;
  const float l = __builtin_amdgcn_logf(w) * 0.69314718f;  // ln w
  return w == __builtin_huge_valf() ? w : l * (z * __builtin_amdgcn_rcpf(w - 1.0f));
}

__device__ __forceinline__ float fast_expm1(float t) {  // t >= 0
  const float e = __builtin_amdgcn_exp2f(t * 1.44269504f);  // e^t
  if (e == 1.0f) return t;
  if (e == __builtin_huge_valf()) return e;
  return (e - 1.0f) * (t * __builtin_amdgcn_rcpf(__builtin_amdgcn_logf(e) * 0.69314718f));
}

__device__ __forceinline__ V3 unit_fast(V3 v) {
  const float n2 = v.x * v.x + v.y * v.y + v.z * v.z;
  if (!(n2 > 1e-30f)) return v;
  const float inv = __builtin_amdgcn_rsqf(n2);
  return V3{v.x * inv, v.y * inv, v.z * inv};
}

__device__ __forceinline__ V3 heat_bath_site(V3 h, const ReplicaParams &P, uint64_t j, uint64_t sweep,
                                             uint32_t colour, uint32_t w0, uint32_t w1) {
  const float beta = (float)P.beta;
  const float n2 = h.x * h.x + h.y * h.y + h.z * h.z;
  const float inv_n = n2 > 0.0f ? __builtin_amdgcn_rsqf(n2) : 0.0f;
  const float norm = n2 * inv_n;
  const float kappa = beta * norm;
  if (kappa < 1e-12f)
    return random_unit(philox4x32_10((uint32_t)j, (uint32_t)sweep, ctr2(sweep, colour, kStreamContFallback),
                                     (uint32_t)(j >> 32), (uint32_t)P.key, (uint32_t)(P.key >> 32)));
  const float u = fmaxf((float)w0 * 2.3283064365386963e-10f, 1.17549435e-38f);
  // With E = e^{2 kappa} - 1 and p = u + (1 - u) e^{-2 kappa} = (1 + u E) / (1 + E),
  // cos(theta) = 1 + ln(p) / kappa.  The small one of a = 1 - cos and
  // b = 1 + cos is formed, so cos(theta) and sin(theta) = sqrt(a b) keep
  // their relative precision at both ends:
  //   cos <  0:             b = log1p(u E) / kappa;
  //   cos >= 0, p >= 1/2:   a = -log1p((1 - u) expm1(-2 kappa)) / kappa;
  //   cos >= 0, p <  1/2:   a = 2 - log1p(u E) / kappa (the argument above
  //                         would be near -1, where it cancels; a > ln 2 / kappa
  //                         here, so the difference does not).
  // cos < 0 exactly when u < 1 / (1 + e^kappa); the choice needs no
  // precision, both sides being accurate near cos = 0.  Every lane takes one
  // log1p.
  const float E = fast_expm1(2.0f * kappa);  // inf from kappa ~ 44.4
  const float ik = __builtin_amdgcn_rcpf(kappa);
  const bool tail = u * (1.0f + __expf(kappa)) < 1.0f;
  const float em = kappa < 40.0f ? -E * __builtin_amdgcn_rcpf(1.0f + E) : -1.0f;  // expm1(-2 kappa)
  const float y = (1.0f - u) * em;
  const bool via_e = tail || y < -0.5f;
  const float l = fast_log1p(via_e ? u * E : y) * ik;
  float x = tail ? l : (via_e ? 2.0f - l : -l);  // the small one of a and b
  if (via_e && !(E <= 3.4028235e38f)) {
    // e^{2 kappa} beyond FP32 range (kappa > 44.3): ln p formed directly;
    // u dominates p here unless u is zero's stand-in
    const float lp = logf(u + (1.0f - u) * expf(-2.0f * kappa)) * ik;
    x = tail ? 2.0f + lp : -lp;
  }
  x = fminf(fmaxf(x, 0.0f), 2.0f);
  const float c = tail ? x - 1.0f : 1.0f - x;
  const float sn = __builtin_amdgcn_sqrtf(x * (2.0f - x));
  float sphi, cphi;
  sincos_turns((float)w1 * 2.3283064365386963e-10f, &sphi, &cphi);
  const float lx = sn * cphi, ly = sn * sphi;
  const float ux = h.x * inv_n, uy = h.y * inv_n, uz = h.z * inv_n;
  const float t2 = ux * ux + uy * uy;
  if (t2 < 1e-12f) return unit_fast(V3{lx, ly, uz >= 0.0f ? c : -c});
  const float it = __builtin_amdgcn_rsqf(t2), tr = t2 * it;
  return unit_fast(V3{-uy * it * lx - ux * uz * it * ly + ux * c, ux * it * lx - uy * uz * it * ly + uy * c,
                      tr * ly + uz * c});
}

// Over-relaxation (MicrocanonicalCore, microcanonical.rs:54-75): s reflected
// about h, unchanged where h vanishes.
__device__ __forceinline__ V3 reflect_site(V3 s, V3 h) {
  const float nsq = h.x * h.x + h.y * h.y + h.z * h.z;
  if (!(nsq >= 1e-28f)) return s;
  const float proj = (s.x * h.x + s.y * h.y + s.z * h.z) / nsq;
  return normalised(V3{2.0f * proj * h.x - s.x, 2.0f * proj * h.y - s.y, 2.0f * proj * h.z - s.z});
}

// The grid is one-dimensional over (replica, block of ranks), nbx blocks per
// replica.  Hardware hands consecutive blocks to the eight XCDs in turn, so
// block b works on item (b % 8) (total / 8) + b / 8 when 8 | total: each XCD
// sweeps one contiguous eighth, and the neighbour rows and planes a block
// reads sit in its own XCD's L2.
// (A grid too large for one dimension keeps (block, replica) in x and y.)
__device__ __forceinline__ void xcd_item(uint32_t nbx, uint32_t *rep, uint32_t *bx) {
  if (gridDim.y > 1) {
    *rep = blockIdx.y;
    *bx = blockIdx.x;
    return;
  }
  const uint32_t total = gridDim.x, b = blockIdx.x;
  const uint32_t item = total % 8 == 0 ? (b % 8) * (total / 8) + b / 8 : b;
  *rep = item / nbx;
  *bx = item - *rep * nbx;
}

template <int ND, int ALG, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_on3(float *__restrict__ spins, const ReplicaParams *__restrict__ reps,
                                                StencilGeom g, uint32_t nbx, uint32_t colour, uint64_t sweep,
                                                float jc, double *__restrict__ em, const StepClock *__restrict__ clk) {
  if (clk) sweep = clk->sweep;
  uint32_t rep, bx;
  xcd_item(nbx, &rep, &bx);
  const uint64_t j = (uint64_t)bx * kBlock + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (j < g.nh) {
    float *__restrict__ own = spins + (2ull * rep + colour) * 3ull * g.nh;
    const float *__restrict__ oth = spins + (2ull * rep + 1 - colour) * 3ull * g.nh;
    const SiteCoords sc_ = site_coords(j, g);
    const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
    const uint32_t p = (colour + y + zz) & 1u;
    V3 centre{0.f, 0.f, 0.f};
    const V3 h = local_field<ND>(oth, g, j, row, i, y, zz, p, jc, &centre);
    const V3 s = load3(own, g.nh, j);
    V3 ns;
    if (ALG == kContinuousHeatBath) {
      const ReplicaParams &P = reps[rep];
      const U4 blk = philox4x32_10((uint32_t)(j >> 1), (uint32_t)sweep, ctr2(sweep, colour, kStreamCont),
                                   (uint32_t)(j >> 33), (uint32_t)P.key, (uint32_t)(P.key >> 32));
      ns = heat_bath_site(h, P, j, sweep, colour, (j & 1) ? blk.z : blk.x, (j & 1) ? blk.w : blk.y);
    } else {
      ns = reflect_site(s, h);
    }
    store3(own, g.nh, j, ns);
    if (MEAS) {
      acc[0] = (double)ns.x * (double)h.x + (double)ns.y * (double)h.y + (double)ns.z * (double)h.z;
      acc[1] = (double)ns.x + (double)centre.x;
      acc[2] = (double)ns.y + (double)centre.y;
      acc[3] = (double)ns.z + (double)centre.z;
    }
  }
  if (MEAS) {
    __shared__ double red[4];
    if (threadIdx.x < 4) red[threadIdx.x] = 0.0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double v = acc[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if ((threadIdx.x & 63) == 0) atomicAdd(&red[k], v);
    }
    __syncthreads();
    if (threadIdx.x < 4) atomicAdd(&em[(uint64_t)rep * kEmStride + threadIdx.x], red[threadIdx.x]);
  }
}

// Two ranks per thread, j = 2t and j + 1, when Lx/2 is even (so both lie in
// one row and every pair is 8-byte aligned): one Philox block (j >> 1) serves
// both heat-bath draws, each component of a neighbour row is one float2
// load, and the left/right neighbours are that row's pair plus one more rank
// (i + 2 when the row's parity p = 1, else i - 1).
template <int ND, int ALG, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_on3_pair(float *__restrict__ spins, const ReplicaParams *__restrict__ reps,
                                                     StencilGeom g, uint32_t nbx, uint32_t colour, uint64_t sweep,
                                                     float jc, double *__restrict__ em, const StepClock *__restrict__ clk) {
  if (clk) sweep = clk->sweep;
  uint32_t rep, bx;
  xcd_item(nbx, &rep, &bx);
  const uint64_t t = (uint64_t)bx * kBlock + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (t < g.nh / 2) {
    const uint64_t j = 2 * t;
    float *__restrict__ own = spins + (2ull * rep + colour) * 3ull * g.nh;
    const float *__restrict__ oth = spins + (2ull * rep + 1 - colour) * 3ull * g.nh;
    const SiteCoords c = site_coords(j, g);
    const uint32_t p = (colour + c.y + c.z) & 1u;
    const uint64_t rb = (uint64_t)c.row * g.lx2;
    const uint64_t ex = rb + (p ? wrap_inc(c.i + 1, g.lx2) : wrap_dec(c.i, g.lx2));
    uint64_t nb[2 * ND > 2 ? 2 * ND - 2 : 1];  // first ranks of the neighbour rows' pairs: up, down, back, front
    if (ND >= 2) {
      nb[0] = ((uint64_t)(wrap_dec(c.y, g.ly) + g.ly * c.z)) * g.lx2 + c.i;
      nb[1] = ((uint64_t)(wrap_inc(c.y, g.ly) + g.ly * c.z)) * g.lx2 + c.i;
    }
    if (ND >= 3) {
      nb[2] = ((uint64_t)(c.y + g.ly * wrap_dec(c.z, g.lz))) * g.lx2 + c.i;
      nb[3] = ((uint64_t)(c.y + g.ly * wrap_inc(c.z, g.lz))) * g.lx2 + c.i;
    }
    float ha[3], hb[3], ca[3], cb[3], sa[3], sb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float *base = oth + (uint64_t)k * g.nh;
      const float2 mid = *reinterpret_cast<const float2 *>(base + j);
      const float e = base[ex];
      // left then right, as nb[0], nb[1] of the one-rank kernel
      float a = p ? mid.x : e, b = p ? mid.y : mid.x;
      a += p ? mid.y : mid.x;
      b += p ? e : mid.y;
#pragma unroll
      for (int n = 0; n < 2 * ND - 2; ++n) {
        const float2 v = *reinterpret_cast<const float2 *>(base + nb[n]);
        a += v.x;
        b += v.y;
      }
      ha[k] = jc * a;
      hb[k] = jc * b;
      ca[k] = mid.x;
      cb[k] = mid.y;
      const float2 sv = *reinterpret_cast<const float2 *>(own + (uint64_t)k * g.nh + j);
      sa[k] = sv.x;
      sb[k] = sv.y;
    }
    const V3 hA{ha[0], ha[1], ha[2]}, hB{hb[0], hb[1], hb[2]};
    V3 nA, nB;
    if (ALG == kContinuousHeatBath) {
      const ReplicaParams &P = reps[rep];
      const U4 blk = philox4x32_10((uint32_t)t, (uint32_t)sweep, ctr2(sweep, colour, kStreamCont), (uint32_t)(t >> 32),
                                   (uint32_t)P.key, (uint32_t)(P.key >> 32));
      nA = heat_bath_site(hA, P, j, sweep, colour, blk.x, blk.y);
      nB = heat_bath_site(hB, P, j + 1, sweep, colour, blk.z, blk.w);
    } else {
      nA = reflect_site(V3{sa[0], sa[1], sa[2]}, hA);
      nB = reflect_site(V3{sb[0], sb[1], sb[2]}, hB);
    }
    *reinterpret_cast<float2 *>(own + j) = make_float2(nA.x, nB.x);
    *reinterpret_cast<float2 *>(own + g.nh + j) = make_float2(nA.y, nB.y);
    *reinterpret_cast<float2 *>(own + 2 * g.nh + j) = make_float2(nA.z, nB.z);
    if (MEAS) {
      acc[0] = (double)nA.x * (double)hA.x + (double)nA.y * (double)hA.y + (double)nA.z * (double)hA.z +
               ((double)nB.x * (double)hB.x + (double)nB.y * (double)hB.y + (double)nB.z * (double)hB.z);
      acc[1] = ((double)nA.x + (double)ca[0]) + ((double)nB.x + (double)cb[0]);
      acc[2] = ((double)nA.y + (double)ca[1]) + ((double)nB.y + (double)cb[1]);
      acc[3] = ((double)nA.z + (double)ca[2]) + ((double)nB.z + (double)cb[2]);
    }
  }
  if (MEAS) {
    __shared__ double red[4];
    if (threadIdx.x < 4) red[threadIdx.x] = 0.0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double v = acc[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if ((threadIdx.x & 63) == 0) atomicAdd(&red[k], v);
    }
    __syncthreads();
    if (threadIdx.x < 4) atomicAdd(&em[(uint64_t)rep * kEmStride + threadIdx.x], red[threadIdx.x]);
  }
}

template <int ND>
__global__ __launch_bounds__(kBlock) void k_on3_measure(const float *__restrict__ spins, StencilGeom g, float jc,
                                                        double *__restrict__ em) {
  const uint32_t rep = blockIdx.y;
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (j < g.nh) {
    const float *own = spins + (2ull * rep + 1) * 3ull * g.nh;
    const float *oth = spins + (2ull * rep) * 3ull * g.nh;
    const SiteCoords sc_ = site_coords(j, g);
    const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
    V3 centre{0.f, 0.f, 0.f};
    const V3 h = local_field<ND>(oth, g, j, row, i, y, zz, (1 + y + zz) & 1u, jc, &centre);
    const V3 s = load3(own, g.nh, j);
    acc[0] = (double)s.x * (double)h.x + (double)s.y * (double)h.y + (double)s.z * (double)h.z;
    acc[1] = (double)s.x + (double)centre.x;
    acc[2] = (double)s.y + (double)centre.y;
    acc[3] = (double)s.z + (double)centre.z;
  }
  __shared__ double red[4];
  if (threadIdx.x < 4) red[threadIdx.x] = 0.0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&red[k], v);
  }
  __syncthreads();
  if (threadIdx.x < 4) atomicAdd(&em[(uint64_t)rep * kEmStride + threadIdx.x], red[threadIdx.x]);
}

__global__ __launch_bounds__(kBlock) void k_on3_init(float *__restrict__ spins, StencilGeom g, uint64_t key,
                                                     uint32_t rep, int random) {
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint32_t colour = blockIdx.y;
  if (j >= g.nh) return;
  const SiteCoords sc_ = site_coords(j, g);
  const uint32_t row = sc_.row, i = sc_.i, y = sc_.y, zz = sc_.z;
  const uint64_t site = 2ull * i + ((colour + y + zz) & 1u) + 2ull * g.lx2 * row;
  V3 v{1.f, 0.f, 0.f};  // ordered_spin (models.rs:479-484)
  if (random)
    v = random_unit(philox4x32_10((uint32_t)site, 0u, ctr2(0, 0, kStreamContInit), (uint32_t)(site >> 32),
                                  (uint32_t)key, (uint32_t)(key >> 32)));
  store3(spins + (2ull * rep + colour) * 3ull * g.nh, g.nh, j, v);
}

// ------------------------------------------------------------ O(2) (XY)
// Two component planes per colour ([c][x, y][rank]).  The heat bath draws the
// angle about the local field from the Best-Fisher von Mises sampler
// (ContinuousHeatBathable for ONModel<2>, models.rs:596-614, and
// sample_von_mises, models.rs:666-689) in f64, so that its rejection
// branches follow the f64 rule: try k of rank j takes Philox block (j, k) of
// stream kStreamVonMises (word 0 the angle pi u, word 1 the acceptance
// uniform, word 2 the sign of the accepted try); kappa < 1e-8 takes 2 pi u
// from word 0 of try 0; a 64th rejection in a row keeps that try's value.
// The new spin (cos theta, sin theta), theta = a + atan2(hy, hx), is formed
// as the field's unit direction rotated by a, from cos a and sin a, which
// equals the angle form to f64 rounding and costs no acos, atan2, cos or sin.

struct V2 {
  float x, y;
};

__device__ __forceinline__ V2 load2(const float *base, uint64_t nh, uint64_t j) { return V2{base[j], base[nh + j]}; }
__device__ __forceinline__ void store2(float *base, uint64_t nh, uint64_t j, V2 v) {
  base[j] = v.x;
  base[nh + j] = v.y;
}

constexpr double kTwoM32 = 2.3283064365386963e-10, kTau = 6.283185307179586, kPi = 3.141592653589793;

__device__ __forceinline__ U4 vm_block(const ReplicaParams &P, uint64_t j, uint64_t sweep, uint32_t colour,
                                       uint32_t k) {
  return philox4x32_10((uint32_t)j, (uint32_t)sweep, ctr2(sweep, colour, kStreamVonMises),
                       (uint32_t)(j >> 32) | (k << 20), (uint32_t)P.key, (uint32_t)(P.key >> 32));
}

// The sampler's second test, ln(c / u) + 1 >= c, decided on the hardware's
// FP32 log2 where its result is certain and in f64 elsewhere.  With c and u
// rounded to FP32 (both normal; relative error 2^-24 each) and v_log_f32
// within 2 ulp of log2, ln c - ln u is within 2.4e-7 (|ln c| + |ln u|) +
// 3e-7 of ln(c / u); a margin above 1e-6 (1 + |ln c| + |ln u|) gives the f64
// answer, and below it the f64 path decides.
__device__ __forceinline__ bool log_test(double c, double u) {
  const float cf = (float)c, uf = (float)u;
  if (cf >= 1.17549435e-38f && uf >= 1.17549435e-38f && cf <= 3.0e38f) {
    const float lc = __builtin_amdgcn_logf(cf) * 0.69314718f, lu = __builtin_amdgcn_logf(uf) * 0.69314718f;
    const double d = (double)lc - (double)lu + 1.0 - c;
    if (fabs(d) > 1e-6 * (1.0 + (double)fabsf(lc) + (double)fabsf(lu))) return d > 0.0;
  }
  return log(c / u) + 1.0 >= c;
}

// The von Mises draw as (cos a, sin a): the accepted try's value is cos a
// itself, and sin a = +-sqrt((1 - v)(1 + v)) with the sign word's sign, so
// the angle a = +-acos(v) is never formed; kappa < 1e-8 gives a = 2 pi u.
__device__ void von_mises_cs(const ReplicaParams &P, uint64_t j, uint64_t sweep, uint32_t colour, double kappa,
                             double *ca, double *sa) {
  if (kappa < 1e-8) {
    sincos((double)vm_block(P, j, sweep, colour, 0).x * kTwoM32 * kTau, sa, ca);
    return;
  }
  const double tau = 1.0 + sqrt(1.0 + 4.0 * kappa * kappa);
  const double rho = (tau - sqrt(2.0 * tau)) / (2.0 * kappa);
  const double r = (1.0 + rho * rho) / (2.0 * rho);
  double value = 0.0;
  uint32_t sign = 0;
  for (uint32_t k = 0; k < 64; ++k) {
    const U4 w = vm_block(P, j, sweep, colour, k);
    const double z = cos(kPi * ((double)w.x * kTwoM32));
    value = (1.0 + r * z) / (r + z);
    const double c = kappa * (r - value);
    const double u = fmax((double)w.y * kTwoM32, 2.2250738585072014e-308);
    sign = w.z;
    if (c * (2.0 - c) > u || log_test(c, u)) break;
  }
  const double v = fmin(fmax(value, -1.0), 1.0);
  const double sn = sqrt((1.0 - v) * (1.0 + v));
  *ca = v;
  *sa = sign < 0x80000000u ? sn : -sn;
}

// ONModel<2>::random_unit_vector (models.rs:405-422) in f64 from two words.
__device__ __forceinline__ void random_unit2(uint32_t w0, uint32_t w1, double *out) {
  const double u1 = fmax((double)w0 * kTwoM32, 2.2250738585072014e-308), u2 = (double)w1 * kTwoM32;
  const double radius = sqrt(-2.0 * log(u1)), angle = kTau * u2;
  const double x = radius * cos(angle), y = radius * sin(angle);
  const double norm = sqrt(x * x + y * y);
  out[0] = norm > 1e-15 ? x / norm : x;
  out[1] = norm > 1e-15 ? y / norm : y;
}

template <int ND, int ALG, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_on2(float *__restrict__ spins, const ReplicaParams *__restrict__ reps,
                                                StencilGeom g, uint32_t nbx, uint32_t colour, uint64_t sweep,
                                                float jc, double *__restrict__ em, const StepClock *__restrict__ clk) {
  static_assert(ALG == kOverRelaxation, "the O(2) heat bath is k_on2_hb");
  if (clk) sweep = clk->sweep;
  uint32_t rep, bx;
  xcd_item(nbx, &rep, &bx);
  const uint64_t j = (uint64_t)bx * kBlock + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (j < g.nh) {
    float *__restrict__ own = spins + (2ull * rep + colour) * 2ull * g.nh;
    const float *__restrict__ oth = spins + (2ull * rep + 1 - colour) * 2ull * g.nh;
    const SiteCoords c = site_coords(j, g);
    const uint32_t p = (colour + c.y + c.z) & 1u;
    const uint64_t rb = (uint64_t)c.row * g.lx2;
    uint64_t nb[2 * ND];
    nb[0] = rb + (p ? c.i : wrap_dec(c.i, g.lx2));
    nb[1] = rb + (p ? wrap_inc(c.i, g.lx2) : c.i);
    if (ND >= 2) {
      nb[2] = ((uint64_t)(wrap_dec(c.y, g.ly) + g.ly * c.z)) * g.lx2 + c.i;
      nb[3] = ((uint64_t)(wrap_inc(c.y, g.ly) + g.ly * c.z)) * g.lx2 + c.i;
    }
    if (ND >= 3) {
      nb[4] = ((uint64_t)(c.y + g.ly * wrap_dec(c.z, g.lz))) * g.lx2 + c.i;
      nb[5] = ((uint64_t)(c.y + g.ly * wrap_inc(c.z, g.lz))) * g.lx2 + c.i;
    }
    // local field in f64 (models.rs:579-595), the centre neighbour kept for M
    double hx = 0.0, hy = 0.0;
    V2 centre{0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 2 * ND; ++k) {
      const V2 v = load2(oth, g.nh, nb[k]);
      hx += (double)jc * (double)v.x;
      hy += (double)jc * (double)v.y;
      if (nb[k] == j) centre = v;
    }
    const V2 s = load2(own, g.nh, j);
    V2 ns = s;
    // over-relaxation (microcanonical.rs:54-75): s reflected about h
    const float fx = (float)hx, fy = (float)hy;
    const float nsq = fx * fx + fy * fy;
    if (nsq >= 1e-28f) {
      const float proj = (s.x * fx + s.y * fy) / nsq;
      float rx = 2.0f * proj * fx - s.x, ry = 2.0f * proj * fy - s.y;
      const float inv = rsqrtf(rx * rx + ry * ry);
      ns = V2{rx * inv, ry * inv};
    }
    store2(own, g.nh, j, ns);
    if (MEAS) {
      acc[0] = (double)ns.x * hx + (double)ns.y * hy;
      acc[1] = (double)ns.x + (double)centre.x;
      acc[2] = (double)ns.y + (double)centre.y;
    }
  }
  if (MEAS) {
    __shared__ double red[4];
    if (threadIdx.x < 4) red[threadIdx.x] = 0.0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double v = acc[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if ((threadIdx.x & 63) == 0) atomicAdd(&red[k], v);
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicAdd(&em[(uint64_t)rep * kEmStride + threadIdx.x], red[threadIdx.x]);
  }
}

// The O(2) heat bath with four sites per lane.  A wave of the one-site
// kernel runs the Best-Fisher loop until its slowest lane accepts.  Here a
// lane owns four sites (ranks s * kBlock + lane of the block's 4 * kBlock,
// s = 0 .. 3): it forms their fields and sampler constants first, then
// takes one try per iteration and moves to its next pending site as soon
// as one accepts, writing that site's new spin there, so a wave runs about
// the sum of its lanes' tries over four sites instead of four maxima.
// Every site's arithmetic (its field, the sampler constants, try k's draws
// and tests, the rotation) is the one-site kernel's, in the same order: the
// spins are the same bit for bit.
constexpr int kHbSites = 4;

// Local field of rank j of the colour `colour` in f64 (models.rs:579-595),
// and the other colour's spin at rank j (for M).
template <int ND>
__device__ __forceinline__ void on2_field(const float *__restrict__ oth, const StencilGeom &g, uint32_t colour,
                                          uint64_t j, float jc, double *hx, double *hy, V2 *centre) {
  const SiteCoords c = site_coords(j, g);
  const uint32_t p = (colour + c.y + c.z) & 1u;
  const uint64_t rb = (uint64_t)c.row * g.lx2;
  uint64_t nb[2 * ND];
  nb[0] = rb + (p ? c.i : wrap_dec(c.i, g.lx2));
  nb[1] = rb + (p ? wrap_inc(c.i, g.lx2) : c.i);
  if (ND >= 2) {
    nb[2] = ((uint64_t)(wrap_dec(c.y, g.ly) + g.ly * c.z)) * g.lx2 + c.i;
    nb[3] = ((uint64_t)(wrap_inc(c.y, g.ly) + g.ly * c.z)) * g.lx2 + c.i;
  }
  if (ND >= 3) {
    nb[4] = ((uint64_t)(c.y + g.ly * wrap_dec(c.z, g.lz))) * g.lx2 + c.i;
    nb[5] = ((uint64_t)(c.y + g.ly * wrap_inc(c.z, g.lz))) * g.lx2 + c.i;
  }
  double x = 0.0, y = 0.0;
  V2 ce{0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 2 * ND; ++k) {
    const V2 v = load2(oth, g.nh, nb[k]);
    x += (double)jc * (double)v.x;
    y += (double)jc * (double)v.y;
    if (nb[k] == j) ce = v;
  }
  *hx = x;
  *hy = y;
  *centre = ce;
}

template <int ND, bool MEAS>
__global__ __launch_bounds__(kBlock) void k_on2_hb(float *__restrict__ spins, const ReplicaParams *__restrict__ reps,
                                                   StencilGeom g, uint32_t nbx, uint32_t colour, uint64_t sweep,
                                                   float jc, double *__restrict__ em, const StepClock *__restrict__ clk) {
  if (clk) sweep = clk->sweep;
  uint32_t rep, bx;
  xcd_item(nbx, &rep, &bx);
  // per site of the lane: kappa and r of the sampler, and the field's unit
  // direction (ux, uy)
  __shared__ double skr[kHbSites][2][kBlock];
  __shared__ double sdir[kHbSites][2][kBlock];
  const ReplicaParams P = reps[rep];
  float *__restrict__ own = spins + (2ull * rep + colour) * 2ull * g.nh;
  const float *__restrict__ oth = spins + (2ull * rep + 1 - colour) * 2ull * g.nh;
  const uint64_t jb = (uint64_t)bx * (kBlock * kHbSites) + threadIdx.x;

  uint32_t pending = 0;  // bit s: site s still draws
#pragma unroll
  for (int s = 0; s < kHbSites; ++s) {
    const uint64_t j = jb + (uint64_t)s * kBlock;
    if (j >= g.nh) continue;
    double hx, hy;
    V2 ce;
    on2_field<ND>(oth, g, colour, j, jc, &hx, &hy, &ce);
    const double norm = sqrt(hx * hx + hy * hy);
    if (norm < 1e-15) {
      const U4 fb = philox4x32_10((uint32_t)j, (uint32_t)sweep, ctr2(sweep, colour, kStreamContFallback),
                                  (uint32_t)(j >> 32), (uint32_t)P.key, (uint32_t)(P.key >> 32));
      double out[2];
      random_unit2(fb.x, fb.y, out);
      store2(own, g.nh, j, V2{(float)out[0], (float)out[1]});
      continue;
    }
    const double kappa = P.beta * norm;
    const double ux = hx / norm, uy = hy / norm;
    if (kappa < 1e-8) {
      double ca, sa;
      sincos((double)vm_block(P, j, sweep, colour, 0).x * kTwoM32 * kTau, &sa, &ca);
      store2(own, g.nh, j, V2{(float)(ca * ux - sa * uy), (float)(sa * ux + ca * uy)});
      continue;
    }
    const double tau = 1.0 + sqrt(1.0 + 4.0 * kappa * kappa);
    const double rho = (tau - sqrt(2.0 * tau)) / (2.0 * kappa);
    skr[s][0][threadIdx.x] = kappa;
    skr[s][1][threadIdx.x] = (1.0 + rho * rho) / (2.0 * rho);
    sdir[s][0][threadIdx.x] = ux;
    sdir[s][1][threadIdx.x] = uy;
    pending |= 1u << s;
  }

  // Best-Fisher tries (sample_von_mises, models.rs:666-689): try k of rank j
  // takes Philox block (j, k); the 64th try is kept whatever its test says
  uint32_t s = pending ? (uint32_t)__builtin_ctz(pending) : (uint32_t)kHbSites, k = 0;
  double kappa = s < kHbSites ? skr[s][0][threadIdx.x] : 0.0, r = s < kHbSites ? skr[s][1][threadIdx.x] : 0.0;
  while (__any(s < kHbSites)) {
    if (s < kHbSites) {
      const uint64_t j = jb + (uint64_t)s * kBlock;
      const U4 w = vm_block(P, j, sweep, colour, k);
      const double zc = cos(kPi * ((double)w.x * kTwoM32));
      const double value = (1.0 + r * zc) / (r + zc);
      const double c = kappa * (r - value);
      const double u = fmax((double)w.y * kTwoM32, 2.2250738585072014e-308);
      if (c * (2.0 - c) > u || log_test(c, u) || k == 63) {
        // theta = a + atan2(hy, hx) as the field's direction rotated by a
        const double ca = fmin(fmax(value, -1.0), 1.0);
        const double sn = sqrt((1.0 - ca) * (1.0 + ca));
        const double sa = w.z < 0x80000000u ? sn : -sn;
        const double ux = sdir[s][0][threadIdx.x], uy = sdir[s][1][threadIdx.x];
        store2(own, g.nh, j, V2{(float)(ca * ux - sa * uy), (float)(sa * ux + ca * uy)});
        pending &= pending - 1;  // s is the lowest pending site
        s = pending ? (uint32_t)__builtin_ctz(pending) : (uint32_t)kHbSites;
        k = 0;
        if (s < kHbSites) {
          kappa = skr[s][0][threadIdx.x];
          r = skr[s][1][threadIdx.x];
        }
      } else {
        ++k;
      }
    }
  }

  if (MEAS) {
    // the new spins (written by this lane) against their fields, and the
    // other colour's spins at the same ranks
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < kHbSites; ++t) {
      const uint64_t j = jb + (uint64_t)t * kBlock;
      if (j >= g.nh) continue;
      double hx, hy;
      V2 ce;
      on2_field<ND>(oth, g, colour, j, jc, &hx, &hy, &ce);
      const V2 ns = load2(own, g.nh, j);
      acc[0] += (double)ns.x * hx + (double)ns.y * hy;
      acc[1] += (double)ns.x + (double)ce.x;
      acc[2] += (double)ns.y + (double)ce.y;
    }
    __shared__ double red[4];
    if (threadIdx.x < 4) red[threadIdx.x] = 0.0;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = acc[c];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if ((threadIdx.x & 63) == 0) atomicAdd(&red[c], v);
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicAdd(&em[(uint64_t)rep * kEmStride + threadIdx.x], red[threadIdx.x]);
  }
}

template <int ND>
__global__ __launch_bounds__(kBlock) void k_on2_measure(const float *__restrict__ spins, StencilGeom g, float jc,
                                                        double *__restrict__ em) {
  const uint32_t rep = blockIdx.y;
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (j < g.nh) {
    const float *own = spins + (2ull * rep + 1) * 2ull * g.nh;
    const float *oth = spins + (2ull * rep) * 2ull * g.nh;
    const SiteCoords c = site_coords(j, g);
    const uint32_t p = (1 + c.y + c.z) & 1u;
    const uint64_t rb = (uint64_t)c.row * g.lx2;
    uint64_t nb[2 * ND];
    nb[0] = rb + (p ? c.i : wrap_dec(c.i, g.lx2));
    nb[1] = rb + (p ? wrap_inc(c.i, g.lx2) : c.i);
    if (ND >= 2) {
      nb[2] = ((uint64_t)(wrap_dec(c.y, g.ly) + g.ly * c.z)) * g.lx2 + c.i;
      nb[3] = ((uint64_t)(wrap_inc(c.y, g.ly) + g.ly * c.z)) * g.lx2 + c.i;
    }
    if (ND >= 3) {
      nb[4] = ((uint64_t)(c.y + g.ly * wrap_dec(c.z, g.lz))) * g.lx2 + c.i;
      nb[5] = ((uint64_t)(c.y + g.ly * wrap_inc(c.z, g.lz))) * g.lx2 + c.i;
    }
    double hx = 0.0, hy = 0.0;
    V2 centre{0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 2 * ND; ++k) {
      const V2 v = load2(oth, g.nh, nb[k]);
      hx += (double)jc * (double)v.x;
      hy += (double)jc * (double)v.y;
      if (nb[k] == j) centre = v;
    }
    const V2 s = load2(own, g.nh, j);
    acc[0] = (double)s.x * hx + (double)s.y * hy;
    acc[1] = (double)s.x + (double)centre.x;
    acc[2] = (double)s.y + (double)centre.y;
  }
  __shared__ double red[3];
  if (threadIdx.x < 3) red[threadIdx.x] = 0.0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&red[k], v);
  }
  __syncthreads();
  if (threadIdx.x < 3) atomicAdd(&em[(uint64_t)rep * kEmStride + threadIdx.x], red[threadIdx.x]);
}

__global__ __launch_bounds__(kBlock) void k_on2_init(float *__restrict__ spins, StencilGeom g, uint64_t key,
                                                     uint32_t rep, int random) {
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint32_t colour = blockIdx.y;
  if (j >= g.nh) return;
  const SiteCoords c = site_coords(j, g);
  const uint64_t site = 2ull * c.i + ((colour + c.y + c.z) & 1u) + 2ull * g.lx2 * c.row;
  V2 v{1.f, 0.f};  // ordered_spin (models.rs:479-484)
  if (random) {
    const U4 w = philox4x32_10((uint32_t)site, 0u, ctr2(0, 0, kStreamContInit), (uint32_t)(site >> 32),
                               (uint32_t)key, (uint32_t)(key >> 32));
    double out[2];
    random_unit2(w.x, w.y, out);
    v = V2{(float)out[0], (float)out[1]};
  }
  store2(spins + (2ull * rep + colour) * 2ull * g.nh, g.nh, j, v);
}

// ------------------------------------------------ O(N) on graph lattices
// Any CsrLattice: one replica's spins are FP32 planes [component][site],
// updated one colour class at a time in the lattice's class order; the rank
// j of a site inside its class numbers its draws as on the stencil, and the
// local field sums J w s over the non-loop incidences in CSR order
// (local_field, models.rs:579-595).  O(3) sums in FP32 like the stencil
// kernels; O(2) in f64 like k_on2.

template <int D, int ALG>
__global__ __launch_bounds__(kBlock) void k_on_csr(float *__restrict__ spins, const ReplicaParams *__restrict__ reps,
                                                   OnCsrGeom g, uint64_t begin, uint64_t count, uint32_t colour,
                                                   uint64_t sweep, double jd) {
  const uint32_t rep = blockIdx.y;
  const uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= count) return;
  const uint64_t n = g.n_sites;
  float *__restrict__ v = spins + (uint64_t)rep * D * n;
  const uint32_t site = g.class_sites[begin + j];
  const uint32_t e0 = g.offsets[site], e1 = g.offsets[site + 1];
  const ReplicaParams &P = reps[rep];
  if (D == 3) {
    V3 h{0.f, 0.f, 0.f};
    for (uint32_t e = e0; e < e1; ++e) {
      const uint32_t t = g.neighbors[e];
      const float sc = (float)(jd * g.weights[e]);
      h.x += sc * v[t];
      h.y += sc * v[n + t];
      h.z += sc * v[2 * n + t];
    }
    V3 ns;
    if (ALG == kContinuousHeatBath) {
      const U4 blk = philox4x32_10((uint32_t)(j >> 1), (uint32_t)sweep, ctr2(sweep, colour, kStreamCont),
                                   (uint32_t)(j >> 33), (uint32_t)P.key, (uint32_t)(P.key >> 32));
      ns = heat_bath_site(h, P, j, sweep, colour, (j & 1) ? blk.z : blk.x, (j & 1) ? blk.w : blk.y);
    } else {
      ns = reflect_site(V3{v[site], v[n + site], v[2 * n + site]}, h);
    }
    v[site] = ns.x;
    v[n + site] = ns.y;
    v[2 * n + site] = ns.z;
  } else {
    double hx = 0.0, hy = 0.0;
    for (uint32_t e = e0; e < e1; ++e) {
      const uint32_t t = g.neighbors[e];
      const double sc = jd * g.weights[e];
      hx += sc * (double)v[t];
      hy += sc * (double)v[n + t];
    }
    V2 ns{v[site], v[n + site]};
    if (ALG == kContinuousHeatBath) {
      const double norm = sqrt(hx * hx + hy * hy);
      double out[2];
      if (norm < 1e-15) {
        const U4 fb = philox4x32_10((uint32_t)j, (uint32_t)sweep, ctr2(sweep, colour, kStreamContFallback),
                                    (uint32_t)(j >> 32), (uint32_t)P.key, (uint32_t)(P.key >> 32));
        random_unit2(fb.x, fb.y, out);
      } else {
        double ca, sa;
        von_mises_cs(P, j, sweep, colour, P.beta * norm, &ca, &sa);
        const double ux = hx / norm, uy = hy / norm;
        out[0] = ca * ux - sa * uy;
        out[1] = sa * ux + ca * uy;
      }
      ns = V2{(float)out[0], (float)out[1]};
    } else {
      const float fx = (float)hx, fy = (float)hy;
      const float nsq = fx * fx + fy * fy;
      if (nsq >= 1e-28f) {
        const float proj = (ns.x * fx + ns.y * fy) / nsq;
        const float rx = 2.0f * proj * fx - ns.x, ry = 2.0f * proj * fy - ns.y;
        const float inv = rsqrtf(rx * rx + ry * ry);
        ns = V2{rx * inv, ry * inv};
      }
    }
    v[site] = ns.x;
    v[n + site] = ns.y;
  }
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_on_csr_measure(const float *__restrict__ spins, OnCsrGeom g, double jd,
                                                           double *__restrict__ em) {
  const uint32_t rep = blockIdx.y;
  const uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t n = g.n_sites;
  double acc[1 + D];
#pragma unroll
  for (int k = 0; k <= D; ++k) acc[k] = 0.0;
  if (s < n) {
    const float *__restrict__ v = spins + (uint64_t)rep * D * n;
    double h[D];
#pragma unroll
    for (int k = 0; k < D; ++k) h[k] = 0.0;
    for (uint32_t e = g.offsets[s]; e < g.offsets[s + 1]; ++e) {
      const uint32_t t = g.neighbors[e];
      const double sc = jd * g.weights[e];
#pragma unroll
      for (int k = 0; k < D; ++k) h[k] += sc * (double)v[(uint64_t)k * n + t];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double x = (double)v[(uint64_t)k * n + s];
      acc[0] += x * h[k];
      acc[1 + k] = x;
    }
  }
  __shared__ double red[1 + D];
  if (threadIdx.x <= D) red[threadIdx.x] = 0.0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k <= D; ++k) {
    double x = acc[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&red[k], x);
  }
  __syncthreads();
  if (threadIdx.x <= D) atomicAdd(&em[(uint64_t)rep * kEmStride + threadIdx.x], red[threadIdx.x]);
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_on_csr_init(float *__restrict__ spins, uint64_t n, uint64_t key,
                                                        uint32_t rep, int random) {
  const uint64_t site = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (site >= n) return;
  float *__restrict__ v = spins + (uint64_t)rep * D * n;
  const U4 w = random ? philox4x32_10((uint32_t)site, 0u, ctr2(0, 0, kStreamContInit), (uint32_t)(site >> 32),
                                      (uint32_t)key, (uint32_t)(key >> 32))
                      : U4{0u, 0u, 0u, 0u};
  if (D == 3) {
    const V3 u = random ? random_unit(w) : V3{1.f, 0.f, 0.f};  // ordered_spin (models.rs:479-484)
    v[site] = u.x;
    v[n + site] = u.y;
    v[2 * n + site] = u.z;
  } else {
    double out[2] = {1.0, 0.0};
    if (random) random_unit2(w.x, w.y, out);
    v[site] = (float)out[0];
    v[n + site] = (float)out[1];
  }
}

}  // namespace

hipError_t launch_on3_sweep(const StencilGeom &g, int dim, int alg, float *spins, const ReplicaParams *reps,
                            uint32_t n_rep, uint32_t colour, uint64_t sweep, float j, double *em, hipStream_t st,
                            const StepClock *clk) {
  if (dim == 2) {
    const uint32_t nbx = (uint32_t)((g.nh + kBlock - 1) / kBlock);
    const bool flat = (uint64_t)nbx * n_rep * kBlock < (1ull << 32);
    const dim3 grid = flat ? dim3(nbx * n_rep) : dim3(nbx, n_rep);
#define SC_ON2(ND, A)                                                                                                 \
  if (em)                                                                                                             \
    hipLaunchKernelGGL((k_on2<ND, A, true>), grid, dim3(kBlock), 0, st, spins, reps, g, nbx, colour, sweep, j, em,    \
                       clk);                                                                                          \
  else                                                                                                                \
    hipLaunchKernelGGL((k_on2<ND, A, false>), grid, dim3(kBlock), 0, st, spins, reps, g, nbx, colour, sweep, j, em,   \
                       clk);
    const uint32_t nbh = (uint32_t)((g.nh + kBlock * kHbSites - 1) / (kBlock * kHbSites));
    const bool flat_hb = (uint64_t)nbh * n_rep * kBlock < (1ull << 32);
    const dim3 grid_hb = flat_hb ? dim3(nbh * n_rep) : dim3(nbh, n_rep);
#define SC_ON2_HB(ND)                                                                                               \
  if (em)                                                                                                           \
    hipLaunchKernelGGL((k_on2_hb<ND, true>), grid_hb, dim3(kBlock), 0, st, spins, reps, g, nbh, colour, sweep, j, em, \
                       clk);                                                                                        \
  else                                                                                                              \
    hipLaunchKernelGGL((k_on2_hb<ND, false>), grid_hb, dim3(kBlock), 0, st, spins, reps, g, nbh, colour, sweep, j,   \
                       em, clk);
#define SC_ON2_ND(ND)                                \
  if (alg == kContinuousHeatBath) { SC_ON2_HB(ND) } \
  else { SC_ON2(ND, kOverRelaxation) }
    switch (g.nd) {
      case 1: SC_ON2_ND(1) break;
      case 2: SC_ON2_ND(2) break;
      case 3: SC_ON2_ND(3) break;
      default: return hipErrorInvalidValue;
    }
#undef SC_ON2_ND
#undef SC_ON2_HB
#undef SC_ON2
    return hipGetLastError();
  }
  // two ranks per thread for the heat bath (one Philox block for both);
  // over-relaxation has no draws to share and runs faster one rank per thread
  const bool pairs = alg == kContinuousHeatBath && g.lx2 % 2 == 0;
  const uint64_t per_rep = pairs ? g.nh / 2 : g.nh;
  const uint32_t nbx = (uint32_t)((per_rep + kBlock - 1) / kBlock);
  const bool flat = (uint64_t)nbx * n_rep * kBlock < (1ull << 32);
  const dim3 grid = flat ? dim3(nbx * n_rep) : dim3(nbx, n_rep);
#define SC_ON3_K(K, ND, A)                                                                                             \
  if (em)                                                                                                              \
    hipLaunchKernelGGL((K<ND, A, true>), grid, dim3(kBlock), 0, st, spins, reps, g, nbx, colour, sweep, j, em, clk);   \
  else                                                                                                                 \
    hipLaunchKernelGGL((K<ND, A, false>), grid, dim3(kBlock), 0, st, spins, reps, g, nbx, colour, sweep, j, em, clk);
#define SC_ON3(ND, A)              \
  if (pairs) {                     \
    SC_ON3_K(k_on3_pair, ND, A)    \
  } else {                         \
    SC_ON3_K(k_on3, ND, A)         \
  }
#define SC_ON3_ND(ND)                                   \
  if (alg == kContinuousHeatBath) { SC_ON3(ND, kContinuousHeatBath) } \
  else { SC_ON3(ND, kOverRelaxation) }
  switch (g.nd) {
    case 1: SC_ON3_ND(1) break;
    case 2: SC_ON3_ND(2) break;
    case 3: SC_ON3_ND(3) break;
    default: return hipErrorInvalidValue;
  }
#undef SC_ON3_ND
#undef SC_ON3
#undef SC_ON3_K
  return hipGetLastError();
}

hipError_t launch_on3_measure(const StencilGeom &g, int dim, const float *spins, uint32_t n_rep, float j, double *em,
                              hipStream_t st) {
  dim3 grid((unsigned)((g.nh + kBlock - 1) / kBlock), n_rep);
  if (dim == 2) {
    switch (g.nd) {
      case 1: hipLaunchKernelGGL((k_on2_measure<1>), grid, dim3(kBlock), 0, st, spins, g, j, em); break;
      case 2: hipLaunchKernelGGL((k_on2_measure<2>), grid, dim3(kBlock), 0, st, spins, g, j, em); break;
      case 3: hipLaunchKernelGGL((k_on2_measure<3>), grid, dim3(kBlock), 0, st, spins, g, j, em); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  switch (g.nd) {
    case 1: hipLaunchKernelGGL((k_on3_measure<1>), grid, dim3(kBlock), 0, st, spins, g, j, em); break;
    case 2: hipLaunchKernelGGL((k_on3_measure<2>), grid, dim3(kBlock), 0, st, spins, g, j, em); break;
    case 3: hipLaunchKernelGGL((k_on3_measure<3>), grid, dim3(kBlock), 0, st, spins, g, j, em); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_on3_init(const StencilGeom &g, int dim, float *spins, uint64_t key, uint32_t rep, int random,
                           hipStream_t st) {
  if (dim == 2) {
    hipLaunchKernelGGL(k_on2_init, dim3((unsigned)((g.nh + kBlock - 1) / kBlock), 2), dim3(kBlock), 0, st, spins, g,
                       key, rep, random);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_on3_init, dim3((unsigned)((g.nh + kBlock - 1) / kBlock), 2), dim3(kBlock), 0, st, spins, g, key,
                     rep, random);
  return hipGetLastError();
}

hipError_t launch_on_csr_sweep(const OnCsrGeom &g, int dim, int alg, float *spins, const ReplicaParams *reps,
                               uint32_t n_rep, const uint64_t *class_start_host, uint32_t n_colours, uint64_t sweep,
                               double j, hipStream_t st) {
  if (dim != 2 && dim != 3) return hipErrorInvalidValue;
  const bool hb = alg == kContinuousHeatBath;
  for (uint32_t c = 0; c < n_colours; ++c) {
    const uint64_t begin = class_start_host[c], count = class_start_host[c + 1] - begin;
    if (count == 0) continue;
    const dim3 grid((unsigned)((count + kBlock - 1) / kBlock), n_rep);
    if (dim == 3 && hb)
      hipLaunchKernelGGL((k_on_csr<3, kContinuousHeatBath>), grid, dim3(kBlock), 0, st, spins, reps, g, begin, count,
                         c, sweep, j);
    else if (dim == 3)
      hipLaunchKernelGGL((k_on_csr<3, kOverRelaxation>), grid, dim3(kBlock), 0, st, spins, reps, g, begin, count, c,
                         sweep, j);
    else if (hb)
      hipLaunchKernelGGL((k_on_csr<2, kContinuousHeatBath>), grid, dim3(kBlock), 0, st, spins, reps, g, begin, count,
                         c, sweep, j);
    else
      hipLaunchKernelGGL((k_on_csr<2, kOverRelaxation>), grid, dim3(kBlock), 0, st, spins, reps, g, begin, count, c,
                         sweep, j);
  }
  return hipGetLastError();
}

hipError_t launch_on_csr_measure(const OnCsrGeom &g, int dim, const float *spins, uint32_t n_rep, double j,
                                 double *em, hipStream_t st) {
  const dim3 grid((unsigned)((g.n_sites + kBlock - 1) / kBlock), n_rep);
  if (dim == 3) hipLaunchKernelGGL(k_on_csr_measure<3>, grid, dim3(kBlock), 0, st, spins, g, j, em);
  else if (dim == 2) hipLaunchKernelGGL(k_on_csr_measure<2>, grid, dim3(kBlock), 0, st, spins, g, j, em);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_on_csr_init(const OnCsrGeom &g, int dim, float *spins, uint64_t key, uint32_t rep, int random,
                              hipStream_t st) {
  const dim3 grid((unsigned)((g.n_sites + kBlock - 1) / kBlock));
  if (dim == 3) hipLaunchKernelGGL(k_on_csr_init<3>, grid, dim3(kBlock), 0, st, spins, g.n_sites, key, rep, random);
  else if (dim == 2) hipLaunchKernelGGL(k_on_csr_init<2>, grid, dim3(kBlock), 0, st, spins, g.n_sites, key, rep, random);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace scamper
