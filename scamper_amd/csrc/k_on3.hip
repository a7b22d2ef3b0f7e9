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
// The inverse-CDF draw of cos(theta) is written in the form
//   cos = 1 + log1p((1 - u) * expm1(-2 kappa)) / kappa,
// algebraically the reference's 1 + ln(u + (1 - u) e^{-2 kappa}) / kappa but
// without its FP32 cancellation at small kappa.
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
__device__ __forceinline__ V3 heat_bath_site(V3 h, const ReplicaParams &P, uint64_t j, uint64_t sweep,
                                             uint32_t colour, uint32_t w0, uint32_t w1) {
  const float beta = (float)P.beta;
  const float norm = sqrtf(h.x * h.x + h.y * h.y + h.z * h.z);
  const float kappa = beta * norm;
  if (kappa < 1e-12f)
    return random_unit(philox4x32_10((uint32_t)j, (uint32_t)sweep, ctr2(sweep, colour, kStreamContFallback),
                                     (uint32_t)(j >> 32), (uint32_t)P.key, (uint32_t)(P.key >> 32)));
  const float u = fmaxf((float)w0 * 2.3283064365386963e-10f, 1.17549435e-38f);
  float c = 1.0f + log1pf((1.0f - u) * expm1f(-2.0f * kappa)) / kappa;
  c = fminf(fmaxf(c, -1.0f), 1.0f);
  const float sn = sqrtf(fmaxf(1.0f - c * c, 0.0f));
  float sphi, cphi;
  sincos_turns((float)w1 * 2.3283064365386963e-10f, &sphi, &cphi);
  const float lx = sn * cphi, ly = sn * sphi;
  const float inv_n = 1.0f / norm;
  const float ux = h.x * inv_n, uy = h.y * inv_n, uz = h.z * inv_n;
  const float tr = sqrtf(ux * ux + uy * uy);
  if (tr < 1e-6f) return normalised(V3{lx, ly, uz >= 0.0f ? c : -c});
  const float it = 1.0f / tr;
  return normalised(V3{-uy * it * lx - ux * uz * it * ly + ux * c, ux * it * lx - uy * uz * it * ly + uy * c,
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
                                                float jc, double *__restrict__ em) {
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
                                                     float jc, double *__restrict__ em) {
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

}  // namespace

hipError_t launch_on3_sweep(const StencilGeom &g, int alg, float *spins, const ReplicaParams *reps, uint32_t n_rep,
                            uint32_t colour, uint64_t sweep, float j, double *em, hipStream_t st) {
  // two ranks per thread for the heat bath (one Philox block for both);
  // over-relaxation has no draws to share and runs faster one rank per thread
  const bool pairs = alg == kContinuousHeatBath && g.lx2 % 2 == 0;
  const uint64_t per_rep = pairs ? g.nh / 2 : g.nh;
  const uint32_t nbx = (uint32_t)((per_rep + kBlock - 1) / kBlock);
  const bool flat = (uint64_t)nbx * n_rep * kBlock < (1ull << 32);
  const dim3 grid = flat ? dim3(nbx * n_rep) : dim3(nbx, n_rep);
#define SC_ON3_K(K, ND, A)                                                                                             \
  if (em)                                                                                                              \
    hipLaunchKernelGGL((K<ND, A, true>), grid, dim3(kBlock), 0, st, spins, reps, g, nbx, colour, sweep, j, em);        \
  else                                                                                                                 \
    hipLaunchKernelGGL((K<ND, A, false>), grid, dim3(kBlock), 0, st, spins, reps, g, nbx, colour, sweep, j, em);
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

hipError_t launch_on3_measure(const StencilGeom &g, const float *spins, uint32_t n_rep, float j, double *em,
                              hipStream_t st) {
  dim3 grid((unsigned)((g.nh + kBlock - 1) / kBlock), n_rep);
  switch (g.nd) {
    case 1: hipLaunchKernelGGL((k_on3_measure<1>), grid, dim3(kBlock), 0, st, spins, g, j, em); break;
    case 2: hipLaunchKernelGGL((k_on3_measure<2>), grid, dim3(kBlock), 0, st, spins, g, j, em); break;
    case 3: hipLaunchKernelGGL((k_on3_measure<3>), grid, dim3(kBlock), 0, st, spins, g, j, em); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_on3_init(const StencilGeom &g, float *spins, uint64_t key, uint32_t rep, int random, hipStream_t st) {
  hipLaunchKernelGGL(k_on3_init, dim3((unsigned)((g.nh + kBlock - 1) / kBlock), 2), dim3(kBlock), 0, st, spins, g, key,
                     rep, random);
  return hipGetLastError();
}

}  // namespace scamper
