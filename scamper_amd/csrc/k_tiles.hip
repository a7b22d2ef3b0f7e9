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
// one job; wave w of the 256-thread block owns ranks [1024 w, 1024 w + 1024)
// of the tile.  Lane t evaluates the Philox blocks of group g = j0/16 + t
// (16 ranks each: top and mid bytes of the acceptance draw, and of the
// proposal draw for Potts), parks the bytes in LDS, and then updates ranks
// j0 + 64 k + t (k < 16) so that spin loads and stores are coalesced.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace scamper {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kWaveSites = 1024;
static_assert(kWaves * kWaveSites == (int)kTileSites, "tile shape");

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
  uint32_t acc_top[kWaves][kWaveSites / 4];
  uint32_t acc_mid[kWaves][kWaveSites / 4];
  uint32_t pro_top[kWaves][kWaveSites / 4];
  uint32_t pro_mid[kWaves][kWaveSites / 4];
  Threshold thr[kStencilTable];
  unsigned long long red[2 + kMaxQ];
};

__device__ __forceinline__ uint32_t lds_byte(const uint32_t *a, uint32_t idx) {
  return reinterpret_cast<const uint8_t *>(a)[idx];
}

__device__ __forceinline__ void put_block(uint32_t *dst, const U4 &b) {
  *reinterpret_cast<uint4 *>(dst) = make_uint4(b.x, b.y, b.z, b.w);
}

// One tile of one job.  Returns after the block's counters (if any) have been
// added to `em_rec`.
template <int ND, int KIND>
__device__ __forceinline__ void run_tile(TileLds &L, const TileJob &J, uint32_t tile, uint64_t sweep, uint32_t colour, bool meas,
                         long long *em_rec) {
  constexpr int z = 2 * ND;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t nh = J.nh;
  const uint32_t jw0 = tile * kTileSites + wave * kWaveSites;
  const uint64_t key = J.key;
  const uint32_t q = J.q;
  const uint32_t m = q - 1;
  const bool need_prop = KIND == kPotts && q > 2;
  const bool need_pmid = need_prop && (m & (m - 1)) != 0;

  if (threadIdx.x < kStencilTable) L.thr[threadIdx.x] = J.thr[threadIdx.x];
  if (threadIdx.x < 2 + kMaxQ) L.red[threadIdx.x] = 0;
  {
    const uint32_t g = (jw0 >> 4) + lane;
    if (16u * g < nh) {
      put_block(&L.acc_top[wave][4 * lane], byte_block(key, g, sweep, colour, kStreamAccept, 0));
      put_block(&L.acc_mid[wave][4 * lane], byte_block(key, g, sweep, colour, kStreamAccept, 1));
      if (need_prop) put_block(&L.pro_top[wave][4 * lane], byte_block(key, g, sweep, colour, kStreamPropose, 0));
      if (need_pmid) put_block(&L.pro_mid[wave][4 * lane], byte_block(key, g, sweep, colour, kStreamPropose, 1));
    }
  }
  __syncthreads();

  int8_t *__restrict__ own = J.spins + (uint64_t)colour * nh;
  const int8_t *__restrict__ oth = J.spins + (uint64_t)(1 - colour) * nh;
  const uint32_t lx2 = J.lx2, ly = J.ly, lz = J.lz;
  int acc_e = 0, acc_m = 0;
  uint64_t cnt_lo = 0, cnt_hi = 0;  // byte v of (lo, hi) counts sites in Potts state v (< 16)

#pragma unroll 2
  for (int k = 0; k < 16; ++k) {
    const uint32_t idx = 64u * k + lane;
    const uint32_t j = jw0 + idx;
    if (j >= nh) break;
    const uint32_t row = div_by(j, J.div_lx2);
    const uint32_t i = j - row * lx2;
    const uint32_t zz = ND == 3 ? div_by(row, J.div_ly) : 0u;
    const uint32_t y = ND == 1 ? 0u : row - zz * ly;
    const uint32_t p = (colour + y + zz) & 1u;
    int8_t n[z];
    {
      const uint32_t rb = row * lx2;
      n[0] = oth[rb + (p ? i : wrap_dec(i, lx2))];
      n[1] = oth[rb + (p ? wrap_inc(i, lx2) : i)];
      if (ND >= 2) {
        n[2] = oth[(wrap_dec(y, ly) + ly * zz) * lx2 + i];
        n[3] = oth[(wrap_inc(y, ly) + ly * zz) * lx2 + i];
      }
      if (ND >= 3) {
        n[4] = oth[(y + ly * wrap_dec(zz, lz)) * lx2 + i];
        n[5] = oth[(y + ly * wrap_inc(zz, lz)) * lx2 + i];
      }
    }
    const int8_t s = own[j];
    int8_t ns = s;
    const uint32_t top = lds_byte(L.acc_top[wave], idx);
    const uint32_t mid = lds_byte(L.acc_mid[wave], idx);
    if (KIND == kIsing) {
      int anti = 0;
#pragma unroll
      for (int c = 0; c < z; ++c) anti += (n[c] != s);
      if (decide2(top, mid, L.thr[anti], key, j, sweep, colour)) ns = (int8_t)-s;
    } else {
      uint32_t r;
      if (q == 2) {
        r = 1u - (uint32_t)s;
      } else {
        const uint32_t pt = lds_byte(L.pro_top[wave], idx);
        const uint32_t pm = need_pmid ? lds_byte(L.pro_mid[wave], idx) : 0u;
        r = propose2(pt, pm, m, key, j, sweep, colour);
        r += (r >= (uint32_t)s);
      }
      int d = 0;
#pragma unroll
      for (int c = 0; c < z; ++c) d += ((uint32_t)n[c] == r) - (n[c] == s);
      if (decide2(top, mid, L.thr[d + z], key, j, sweep, colour)) ns = (int8_t)r;
    }
    if (ns != s) own[j] = ns;
    if (meas) {
      const int8_t o = oth[j];
      if (KIND == kIsing) {
        int sum = 0;
#pragma unroll
        for (int c = 0; c < z; ++c) sum += n[c];
        acc_e += ns * sum;
        acc_m += ns + o;
      } else {
        int same = 0;
#pragma unroll
        for (int c = 0; c < z; ++c) same += (n[c] == ns);
        acc_e += same;
        const uint64_t one_ns = 1ull << (8 * (ns & 7)), one_o = 1ull << (8 * (o & 7));
        if (ns < 8) cnt_lo += one_ns; else cnt_hi += one_ns;
        if (o < 8) cnt_lo += one_o; else cnt_hi += one_o;
      }
    }
  }

  if (meas) {
    const int nc = KIND == kIsing ? 2 : 2 + (int)q;
    for (int c = 0; c < nc; ++c) {
      long long v;
      if (c == 0) v = acc_e;
      else if (c == 1) v = KIND == kIsing ? acc_m : 0;
      else v = (long long)(((c - 2 < 8 ? cnt_lo : cnt_hi) >> (8 * ((c - 2) & 7))) & 0xFFu);
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0 && v != 0) atomicAdd(&L.red[c], (unsigned long long)v);
    }
    __syncthreads();
    if ((int)threadIdx.x < nc && L.red[threadIdx.x] != 0)
      atomicAdd(reinterpret_cast<unsigned long long *>(em_rec) + threadIdx.x, L.red[threadIdx.x]);
  }
  __syncthreads();  // LDS is reused by the block's next tile
}

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
    long long *rec = meas ? em.base + J.em_row * em.row_stride + step * em.step_stride +
                                (uint64_t)(blockIdx.x & em.sub_mask) * em.sub_stride
                          : nullptr;
    const uint32_t tile = t - J.tile_start;
    const uint32_t sel = J.nd * 2 + (J.kind == kPotts ? 1 : 0);
    switch (sel) {
      case 2: run_tile<1, kIsing>(L, J, tile, sweep, colour, meas, rec); break;
      case 3: run_tile<1, kPotts>(L, J, tile, sweep, colour, meas, rec); break;
      case 4: run_tile<2, kIsing>(L, J, tile, sweep, colour, meas, rec); break;
      case 5: run_tile<2, kPotts>(L, J, tile, sweep, colour, meas, rec); break;
      case 6: run_tile<3, kIsing>(L, J, tile, sweep, colour, meas, rec); break;
      default: run_tile<3, kPotts>(L, J, tile, sweep, colour, meas, rec); break;
    }
  }
}

__global__ void k_tile_zero(const TileHeader *__restrict__ hdr, long long *__restrict__ p, uint64_t per_row) {
  const uint64_t n = (uint64_t)hdr->n_jobs * per_row;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
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

hipError_t launch_tile_sweep(const TileHeader *hdr, const TileJob *jobs, uint32_t grid, uint64_t step,
                             uint32_t colour, int measure, EmLayout em, hipStream_t stream) {
  if (grid == 0 || (grid & 7u)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tile_sweep, dim3(grid), dim3(kThreads), 0, stream, hdr, jobs, step, colour, measure, em);
  return hipGetLastError();
}

hipError_t launch_tile_zero(const TileHeader *hdr, long long *p, uint64_t per_row, hipStream_t stream) {
  hipLaunchKernelGGL(k_tile_zero, dim3(256), dim3(256), 0, stream, hdr, p, per_row);
  return hipGetLastError();
}

hipError_t launch_tile_fold(const TileHeader *hdr, const long long *ring, long long *compact, uint32_t steps,
                            hipStream_t stream) {
  hipLaunchKernelGGL(k_tile_fold, dim3(64), dim3(256), 0, stream, hdr, ring, compact, steps);
  return hipGetLastError();
}

}  // namespace scamper
