// Internal interfaces between the host library (capi.cpp, host_*.cpp) and
// the HIP kernels (k_*.hip).  Nothing here crosses the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.h"

namespace scamper {

// Model kinds (scamper_model.kind).
enum ModelKind : int32_t { kIsing = 0, kPotts = 1, kON = 2 };

// Update algorithms (scamper_sweep's `algorithm`).
enum Algorithm : int32_t {
  kMetropolis = 0,
  kHeatBath = 1,
  kContinuousHeatBath = 2,
  kOverRelaxation = 3,
  kHybridHeatBathOverRelaxation = 4,
  kSwendsenWang = 5,
  kWolff = 6,
};

constexpr int kMaxZ = 32;          // largest coordination the discrete kernels accept
constexpr int kMaxTable = 2 * kMaxZ + 1;
constexpr int kMaxQ = 16;          // largest Potts q on the GPU path

// Per-replica device parameters.  `thr` is indexed by the kernel's local
// energy class (see each kernel); `hb_exp` holds the heat-bath Potts weights
// exp(lw(n) - lw(m)) indexed [n * (z+1) + m], where the kernels take m as
// the count of largest log weight: the largest count when J > 0, the
// smallest when J < 0 (hb_min).
struct alignas(16) ReplicaParams {
  uint64_t key;
  double beta;
  Threshold thr[kMaxTable];
};

// Regular periodic hypercubic geometry with even sides, stored colour-split:
// colour c of replica r occupies spins[(2r + c) * cs ...] (int8 spins; cs is
// nh rounded up to kColourAlign bytes, so that every colour array starts
// aligned), rows of lx2 bytes, row index = y + ly * z, element i of a row is
// site x = 2i + ((c+y+z)&1).  Continuous spins keep their own plane layout
// over nh.
struct StencilGeom {
  int nd;        // 1, 2 or 3
  uint32_t lx2;  // Lx / 2
  uint32_t ly;
  uint32_t lz;
  uint64_t nh;   // sites per colour = lx2 * ly * lz
  uint64_t div_lx2 = 0, div_ly = 0;  // division magics (set_divisors)
  uint64_t cs = 0;                   // int8 spins: bytes from a colour array to the next (>= nh)
};

// Colour arrays of int8 spins start at multiples of kColourAlign bytes, so
// that 16-byte windows at multiples of 16 ranks are aligned, and each is
// followed by a gap of at least 16 bytes, which the tiled kernel's last
// window of a colour class reads and writes (bytes of no site).
constexpr uint64_t kColourAlign = 256;
inline uint64_t aligned_colour(uint64_t nh) { return (nh + 16 + kColourAlign - 1) / kColourAlign * kColourAlign; }

// ceil(2^64 / d) for d >= 2, 0 for d = 1: floor(n / d) = mulhi64(n, magic)
// for every 32-bit n (the error term n * (magic * d - 2^64) / 2^64 stays
// below n * d / 2^64 < 1 / d).
uint64_t div_magic(uint32_t d);
inline void set_divisors(StencilGeom &g) {
  g.div_lx2 = div_magic(g.lx2);
  g.div_ly = div_magic(g.ly);
}

__device__ __forceinline__ uint32_t magic_div(uint32_t n, uint64_t magic) {
  return magic ? (uint32_t)__umul64hi((uint64_t)n, magic) : n;
}

// Row (y + ly z), column i, y and z of colour-class rank j.
struct SiteCoords {
  uint32_t row, i, y, z;
};
__device__ __forceinline__ SiteCoords site_coords(uint64_t j, const StencilGeom &g) {
  const uint32_t row = g.nh < (1ull << 32) ? magic_div((uint32_t)j, g.div_lx2) : (uint32_t)(j / g.lx2);
  const uint32_t zz = magic_div(row, g.div_ly);
  return SiteCoords{row, (uint32_t)(j - (uint64_t)row * g.lx2), row - zz * g.ly, zz};
}

// Graph lattice (any CSR) with a greedy colouring.  Spins are site-major
// int8 per replica; class_sites lists the sites of every colour class,
// ascending, starting at class_start[c].
struct CsrGeom {
  uint64_t n_sites;
  uint32_t n_colours;
  int32_t weight_max;        // max sum of |w| over one site's non-loop incidences
  const uint32_t *offsets;   // [n_sites+1]
  const uint32_t *neighbors; // non-loop incidences only
  const int8_t *weights;     // integer weights per incidence
  const uint32_t *class_sites;
};

// Accumulated observables of one measured sweep of one replica:
// em[0] = sum over bonds of w * s_i s_j (Ising) or w * delta(s_i,s_j) (Potts)
// em[1] = sum of spins (Ising) or number of sites in state 0 .. q-1 at
//         em[1 .. q] (Potts).
constexpr int kEmStride = 2 + kMaxQ;

// Step counters kept on the device, so that a captured graph of several
// sweeps or tempering steps replays for any starting point: kernels given a
// clock take their sweep number (and exchange round, series slot) from it
// instead of their by-value arguments, and a tick kernel advances it.
struct alignas(16) StepClock {
  uint64_t sweep;     // sweep number of the step being run
  uint64_t exchange;  // exchange round of the step (as the exchange kernels number it)
  uint64_t slot;      // series slot the step's measurement goes to
  void *series;       // series buffer (its records per slot as the owner lays them out)
};
hipError_t launch_clock_set(StepClock *clk, uint64_t sweep, uint64_t exchange, uint64_t slot, void *series,
                            hipStream_t stream);
hipError_t launch_clock_tick(StepClock *clk, uint64_t d_exchange, uint64_t d_slot, hipStream_t stream);
// Copy `count` counters from em into the clock's series at its slot.
hipError_t launch_series_from_em(const long long *em, const StepClock *clk, uint64_t count, hipStream_t stream);

// ------------------------------------------------------------ launchers
hipError_t launch_stencil_discrete(const StencilGeom &g, int model_kind, int q, int algorithm,
                                   int8_t *spins, const ReplicaParams *reps, uint32_t n_rep,
                                   uint32_t colour, uint64_t sweep, const double *hb_exp, int hb_min,
                                   long long *em, hipStream_t stream);
hipError_t launch_stencil_measure(const StencilGeom &g, int model_kind, int q,
                                  const int8_t *spins, uint32_t n_rep, long long *em,
                                  hipStream_t stream);
hipError_t launch_stencil_init(const StencilGeom &g, int model_kind, int q, int8_t *spins,
                               uint64_t key, uint32_t rep, int random, hipStream_t stream);
hipError_t launch_csr_discrete(const CsrGeom &g, const uint64_t *class_start_host, int kind, int q,
                               int alg, int8_t *spins, const ReplicaParams *reps, uint32_t n_rep,
                               uint64_t sweep, const double *hb_exp, int hb_min, hipStream_t stream);
hipError_t launch_csr_measure(const CsrGeom &g, int kind, int q, const int8_t *spins, uint32_t n_rep,
                              long long *em, hipStream_t stream);
// Ising / Potts on a graph lattice with real weights: every incidence in CSR
// order, self-loops included (neighbour == site), with its f64 weight.
struct RealCsrGeom {
  uint64_t n_sites;
  const uint32_t *offsets;      // [n_sites+1]
  const uint32_t *neighbors;    // all incidences
  const double *weights;        // per incidence
  const uint32_t *class_sites;  // sites of class 0, then class 1, ...
};
hipError_t launch_csr_real(const RealCsrGeom &g, const uint64_t *class_start_host, uint32_t n_colours, int kind, int q,
                           int algorithm, int8_t *spins, const ReplicaParams *reps, uint32_t n_rep, uint64_t sweep,
                           double j, hipStream_t st);
// em[0] holds a double (the weighted bond sum, every bond twice); the
// magnetisation counters are integers as for launch_csr_measure.
hipError_t launch_csr_real_measure(const RealCsrGeom &g, int kind, int q, const int8_t *spins, uint32_t n_rep,
                                   long long *em, hipStream_t st);
hipError_t launch_csr_init(uint64_t n_sites, int kind, int q, int8_t *spins, uint64_t key, uint32_t rep,
                           int random, hipStream_t stream);

// O(N) spins, N = dim = 2 or 3 (FP32 component planes per colour, dim per
// colour per replica); em gets double sums {s.h, Mx, My, Mz} per replica.
// O(N) spins on a graph lattice: FP32 planes [component][site] per replica
// (dim n_sites floats), updated colour class by colour class; `weights` are
// the bond weights of the non-loop incidences, in CSR order.
struct OnCsrGeom {
  uint64_t n_sites;
  const uint32_t *offsets;      // [n_sites+1]
  const uint32_t *neighbors;    // non-loop incidences only
  const double *weights;        // per incidence
  const uint32_t *class_sites;  // sites of class 0, then class 1, ...
};
hipError_t launch_on_csr_sweep(const OnCsrGeom &g, int dim, int alg, float *spins, const ReplicaParams *reps,
                               uint32_t n_rep, const uint64_t *class_start_host, uint32_t n_colours, uint64_t sweep,
                               double j, hipStream_t st);
// em[0] = sum over sites of s.h (J w included: every bond twice), em[1..dim]
// = the summed components, one kEmStride record per replica (doubles).
hipError_t launch_on_csr_measure(const OnCsrGeom &g, int dim, const float *spins, uint32_t n_rep, double j,
                                 double *em, hipStream_t st);
hipError_t launch_on_csr_init(const OnCsrGeom &g, int dim, float *spins, uint64_t key, uint32_t rep, int random,
                              hipStream_t st);

hipError_t launch_on3_sweep(const StencilGeom &g, int dim, int alg, float *spins, const ReplicaParams *reps,
                            uint32_t n_rep, uint32_t colour, uint64_t sweep, float j, double *em, hipStream_t stream,
                            const StepClock *clk = nullptr);
hipError_t launch_on3_measure(const StencilGeom &g, int dim, const float *spins, uint32_t n_rep, float j, double *em,
                              hipStream_t stream);
hipError_t launch_on3_init(const StencilGeom &g, int dim, float *spins, uint64_t key, uint32_t rep, int random,
                           hipStream_t stream);

// Multi-spin coded Ising (u64 words, bit r = replica 64b + r).  A measured
// sweep's counters are kMsCountSlots copies per bundle of 128 u64 (per bit:
// anti-aligned bonds, then +1 spins); block b adds into copy b % kMsCountSlots,
// which keeps the global atomics on one address few, and the exchange kernel
// sums and clears the copies.
constexpr uint32_t kMsCountSlots = 16;
hipError_t launch_ms_init(uint64_t *words, const uint64_t *keys, const StencilGeom &g, uint32_t bundle, int random,
                          hipStream_t stream);
hipError_t launch_ms_sweep(uint64_t *words, const Threshold *thr_by_label, const int32_t *labels, const uint64_t *keys,
                           const StencilGeom &g, uint32_t n_bundles, uint32_t colour, uint64_t sweep,
                           unsigned long long *counts, hipStream_t stream, const StepClock *clk = nullptr);
// With a clock, the round is clk->exchange and, when measure_to_clock is
// set, the energies and |m| go to the clock's series slot ([slot][2][n]);
// with tick set the kernel then advances the clock as launch_clock_tick
// would (sweep + 1, exchange + do_exchange, slot + measure_to_clock).
hipError_t launch_ms_exchange(int32_t *labels, unsigned long long *counts, const double *beta_by_label, uint32_t n,
                              double n_bonds, double n_sites, double j, uint64_t ex_key, uint64_t step, int do_exchange,
                              double *energy_out, double *mag_out, unsigned long long *stats, hipStream_t stream,
                              StepClock *clk = nullptr, int measure_to_clock = 0, int tick = 0);

// ------------------------------------------------------------ tiled jobs
// Checkerboard Metropolis of Ising and Potts spins for any number of
// independent periodic lattices ("jobs") in one launch.  A tile is
// 256 * lane_sites consecutive sites of one colour class of one job (a lane
// updates 4 or 16 sites); block b of a launch walks the tiles of its XCD's
// share of the flattened tile list.  Both shapes give identical results;
// 16 sites per lane amortises more per-lane work, 4 gives more waves.
uint32_t tile_lane_sites(uint64_t total_sites);  // the shape for a launch over that many sites
inline uint32_t tile_sites(uint32_t lane_sites) { return 256u * lane_sites; }
constexpr int kStencilTable = 13;  // 2z+1 for z = 6

// Allocated bytes kept before and after every int8 spin buffer the tiled
// kernel reads: its windows are unaligned 16-byte loads that start up to 16
// bytes before a colour array and end up to 16 bytes after it (the bytes
// outside the array are loaded and never used).
constexpr uint32_t kSpinSlack = 64;

struct alignas(16) TileJob {
  uint64_t key;          // Philox key of the job's replica 0
  int8_t *spins;         // colour 0 at spins, colour 1 at spins + cs
  uint64_t sweep_base;   // sweep number of step 0
  uint64_t sweep_end;    // steps whose sweep is >= sweep_end do nothing
  uint64_t therm;        // sweeps >= therm are measured (when asked)
  uint64_t div_lx2;      // ceil(2^64 / lx2), 0 when lx2 == 1
  uint64_t div_ly;       // ceil(2^64 / ly), 0 when ly == 1
  uint32_t nd, lx2, ly, lz;
  uint32_t nh, kind, q, tile_start;
  uint32_t n_tiles, em_row, cs, pad1;  // cs: bytes from colour 0 to colour 1
  Threshold thr[kStencilTable];  // by anti-aligned count (Ising) or d + z (Potts)
  // byte look-up tables over the same index (v_perm operands, 4 entries per
  // word): top and second byte of K, both 0xff where K = 2^32 (a draw whose
  // top 16 bits are all ones then ties, and the tie is settled exactly)
  uint32_t lut_k8[4], lut_m8[4];
  // Bit-plane form (Ising, Potts q <= 4; null bits: int8 only): four planes
  // (colour, bit) of nwp words each, a spare word before and after
  // ceil(nh / 32) words of 32 ranks; the job's first tile of 256 words in
  // the bit kernel's list; the top byte of K over t = index + 7 - z.
  uint32_t *bits;
  uint32_t nwp, bit_tile_start;
  uint32_t lut_t8[4];
  // the rest of K over t for the sites settled one by one: second byte, low
  // 16 bits as two bytes, and bit t set where K = 2^32
  uint32_t lut_tm8[4], lut_tk16lo[4], lut_tk16hi[4], tal;
  // bit 0: pack the int8 spins into the planes at this chunk's start (the
  // job's first chunk), bit 1: unpack them at its end (its last); between
  // them the planes are the job's state
  uint32_t bit_io;
};

// Fill a row's byte look-up tables from its thresholds.
inline void fill_tile_luts(TileJob &t) {
  for (int w = 0; w < 4; ++w) t.lut_k8[w] = t.lut_m8[w] = 0;
  for (int a = 0; a < kStencilTable; ++a) {
    const uint32_t k8 = t.thr[a].k8;
    const uint32_t sh = 8u * (uint32_t)(a & 3);
    t.lut_k8[a >> 2] |= (k8 > 0xFFu ? 0xFFu : k8) << sh;
    t.lut_m8[a >> 2] |= (k8 > 0xFFu ? 0xFFu : (uint32_t)t.thr[a].m8) << sh;
  }
}

struct alignas(16) TileHeader {
  uint32_t n_jobs;
  uint32_t n_tiles;  // tiles of the int8 kernel
  uint32_t pad[2];   // pad[0]: tiles of 256 words of the bit-plane kernel; pad[1] bit 0: the int8
                     // kernel deals the tiles over all XCDs (the pool sets it for several jobs)
};


// Where a launch adds its measured counters: for job row r, step k and
// block b the record (em[0], em[1], em[2 .. 2+q) as for the stencil) at
//   base + r * row_stride + k * step_stride + (b & sub_mask) * sub_stride.
struct EmLayout {
  long long *base;
  uint64_t row_stride;
  uint64_t step_stride;
  uint32_t sub_mask;
  uint32_t sub_stride;
};

// Bit-plane jobs of a table: the tile kernel on their planes (the same
// draws, decisions and measured records as launch_tile_sweep), and the
// conversion of every bit-plane job's int8 colour arrays to planes (unpack 0)
// or back (unpack 1).  The planes of a job must hold 16 (ceil(nh / 32) + 2)
// bytes at `bits`; job_of_tile[t] is the row of bit tile t.
void fill_bit_lut(TileJob &t);
hipError_t launch_tile_bits(const TileHeader *hdr, const TileJob *jobs, const uint32_t *job_of_tile, uint32_t grid,
                            uint64_t step, uint32_t colour, int measure, EmLayout em, hipStream_t stream);
hipError_t launch_bits_pack(const TileHeader *hdr, const TileJob *jobs, const uint32_t *job_of_tile, int unpack,
                            hipStream_t stream);

hipError_t launch_tile_sweep(const TileHeader *hdr, const TileJob *jobs, uint32_t grid, uint32_t lane_sites,
                             uint64_t step, uint32_t colour, int measure, EmLayout em, hipStream_t stream);
// Zero hdr->n_jobs * per_row counters at p (a kernel: safe inside graphs).
hipError_t launch_tile_zero(const TileHeader *hdr, long long *p, uint64_t per_row, hipStream_t stream);
// Zero n counters; out[rec * width + c] = sum over s < 8 of
// ring[(rec * 8 + s) * width + c] for rec < n_rec (kernels: safe inside graphs).
hipError_t launch_em_zero(long long *p, uint64_t n, hipStream_t stream);
hipError_t launch_em_fold(const long long *ring, long long *out, uint64_t n_rec, uint32_t width, hipStream_t stream);
// compact[(r * steps + k) * 8 + c] = sum over s < 8 of ring[((r * steps + k) * 8 + s) * 8 + c].
hipError_t launch_tile_fold(const TileHeader *hdr, const long long *ring, long long *compact, uint32_t steps,
                            hipStream_t stream);

// ------------------------------------------------------------ cluster updates
// Swendsen-Wang and Wolff updates of one replica (Ising +-1 or Potts states
// in int8).  Sites are numbered site-major (x + Lx (y + Ly z) on stencils,
// the lattice's numbering on graphs); a stencil replica's spins are its two
// colour arrays at `spins` (site s at colour (x+y+z)&1, rank s >> 1), a
// graph replica's are site-major.  `parent` holds one uint32 per site.
struct ClusterEdge {
  uint32_t source, target;
};
// Bond thresholds: an aligned bond of integer weight w is active when its
// draw is below thr[|w|] (thr[1] for stencils), built with |J| (J w >= 0 on
// every edge, so the coupling J w is |J| |w|).
constexpr int kClusterWeights = 128;
hipError_t launch_cluster_bonds_stencil(const StencilGeom &g, const int8_t *spins, uint32_t *parent,
                                        uint64_t key, uint64_t step, uint64_t thr, hipStream_t stream);
// `edge_thr` (one threshold per edge, or null for thr by |w|) serves graphs
// with real weights.
hipError_t launch_cluster_bonds_csr(uint64_t n_sites, uint64_t n_edges, const ClusterEdge *edges,
                                    const int8_t *weights, const int8_t *spins, uint32_t *parent, uint64_t key,
                                    uint64_t step, const uint64_t *thr, const uint64_t *edge_thr,
                                    hipStream_t stream);
// SW: every cluster takes its smallest site's target; Wolff: the seed's
// cluster takes -seed (Ising) or another state (Potts).  `scratch` holds 2
// uint32 for the Wolff seed's root and target.
hipError_t launch_cluster_transform(const StencilGeom *g, uint64_t n_sites, int8_t *spins, uint32_t *parent,
                                    int kind, int q, int wolff, uint64_t key, uint64_t step, uint32_t *scratch,
                                    hipStream_t stream);
// One Swendsen-Wang or Wolff update of O(N) spins, N = dim = 2 or 3, on a
// stencil lattice (`planes`: one replica's FP32 component planes); scratch
// holds the Wolff seed's root (4 B) and, from byte 8, the update's direction
// (dim doubles).
hipError_t launch_cluster_on(const StencilGeom &g, int dim, float *planes, uint32_t *parent, uint64_t key,
                             uint64_t step, double beta, double j, int wolff, void *scratch, hipStream_t stream);
// The same on a graph lattice (`planes`: one replica's [component][site]
// planes; edges and their f64 weights in edge order, self-loops included).
hipError_t launch_cluster_on_csr(uint64_t n_sites, uint64_t n_edges, const ClusterEdge *edges, const double *weights,
                                 int dim, float *planes, uint32_t *parent, uint64_t key, uint64_t step, double beta,
                                 double j, int wolff, void *scratch, hipStream_t stream);

// ------------------------------------------------------------ replica exchange
// How a replica's energy follows from its recount counters (the expressions
// of the host's observables): stencil and graph Ising / Potts (integer bond
// sums), real-weight graphs (an f64 bond sum, every bond twice), O(N) on
// stencils and graphs (f64 sums of s.h).
enum EnergyKind : int32_t {
  kEnergyStencil = 0,
  kEnergyGraph = 1,
  kEnergyGraphReal = 2,
  kEnergyVecStencil = 3,
  kEnergyVecGraph = 4,
};
struct EnergyForm {
  int32_t kind;
  int32_t pad;
  double j;
  double self_loops;  // sum of the self-loop weights
};
// One exchange round over the labels of n <= 1024 replicas from the
// counters em[r * kEmStride ...]; stats = {accepted, near ties, rounds}.
hipError_t launch_state_exchange(int32_t *labels, const long long *em, uint32_t n, const EnergyForm &form,
                                 const double *ladder, uint64_t ex_key, uint64_t step, double *energy_out,
                                 unsigned long long *stats, hipStream_t stream, const StepClock *clk = nullptr);
// Copy the decision tables of the label each replica holds into its own
// parameters (and tiled rows / heat-bath weights when given).
hipError_t launch_apply_labels(const int32_t *labels, uint32_t n, ReplicaParams *params,
                               const ReplicaParams *label_params, TileJob *rows, const TileJob *label_rows,
                               double *hbexp, const double *label_hbexp, uint32_t hb_per, hipStream_t stream);

// Register-blocked Ising Metropolis for 2-D/3-D with Lx/2 a multiple of 16.
bool ising_fast_path_ok(const StencilGeom &g);
hipError_t launch_ising_fast(const StencilGeom &g, int8_t *spins, const ReplicaParams *reps,
                             uint32_t n_rep, uint32_t colour, uint64_t sweep, long long *em,
                             hipStream_t stream, const StepClock *clk = nullptr);

}  // namespace scamper
