// C ABI of libscamper_hip.so: lattice handles, device-resident replica
// batches and the sweep driver.  Declarations and reference citations are in
// include/scamper_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "host.h"

using namespace scamper;

namespace {
thread_local std::string g_last_error;
}  // namespace

int scamper::fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

namespace {

// RngStreamKey::seed (Carlo.rs/src/rng_stream.rs:126-154): SplitMix64
// rounds fold every logical coordinate into the 64-bit stream seed.
uint64_t mix64(uint64_t v) {
  v += 0x9E3779B97F4A7C15ull;
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
  return v ^ (v >> 31);
}
uint64_t fold(uint64_t s, uint64_t domain, uint64_t value) {
  return mix64(s ^ domain ^ mix64(value + ((domain << 17) | (domain >> 47))));
}
}  // namespace

uint64_t scamper::stream_seed(uint64_t base, uint64_t replica, uint64_t phase) {
  uint64_t s = mix64(base ^ 0x53435554544C4552ull);
  s = fold(s, 0x5441534Bull, 0);               // task
  s = fold(s, 0x52554E00ull, 0);               // run
  s = fold(s, 0x434841494E00ull, 0);           // chain
  s = fold(s, 0x5245504C49434100ull, replica); // replica
  s = fold(s, 0x544852454144ull, 0);           // thread
  s = fold(s, 0x504841534500ull, phase);       // phase
  return fold(s, 0x535542535452ull, 0);        // substream
}

struct scamper_state {
  Lattice lat;
  scamper_model model{};
  uint32_t n_rep = 0;
  uint64_t seed = 0;
  uint64_t sweeps = 0;
  std::vector<double> beta;
  std::vector<uint64_t> keys;
  StencilGeom geom{};
  int z = 0;
  // graph lattices (any CSR): integer weights, greedy colour classes
  bool csr = false;
  CsrGeom cgeom{};
  std::vector<uint64_t> class_start;
  int w_max = 0;
  bool unit_weights = true;
  double self_loop_weight = 0.0;  // sum of w over self-loops (constant energy)
  uint32_t *d_offsets = nullptr, *d_neighbors = nullptr, *d_class_sites = nullptr;
  int8_t *d_weights = nullptr;
  // Ising / Potts on a graph lattice: every incidence with its f64 weight,
  // for real weights (real_weights) and the Potts heat bath on weighted graphs
  bool real_weights = false;
  RealCsrGeom rgeom{};
  uint32_t *d_roffsets = nullptr, *d_rneighbors = nullptr;
  double *d_rweights = nullptr;
  // O(N) on a graph lattice: f64 weights of the non-loop incidences
  OnCsrGeom ogeom{};
  double *d_fweights = nullptr;
  hipStream_t stream = nullptr;
  int8_t *d_spins = nullptr;       // d_spins_alloc + kSpinSlack
  int8_t *d_spins_alloc = nullptr;
  bool continuous = false;  // O(N): FP32 component planes in d_vec
  float *d_vec = nullptr;
  ReplicaParams *d_params = nullptr;
  double *d_hbexp = nullptr;
  long long *d_em = nullptr;      // scratch for exact recomputation
  long long *d_series = nullptr;  // [slot][replica][kEmStride]
  // stencil Metropolis outside the fast Ising path: one tile-table row per
  // replica (header, then the rows)
  void *d_tiles = nullptr;
  uint32_t tile_grid = 0;
  uint32_t tile_lane = 4;  // sites per lane of the tiled launches
  // measured tile sweeps add into 8 per-XCD sub-records per replica in this
  // ring, folded into the series every ring_slots sweeps
  long long *d_ring = nullptr;
  uint32_t ring_slots = 0;
  uint64_t series_cap = 0, series_len = 0;
  int tables_for = -1;            // algorithm the uploaded tables serve
  // cluster updates: union-find parents (one replica at a time), the Wolff
  // seed's root and target, bond thresholds per replica and weight, and a
  // graph lattice's edges with their integer weights
  uint32_t *d_parent = nullptr, *d_cscratch = nullptr;
  uint64_t *d_cthr = nullptr;
  uint64_t *d_ethr = nullptr;  // real-weight graphs: thresholds per (replica, edge)
  ClusterEdge *d_cedges = nullptr;
  int8_t *d_cweights = nullptr;
  double *d_cweights_f64 = nullptr;  // O(N) on a graph lattice: f64 edge weights
  std::vector<uint64_t> cthr;  // host copy of d_cthr, valid while cthr_ready
  bool cthr_ready = false;
  int device = 0;
  // parallel tempering: beta by temperature label (empty: no ladder), the
  // label each replica holds (on the device; the host copy and beta[] follow
  // it when labels_dirty is false), per-label decision tables, the energies
  // of the last exchange round and the counters {accepted, near ties, rounds}
  std::vector<double> ladder;
  std::vector<int32_t> labels;
  bool labels_dirty = false;
  int32_t *d_labels = nullptr;
  double *d_ladder = nullptr;
  double *d_xenergy = nullptr;
  unsigned long long *d_xstats = nullptr;
  ReplicaParams *d_lab_params = nullptr;
  TileJob *d_lab_rows = nullptr;
  double *d_lab_hbexp = nullptr;
  // tempering steps: device step counters and captured chunks of steps by
  // (algorithm, measure), captured on cap_stream and launched on stream
  StepClock *d_clock = nullptr;
  hipStream_t cap_stream = nullptr;
  std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> pt_graphs;
  uint64_t pt_graph_launches = 0;

  ~scamper_state() {
    if (stream) hipStreamSynchronize(stream);
    hipFree(d_spins_alloc);
    hipFree(d_parent);
    hipFree(d_cscratch);
    hipFree(d_cthr);
    hipFree(d_ethr);
    hipFree(d_cedges);
    hipFree(d_cweights);
    hipFree(d_cweights_f64);
    hipFree(d_fweights);
    hipFree(d_roffsets);
    hipFree(d_rneighbors);
    hipFree(d_rweights);
    hipFree(d_vec);
    hipFree(d_params);
    hipFree(d_hbexp);
    hipFree(d_em);
    hipFree(d_series);
    hipFree(d_tiles);
    hipFree(d_ring);
    hipFree(d_offsets);
    hipFree(d_neighbors);
    hipFree(d_class_sites);
    hipFree(d_weights);
    hipFree(d_labels);
    hipFree(d_ladder);
    hipFree(d_xenergy);
    hipFree(d_xstats);
    hipFree(d_lab_params);
    hipFree(d_lab_rows);
    hipFree(d_lab_hbexp);
    hipFree(d_clock);
    for (auto &kv : pt_graphs) {
      hipGraphExecDestroy(kv.second.second);
      hipGraphDestroy(kv.second.first);
    }
    if (cap_stream) hipStreamDestroy(cap_stream);
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

int check_state(const scamper_state *st, uint32_t replica) {
  if (!st) return fail(SCAMPER_EINVAL, "null state");
  if (replica >= st->n_rep) return fail(SCAMPER_EINVAL, "replica %u out of range (%u replicas)", replica, st->n_rep);
  return SCAMPER_OK;
}

// Stencil Metropolis runs on the tiled kernel unless the register-blocked
// Ising kernel takes it.
bool uses_tiles(const scamper_state *st, int alg) {
  if (st->csr || st->continuous || alg != kMetropolis) return false;
  if (st->model.kind == kIsing && ising_fast_path_ok(st->geom)) return false;
  return st->geom.nh < (1ull << 31);
}

// Ising / Potts updates on a graph that run on the f64 kernels: real
// weights, or the Potts heat bath on a weighted graph.
bool uses_real_kernels(const scamper_state *st, int alg) {
  if (!st->csr || st->continuous) return false;
  return st->real_weights || (st->model.kind == kPotts && alg == kHeatBath && !st->unit_weights);
}

// Adds the Ising / Potts counters of every replica on a graph into em.
hipError_t measure_graph(const scamper_state *st, long long *em) {
  return st->real_weights ? launch_csr_real_measure(st->rgeom, st->model.kind, st->model.q, st->d_spins, st->n_rep, em,
                                                    st->stream)
                          : launch_csr_measure(st->cgeom, st->model.kind, st->model.q, st->d_spins, st->n_rep, em,
                                               st->stream);
}

// Copy each replica's decision tables from the label it holds (on the
// device, in stream order), for the tables uploaded last.
hipError_t apply_labels(scamper_state *st) {
  if (st->ladder.empty() || st->tables_for < 0) return hipSuccess;
  TileJob *rows = nullptr;
  if (uses_tiles(st, st->tables_for) && st->d_tiles && st->d_lab_rows)
    rows = reinterpret_cast<TileJob *>(static_cast<TileHeader *>(st->d_tiles) + 1);
  const bool hb = st->d_lab_hbexp && st->d_hbexp && st->model.kind == kPotts && st->tables_for == kHeatBath &&
                  !uses_real_kernels(st, kHeatBath);
  return launch_apply_labels(st->d_labels, st->n_rep, st->d_params, st->d_lab_params, rows,
                             rows ? st->d_lab_rows : nullptr, hb ? st->d_hbexp : nullptr,
                             hb ? st->d_lab_hbexp : nullptr, (uint32_t)((kMaxZ + 1) * (kMaxZ + 1)), st->stream);
}

// Make beta[] and the host copy of the labels follow the device's labels
// (after exchange rounds).
int sync_labels(scamper_state *st) {
  if (st->ladder.empty() || !st->labels_dirty) return SCAMPER_OK;
  SC_HIP(hipStreamSynchronize(st->stream));
  SC_HIP(hipMemcpy(st->labels.data(), st->d_labels, st->n_rep * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (uint32_t r = 0; r < st->n_rep; ++r) st->beta[r] = st->ladder[st->labels[r]];
  st->labels_dirty = false;
  return SCAMPER_OK;
}

// Exact energies of the current spins (change_parameter's recompute), one
// exchange round over the labels, and each replica's tables from its new
// label.  With a clock, the round number is the clock's.
hipError_t enqueue_exchange(scamper_state *st, uint64_t step, const StepClock *clk, hipStream_t s) {
  hipError_t e = launch_em_zero(st->d_em, (uint64_t)kEmStride * st->n_rep, s);
  if (e == hipSuccess) {
    if (st->continuous)
      e = st->csr ? launch_on_csr_measure(st->ogeom, st->model.spin_dim, st->d_vec, st->n_rep, st->model.j,
                                          reinterpret_cast<double *>(st->d_em), s)
                  : launch_on3_measure(st->geom, st->model.spin_dim, st->d_vec, st->n_rep, (float)st->model.j,
                                       reinterpret_cast<double *>(st->d_em), s);
    else if (st->csr)
      e = st->real_weights ? launch_csr_real_measure(st->rgeom, st->model.kind, st->model.q, st->d_spins, st->n_rep,
                                                     st->d_em, s)
                           : launch_csr_measure(st->cgeom, st->model.kind, st->model.q, st->d_spins, st->n_rep,
                                                st->d_em, s);
    else
      e = launch_stencil_measure(st->geom, st->model.kind, st->model.q, st->d_spins, st->n_rep, st->d_em, s);
  }
  EnergyForm form{};
  form.j = st->model.j;
  form.self_loops = st->self_loop_weight;
  form.kind = st->continuous ? (st->csr ? kEnergyVecGraph : kEnergyVecStencil)
              : st->csr      ? (st->real_weights ? kEnergyGraphReal : kEnergyGraph)
                             : kEnergyStencil;
  if (e == hipSuccess)
    e = launch_state_exchange(st->d_labels, st->d_em, st->n_rep, form, st->d_ladder, stream_seed(st->seed, 0, 3), step,
                              st->d_xenergy, st->d_xstats, s, clk);
  if (e == hipSuccess && !st->ladder.empty() && st->tables_for >= 0) {
    TileJob *rows = nullptr;
    if (uses_tiles(st, st->tables_for) && st->d_tiles && st->d_lab_rows)
      rows = reinterpret_cast<TileJob *>(static_cast<TileHeader *>(st->d_tiles) + 1);
    const bool hb = st->d_lab_hbexp && st->d_hbexp && st->model.kind == kPotts && st->tables_for == kHeatBath &&
                    !uses_real_kernels(st, kHeatBath);
    e = launch_apply_labels(st->d_labels, st->n_rep, st->d_params, st->d_lab_params, rows,
                            rows ? st->d_lab_rows : nullptr, hb ? st->d_hbexp : nullptr, hb ? st->d_lab_hbexp : nullptr,
                            (uint32_t)((kMaxZ + 1) * (kMaxZ + 1)), s);
  }
  return e;
}

constexpr uint64_t kPtGraphSteps = 16;  // tempering steps per captured chunk

// Tempering steps whose sweep kernels read the device clock: O(N) on
// stencils (heat bath, over-relaxation, hybrid) and the register-blocked
// Ising Metropolis.
bool pt_graph_ok(const scamper_state *st, int alg) {
  if (st->csr) return false;
  if (st->continuous)
    return alg == kContinuousHeatBath || alg == kOverRelaxation || alg == kHybridHeatBathOverRelaxation;
  return st->model.kind == kIsing && alg == kMetropolis && ising_fast_path_ok(st->geom);
}

// One tempering step on `s` from the clock: the sweep, the exchange round
// (its energies are the step's measurement), the series record, the tick.
int enqueue_pt_step(scamper_state *st, int alg, int measure, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (st->continuous) {
    const int parts[2] = {alg == kOverRelaxation ? kOverRelaxation : kContinuousHeatBath,
                          alg == kHybridHeatBathOverRelaxation ? kOverRelaxation : -1};
    for (int pi = 0; pi < 2 && e == hipSuccess; ++pi)
      for (uint32_t c = 0; c < 2 && e == hipSuccess && parts[pi] >= 0; ++c)
        e = launch_on3_sweep(st->geom, st->model.spin_dim, parts[pi], st->d_vec, st->d_params, st->n_rep, c, 0,
                             (float)st->model.j, nullptr, s, st->d_clock);
  } else {
    for (uint32_t c = 0; c < 2 && e == hipSuccess; ++c)
      e = launch_ising_fast(st->geom, st->d_spins, st->d_params, st->n_rep, c, 0, nullptr, s, st->d_clock);
  }
  if (e == hipSuccess) e = enqueue_exchange(st, 0, st->d_clock, s);
  if (e == hipSuccess && measure)
    e = launch_series_from_em(st->d_em, st->d_clock, (uint64_t)st->n_rep * kEmStride, s);
  if (e == hipSuccess) e = launch_clock_tick(st->d_clock, 1, measure ? 1 : 0, s);
  if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "tempering step: %s", hipGetErrorString(e));
  return SCAMPER_OK;
}

// The captured chunk of kPtGraphSteps tempering steps for (algorithm, measure).
int pt_graph(scamper_state *st, int alg, int measure, hipGraphExec_t *out) {
  const int key = alg * 2 + (measure != 0);
  auto it = st->pt_graphs.find(key);
  if (it != st->pt_graphs.end()) {
    *out = it->second.second;
    return SCAMPER_OK;
  }
  if (!st->cap_stream) SC_HIP(hipStreamCreateWithFlags(&st->cap_stream, hipStreamNonBlocking));
  SC_HIP(hipStreamBeginCapture(st->cap_stream, hipStreamCaptureModeThreadLocal));
  int rc = SCAMPER_OK;
  for (uint64_t t = 0; t < kPtGraphSteps && rc == SCAMPER_OK; ++t) rc = enqueue_pt_step(st, alg, measure, st->cap_stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st->cap_stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "graph capture: %s", hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    hipGraphDestroy(g);
    return fail(SCAMPER_EDEVICE, "graph instantiation: %s", hipGetErrorString(ei));
  }
  st->pt_graphs[key] = {g, x};
  *out = x;
  return SCAMPER_OK;
}

// Build and upload the per-replica decision tables for `alg`.  With a
// temperature ladder the tables are built per label, and the device copies
// each replica's from the label it holds.
int upload_tables(scamper_state *st, int alg) {
  const bool ladder = !st->ladder.empty();
  std::vector<ReplicaParams> p(st->n_rep);
  std::vector<double> hb;
  const int z = st->csr ? st->w_max : st->z;
  const size_t hb_per = (size_t)(kMaxZ + 1) * (kMaxZ + 1);
  // the f64 graph kernels read only key and beta
  const bool f64_graph = uses_real_kernels(st, alg);
  if (st->model.kind == kPotts && alg == kHeatBath && !f64_graph) hb.assign((size_t)st->n_rep * hb_per, 0.0);
  // tables of replica r, or of label r with a ladder
  for (uint32_t r = 0; r < st->n_rep; ++r) {
    const double beta = ladder ? st->ladder[r] : st->beta[r];
    memset(&p[r], 0, sizeof(ReplicaParams));
    p[r].key = st->keys[r];
    p[r].beta = beta;
    if (f64_graph) {
    } else if (st->csr && st->model.kind == kIsing) {
      if (alg == kMetropolis) csr_ising_metropolis_table(beta, st->model.j, z, p[r].thr);
      else csr_ising_heat_bath_table(beta, st->model.j, z, p[r].thr);
    } else if (st->model.kind == kIsing) {
      if (alg == kMetropolis) ising_metropolis_table(beta, st->model.j, z, p[r].thr);
      else ising_heat_bath_table(beta, st->model.j, z, p[r].thr);
    } else if (st->continuous) {
      // the O(N) kernels read only key and beta
    } else if (st->model.kind == kPotts) {
      if (alg == kMetropolis) potts_metropolis_table(beta, st->model.j, z, p[r].thr);
      else potts_heat_bath_exp(beta, st->model.j, z, hb.data() + (size_t)r * hb_per);
    }
  }
  SC_HIP(hipStreamSynchronize(st->stream));
  SC_HIP(hipMemcpy(st->d_params, p.data(), p.size() * sizeof(ReplicaParams), hipMemcpyHostToDevice));
  if (ladder) {
    if (!st->d_lab_params) SC_HIP(hipMalloc(&st->d_lab_params, p.size() * sizeof(ReplicaParams)));
    SC_HIP(hipMemcpy(st->d_lab_params, p.data(), p.size() * sizeof(ReplicaParams), hipMemcpyHostToDevice));
  }
  std::vector<TileJob> rows;
  if (uses_tiles(st, alg)) {
    // one row per replica: every sweep of every replica, measured from 0
    st->tile_lane = tile_lane_sites((uint64_t)st->n_rep * st->geom.nh);
    const uint32_t per = (uint32_t)((st->geom.nh + tile_sites(st->tile_lane) - 1) / tile_sites(st->tile_lane));
    rows.resize(st->n_rep);
    for (uint32_t r = 0; r < st->n_rep; ++r) {
      TileJob &t = rows[r];
      memset(&t, 0, sizeof t);
      t.key = st->keys[r];
      t.spins = st->d_spins + 2ull * r * st->geom.cs;
      t.sweep_base = 0;
      t.sweep_end = ~0ull;
      t.therm = 0;
      t.div_lx2 = div_magic(st->geom.lx2);
      t.div_ly = div_magic(st->geom.ly);
      t.nd = (uint32_t)st->geom.nd;
      t.lx2 = st->geom.lx2;
      t.ly = st->geom.ly;
      t.lz = st->geom.lz;
      t.nh = (uint32_t)st->geom.nh;
      t.cs = (uint32_t)st->geom.cs;
      t.kind = (uint32_t)st->model.kind;
      t.q = (uint32_t)st->model.q;
      t.tile_start = r * per;
      t.n_tiles = per;
      t.em_row = r;
      std::memcpy(t.thr, p[r].thr, sizeof t.thr);
      fill_tile_luts(t);
    }
    TileHeader h{st->n_rep, st->n_rep * per, {0, 0}};
    if (!st->d_tiles) SC_HIP(hipMalloc(&st->d_tiles, sizeof(TileHeader) + rows.size() * sizeof(TileJob)));
    SC_HIP(hipMemcpy(st->d_tiles, &h, sizeof h, hipMemcpyHostToDevice));
    SC_HIP(hipMemcpy(static_cast<char *>(st->d_tiles) + sizeof(TileHeader), rows.data(),
                     rows.size() * sizeof(TileJob), hipMemcpyHostToDevice));
    st->tile_grid = std::min<uint32_t>((h.n_tiles + 7u) & ~7u, 2048u);
    if (ladder) {
      if (!st->d_lab_rows) SC_HIP(hipMalloc(&st->d_lab_rows, rows.size() * sizeof(TileJob)));
      SC_HIP(hipMemcpy(st->d_lab_rows, rows.data(), rows.size() * sizeof(TileJob), hipMemcpyHostToDevice));
    }
  }
  if (!hb.empty()) {
    if (!st->d_hbexp) SC_HIP(hipMalloc(&st->d_hbexp, hb.size() * sizeof(double)));
    SC_HIP(hipMemcpy(st->d_hbexp, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice));
    if (ladder) {
      if (!st->d_lab_hbexp) SC_HIP(hipMalloc(&st->d_lab_hbexp, hb.size() * sizeof(double)));
      SC_HIP(hipMemcpy(st->d_lab_hbexp, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  st->tables_for = alg;
  if (ladder) SC_HIP(apply_labels(st));
  return SCAMPER_OK;
}

// The measurement ring of the tiled path: up to 64 sweeps of per-XCD
// sub-records within 64 MiB.
int ensure_ring(scamper_state *st, int alg) {
  if (!uses_tiles(st, alg) || st->d_ring) return SCAMPER_OK;
  const uint64_t per_sweep = (uint64_t)st->n_rep * 8 * kEmStride * sizeof(long long);
  const uint64_t slots = std::max<uint64_t>(1, std::min<uint64_t>(64, (64ull << 20) / per_sweep));
  if (hipMalloc(&st->d_ring, slots * per_sweep) != hipSuccess) return fail(SCAMPER_ENOMEM, "measurement ring");
  st->ring_slots = (uint32_t)slots;
  return SCAMPER_OK;
}

int ensure_series(scamper_state *st, uint64_t extra) {
  const uint64_t need = st->series_len + extra;
  if (need <= st->series_cap) return SCAMPER_OK;
  uint64_t cap = std::max<uint64_t>(need, st->series_cap * 2);
  cap = std::max<uint64_t>(cap, 1024);
  long long *fresh = nullptr;
  const size_t slot = (size_t)st->n_rep * kEmStride * sizeof(long long);
  if (hipMalloc(&fresh, cap * slot) != hipSuccess) return fail(SCAMPER_ENOMEM, "series buffer of %llu sweeps", (unsigned long long)cap);
  if (st->series_len) SC_HIP(hipMemcpyAsync(fresh, st->d_series, st->series_len * slot, hipMemcpyDeviceToDevice, st->stream));
  SC_HIP(hipStreamSynchronize(st->stream));
  hipFree(st->d_series);
  st->d_series = fresh;
  st->series_cap = cap;
  return SCAMPER_OK;
}

// Site (x, row) of the stencil geometry -> colour-split byte offset.
inline uint64_t split_index(const StencilGeom &g, uint64_t site, uint32_t *colour) {
  const uint64_t lx = 2ull * g.lx2;
  const uint64_t x = site % lx, row = site / lx;
  const uint64_t y = row % g.ly, zz = row / g.ly;
  *colour = (uint32_t)((x + y + zz) & 1);
  return row * g.lx2 + (x >> 1);
}

int upload_i8_impl(scamper_state *st, uint32_t replica, const int8_t *spins) {
  const uint64_t n = st->lat.n_sites;
  for (uint64_t s = 0; s < n; ++s) {
    const int8_t v = spins[s];
    if (st->model.kind == kIsing ? (v != 1 && v != -1) : (v < 0 || v >= st->model.q))
      return fail(SCAMPER_EINVAL, "invalid spin at site %llu: %d", (unsigned long long)s, (int)v);
  }
  SC_HIP(hipStreamSynchronize(st->stream));
  if (st->csr) {
    SC_HIP(hipMemcpy(st->d_spins + (uint64_t)replica * n, spins, n, hipMemcpyHostToDevice));
    return SCAMPER_OK;
  }
  std::vector<int8_t> split(n);
  for (uint64_t s = 0; s < n; ++s) {
    uint32_t c;
    const uint64_t i = split_index(st->geom, s, &c);
    split[c * st->geom.nh + i] = spins[s];
  }
  int8_t *base = st->d_spins + 2ull * replica * st->geom.cs;
  SC_HIP(hipMemcpy(base, split.data(), st->geom.nh, hipMemcpyHostToDevice));
  SC_HIP(hipMemcpy(base + st->geom.cs, split.data() + st->geom.nh, st->geom.nh, hipMemcpyHostToDevice));
  return SCAMPER_OK;
}

int download_i8_impl(scamper_state *st, uint32_t replica, int8_t *spins) {
  const uint64_t n = st->lat.n_sites;
  SC_HIP(hipStreamSynchronize(st->stream));
  if (st->csr) {
    SC_HIP(hipMemcpy(spins, st->d_spins + (uint64_t)replica * n, n, hipMemcpyDeviceToHost));
    return SCAMPER_OK;
  }
  std::vector<int8_t> split(n);
  const int8_t *base = st->d_spins + 2ull * replica * st->geom.cs;
  SC_HIP(hipMemcpy(split.data(), base, st->geom.nh, hipMemcpyDeviceToHost));
  SC_HIP(hipMemcpy(split.data() + st->geom.nh, base + st->geom.cs, st->geom.nh, hipMemcpyDeviceToHost));
  for (uint64_t s = 0; s < n; ++s) {
    uint32_t c;
    const uint64_t i = split_index(st->geom, s, &c);
    spins[s] = split[c * st->geom.nh + i];
  }
  return SCAMPER_OK;
}

// Adds the O(N) counters of every replica into em (kEmStride doubles each).
hipError_t measure_on(const scamper_state *st, double *em) {
  return st->csr ? launch_on_csr_measure(st->ogeom, st->model.spin_dim, st->d_vec, st->n_rep, st->model.j, em,
                                         st->stream)
                 : launch_on3_measure(st->geom, st->model.spin_dim, st->d_vec, st->n_rep, (float)st->model.j, em,
                                      st->stream);
}

// Energy and magnetisation from the accumulated counters of one replica.
// On the stencil every bond is counted once (from its colour-1 end); on a
// graph both ends count it, and self-loops add the constant -J w.
void observables_from_em(const scamper_state *st, const long long *em, double *e, double *m) {
  const double n = (double)st->lat.n_sites;
  if (st->continuous) {
    // double sums {s.h (J included), Mx, My, Mz}; ONModel::magnetization
    // is |sum s| / N (models.rs:500-514)
    double d[4];
    std::memcpy(d, em, sizeof d);
    // a graph counts every bond from both ends; self-loops add -J w (s.s = 1)
    if (e) *e = st->csr ? -0.5 * d[0] - st->model.j * st->self_loop_weight : -d[0];
    if (m) *m = std::sqrt(d[1] * d[1] + d[2] * d[2] + d[3] * d[3]) / n;
    return;
  }
  if (e) {
    if (st->csr && st->real_weights) {
      double bonds;  // f64 weighted bond sum, every bond twice
      std::memcpy(&bonds, em, sizeof bonds);
      *e = -st->model.j * (0.5 * bonds + st->self_loop_weight);
    } else if (st->csr)
      *e = -st->model.j * ((double)(em[0] / 2) + st->self_loop_weight);
    else
      *e = -st->model.j * (double)em[0];
  }
  if (!m) return;
  if (st->model.kind == kIsing) {
    *m = std::fabs((double)em[1]) / n;
  } else {
    long long largest = 0;
    for (int v = 0; v < st->model.q; ++v) largest = std::max(largest, em[2 + v]);
    *m = ((double)st->model.q * (double)largest - n) / (n * (double)(st->model.q - 1));
  }
}

}  // namespace

extern "C" {

size_t scamper_last_error(char *buf, size_t len) {
  if (buf && len) {
    const size_t k = std::min(len - 1, g_last_error.size());
    memcpy(buf, g_last_error.data(), k);
    buf[k] = 0;
  }
  return g_last_error.size();
}

const char *scamper_version(void) { return "0.1.0 gfx950"; }

// ------------------------------------------------------------- lattices
static int wrap_lattice(Lattice &&l, scamper_lattice **out) {
  auto *h = new (std::nothrow) scamper_lattice;
  if (!h) return fail(SCAMPER_ENOMEM, "lattice handle");
  h->lat = std::move(l);
  *out = h;
  return SCAMPER_OK;
}

int scamper_lattice_hypercubic(const uint64_t *dims, int n_dims, int pbc, scamper_lattice **out) {
  if (!dims || !out) return fail(SCAMPER_EINVAL, "null argument");
  Lattice l;
  std::string err;
  if (!lattice_hypercubic(dims, n_dims, pbc != 0, l, err)) return fail(SCAMPER_EINVAL, "%s", err.c_str());
  return wrap_lattice(std::move(l), out);
}

int scamper_lattice_triangular(uint64_t lx, uint64_t ly, scamper_lattice **out) {
  if (!out) return fail(SCAMPER_EINVAL, "null argument");
  Lattice l;
  std::string err;
  if (!lattice_triangular(lx, ly, l, err)) return fail(SCAMPER_EINVAL, "%s", err.c_str());
  return wrap_lattice(std::move(l), out);
}

int scamper_lattice_honeycomb(uint64_t lx, uint64_t ly, scamper_lattice **out) {
  if (!out) return fail(SCAMPER_EINVAL, "null argument");
  Lattice l;
  std::string err;
  if (!lattice_honeycomb(lx, ly, l, err)) return fail(SCAMPER_EINVAL, "%s", err.c_str());
  return wrap_lattice(std::move(l), out);
}

int scamper_lattice_kagome(uint64_t lx, uint64_t ly, scamper_lattice **out) {
  if (!out) return fail(SCAMPER_EINVAL, "null argument");
  Lattice l;
  std::string err;
  if (!lattice_kagome(lx, ly, l, err)) return fail(SCAMPER_EINVAL, "%s", err.c_str());
  return wrap_lattice(std::move(l), out);
}

int scamper_lattice_from_edges(uint64_t n_sites, const uint64_t *src, const uint64_t *dst,
                               const double *weight, uint64_t n_edges, scamper_lattice **out) {
  if (!out || (n_edges && (!src || !dst))) return fail(SCAMPER_EINVAL, "null argument");
  std::vector<Bond> e(n_edges);
  for (uint64_t k = 0; k < n_edges; ++k) e[k] = Bond{src[k], dst[k], kBondGeneric, weight ? weight[k] : 1.0};
  Lattice l;
  std::string err;
  if (!lattice_from_edges(n_sites, std::move(e), l, err)) return fail(SCAMPER_EINVAL, "%s", err.c_str());
  return wrap_lattice(std::move(l), out);
}

int scamper_lattice_info(const scamper_lattice *h, uint64_t *n_sites, uint64_t *n_edges,
                         uint64_t *n_bonds, uint32_t *n_colours) {
  if (!h) return fail(SCAMPER_EINVAL, "null lattice");
  if (n_sites) *n_sites = h->lat.n_sites;
  if (n_edges) *n_edges = h->lat.n_edge_count;
  if (n_bonds) *n_bonds = h->lat.n_bond_count;
  if (n_colours) *n_colours = h->lat.n_colours;
  return SCAMPER_OK;
}

int scamper_lattice_csr(const scamper_lattice *h, uint64_t *offsets, uint64_t *neighbors, uint64_t *edge_ids) {
  if (!h) return fail(SCAMPER_EINVAL, "null lattice");
  std::string err;
  if (!materialize(const_cast<scamper_lattice *>(h)->lat, err)) return fail(SCAMPER_ENOMEM, "%s", err.c_str());
  const Lattice &l = h->lat;
  if (offsets) std::copy(l.offsets.begin(), l.offsets.end(), offsets);
  if (neighbors) std::copy(l.neighbors.begin(), l.neighbors.end(), neighbors);
  if (edge_ids) std::copy(l.edge_ids.begin(), l.edge_ids.end(), edge_ids);
  return SCAMPER_OK;
}

int scamper_lattice_edges(const scamper_lattice *h, uint64_t *source, uint64_t *target, uint8_t *kind,
                          double *weight) {
  if (!h) return fail(SCAMPER_EINVAL, "null lattice");
  std::string err;
  if (!materialize(const_cast<scamper_lattice *>(h)->lat, err)) return fail(SCAMPER_ENOMEM, "%s", err.c_str());
  const std::vector<Bond> &e = h->lat.edges;
  for (size_t k = 0; k < e.size(); ++k) {
    if (source) source[k] = e[k].source;
    if (target) target[k] = e[k].target;
    if (kind) kind[k] = e[k].kind;
    if (weight) weight[k] = e[k].weight;
  }
  return SCAMPER_OK;
}

int scamper_lattice_colours(const scamper_lattice *h, uint32_t *colour, uint64_t *class_index) {
  if (!h) return fail(SCAMPER_EINVAL, "null lattice");
  std::string err;
  if (!materialize(const_cast<scamper_lattice *>(h)->lat, err)) return fail(SCAMPER_ENOMEM, "%s", err.c_str());
  if (colour) std::copy(h->lat.colour.begin(), h->lat.colour.end(), colour);
  if (class_index) std::copy(h->lat.class_index.begin(), h->lat.class_index.end(), class_index);
  return SCAMPER_OK;
}

void scamper_lattice_free(scamper_lattice *h) { delete h; }

// ---------------------------------------------------------------- state
int scamper_state_create(const scamper_lattice *lh, const scamper_model *model, uint32_t n_replicas,
                         uint64_t seed, scamper_state **out) {
  if (!lh || !model || !out) return fail(SCAMPER_EINVAL, "null argument");
  if (n_replicas == 0 || n_replicas > 65535) return fail(SCAMPER_EINVAL, "n_replicas must be in 1..65535");
  if (!std::isfinite(model->j)) return fail(SCAMPER_EINVAL, "coupling must be finite");
  if (model->kind == kPotts && (model->q < 2 || model->q > kMaxQ))
    return fail(model->q < 2 ? SCAMPER_EINVAL : SCAMPER_EUNSUPPORTED, "Potts q=%d (GPU path supports 2..%d)", model->q, kMaxQ);
  if (model->kind != kIsing && model->kind != kPotts && model->kind != kON)
    return fail(SCAMPER_EINVAL, "unknown model kind %d", model->kind);
  const Lattice &l = lh->lat;
  if (model->kind == kON && model->spin_dim != 2 && model->spin_dim != 3)
    return fail(SCAMPER_EUNSUPPORTED, "O(N) spins run with N = 2 or 3 (got N = %d)", model->spin_dim);
  auto st = std::unique_ptr<scamper_state>(new (std::nothrow) scamper_state);
  if (!st) return fail(SCAMPER_ENOMEM, "state handle");
  st->model = *model;
  if (st->model.kind != kON) st->model.spin_dim = 1;
  st->n_rep = n_replicas;
  st->seed = seed;
  st->beta.assign(n_replicas, 1.0);
  st->keys.resize(n_replicas);
  for (uint32_t r = 0; r < n_replicas; ++r) st->keys[r] = stream_seed(seed, r, 0);
  SC_HIP(hipGetDevice(&st->device));
  SC_HIP(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
  if (l.stencil) {
    st->lat = l;
    st->geom.nd = l.nd;
    st->geom.lx2 = (uint32_t)(l.dims[0] / 2);
    st->geom.ly = (uint32_t)l.dims[1];
    st->geom.lz = (uint32_t)l.dims[2];
    st->geom.nh = l.n_sites / 2;
    st->geom.cs = aligned_colour(st->geom.nh);
    set_divisors(st->geom);
    st->z = 2 * l.nd;
  } else if (model->kind == kON) {
    // Graph lattice with O(N) spins: any finite weights, self-loops kept
    // aside (constant energy, never in a local field).
    if (l.n_sites >= (1ull << 32) || l.neighbors.size() >= (1ull << 32))
      return fail(SCAMPER_EUNSUPPORTED, "graph lattices are limited to 2^32 sites and incidences");
    st->csr = true;
    st->lat = l;
    std::vector<uint32_t> off(l.n_sites + 1, 0), nb;
    std::vector<double> w;
    nb.reserve(l.neighbors.size());
    w.reserve(l.neighbors.size());
    for (uint64_t s = 0; s < l.n_sites; ++s) {
      for (uint64_t i = l.offsets[s]; i < l.offsets[s + 1]; ++i) {
        const Bond &b = l.edges[l.edge_ids[i]];
        if (b.source == b.target) continue;
        nb.push_back((uint32_t)l.neighbors[i]);
        w.push_back(b.weight);
      }
      off[s + 1] = (uint32_t)nb.size();
    }
    for (const Bond &b : l.edges)
      if (b.source == b.target) st->self_loop_weight += b.weight;
    std::vector<uint32_t> cls(l.class_sites.begin(), l.class_sites.end());
    st->class_start = l.class_start;
    SC_HIP(hipMalloc(&st->d_offsets, off.size() * 4));
    SC_HIP(hipMalloc(&st->d_neighbors, std::max<size_t>(nb.size(), 1) * 4));
    SC_HIP(hipMalloc(&st->d_fweights, std::max<size_t>(w.size(), 1) * sizeof(double)));
    SC_HIP(hipMalloc(&st->d_class_sites, std::max<size_t>(cls.size(), 1) * 4));
    SC_HIP(hipMemcpy(st->d_offsets, off.data(), off.size() * 4, hipMemcpyHostToDevice));
    if (!nb.empty()) {
      SC_HIP(hipMemcpy(st->d_neighbors, nb.data(), nb.size() * 4, hipMemcpyHostToDevice));
      SC_HIP(hipMemcpy(st->d_fweights, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (!cls.empty()) SC_HIP(hipMemcpy(st->d_class_sites, cls.data(), cls.size() * 4, hipMemcpyHostToDevice));
    st->ogeom = OnCsrGeom{l.n_sites, st->d_offsets, st->d_neighbors, st->d_fweights, st->d_class_sites};
  } else {
    // Graph lattice.  Integer weights |w| <= 127 with weighted degree <= 32
    // take the integer-threshold kernels (self-loops kept aside: constant
    // energy, never in a local field); any other finite weights take the
    // f64 kernels, which also serve the Potts heat bath on weighted graphs.
    if (l.n_sites >= (1ull << 32) || l.neighbors.size() >= (1ull << 32))
      return fail(SCAMPER_EUNSUPPORTED, "graph lattices are limited to 2^32 sites and incidences");
    st->csr = true;
    st->lat = l;
    std::vector<uint32_t> off(l.n_sites + 1, 0), nb, roff(l.n_sites + 1, 0), rnb;
    std::vector<int8_t> w;
    std::vector<double> rw;
    nb.reserve(l.neighbors.size());
    w.reserve(l.neighbors.size());
    rnb.reserve(l.neighbors.size());
    rw.reserve(l.neighbors.size());
    for (uint64_t s = 0; s < l.n_sites; ++s) {
      int wsum = 0;
      for (uint64_t i = l.offsets[s]; i < l.offsets[s + 1]; ++i) {
        const Bond &b = l.edges[l.edge_ids[i]];
        if (!std::isfinite(b.weight)) return fail(SCAMPER_EINVAL, "bond weights must be finite");
        rnb.push_back((uint32_t)l.neighbors[i]);
        rw.push_back(b.weight);
        if (b.weight != std::floor(b.weight) || std::fabs(b.weight) > 127.0) st->real_weights = true;
        if (b.weight != 1.0) st->unit_weights = false;
        if (b.source == b.target) continue;
        nb.push_back((uint32_t)l.neighbors[i]);
        w.push_back(st->real_weights ? (int8_t)0 : (int8_t)b.weight);
        wsum += (int)std::fmin(std::fabs(b.weight), 1e6);
      }
      off[s + 1] = (uint32_t)nb.size();
      roff[s + 1] = (uint32_t)rnb.size();
      st->w_max = std::max(st->w_max, wsum);
    }
    if (st->w_max > kMaxZ) st->real_weights = true;
    for (const Bond &b : l.edges)
      if (b.source == b.target) st->self_loop_weight += b.weight;
    std::vector<uint32_t> cls(l.class_sites.begin(), l.class_sites.end());
    st->class_start = l.class_start;
    SC_HIP(hipMalloc(&st->d_class_sites, std::max<size_t>(cls.size(), 1) * 4));
    if (!cls.empty()) SC_HIP(hipMemcpy(st->d_class_sites, cls.data(), cls.size() * 4, hipMemcpyHostToDevice));
    SC_HIP(hipMalloc(&st->d_roffsets, roff.size() * 4));
    SC_HIP(hipMalloc(&st->d_rneighbors, std::max<size_t>(rnb.size(), 1) * 4));
    SC_HIP(hipMalloc(&st->d_rweights, std::max<size_t>(rw.size(), 1) * sizeof(double)));
    SC_HIP(hipMemcpy(st->d_roffsets, roff.data(), roff.size() * 4, hipMemcpyHostToDevice));
    if (!rnb.empty()) {
      SC_HIP(hipMemcpy(st->d_rneighbors, rnb.data(), rnb.size() * 4, hipMemcpyHostToDevice));
      SC_HIP(hipMemcpy(st->d_rweights, rw.data(), rw.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    st->rgeom = RealCsrGeom{l.n_sites, st->d_roffsets, st->d_rneighbors, st->d_rweights, st->d_class_sites};
    if (!st->real_weights) {
      SC_HIP(hipMalloc(&st->d_offsets, off.size() * 4));
      SC_HIP(hipMalloc(&st->d_neighbors, std::max<size_t>(nb.size(), 1) * 4));
      SC_HIP(hipMalloc(&st->d_weights, std::max<size_t>(w.size(), 1)));
      SC_HIP(hipMemcpy(st->d_offsets, off.data(), off.size() * 4, hipMemcpyHostToDevice));
      if (!nb.empty()) {
        SC_HIP(hipMemcpy(st->d_neighbors, nb.data(), nb.size() * 4, hipMemcpyHostToDevice));
        SC_HIP(hipMemcpy(st->d_weights, w.data(), w.size(), hipMemcpyHostToDevice));
      }
      st->cgeom = CsrGeom{l.n_sites, l.n_colours, st->w_max, st->d_offsets, st->d_neighbors, st->d_weights,
                          st->d_class_sites};
    }
  }
  st->continuous = model->kind == kON;
  if (st->continuous) {
    const size_t bytes = (size_t)n_replicas * l.n_sites * model->spin_dim * sizeof(float);
    if (hipMalloc(&st->d_vec, bytes) != hipSuccess) return fail(SCAMPER_ENOMEM, "spins: %zu bytes", bytes);
  } else {
    // graphs: site-major replicas; stencils: two colour arrays of cs bytes each
    const size_t bytes = (size_t)n_replicas * (st->csr ? l.n_sites : 2 * st->geom.cs);
    if (hipMalloc(&st->d_spins_alloc, bytes + 2 * kSpinSlack) != hipSuccess)
      return fail(SCAMPER_ENOMEM, "spins: %zu bytes", bytes);
    st->d_spins = st->d_spins_alloc + kSpinSlack;
  }
  SC_HIP(hipMalloc(&st->d_params, sizeof(ReplicaParams) * n_replicas));
  SC_HIP(hipMalloc(&st->d_em, sizeof(long long) * kEmStride * n_replicas));
  for (uint32_t r = 0; r < n_replicas; ++r) {
    hipError_t e = st->continuous ? (st->csr ? launch_on_csr_init(st->ogeom, st->model.spin_dim, st->d_vec, st->keys[r], r, 0, st->stream)
                                             : launch_on3_init(st->geom, st->model.spin_dim, st->d_vec, st->keys[r], r, 0, st->stream))
                   : st->csr ? launch_csr_init(l.n_sites, st->model.kind, st->model.q, st->d_spins, st->keys[r], r, 0, st->stream)
                             : launch_stencil_init(st->geom, st->model.kind, st->model.q, st->d_spins, st->keys[r], r, 0, st->stream);
    if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "init: %s", hipGetErrorString(e));
  }
  SC_HIP(hipStreamSynchronize(st->stream));
  *out = st.release();
  return SCAMPER_OK;
}

void scamper_state_free(scamper_state *st) { delete st; }

int scamper_state_set_beta(scamper_state *st, uint32_t replica, double beta) {
  if (int rc = check_state(st, replica)) return rc;
  if (!std::isfinite(beta) || beta < 0.0) return fail(SCAMPER_EINVAL, "beta must be finite and non-negative");
  // a beta set by hand ends tempering: every replica keeps the beta of the
  // label it holds now
  if (int rc = sync_labels(st)) return rc;
  st->ladder.clear();
  st->beta[replica] = beta;
  st->tables_for = -1;
  st->cthr_ready = false;
  return SCAMPER_OK;
}

int scamper_state_get_beta(const scamper_state *st, uint32_t replica, double *beta) {
  if (int rc = check_state(st, replica)) return rc;
  if (!beta) return fail(SCAMPER_EINVAL, "null argument");
  if (int rc = sync_labels(const_cast<scamper_state *>(st))) return rc;
  *beta = st->beta[replica];
  return SCAMPER_OK;
}

int scamper_state_set_ladder(scamper_state *st, const double *beta_by_label) {
  if (!st || !beta_by_label) return fail(SCAMPER_EINVAL, "null argument");
  if (st->n_rep > 1024) return fail(SCAMPER_EUNSUPPORTED, "replica exchange runs over at most 1024 replicas");
  for (uint32_t l = 0; l < st->n_rep; ++l)
    if (!std::isfinite(beta_by_label[l]) || beta_by_label[l] < 0.0)
      return fail(SCAMPER_EINVAL, "beta of label %u must be finite and non-negative", l);
  if (!st->d_labels) {
    SC_HIP(hipMalloc(&st->d_labels, st->n_rep * sizeof(int32_t)));
    SC_HIP(hipMalloc(&st->d_ladder, st->n_rep * sizeof(double)));
    SC_HIP(hipMalloc(&st->d_xenergy, st->n_rep * sizeof(double)));
    SC_HIP(hipMalloc(&st->d_xstats, 4 * sizeof(unsigned long long)));
  }
  st->ladder.assign(beta_by_label, beta_by_label + st->n_rep);
  st->labels.resize(st->n_rep);
  for (uint32_t r = 0; r < st->n_rep; ++r) {
    st->labels[r] = (int32_t)r;
    st->beta[r] = st->ladder[r];
  }
  SC_HIP(hipStreamSynchronize(st->stream));
  SC_HIP(hipMemcpy(st->d_labels, st->labels.data(), st->n_rep * sizeof(int32_t), hipMemcpyHostToDevice));
  SC_HIP(hipMemcpy(st->d_ladder, st->ladder.data(), st->n_rep * sizeof(double), hipMemcpyHostToDevice));
  SC_HIP(hipMemset(st->d_xenergy, 0, st->n_rep * sizeof(double)));
  SC_HIP(hipMemset(st->d_xstats, 0, 4 * sizeof(unsigned long long)));
  st->labels_dirty = false;
  st->tables_for = -1;
  st->cthr_ready = false;
  return SCAMPER_OK;
}

int scamper_state_exchange(scamper_state *st, uint64_t step) {
  if (!st) return fail(SCAMPER_EINVAL, "null state");
  if (st->ladder.empty()) return fail(SCAMPER_EINVAL, "no temperature ladder: call scamper_state_set_ladder first");
  const hipError_t e = enqueue_exchange(st, step, nullptr, st->stream);
  if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "exchange: %s", hipGetErrorString(e));
  st->labels_dirty = true;
  st->cthr_ready = false;
  return SCAMPER_OK;
}

int scamper_state_pt_step(scamper_state *st, int algorithm, uint64_t n_steps, int measure) {
  if (!st) return fail(SCAMPER_EINVAL, "null state");
  if (st->ladder.empty()) return fail(SCAMPER_EINVAL, "no temperature ladder: call scamper_state_set_ladder first");
  if (n_steps == 0) return SCAMPER_OK;
  if (measure)
    if (int rc = ensure_series(st, n_steps)) return rc;
  if (!st->d_clock) SC_HIP(hipMalloc(&st->d_clock, sizeof(StepClock)));
  const bool graphs = pt_graph_ok(st, algorithm);
  if (graphs) {
    const int table_alg = st->continuous ? kContinuousHeatBath : algorithm;
    if (st->tables_for != table_alg)
      if (int rc = upload_tables(st, table_alg)) return rc;
  }
  // the clock starts at the host's counters: sweep s, exchange round s + 1
  SC_HIP(launch_clock_set(st->d_clock, st->sweeps, st->sweeps + 1, st->series_len, st->d_series, st->stream));
  const uint64_t chunks = graphs ? n_steps / kPtGraphSteps : 0;
  if (chunks) {
    hipGraphExec_t x = nullptr;
    if (int rc = pt_graph(st, algorithm, measure, &x)) return rc;
    for (uint64_t k = 0; k < chunks; ++k) SC_HIP(hipGraphLaunch(x, st->stream));
    st->pt_graph_launches += chunks;
    st->sweeps += chunks * kPtGraphSteps;
    if (measure) st->series_len += chunks * kPtGraphSteps;
  }
  for (uint64_t t = chunks * kPtGraphSteps; t < n_steps; ++t) {
    if (graphs) {
      if (int rc = enqueue_pt_step(st, algorithm, measure, st->stream)) return rc;
      ++st->sweeps;
      if (measure) ++st->series_len;
      continue;
    }
    // kernels without the clock: the sweep by value, then the clock's round
    if (int rc = scamper_sweep(st, algorithm, 1, 0)) return rc;
    hipError_t e = enqueue_exchange(st, 0, st->d_clock, st->stream);
    if (e == hipSuccess && measure)
      e = launch_series_from_em(st->d_em, st->d_clock, (uint64_t)st->n_rep * kEmStride, st->stream);
    if (e == hipSuccess) e = launch_clock_tick(st->d_clock, 1, measure ? 1 : 0, st->stream);
    if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "tempering step: %s", hipGetErrorString(e));
    if (measure) ++st->series_len;
  }
  st->labels_dirty = true;
  st->cthr_ready = false;
  return SCAMPER_OK;
}

int scamper_state_pt_graph_launches(const scamper_state *st, uint64_t *n) {
  if (!st || !n) return fail(SCAMPER_EINVAL, "null argument");
  *n = st->pt_graph_launches;
  return SCAMPER_OK;
}

int scamper_state_labels(scamper_state *st, int32_t *labels) {
  if (!st || !labels) return fail(SCAMPER_EINVAL, "null argument");
  if (st->ladder.empty()) return fail(SCAMPER_EINVAL, "no temperature ladder");
  if (int rc = sync_labels(st)) return rc;
  std::copy(st->labels.begin(), st->labels.end(), labels);
  return SCAMPER_OK;
}

int scamper_state_exchange_energies(scamper_state *st, double *energy) {
  if (!st || !energy) return fail(SCAMPER_EINVAL, "null argument");
  if (st->ladder.empty()) return fail(SCAMPER_EINVAL, "no temperature ladder");
  SC_HIP(hipStreamSynchronize(st->stream));
  SC_HIP(hipMemcpy(energy, st->d_xenergy, st->n_rep * sizeof(double), hipMemcpyDeviceToHost));
  return SCAMPER_OK;
}

int scamper_state_exchange_stats(scamper_state *st, uint64_t *accepted, uint64_t *near_ties, uint64_t *rounds) {
  if (!st) return fail(SCAMPER_EINVAL, "null state");
  if (st->ladder.empty()) return fail(SCAMPER_EINVAL, "no temperature ladder");
  unsigned long long v[4];
  SC_HIP(hipStreamSynchronize(st->stream));
  SC_HIP(hipMemcpy(v, st->d_xstats, sizeof v, hipMemcpyDeviceToHost));
  if (accepted) *accepted = v[0];
  if (near_ties) *near_ties = v[1];
  if (rounds) *rounds = v[2];
  return SCAMPER_OK;
}

int scamper_state_init(scamper_state *st, uint32_t replica, int random) {
  if (int rc = check_state(st, replica)) return rc;
  hipError_t e = st->continuous ? (st->csr ? launch_on_csr_init(st->ogeom, st->model.spin_dim, st->d_vec, st->keys[replica], replica, random != 0, st->stream)
                                           : launch_on3_init(st->geom, st->model.spin_dim, st->d_vec, st->keys[replica], replica, random != 0, st->stream))
                 : st->csr ? launch_csr_init(st->lat.n_sites, st->model.kind, st->model.q, st->d_spins, st->keys[replica], replica, random != 0, st->stream)
                           : launch_stencil_init(st->geom, st->model.kind, st->model.q, st->d_spins, st->keys[replica], replica, random != 0, st->stream);
  if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "init: %s", hipGetErrorString(e));
  return SCAMPER_OK;
}

// O(N) replicas: D planes per colour, [c][component][rank], 2 D nh floats.
static int vec_upload(scamper_state *st, uint32_t replica, const float *xyz) {
  const uint64_t n = st->lat.n_sites, nh = st->geom.nh;
  const int d = st->model.spin_dim;
  if (st->csr) {  // graph lattice: planes [component][site]
    std::vector<float> planes((uint64_t)d * n);
    for (uint64_t s = 0; s < n; ++s)
      for (int k = 0; k < d; ++k) planes[(uint64_t)k * n + s] = xyz[(uint64_t)d * s + k];
    SC_HIP(hipStreamSynchronize(st->stream));
    SC_HIP(hipMemcpy(st->d_vec + (uint64_t)d * replica * n, planes.data(), planes.size() * sizeof(float),
                     hipMemcpyHostToDevice));
    return SCAMPER_OK;
  }
  std::vector<float> planes(2ull * d * nh);
  for (uint64_t s = 0; s < n; ++s) {
    uint32_t c;
    const uint64_t i = split_index(st->geom, s, &c);
    for (int k = 0; k < d; ++k) planes[((uint64_t)d * c + k) * nh + i] = xyz[(uint64_t)d * s + k];
  }
  SC_HIP(hipStreamSynchronize(st->stream));
  SC_HIP(hipMemcpy(st->d_vec + 2ull * d * replica * nh, planes.data(), planes.size() * sizeof(float),
                   hipMemcpyHostToDevice));
  return SCAMPER_OK;
}

static int vec_download(scamper_state *st, uint32_t replica, float *xyz) {
  const uint64_t n = st->lat.n_sites, nh = st->geom.nh;
  const int d = st->model.spin_dim;
  if (st->csr) {
    std::vector<float> planes((uint64_t)d * n);
    SC_HIP(hipStreamSynchronize(st->stream));
    SC_HIP(hipMemcpy(planes.data(), st->d_vec + (uint64_t)d * replica * n, planes.size() * sizeof(float),
                     hipMemcpyDeviceToHost));
    for (uint64_t s = 0; s < n; ++s)
      for (int k = 0; k < d; ++k) xyz[(uint64_t)d * s + k] = planes[(uint64_t)k * n + s];
    return SCAMPER_OK;
  }
  std::vector<float> planes(2ull * d * nh);
  SC_HIP(hipStreamSynchronize(st->stream));
  SC_HIP(hipMemcpy(planes.data(), st->d_vec + 2ull * d * replica * nh, planes.size() * sizeof(float),
                   hipMemcpyDeviceToHost));
  for (uint64_t s = 0; s < n; ++s) {
    uint32_t c;
    const uint64_t i = split_index(st->geom, s, &c);
    for (int k = 0; k < d; ++k) xyz[(uint64_t)d * s + k] = planes[((uint64_t)d * c + k) * nh + i];
  }
  return SCAMPER_OK;
}

// ONModel::validate_spin (models.rs:454-474): finite, |s|^2 within tol of
// one (the reference's 1e-10 D for f64 input; f32 input is checked to
// single-precision accuracy).
static int vec_validate(const double *xyz, uint64_t n, int d, double tol) {
  for (uint64_t s = 0; s < n; ++s) {
    double sq = 0.0;
    bool finite = true;
    for (int k = 0; k < d; ++k) {
      const double v = xyz[(uint64_t)d * s + k];
      finite = finite && std::isfinite(v);
      sq += v * v;
    }
    if (!finite || std::fabs(sq - 1.0) > tol)
      return fail(SCAMPER_EINVAL, "O(%d) spin at site %llu is not a finite unit vector", d, (unsigned long long)s);
  }
  return SCAMPER_OK;
}

int scamper_state_upload_i8(scamper_state *st, uint32_t replica, const int8_t *spins) {
  if (int rc = check_state(st, replica)) return rc;
  if (!spins) return fail(SCAMPER_EINVAL, "null spins");
  if (st->continuous) return fail(SCAMPER_EUNSUPPORTED, "int8 spins belong to Ising and Potts states");
  return upload_i8_impl(st, replica, spins);
}

int scamper_state_download_i8(scamper_state *st, uint32_t replica, int8_t *spins) {
  if (int rc = check_state(st, replica)) return rc;
  if (!spins) return fail(SCAMPER_EINVAL, "null spins");
  if (st->continuous) return fail(SCAMPER_EUNSUPPORTED, "int8 spins belong to Ising and Potts states");
  return download_i8_impl(st, replica, spins);
}

int scamper_state_upload_f64(scamper_state *st, uint32_t replica, const double *spins) {
  if (int rc = check_state(st, replica)) return rc;
  if (!spins) return fail(SCAMPER_EINVAL, "null spins");
  const uint64_t n = st->lat.n_sites;
  if (st->continuous) {
    const int d = st->model.spin_dim;
    if (int rc = vec_validate(spins, n, d, 1e-10 * d)) return rc;
    std::vector<float> f((uint64_t)d * n);
    for (uint64_t k = 0; k < f.size(); ++k) f[k] = (float)spins[k];
    return vec_upload(st, replica, f.data());
  }
  std::vector<int8_t> v(n);
  for (uint64_t s = 0; s < n; ++s) {
    const double x = spins[s];
    // IsingModel::validate_spin / PottsModel::validate_spin (models.rs:62-77, 238-255)
    const bool ok = st->model.kind == kIsing ? (x == 1.0 || x == -1.0)
                                             : (std::isfinite(x) && x == std::floor(x) && x >= 0.0 && x < st->model.q);
    if (!ok) return fail(SCAMPER_EINVAL, "invalid spin at site %llu: %g", (unsigned long long)s, x);
    v[s] = (int8_t)x;
  }
  return upload_i8_impl(st, replica, v.data());
}

int scamper_state_download_f64(scamper_state *st, uint32_t replica, double *spins) {
  if (int rc = check_state(st, replica)) return rc;
  if (!spins) return fail(SCAMPER_EINVAL, "null spins");
  if (st->continuous) {
    std::vector<float> f((uint64_t)st->model.spin_dim * st->lat.n_sites);
    if (int rc = vec_download(st, replica, f.data())) return rc;
    for (uint64_t k = 0; k < f.size(); ++k) spins[k] = (double)f[k];
    return SCAMPER_OK;
  }
  std::vector<int8_t> v(st->lat.n_sites);
  if (int rc = download_i8_impl(st, replica, v.data())) return rc;
  for (uint64_t s = 0; s < st->lat.n_sites; ++s) spins[s] = (double)v[s];
  return SCAMPER_OK;
}

int scamper_state_upload_f32(scamper_state *st, uint32_t replica, const float *spins) {
  if (int rc = check_state(st, replica)) return rc;
  if (!spins) return fail(SCAMPER_EINVAL, "null spins");
  if (!st->continuous) return fail(SCAMPER_EUNSUPPORTED, "f32 spins belong to O(N) states");
  std::vector<double> d((uint64_t)st->model.spin_dim * st->lat.n_sites);
  for (uint64_t k = 0; k < d.size(); ++k) d[k] = spins[k];
  if (int rc = vec_validate(d.data(), st->lat.n_sites, st->model.spin_dim, 1e-5)) return rc;
  return vec_upload(st, replica, spins);
}

int scamper_state_download_f32(scamper_state *st, uint32_t replica, float *spins) {
  if (int rc = check_state(st, replica)) return rc;
  if (!spins) return fail(SCAMPER_EINVAL, "null spins");
  if (!st->continuous) return fail(SCAMPER_EUNSUPPORTED, "f32 spins belong to O(N) states");
  return vec_download(st, replica, spins);
}

int scamper_state_sweeps(const scamper_state *st, uint64_t *sweeps) {
  if (!st || !sweeps) return fail(SCAMPER_EINVAL, "null argument");
  *sweeps = st->sweeps;
  return SCAMPER_OK;
}

int scamper_state_set_sweeps(scamper_state *st, uint64_t sweeps) {
  if (!st) return fail(SCAMPER_EINVAL, "null state");
  st->sweeps = sweeps;
  return SCAMPER_OK;
}

namespace {

// Buffers of the cluster updates, allocated on first use.
int ensure_cluster(scamper_state *st) {
  if (st->d_parent) return SCAMPER_OK;
  const uint64_t n = st->lat.n_sites;
  if (hipMalloc(&st->d_parent, std::max<uint64_t>(n, 1) * sizeof(uint32_t)) != hipSuccess)
    return fail(SCAMPER_ENOMEM, "cluster labels: %llu sites", (unsigned long long)n);
  SC_HIP(hipMalloc(&st->d_cscratch, 64));  // Wolff seed root and target; O(N): direction from byte 8
  SC_HIP(hipMalloc(&st->d_cthr, (size_t)st->n_rep * kClusterWeights * sizeof(uint64_t)));
  if (st->csr) {
    const std::vector<Bond> &e = st->lat.edges;
    std::vector<ClusterEdge> ce(e.size());
    std::vector<int8_t> cw(e.size());
    std::vector<double> cwf(e.size());
    for (size_t k = 0; k < e.size(); ++k) {
      ce[k] = ClusterEdge{(uint32_t)e[k].source, (uint32_t)e[k].target};
      // integer weights in [-127, 127] on the discrete graph path; graphs
      // with real weights read per-edge thresholds instead
      cw[k] = st->real_weights ? (int8_t)0 : (int8_t)e[k].weight;
      cwf[k] = e[k].weight;         // any weight for O(N)
    }
    if (st->continuous) {
      SC_HIP(hipMalloc(&st->d_cweights_f64, std::max<size_t>(cwf.size(), 1) * sizeof(double)));
      if (!cwf.empty())
        SC_HIP(hipMemcpy(st->d_cweights_f64, cwf.data(), cwf.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    SC_HIP(hipMalloc(&st->d_cedges, std::max<size_t>(ce.size(), 1) * sizeof(ClusterEdge)));
    SC_HIP(hipMalloc(&st->d_cweights, std::max<size_t>(cw.size(), 1)));
    if (!ce.empty()) {
      SC_HIP(hipMemcpy(st->d_cedges, ce.data(), ce.size() * sizeof(ClusterEdge), hipMemcpyHostToDevice));
      SC_HIP(hipMemcpy(st->d_cweights, cw.data(), cw.size(), hipMemcpyHostToDevice));
    }
  }
  return SCAMPER_OK;
}

// Buffers and bond thresholds (per replica and integer weight) of the
// cluster updates; the thresholds are rebuilt after a beta changes.
int prepare_cluster(scamper_state *st) {
  if (int rc = ensure_cluster(st)) return rc;
  if (int rc = sync_labels(st)) return rc;
  if (st->cthr_ready) return SCAMPER_OK;
  st->cthr.assign((size_t)st->n_rep * kClusterWeights, 0);
  for (uint32_t r = 0; r < st->n_rep; ++r)
    for (int w = 0; w < kClusterWeights; ++w)
      // by |w| with |J|: cluster updates require J w >= 0, so J w = |J| |w|
      st->cthr[(size_t)r * kClusterWeights + w] =
          cluster_bond_threshold(st->model.kind, st->beta[r], std::fabs(st->model.j), (double)w);
  SC_HIP(hipStreamSynchronize(st->stream));
  SC_HIP(hipMemcpy(st->d_cthr, st->cthr.data(), st->cthr.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  if (st->csr && st->real_weights) {
    // one threshold per edge: K(beta, |J| |w|), the same rule as by integer |w|
    const std::vector<Bond> &edges = st->lat.edges;
    const size_t ne = edges.size();
    if (!st->d_ethr && hipMalloc(&st->d_ethr, std::max<size_t>(ne * st->n_rep, 1) * sizeof(uint64_t)) != hipSuccess)
      return fail(SCAMPER_ENOMEM, "edge thresholds: %zu edges x %u replicas", ne, st->n_rep);
    std::vector<uint64_t> et(ne * st->n_rep);
    for (uint32_t r = 0; r < st->n_rep; ++r)
      for (size_t e = 0; e < ne; ++e)
        et[(size_t)r * ne + e] = cluster_bond_threshold(st->model.kind, st->beta[r], std::fabs(st->model.j),
                                                        std::fabs(edges[e].weight));
    if (!et.empty()) SC_HIP(hipMemcpy(st->d_ethr, et.data(), et.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  }
  st->cthr_ready = true;
  return SCAMPER_OK;
}

// n Swendsen-Wang or Wolff updates of every replica (SWCore / WolffCore,
// algorithms/swendsen_wang.rs:73-135, wolff.rs:72-134); each counts as one
// sweep of the draw counter.
int cluster_sweeps(scamper_state *st, int algorithm, uint64_t n_sweeps, int measure) {
  const uint64_t n = st->lat.n_sites;
  if (n >= (1ull << 32)) return fail(SCAMPER_EUNSUPPORTED, "cluster updates need fewer than 2^32 sites");
  // the reference asserts J * w >= 0 (models.rs:139-142, 321-324)
  if (st->csr) {
    for (const Bond &b : st->lat.edges)
      if (st->model.j * b.weight < 0.0)
        return fail(SCAMPER_EUNSUPPORTED, "cluster updates require J * bond.weight >= 0");
  } else if (st->model.j < 0.0) {
    return fail(SCAMPER_EUNSUPPORTED, "cluster updates require J * bond.weight >= 0");
  }
  // O(N) bonds depend on the spins: no threshold tables (their beta is
  // passed per launch)
  if (int rc = st->continuous ? ensure_cluster(st) : prepare_cluster(st)) return rc;
  if (int rc = sync_labels(st)) return rc;
  const std::vector<uint64_t> &thr = st->cthr;
  if (measure) {
    if (int rc = ensure_series(st, n_sweeps)) return rc;
    const size_t slot = (size_t)st->n_rep * kEmStride;
    SC_HIP(hipMemsetAsync(st->d_series + st->series_len * slot, 0, n_sweeps * slot * sizeof(long long), st->stream));
  }
  const int wolff = algorithm == kWolff;
  for (uint64_t t = 0; t < n_sweeps; ++t) {
    const uint64_t step = st->sweeps + t;
    for (uint32_t r = 0; r < st->n_rep; ++r) {
      hipError_t e;
      if (st->continuous && st->csr) {
        const int d = st->model.spin_dim;
        e = launch_cluster_on_csr(n, st->lat.edges.size(), st->d_cedges, st->d_cweights_f64, d,
                                  st->d_vec + (uint64_t)d * r * n, st->d_parent, st->keys[r], step, st->beta[r],
                                  st->model.j, wolff, st->d_cscratch, st->stream);
      } else if (st->continuous) {
        e = launch_cluster_on(st->geom, st->model.spin_dim, st->d_vec + 2ull * st->model.spin_dim * r * st->geom.nh,
                              st->d_parent, st->keys[r], step, st->beta[r], st->model.j, wolff, st->d_cscratch,
                              st->stream);
      } else if (st->csr) {
        int8_t *spins = st->d_spins + (uint64_t)r * n;
        e = launch_cluster_bonds_csr(n, st->lat.edges.size(), st->d_cedges, st->d_cweights, spins, st->d_parent,
                                     st->keys[r], step, st->d_cthr + (size_t)r * kClusterWeights,
                                     st->real_weights ? st->d_ethr + (size_t)r * st->lat.edges.size() : nullptr,
                                     st->stream);
        if (e == hipSuccess)
          e = launch_cluster_transform(nullptr, n, spins, st->d_parent, st->model.kind, st->model.q, wolff,
                                       st->keys[r], step, st->d_cscratch, st->stream);
      } else {
        int8_t *spins = st->d_spins + 2ull * r * st->geom.cs;
        e = launch_cluster_bonds_stencil(st->geom, spins, st->d_parent, st->keys[r], step,
                                         thr[(size_t)r * kClusterWeights + 1], st->stream);
        if (e == hipSuccess)
          e = launch_cluster_transform(&st->geom, n, spins, st->d_parent, st->model.kind, st->model.q, wolff,
                                       st->keys[r], step, st->d_cscratch, st->stream);
      }
      if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "cluster update launch: %s", hipGetErrorString(e));
    }
    if (measure) {
      long long *em = st->d_series + (st->series_len + t) * (size_t)st->n_rep * kEmStride;
      const hipError_t e =
          st->continuous ? measure_on(st, reinterpret_cast<double *>(em))
          : st->csr      ? measure_graph(st, em)
                         : launch_stencil_measure(st->geom, st->model.kind, st->model.q, st->d_spins, st->n_rep, em, st->stream);
      if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "measure launch: %s", hipGetErrorString(e));
    }
  }
  st->sweeps += n_sweeps;
  if (measure) st->series_len += n_sweeps;
  return SCAMPER_OK;
}

}  // namespace

// Bytes of O(N) stencil spins swept as one group of replicas (see
// scamper_sweep), inside the 256 MB Infinity Cache: O(3) at 96^3 x 32 (340 MB) runs as two
// groups of 16 replicas (123 -> 147 site-updates/ns); a state that already
// fits (XY at 96^3 x 32, 226 MB) stays one group, splitting it cost 2%.
constexpr uint64_t kReplicaBlockBytes = 240ull << 20;

int scamper_sweep(scamper_state *st, int algorithm, uint64_t n_sweeps, int measure) {
  if (!st) return fail(SCAMPER_EINVAL, "null state");
  if (algorithm == kSwendsenWang || algorithm == kWolff) return cluster_sweeps(st, algorithm, n_sweeps, measure);
  if (st->continuous) {
    if (algorithm != kContinuousHeatBath && algorithm != kOverRelaxation && algorithm != kHybridHeatBathOverRelaxation)
      return fail(SCAMPER_EUNSUPPORTED, "O(N) spins take the continuous heat bath, over-relaxation or their hybrid");
  } else if (algorithm != kMetropolis && algorithm != kHeatBath) {
    return fail(SCAMPER_EUNSUPPORTED, "algorithm %d needs continuous spins", algorithm);
  }
  const int table_alg = st->continuous ? kContinuousHeatBath : algorithm;
  if (st->tables_for != table_alg)
    if (int rc = upload_tables(st, table_alg)) return rc;
  if (measure) {
    if (int rc = ensure_series(st, n_sweeps)) return rc;
    const size_t slot = (size_t)st->n_rep * kEmStride;
    SC_HIP(hipMemsetAsync(st->d_series + st->series_len * slot, 0, n_sweeps * slot * sizeof(long long), st->stream));
    if (int rc = ensure_ring(st, algorithm)) return rc;
  }
  if (st->continuous && !st->csr) {
    // O(N) on a stencil: every sweep of a group of replicas before the next
    // group, each group small enough to stay in the 256 MB Infinity Cache
    // across its half-sweeps (32 O(3) replicas at 96^3 are 340 MB, which
    // would stream from HBM at every launch).  Draws are keyed by (replica,
    // sweep, colour, rank), so the order of groups changes no result.
    const int d = st->model.spin_dim;
    const uint64_t per_rep = 2ull * d * st->geom.nh * sizeof(float);
    const uint64_t groups = std::max<uint64_t>(1, (per_rep * st->n_rep + kReplicaBlockBytes - 1) / kReplicaBlockBytes);
    const uint32_t g_size = (uint32_t)((st->n_rep + groups - 1) / groups);
    // HybridCore: heat bath then over-relaxation each sweep (hybrid.rs:46-59)
    const int parts[2] = {algorithm == kOverRelaxation ? kOverRelaxation : kContinuousHeatBath,
                          algorithm == kHybridHeatBathOverRelaxation ? kOverRelaxation : -1};
    for (uint32_t r0 = 0; r0 < st->n_rep; r0 += g_size) {
      const uint32_t nr = std::min(g_size, st->n_rep - r0);
      float *spins = st->d_vec + (uint64_t)r0 * 2 * d * st->geom.nh;
      for (uint64_t t = 0; t < n_sweeps; ++t) {
        for (int pi = 0; pi < 2; ++pi) {
          if (parts[pi] < 0) continue;
          const bool last = pi == 1 || parts[1] < 0;
          for (uint32_t c = 0; c < 2; ++c) {
            double *em = (measure && last && c == 1)
                             ? reinterpret_cast<double *>(st->d_series + (st->series_len + t) * (size_t)st->n_rep * kEmStride) +
                                   (size_t)r0 * kEmStride
                             : nullptr;
            hipError_t e = launch_on3_sweep(st->geom, d, parts[pi], spins, st->d_params + r0, nr, c, st->sweeps + t,
                                            (float)st->model.j, em, st->stream);
            if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "sweep launch: %s", hipGetErrorString(e));
          }
        }
      }
    }
    st->sweeps += n_sweeps;
    if (measure) st->series_len += n_sweeps;
    return SCAMPER_OK;
  }
  for (uint64_t t = 0; t < n_sweeps; ++t) {
    const uint64_t sweep = st->sweeps + t;
    if (st->continuous) {  // a graph lattice (stencils took the grouped loop above)
      // HybridCore: heat bath then over-relaxation each sweep (hybrid.rs:46-59)
      const int parts[2] = {algorithm == kOverRelaxation ? kOverRelaxation : kContinuousHeatBath,
                            algorithm == kHybridHeatBathOverRelaxation ? kOverRelaxation : -1};
      // colour classes in order, then the measurement
      hipError_t e = hipSuccess;
      for (int pi = 0; pi < 2 && e == hipSuccess; ++pi)
        if (parts[pi] >= 0)
          e = launch_on_csr_sweep(st->ogeom, st->model.spin_dim, parts[pi], st->d_vec, st->d_params, st->n_rep,
                                  st->class_start.data(), (uint32_t)(st->class_start.size() - 1), sweep,
                                  st->model.j, st->stream);
      if (e == hipSuccess && measure)
        e = measure_on(st, reinterpret_cast<double *>(st->d_series + (st->series_len + t) * (size_t)st->n_rep * kEmStride));
      if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "sweep launch: %s", hipGetErrorString(e));
      continue;
    }
    if (st->csr) {
      hipError_t e =
          uses_real_kernels(st, algorithm)
              ? launch_csr_real(st->rgeom, st->class_start.data(), (uint32_t)(st->class_start.size() - 1),
                                st->model.kind, st->model.q, algorithm, st->d_spins, st->d_params, st->n_rep, sweep,
                                st->model.j, st->stream)
              : launch_csr_discrete(st->cgeom, st->class_start.data(), st->model.kind, st->model.q, algorithm,
                                    st->d_spins, st->d_params, st->n_rep, sweep, st->d_hbexp, st->model.j < 0.0,
                                    st->stream);
      if (e == hipSuccess && measure)
        e = measure_graph(st, st->d_series + (st->series_len + t) * (size_t)st->n_rep * kEmStride);
      if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "sweep launch: %s", hipGetErrorString(e));
      continue;
    }
    if (uses_tiles(st, algorithm)) {
      const TileHeader *hdr = static_cast<const TileHeader *>(st->d_tiles);
      const TileJob *rows = reinterpret_cast<const TileJob *>(hdr + 1);
      const uint64_t rec = 8ull * kEmStride;  // one replica's sub-records of one sweep
      const uint32_t slot = measure ? (uint32_t)(t % st->ring_slots) : 0u;
      hipError_t e = hipSuccess;
      if (measure && slot == 0)
        e = launch_em_zero(st->d_ring, std::min<uint64_t>(st->ring_slots, n_sweeps - t) * st->n_rep * rec, st->stream);
      const EmLayout em{st->d_ring + (uint64_t)slot * st->n_rep * rec, rec, 0, 7, (uint32_t)kEmStride};
      for (uint32_t c = 0; c < 2 && e == hipSuccess; ++c)
        e = launch_tile_sweep(hdr, rows, st->tile_grid, st->tile_lane, sweep, c, measure && c == 1, em, st->stream);
      if (e == hipSuccess && measure && (slot + 1 == st->ring_slots || t + 1 == n_sweeps))
        e = launch_em_fold(st->d_ring, st->d_series + (st->series_len + t - slot) * (size_t)st->n_rep * kEmStride,
                           (uint64_t)(slot + 1) * st->n_rep, (uint32_t)kEmStride, st->stream);
      if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "sweep launch: %s", hipGetErrorString(e));
      continue;
    }
    for (uint32_t c = 0; c < 2; ++c) {
      long long *em = (measure && c == 1) ? st->d_series + (st->series_len + t) * (size_t)st->n_rep * kEmStride : nullptr;
      hipError_t e = launch_stencil_discrete(st->geom, st->model.kind, st->model.q, algorithm, st->d_spins,
                                             st->d_params, st->n_rep, c, sweep, st->d_hbexp, st->model.j < 0.0, em,
                                             st->stream);
      if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "sweep launch: %s", hipGetErrorString(e));
    }
  }
  st->sweeps += n_sweeps;
  if (measure) st->series_len += n_sweeps;
  return SCAMPER_OK;
}

int scamper_sweep_with_phase(scamper_state *st, int algorithm, uint64_t n_sweeps, int measure, int phase) {
  // SimulationPhase (algorithms/common.rs:9-12): only adaptive proposals act
  // on it (allows_adaptation); every device algorithm is a fixed kernel
  if (phase != 0 && phase != 1) return fail(SCAMPER_EINVAL, "phase must be 0 (thermalization) or 1 (measurement)");
  return scamper_sweep(st, algorithm, n_sweeps, measure);
}

int scamper_sweep_timed(scamper_state *st, int algorithm, uint64_t n_sweeps, int measure, float *ms) {
  if (!st || !ms) return fail(SCAMPER_EINVAL, "null argument");
  // table uploads and series growth synchronise; do them before the events
  const int table_alg = st->continuous ? kContinuousHeatBath : algorithm;
  if (st->tables_for != table_alg && (st->continuous || algorithm == kMetropolis || algorithm == kHeatBath))
    if (int rc = upload_tables(st, table_alg)) return rc;
  if ((algorithm == kSwendsenWang || algorithm == kWolff) && st->lat.n_sites < (1ull << 32))
    if (int rc = st->continuous ? ensure_cluster(st) : prepare_cluster(st)) return rc;
  if (measure) {
    if (int rc = ensure_series(st, n_sweeps)) return rc;
    if (int rc = ensure_ring(st, algorithm)) return rc;
  }
  hipEvent_t a, b;
  SC_HIP(hipEventCreate(&a));
  SC_HIP(hipEventCreate(&b));
  SC_HIP(hipEventRecord(a, st->stream));
  int rc = scamper_sweep(st, algorithm, n_sweeps, measure);
  if (rc == SCAMPER_OK) {
    SC_HIP(hipEventRecord(b, st->stream));
    SC_HIP(hipEventSynchronize(b));
    SC_HIP(hipEventElapsedTime(ms, a, b));
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  return rc;
}

int scamper_synchronize(scamper_state *st) {
  if (!st) return fail(SCAMPER_EINVAL, "null state");
  SC_HIP(hipStreamSynchronize(st->stream));
  return SCAMPER_OK;
}

int scamper_state_observables(scamper_state *st, uint32_t replica, double *energy, double *magnetization) {
  if (int rc = check_state(st, replica)) return rc;
  SC_HIP(hipMemsetAsync(st->d_em, 0, sizeof(long long) * kEmStride * st->n_rep, st->stream));
  hipError_t e = st->continuous ? measure_on(st, reinterpret_cast<double *>(st->d_em))
                 : st->csr ? measure_graph(st, st->d_em)
                           : launch_stencil_measure(st->geom, st->model.kind, st->model.q, st->d_spins, st->n_rep, st->d_em, st->stream);
  if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "measure: %s", hipGetErrorString(e));
  long long em[kEmStride];
  SC_HIP(hipMemcpyAsync(em, st->d_em + (size_t)replica * kEmStride, sizeof em, hipMemcpyDeviceToHost, st->stream));
  SC_HIP(hipStreamSynchronize(st->stream));
  observables_from_em(st, em, energy, magnetization);
  return SCAMPER_OK;
}

int scamper_series_length(const scamper_state *st, uint64_t *n) {
  if (!st || !n) return fail(SCAMPER_EINVAL, "null argument");
  *n = st->series_len;
  return SCAMPER_OK;
}

int scamper_series_get(scamper_state *st, uint32_t replica, double *energy, double *magnetization, uint64_t max_n) {
  if (int rc = check_state(st, replica)) return rc;
  const uint64_t n = std::min(max_n, st->series_len);
  if (n == 0) return SCAMPER_OK;
  const size_t slot = (size_t)st->n_rep * kEmStride;
  std::vector<long long> buf(n * slot);
  SC_HIP(hipStreamSynchronize(st->stream));
  SC_HIP(hipMemcpy(buf.data(), st->d_series, buf.size() * sizeof(long long), hipMemcpyDeviceToHost));
  for (uint64_t t = 0; t < n; ++t)
    observables_from_em(st, buf.data() + t * slot + (size_t)replica * kEmStride, energy ? energy + t : nullptr,
                        magnetization ? magnetization + t : nullptr);
  return SCAMPER_OK;
}

int scamper_series_clear(scamper_state *st) {
  if (!st) return fail(SCAMPER_EINVAL, "null state");
  st->series_len = 0;
  return SCAMPER_OK;
}

}  // extern "C"
