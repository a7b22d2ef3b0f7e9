// Job pool: many independent Ising / Potts Metropolis runs advancing together
// on one GPU (continuous batching).
//
// The reference runs independent tasks with Scheduler::run_parallel ->
// Backend::spawn_tasks, one MonteCarlo + Context per task
// (Carlo.rs/src/scheduler.rs:345-405, Carlo.rs/src/backend/rayon.rs:36-49),
// each task being Scheduler::run_one (scheduler.rs:94-155).  Here a task is a
// row of a device-resident job table.  Every step updates colour 0 and then
// colour 1 of every running job in two launches of the tiled kernel; a chunk
// of `chunk_sweeps` steps is one hipGraph launch.  Between chunks the host
// admits queued jobs into freed spin memory and rows, and folds the measured
// counters of the chunk before last into each job's DefaultObservableSet
// accumulators, so the GPU always has one chunk queued while the host works.
//
// Memory safety of reuse: all work is on one stream.  A job's spins are
// released once its last sweep has been enqueued; a new job's initialisation
// is enqueued after that, so it runs after every launch that touched the
// region.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

#include "carlo.h"
#include "host.h"

using namespace scamper;

namespace {

constexpr uint32_t kRec = 8;   // counters per (job, step) record: em[0], em[1], counts of q <= 6 states
constexpr uint32_t kSubs = 8;  // per-XCD partial records folded after the chunk
constexpr uint32_t kMaxPoolQ = kRec - 2;
constexpr uint32_t kGridCandidates[] = {1024, 2048, 4096};
// Sites per lane follow the launch's resident colour-class sites: 16 from a
// crossover on, 4 below it.  The crossover (in Mi sites) is what the tuner
// chooses, so the lane width still follows the work as the set changes.
constexpr uint32_t kLaneCrossovers[] = {2, 1, 4};  // 2 Mi: the measured one (tile_lane_sites)
constexpr int kTuneRounds = 3;  // timed chunks per candidate; the median rate counts
// Tuning waits for a resident set of at least this many colour-class sites:
// on less work the shape hardly matters and its timings would not carry over
// to larger sets.  Crossovers are compared only where they give the set
// different lane widths, and grids only up to the set's tile count (larger
// ones are capped to it).
constexpr uint64_t kTuneMinSites = 1ull << 20;
// The grid is timed only on sets of at least this many sites: a grid that
// wins on a light set loses on the heavy ones that hold most of the work.
constexpr uint64_t kTuneGridSites = 8ull << 20;
// A stage's candidates are compared only on chunks of one set of resident
// jobs: a change of the set restarts the stage's timing, and after this many
// restarts the shape in use is kept.
constexpr int kMaxTuneRestarts = 16;
constexpr uint64_t kAlign = 256;
constexpr uint32_t kMaxBypass = 64;  // later jobs admitted past a waiting head

struct Job {
  uint64_t id = 0;
  scamper_job_spec spec{};
  StencilGeom geom{};
  uint64_t n_sites = 0, bytes = 0, off = 0;
  uint64_t key = 0, total = 0, enqueued = 0;
  int32_t state = SCAMPER_JOB_QUEUED;
  bool released_memory = false;
  Threshold thr[kStencilTable]{};
  std::unique_ptr<DefaultObservables> obs;
  std::vector<double> e_series, m_series;
  std::vector<int8_t> split_spins;  // copy of the final spins' two colour arrays (2 cs bytes)
  int8_t *h_spins = nullptr;        // pinned staging of that copy
  scamper_estimate est[5]{};
  ~Job() {
    if (h_spins) hipHostFree(h_spins);
  }
};

struct RowInfo {
  uint64_t id;
  uint64_t sweep_base;
  uint32_t steps;
};

struct InFlight {
  bool valid = false;
  std::vector<RowInfo> rows;
  uint64_t updates = 0;
  uint32_t steps = 0;  // chunk length of this chunk
  int candidate = -1;  // autotuner candidate timed by this chunk, an index into that stage's list
  int stage = 0;       // the tuning stage it was chosen in
  uint64_t set = 0;    // signature of the resident jobs it ran
};

// Launch shape of a chunk: blocks per launch, the lane-width crossover (Mi
// colour-class sites from which lanes hold 16 sites) and steps per chunk.
struct Shape {
  uint32_t grid, cross, chunk;
};

uint32_t lane_for(const Shape &s, uint64_t sites) { return sites >= ((uint64_t)s.cross << 20) ? 16u : 4u; }

// First-fit allocator over [0, size).
class Arena {
 public:
  void reset(uint64_t size) {
    free_.clear();
    if (size) free_[0] = size;
  }
  bool alloc(uint64_t len, uint64_t *off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < len) continue;
      *off = it->first;
      const uint64_t rest = it->second - len, at = it->first + len;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t len) {
    auto it = free_.emplace(off, len).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<uint64_t, uint64_t> free_;
};

}  // namespace

struct scamper_pool {
  scamper_pool_config cfg{};
  hipStream_t stream = nullptr;
  hipStream_t cap_stream = nullptr;  // graph capture only; graphs run on `stream`
  int8_t *spins = nullptr;        // spins_alloc + kSpinSlack; jobs at spins + offset
  int8_t *spins_alloc = nullptr;
  // bit planes of the jobs that run on them: a job's at byte offset off / 2
  // (its int8 region's half, which holds its four planes)
  uint32_t *bits = nullptr;
  Arena arena;
  uint64_t in_use = 0;
  size_t table_bytes = 0;
  char *d_table[2] = {nullptr, nullptr};
  char *h_table[2] = {nullptr, nullptr};
  long long *d_ring[2] = {nullptr, nullptr};
  long long *d_compact[2] = {nullptr, nullptr};
  long long *h_compact[2] = {nullptr, nullptr};
  // row of every bit-plane tile of the table (one load finds a tile's job)
  uint32_t *d_tmap[2] = {nullptr, nullptr};
  uint32_t *h_tmap[2] = {nullptr, nullptr};
  uint64_t tmap_cap = 0;
  // captured chunks per table buffer, keyed by launch shape (grid << 8 | lane)
  std::map<uint64_t, std::pair<hipGraph_t, hipGraphExec_t>> graphs[2];
  hipEvent_t ev_start[2] = {nullptr, nullptr}, ev_graph[2] = {nullptr, nullptr}, ev_end[2] = {nullptr, nullptr};
  InFlight inflight[2];
  uint64_t chunk = 0, next_id = 1;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  std::deque<uint64_t> queue;
  uint32_t head_bypassed = 0;
  std::vector<uint64_t> running;    // admitted jobs with sweeps left to enqueue
  std::vector<uint64_t> completed;  // finished, not yet reported
  uint64_t n_submitted = 0, n_done = 0, updates = 0;
  // autotuner: launch shapes timed on real chunks, one coordinate at a time
  // (stage 0 the lane width, 1 the grid at that lane width, 2 the chunk
  // length at both); results never depend on the shape
  Shape shape{0, 0, 0};  // in use (settled, or the fallback while tuning)
  bool tuned = false;
  int stage = -1;  // -1 until a resident set large enough to tune on
  std::vector<Shape> cands;
  std::vector<std::array<double, kTuneRounds>> cand_rates;
  std::vector<int> cand_samples;
  int next_candidate = 0;
  uint64_t cand_set = 0;  // signature of the resident jobs the stage's samples were timed on
  int tune_restarts = 0;
  uint64_t cur_sites = 0;  // resident colour-class sites of the latest chunk

  ~scamper_pool() {
    if (stream) hipStreamSynchronize(stream);
    for (int p = 0; p < 2; ++p) {
      for (auto &kv : graphs[p]) {
        hipGraphExecDestroy(kv.second.second);
        hipGraphDestroy(kv.second.first);
      }
      hipFree(d_table[p]);
      hipFree(d_ring[p]);
      hipFree(d_compact[p]);
      hipFree(d_tmap[p]);
      if (h_tmap[p]) hipHostFree(h_tmap[p]);
      if (h_table[p]) hipHostFree(h_table[p]);
      if (h_compact[p]) hipHostFree(h_compact[p]);
      if (ev_start[p]) hipEventDestroy(ev_start[p]);
      if (ev_graph[p]) hipEventDestroy(ev_graph[p]);
      if (ev_end[p]) hipEventDestroy(ev_end[p]);
    }
    hipFree(spins_alloc);
    hipFree(bits);
    if (cap_stream) hipStreamDestroy(cap_stream);
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

// Counters of one job row in the ring of a chunk of K steps (the fold reads
// rows of this length, so the layout follows each chunk's own K).
uint64_t ring_row_counters(uint32_t K) { return (uint64_t)K * kSubs * kRec; }

// The chunk's launches: zero the ring, pack the bit-plane jobs, K steps of
// (colour 0, colour 1) on the int8 and the bit-plane kernels, unpack, fold.
int enqueue_chunk_body(scamper_pool *P, int p, uint32_t grid, uint32_t lane, uint32_t K, bool int8_jobs,
                       bool bit_jobs, hipStream_t st) {
  const TileHeader *hdr = reinterpret_cast<const TileHeader *>(P->d_table[p]);
  const TileJob *rows = reinterpret_cast<const TileJob *>(hdr + 1);
  const EmLayout em{P->d_ring[p], ring_row_counters(K), (uint64_t)kSubs * kRec, kSubs - 1, kRec};
  hipError_t e = launch_tile_zero(hdr, P->d_ring[p], ring_row_counters(K), st);
  if (e == hipSuccess && bit_jobs) e = launch_bits_pack(hdr, rows, P->d_tmap[p], 0, st);
  for (uint32_t k = 0; k < K && e == hipSuccess; ++k) {
    for (uint32_t c = 0; c < 2 && e == hipSuccess; ++c) {
      if (int8_jobs) e = launch_tile_sweep(hdr, rows, grid, lane, k, c, (int)c, em, st);
      if (e == hipSuccess && bit_jobs) e = launch_tile_bits(hdr, rows, P->d_tmap[p], grid, k, c, (int)c, em, st);
    }
  }
  if (e == hipSuccess && bit_jobs) e = launch_bits_pack(hdr, rows, P->d_tmap[p], 1, st);
  if (e == hipSuccess) e = launch_tile_fold(hdr, P->d_ring[p], P->d_compact[p], K, st);
  if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "chunk launch: %s", hipGetErrorString(e));
  return SCAMPER_OK;
}

// The captured chunk for launch shape (grid, lane, K) and the kernels the
// table needs, on table buffer p, captured on first use.
int chunk_graph(scamper_pool *P, int p, uint32_t grid, uint32_t lane, uint32_t K, bool int8_jobs, bool bit_jobs,
                hipGraphExec_t *exec) {
  const uint64_t key = ((uint64_t)grid << 32) | ((uint64_t)K << 8) | lane | (int8_jobs ? 1ull << 5 : 0) |
                       (bit_jobs ? 1ull << 6 : 0);
  auto it = P->graphs[p].find(key);
  if (it != P->graphs[p].end()) {
    *exec = it->second.second;
    return SCAMPER_OK;
  }
  SC_HIP(hipStreamBeginCapture(P->cap_stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_chunk_body(P, p, grid, lane, K, int8_jobs, bit_jobs, P->cap_stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(P->cap_stream, &g);
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
  P->graphs[p][key] = {g, x};
  *exec = x;
  return SCAMPER_OK;
}

// E and M of one measured sweep from its folded counters (the stencil counts
// every bond once, from its colour-1 end).
void observables(const Job &J, const long long *em, double *e, double *m) {
  const double n = (double)J.n_sites;
  *e = -J.spec.model.j * (double)em[0];
  if (J.spec.model.kind == kIsing) {
    *m = std::fabs((double)em[1]) / n;  // IsingModel::magnetization (models.rs:103-110)
  } else {
    const int q = J.spec.model.q;
    long long largest = 0;
    for (int v = 0; v < q; ++v) largest = std::max(largest, em[2 + v]);
    *m = ((double)q * (double)largest - n) / (n * (double)(q - 1));  // models.rs:276-292
  }
}

// Site-major int8 from the colour-split layout.
void unsplit(const Job &J, const int8_t *split, int8_t *out) {
  const StencilGeom &g = J.geom;
  const uint64_t lx = 2ull * g.lx2;
  for (uint64_t s = 0; s < J.n_sites; ++s) {
    const uint64_t x = s % lx, row = s / lx;
    const uint64_t y = row % g.ly, zz = row / g.ly;
    const uint32_t c = (uint32_t)((x + y + zz) & 1);
    out[s] = split[c * g.cs + row * g.lx2 + (x >> 1)];
  }
}

void finish(scamper_pool *P, Job &J) {
  if (J.h_spins) {  // its copy completed before the chunk's end event
    J.split_spins.assign(J.h_spins, J.h_spins + 2 * J.geom.cs);
    hipHostFree(J.h_spins);
    J.h_spins = nullptr;
  }
  J.obs->finalize(J.est);
  J.state = SCAMPER_JOB_DONE;
  P->completed.push_back(J.id);
  ++P->n_done;
}

// Candidates of tuning stage `stage` around the current shape; past the last
// stage the shape is settled.
void start_stage(scamper_pool *P, int stage) {
  P->stage = stage;
  P->cands.clear();
  const Shape s = P->shape;
  if (stage == 0) {
    for (uint32_t c : kLaneCrossovers) {  // one candidate per distinct lane width on this set
      const Shape t{s.grid, c, s.chunk};
      bool seen = false;
      for (const Shape &o : P->cands) seen = seen || lane_for(o, P->cur_sites) == lane_for(t, P->cur_sites);
      if (!seen) P->cands.push_back(t);
    }
  } else if (stage == 1) {
    const uint32_t lane = lane_for(s, P->cur_sites);
    const uint64_t tiles = (P->cur_sites + tile_sites(lane) - 1) / tile_sites(lane);
    for (uint32_t g : kGridCandidates)
      if (g <= tiles) P->cands.push_back(Shape{g, s.cross, s.chunk});
  } else if (stage == 2)
    for (uint32_t k : {std::max(1u, P->cfg.chunk_sweeps / 4), std::max(1u, P->cfg.chunk_sweeps / 2), P->cfg.chunk_sweeps})
      if (P->cands.empty() || P->cands.back().chunk != k) P->cands.push_back(Shape{s.grid, s.cross, k});
  if (P->cands.empty()) {
    P->tuned = true;
    return;
  }
  if (P->cands.size() == 1) {  // nothing to compare
    P->shape = P->cands[0];
    start_stage(P, stage + 1);
    return;
  }
  P->cand_rates.assign(P->cands.size(), {});
  P->cand_samples.assign(P->cands.size(), 0);
  P->next_candidate = 0;
}

int collect(scamper_pool *P, int p, uint32_t *done) {
  InFlight &F = P->inflight[p];
  SC_HIP(hipEventSynchronize(P->ev_end[p]));
  // a sample counts only for the stage and the job set it was chosen for:
  // the other in-flight chunk may have closed that stage meanwhile, and the
  // next stage's list is shorter
  if (F.candidate >= 0 && !P->tuned && F.stage == P->stage && F.set == P->cand_set &&
      (size_t)F.candidate < P->cands.size()) {
    float ms = 0.f;
    SC_HIP(hipEventElapsedTime(&ms, P->ev_start[p], P->ev_graph[p]));
    if (ms > 0.f) {
      const double rate = (double)F.updates / (double)ms;
      if (P->cand_samples[F.candidate] < kTuneRounds) P->cand_rates[F.candidate][P->cand_samples[F.candidate]++] = rate;
    }
  }
  const uint32_t K = F.steps;
  for (size_t r = 0; r < F.rows.size(); ++r) {
    const RowInfo &ri = F.rows[r];
    Job &J = *P->jobs.at(ri.id);
    for (uint32_t k = 0; k < ri.steps; ++k) {
      if (ri.sweep_base + k < J.spec.thermalization_sweeps) continue;
      double e, m;
      observables(J, P->h_compact[p] + ((uint64_t)r * K + k) * kRec, &e, &m);
      J.obs->add(e, m);
      if (J.spec.keep_series) {
        J.e_series.push_back(e);
        J.m_series.push_back(m);
      }
    }
    if (ri.sweep_base + ri.steps == J.total) {
      finish(P, J);
      ++*done;
    }
  }
  F.valid = false;
  F.rows.clear();
  // close the tuning stage once every candidate has been timed
  if (!P->tuned && !P->cands.empty()) {
    bool all = true;
    for (size_t c = 0; c < P->cands.size(); ++c) all = all && P->cand_samples[c] >= kTuneRounds;
    if (all) {
      auto median = [](std::array<double, kTuneRounds> r) {
        std::sort(r.begin(), r.end());
        return r[kTuneRounds / 2];
      };
      size_t best = 0;
      for (size_t c = 1; c < P->cands.size(); ++c)
        if (median(P->cand_rates[c]) > median(P->cand_rates[best])) best = c;
      P->shape = P->cands[best];
      start_stage(P, P->stage + 1);
    }
  }
  return SCAMPER_OK;
}

int admit(scamper_pool *P, Job &J) {
  hipError_t e = launch_stencil_init(J.geom, J.spec.model.kind, J.spec.model.q, P->spins + J.off, J.key, 0,
                                     J.spec.random_start != 0, P->stream);
  if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "job init: %s", hipGetErrorString(e));
  J.state = SCAMPER_JOB_RUNNING;
  P->in_use += J.bytes;
  if (J.total == 0) {  // nothing to run: done at admission
    if (J.spec.keep_spins) {
      J.split_spins.resize(2 * J.geom.cs);
      SC_HIP(hipStreamSynchronize(P->stream));
      SC_HIP(hipMemcpy(J.split_spins.data(), P->spins + J.off, 2 * J.geom.cs, hipMemcpyDeviceToHost));
    }
    P->arena.release(J.off, J.bytes);
    P->in_use -= J.bytes;
    J.released_memory = true;
    finish(P, J);
    return SCAMPER_OK;
  }
  P->running.push_back(J.id);
  return SCAMPER_OK;
}

int admit_queued(scamper_pool *P) {
  for (auto it = P->queue.begin(); it != P->queue.end();) {
    if (P->running.size() >= P->cfg.max_jobs) break;
    Job &J = *P->jobs.at(*it);
    const bool head = it == P->queue.begin();
    if (!head && P->head_bypassed >= kMaxBypass) break;  // let the head through first
    uint64_t off;
    if (!P->arena.alloc(J.bytes, &off)) {
      ++it;
      continue;
    }
    J.off = off;
    if (head) P->head_bypassed = 0; else ++P->head_bypassed;
    if (int rc = admit(P, J)) return rc;
    it = P->queue.erase(it);
  }
  return SCAMPER_OK;
}

// The shape of the next chunk: the settled one, or a candidate of the
// current tuning stage that still needs timing (*candidate = its index).
// `set` identifies the resident jobs: when it differs from the set the
// stage's samples were timed on, their timing starts again (rates of
// different job sets do not compare), up to kMaxTuneRestarts times.
Shape choose_shape(scamper_pool *P, uint64_t set, int *candidate) {
  *candidate = -1;
  if (P->tuned || P->cur_sites < (P->stage == 1 ? kTuneGridSites : kTuneMinSites)) return P->shape;
  if (P->stage < 0) {
    start_stage(P, 0);
    P->cand_set = set;
    if (P->tuned) return P->shape;
  }
  if (set != P->cand_set) {
    const bool had_samples = std::any_of(P->cand_samples.begin(), P->cand_samples.end(), [](int k) { return k > 0; });
    P->cand_set = set;
    if (had_samples) {
      if (++P->tune_restarts > kMaxTuneRestarts) {
        P->tuned = true;
        return P->shape;
      }
      std::fill(P->cand_rates.begin(), P->cand_rates.end(), std::array<double, kTuneRounds>{});
      std::fill(P->cand_samples.begin(), P->cand_samples.end(), 0);
      P->next_candidate = 0;
    }
  }
  const int n = (int)P->cands.size();
  for (int tries = 0; tries < n; ++tries) {
    const int c = P->next_candidate;
    P->next_candidate = (P->next_candidate + 1) % n;
    if (P->cand_samples[c] < kTuneRounds) {
      *candidate = c;
      return P->cands[c];
    }
  }
  return P->shape;
}

int step(scamper_pool *P, uint32_t *n_completed) {
  const uint64_t done_before = P->n_done;
  uint32_t done = 0;
  const int p = (int)(P->chunk & 1);
  if (P->inflight[p].valid)
    if (int rc = collect(P, p, &done)) return rc;
  // release the spins of jobs whose last sweep is already enqueued
  for (auto it = P->running.begin(); it != P->running.end();) {
    Job &J = *P->jobs.at(*it);
    if (J.enqueued < J.total) {
      ++it;
      continue;
    }
    P->arena.release(J.off, J.bytes);
    P->in_use -= J.bytes;
    J.released_memory = true;
    it = P->running.erase(it);
  }
  if (int rc = admit_queued(P)) return rc;

  // the job table of this chunk, in the chunk's launch shape
  uint64_t set = 1469598103934665603ull;  // FNV-1a over the resident job ids, in order
  uint64_t total_sites = 0;
  for (uint64_t id : P->running) {
    set = (set ^ id) * 1099511628211ull;
    total_sites += P->jobs.at(id)->geom.nh;
  }
  P->cur_sites = total_sites;
  int candidate;
  const Shape shape = choose_shape(P, set, &candidate);
  const uint32_t K = shape.chunk;
  TileHeader *hdr = reinterpret_cast<TileHeader *>(P->h_table[p]);
  TileJob *rows = reinterpret_cast<TileJob *>(hdr + 1);
  InFlight &F = P->inflight[p];
  F.rows.clear();
  F.updates = 0;
  uint32_t tiles = 0, bit_tiles = 0;
  const uint32_t lane = lane_for(shape, total_sites);
  for (uint64_t id : P->running) {
    Job &J = *P->jobs.at(id);
    const uint32_t steps = (uint32_t)std::min<uint64_t>(K, J.total - J.enqueued);
    TileJob &t = rows[F.rows.size()];
    memset(&t, 0, sizeof t);
    t.key = J.key;
    t.spins = P->spins + J.off;
    t.sweep_base = J.enqueued;
    t.sweep_end = J.total;
    t.therm = J.spec.thermalization_sweeps;
    t.div_lx2 = div_magic(J.geom.lx2);
    t.div_ly = div_magic(J.geom.ly);
    t.nd = (uint32_t)J.geom.nd;
    t.lx2 = J.geom.lx2;
    t.ly = J.geom.ly;
    t.lz = J.geom.lz;
    t.nh = (uint32_t)J.geom.nh;
    t.cs = (uint32_t)J.geom.cs;
    t.kind = (uint32_t)J.spec.model.kind;
    t.q = (uint32_t)J.spec.model.q;
    t.tile_start = tiles;
    t.bit_tile_start = bit_tiles;
    t.em_row = (uint32_t)F.rows.size();
    std::memcpy(t.thr, J.thr, sizeof t.thr);
    fill_tile_luts(t);
    if (P->bits && (J.spec.model.kind == kIsing || J.spec.model.q <= 4)) {
      const uint32_t words = (uint32_t)((J.geom.nh + 31) / 32);
      t.bits = P->bits + J.off / 8;  // byte offset off / 2
      t.nwp = words + 2;
      // planes are the job's state from its first chunk to its last
      t.bit_io = (J.enqueued == 0 ? 1u : 0u) | (J.enqueued + steps == J.total ? 2u : 0u);
      fill_bit_lut(t);
      const uint32_t n = (words + 255) / 256;
      if (bit_tiles + (uint64_t)n > P->tmap_cap) return fail(SCAMPER_EINVARIANT, "bit tile map overflow");
      for (uint32_t k = 0; k < n; ++k) P->h_tmap[p][bit_tiles + k] = (uint32_t)F.rows.size();
      bit_tiles += n;
    } else {
      t.n_tiles = (uint32_t)((J.geom.nh + tile_sites(lane) - 1) / tile_sites(lane));
      tiles += t.n_tiles;
    }
    F.rows.push_back(RowInfo{id, J.enqueued, steps});
    F.updates += J.n_sites * steps;
  }
  if (F.rows.empty()) {
    F.valid = false;
    ++P->chunk;
    if (n_completed) *n_completed = (uint32_t)(P->n_done - done_before);
    return SCAMPER_OK;
  }
  hdr->n_jobs = (uint32_t)F.rows.size();
  hdr->n_tiles = tiles;
  hdr->pad[0] = bit_tiles;
  hdr->pad[1] = F.rows.size() > 1 ? 1u : 0u;  // the int8 kernel deals several jobs' tiles over all XCDs
  const bool int8_jobs = tiles > 0, bit_jobs = bit_tiles > 0;
  uint32_t grid = shape.grid;
  // no more blocks than tiles (rounded up to a power of two, so that few
  // launch shapes and captured graphs arise as the table shrinks and grows)
  const uint32_t work = std::max(tiles, bit_tiles);
  uint32_t fit = 8;
  while (fit < work && fit < grid) fit <<= 1;
  grid = std::min(grid, fit);
  hipGraphExec_t exec = nullptr;
  if (P->cfg.use_graphs)
    if (int rc = chunk_graph(P, p, grid, lane, K, int8_jobs, bit_jobs, &exec)) return rc;
  if (bit_tiles)
    SC_HIP(hipMemcpyAsync(P->d_tmap[p], P->h_tmap[p], (size_t)bit_tiles * sizeof(uint32_t), hipMemcpyHostToDevice,
                          P->stream));
  SC_HIP(hipMemcpyAsync(P->d_table[p], P->h_table[p], sizeof(TileHeader) + F.rows.size() * sizeof(TileJob),
                        hipMemcpyHostToDevice, P->stream));
  SC_HIP(hipEventRecord(P->ev_start[p], P->stream));
  if (P->cfg.use_graphs) {
    SC_HIP(hipGraphLaunch(exec, P->stream));
  } else if (int rc = enqueue_chunk_body(P, p, grid, lane, K, int8_jobs, bit_jobs, P->stream)) {
    return rc;
  }
  SC_HIP(hipEventRecord(P->ev_graph[p], P->stream));
  SC_HIP(hipMemcpyAsync(P->h_compact[p], P->d_compact[p], F.rows.size() * (size_t)K * kRec * sizeof(long long),
                        hipMemcpyDeviceToHost, P->stream));
  for (const RowInfo &ri : F.rows) {
    Job &J = *P->jobs.at(ri.id);
    J.enqueued += ri.steps;
    if (J.enqueued == J.total && J.spec.keep_spins) {
      // asynchronous copies go to pinned memory only; finish() moves it
      if (hipHostMalloc(&J.h_spins, 2 * J.geom.cs, hipHostMallocDefault) != hipSuccess)
        return fail(SCAMPER_ENOMEM, "pinned staging of %llu final spins", (unsigned long long)J.n_sites);
      SC_HIP(hipMemcpyAsync(J.h_spins, P->spins + J.off, 2 * J.geom.cs, hipMemcpyDeviceToHost, P->stream));
    }
  }
  SC_HIP(hipEventRecord(P->ev_end[p], P->stream));
  F.valid = true;
  F.steps = K;
  F.candidate = candidate;
  F.stage = P->stage;
  F.set = set;
  P->updates += F.updates;
  ++P->chunk;
  if (n_completed) *n_completed = (uint32_t)(P->n_done - done_before);
  return SCAMPER_OK;
}

bool idle(const scamper_pool *P) {
  return P->queue.empty() && P->running.empty() && !P->inflight[0].valid && !P->inflight[1].valid;
}

}  // namespace

extern "C" {

int scamper_pool_create(const scamper_pool_config *cfg, scamper_pool **out) {
  if (!cfg || !out) return fail(SCAMPER_EINVAL, "null argument");
  scamper_pool_config c = *cfg;
  if (c.chunk_sweeps == 0) c.chunk_sweeps = 32;
  if (c.chunk_sweeps > 1024) return fail(SCAMPER_EINVAL, "chunk_sweeps must be in 1..1024");
  if (c.max_jobs == 0 || c.max_jobs > 4096) return fail(SCAMPER_EINVAL, "max_jobs must be in 1..4096");
  if (c.grid & 7u) return fail(SCAMPER_EINVAL, "grid must be a multiple of 8");
  if (c.spin_budget_bytes == 0) return fail(SCAMPER_EINVAL, "spin budget must be positive");
  auto P = std::unique_ptr<scamper_pool>(new (std::nothrow) scamper_pool);
  if (!P) return fail(SCAMPER_ENOMEM, "pool handle");
  c.spin_budget_bytes = (c.spin_budget_bytes + kAlign - 1) / kAlign * kAlign;
  P->cfg = c;
  // a configured grid fixes the shape (lanes by size, the configured
  // chunk); otherwise the autotuner starts from 4096 blocks (capped at the
  // table's tiles: blocks that finish early take the remaining tiles, which
  // balances jobs of different costs)
  P->shape = Shape{c.grid ? c.grid : kGridCandidates[2], kLaneCrossovers[0], c.chunk_sweeps};
  if (c.grid) P->tuned = true;
  SC_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
  SC_HIP(hipStreamCreateWithFlags(&P->cap_stream, hipStreamNonBlocking));
  if (hipMalloc(&P->spins_alloc, c.spin_budget_bytes + 2 * kSpinSlack) != hipSuccess)
    return fail(SCAMPER_ENOMEM, "spin budget of %llu bytes", (unsigned long long)c.spin_budget_bytes);
  P->spins = P->spins_alloc + kSpinSlack;
  if (!c.int8_only && hipMalloc(&P->bits, c.spin_budget_bytes / 2 + 256) != hipSuccess)
    return fail(SCAMPER_ENOMEM, "bit planes of %llu bytes", (unsigned long long)(c.spin_budget_bytes / 2));
  // resident jobs hold at most budget / 2 colour-class sites in all, in
  // words of 32 (one partial word per job) and tiles of 256 words (one
  // partial tile per job)
  // (and at least one entry per block of any launch: the kernel reads the
  // entry of its block's first tile early)
  P->tmap_cap = c.int8_only ? 0
                            : std::max<uint64_t>(c.spin_budget_bytes / (2ull * 8192) + 2ull * c.max_jobs + 16,
                                                 std::max<uint32_t>(c.grid, kGridCandidates[2]));
  for (int p = 0; p < 2 && P->tmap_cap; ++p)
    if (hipMalloc(&P->d_tmap[p], P->tmap_cap * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc(&P->h_tmap[p], P->tmap_cap * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
      return fail(SCAMPER_ENOMEM, "bit tile map");
  P->arena.reset(c.spin_budget_bytes);
  P->table_bytes = sizeof(TileHeader) + (size_t)c.max_jobs * sizeof(TileJob);
  const size_t ring = (size_t)c.max_jobs * c.chunk_sweeps * kSubs * kRec * sizeof(long long);
  const size_t compact = (size_t)c.max_jobs * c.chunk_sweeps * kRec * sizeof(long long);
  for (int p = 0; p < 2; ++p) {
    if (hipMalloc(&P->d_table[p], P->table_bytes) != hipSuccess || hipMalloc(&P->d_ring[p], ring) != hipSuccess ||
        hipMalloc(&P->d_compact[p], compact) != hipSuccess)
      return fail(SCAMPER_ENOMEM, "pool tables");
    if (hipHostMalloc(&P->h_table[p], P->table_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&P->h_compact[p], compact, hipHostMallocDefault) != hipSuccess)
      return fail(SCAMPER_ENOMEM, "pinned pool buffers");
    SC_HIP(hipEventCreate(&P->ev_start[p]));
    SC_HIP(hipEventCreate(&P->ev_graph[p]));
    SC_HIP(hipEventCreate(&P->ev_end[p]));
  }
  *out = P.release();
  return SCAMPER_OK;
}

void scamper_pool_free(scamper_pool *pool) { delete pool; }

int scamper_pool_submit(scamper_pool *P, const scamper_job_spec *spec, uint64_t *job_id) {
  if (!P || !spec || !job_id) return fail(SCAMPER_EINVAL, "null argument");
  const scamper_job_spec &s = *spec;
  if (s.n_dims < 1 || s.n_dims > 3) return fail(SCAMPER_EINVAL, "jobs are 1-, 2- or 3-dimensional");
  uint64_t n = 1;
  for (uint32_t d = 0; d < s.n_dims; ++d) {
    if (s.dims[d] < 2 || (s.dims[d] & 1)) return fail(SCAMPER_EUNSUPPORTED, "job sides must be even and >= 2");
    if (s.dims[d] > (1ull << 31)) return fail(SCAMPER_EINVAL, "job side too large");
    n *= s.dims[d];
    if (n > (1ull << 32)) return fail(SCAMPER_EUNSUPPORTED, "jobs are limited to 2^32 sites");
  }
  if (s.model.kind != kIsing && s.model.kind != kPotts) return fail(SCAMPER_EUNSUPPORTED, "jobs are Ising or Potts");
  if (s.model.kind == kPotts && (s.model.q < 2 || (uint32_t)s.model.q > kMaxPoolQ))
    return fail(SCAMPER_EUNSUPPORTED, "pool jobs take Potts q in 2..%u", kMaxPoolQ);
  if (!std::isfinite(s.model.j) || !std::isfinite(s.beta)) return fail(SCAMPER_EINVAL, "coupling and beta must be finite");
  if (s.algorithm != kMetropolis) return fail(SCAMPER_EUNSUPPORTED, "pool jobs run Metropolis");
  if (s.binsize == 0) return fail(SCAMPER_EINVAL, "binsize must be positive");
  auto J = std::unique_ptr<Job>(new Job);
  J->id = P->next_id++;
  J->spec = s;
  J->spec.model.spin_dim = 1;
  J->geom.nd = (int)s.n_dims;
  J->geom.lx2 = (uint32_t)(s.dims[0] / 2);
  J->geom.ly = s.n_dims >= 2 ? (uint32_t)s.dims[1] : 1u;
  J->geom.lz = s.n_dims >= 3 ? (uint32_t)s.dims[2] : 1u;
  J->geom.nh = n / 2;
  set_divisors(J->geom);
  J->n_sites = n;
  J->geom.cs = aligned_colour(J->geom.nh);  // colour arrays at aligned offsets
  J->bytes = (2 * J->geom.cs + kAlign - 1) / kAlign * kAlign;
  if (J->bytes > P->cfg.spin_budget_bytes)
    return fail(SCAMPER_EBUDGET, "job needs %llu spin bytes, pool budget is %llu", (unsigned long long)J->bytes,
                (unsigned long long)P->cfg.spin_budget_bytes);
  J->key = stream_seed(s.seed, 0, 0);
  J->total = s.thermalization_sweeps + s.measurement_sweeps;
  const int z = 2 * (int)s.n_dims;
  if (s.model.kind == kIsing) ising_metropolis_table(s.beta, s.model.j, z, J->thr);
  else potts_metropolis_table(s.beta, s.model.j, z, J->thr);
  J->obs.reset(new DefaultObservables(s.binsize));
  *job_id = J->id;
  P->queue.push_back(J->id);
  P->jobs.emplace(J->id, std::move(J));
  ++P->n_submitted;
  return SCAMPER_OK;
}

int scamper_pool_step(scamper_pool *P, uint32_t *n_completed) {
  if (!P) return fail(SCAMPER_EINVAL, "null pool");
  return step(P, n_completed);
}

int scamper_pool_run(scamper_pool *P) {
  if (!P) return fail(SCAMPER_EINVAL, "null pool");
  while (!idle(P))
    if (int rc = step(P, nullptr)) return rc;
  return SCAMPER_OK;
}

int scamper_pool_completed(scamper_pool *P, uint64_t *ids, uint64_t cap, uint64_t *n) {
  if (!P || !n || (cap && !ids)) return fail(SCAMPER_EINVAL, "null argument");
  const uint64_t k = std::min<uint64_t>(cap, P->completed.size());
  for (uint64_t i = 0; i < k; ++i) ids[i] = P->completed[i];
  P->completed.erase(P->completed.begin(), P->completed.begin() + (ptrdiff_t)k);
  *n = k;
  return SCAMPER_OK;
}

int scamper_pool_job_state(const scamper_pool *P, uint64_t job_id, int32_t *state, uint64_t *sweeps_done) {
  if (!P) return fail(SCAMPER_EINVAL, "null pool");
  auto it = P->jobs.find(job_id);
  if (it == P->jobs.end()) return fail(SCAMPER_EINVAL, "unknown job %llu", (unsigned long long)job_id);
  if (state) *state = it->second->state;
  if (sweeps_done) *sweeps_done = it->second->state == SCAMPER_JOB_DONE ? it->second->total : it->second->enqueued;
  return SCAMPER_OK;
}

int scamper_pool_result(const scamper_pool *P, uint64_t job_id, scamper_estimate out[5], double *energy,
                        double *magnetization, int8_t *spins) {
  if (!P) return fail(SCAMPER_EINVAL, "null pool");
  auto it = P->jobs.find(job_id);
  if (it == P->jobs.end()) return fail(SCAMPER_EINVAL, "unknown job %llu", (unsigned long long)job_id);
  const Job &J = *it->second;
  if (J.state != SCAMPER_JOB_DONE) return fail(SCAMPER_EINVAL, "job %llu is not finished", (unsigned long long)job_id);
  if ((energy || magnetization) && !J.spec.keep_series) return fail(SCAMPER_EINVAL, "job did not keep its series");
  if (spins && !J.spec.keep_spins) return fail(SCAMPER_EINVAL, "job did not keep its spins");
  if (out) std::memcpy(out, J.est, sizeof J.est);
  if (energy) std::memcpy(energy, J.e_series.data(), J.e_series.size() * sizeof(double));
  if (magnetization) std::memcpy(magnetization, J.m_series.data(), J.m_series.size() * sizeof(double));
  if (spins) unsplit(J, J.split_spins.data(), spins);
  return SCAMPER_OK;
}

int scamper_pool_release(scamper_pool *P, uint64_t job_id) {
  if (!P) return fail(SCAMPER_EINVAL, "null pool");
  auto it = P->jobs.find(job_id);
  if (it == P->jobs.end()) return fail(SCAMPER_EINVAL, "unknown job %llu", (unsigned long long)job_id);
  if (it->second->state != SCAMPER_JOB_DONE) return fail(SCAMPER_EINVAL, "job %llu is not finished", (unsigned long long)job_id);
  P->jobs.erase(it);
  return SCAMPER_OK;
}

int scamper_pool_stats_get(const scamper_pool *P, scamper_pool_stats *out) {
  if (!P || !out) return fail(SCAMPER_EINVAL, "null argument");
  out->submitted = P->n_submitted;
  out->queued = P->queue.size();
  out->running = P->running.size();
  out->done = P->n_done;
  out->spin_bytes_in_use = P->in_use;
  out->site_updates = P->updates;
  out->chunks = P->chunk;
  out->grid = P->tuned ? P->shape.grid : 0;
  out->tuned = P->tuned ? 1u : 0u;
  out->lane16_from_mi = P->tuned ? P->shape.cross : 0;
  out->chunk_sweeps = P->tuned ? P->shape.chunk : 0;
  return SCAMPER_OK;
}

}  // extern "C"
