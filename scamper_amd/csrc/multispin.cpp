// C ABI of the multi-spin coded replica batch with device-side replica
// exchange (MultiSpinIsing, CMC.rs/src/multi_spin.rs:28-376, with one
// temperature label per bit; exchange rule of Carlo.rs/src/
// parallel_tempering.rs:362-419).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "host.h"

using namespace scamper;

struct scamper_multispin {
  StencilGeom geom{};
  int z = 0;
  uint64_t n_sites = 0;
  uint32_t n_bundles = 0, n = 0;
  double j = 1.0;
  uint64_t seed = 0, sweeps = 0, exchange_steps = 0;
  std::vector<double> beta_by_label;
  hipStream_t stream = nullptr;
  uint64_t *d_words = nullptr, *d_keys = nullptr;
  Threshold *d_thr = nullptr;
  int32_t *d_labels = nullptr;
  double *d_beta = nullptr;
  unsigned long long *d_counts = nullptr, *d_stats = nullptr;
  double *d_series = nullptr;  // [step][2][n]: energies then |m|
  uint64_t series_cap = 0, series_len = 0;
  bool tables_ready = false;
  // device step counters and the captured chunks of kGraphSteps steps, by
  // (exchange, measure); captured on cap_stream, launched on stream
  StepClock *d_clock = nullptr;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph[2][2] = {};
  hipGraphExec_t exec[2][2] = {};
  uint64_t graph_launches = 0;

  ~scamper_multispin() {
    if (stream) hipStreamSynchronize(stream);
    hipFree(d_words);
    hipFree(d_keys);
    hipFree(d_thr);
    hipFree(d_labels);
    hipFree(d_beta);
    hipFree(d_counts);
    hipFree(d_stats);
    hipFree(d_series);
    hipFree(d_clock);
    for (int x = 0; x < 2; ++x)
      for (int m = 0; m < 2; ++m) {
        if (exec[x][m]) hipGraphExecDestroy(exec[x][m]);
        if (graph[x][m]) hipGraphDestroy(graph[x][m]);
      }
    if (cap_stream) hipStreamDestroy(cap_stream);
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

int upload_thresholds(scamper_multispin *ms) {
  std::vector<Threshold> t((size_t)ms->n * (ms->z + 1));
  for (uint32_t l = 0; l < ms->n; ++l) ising_metropolis_table(ms->beta_by_label[l], ms->j, ms->z, &t[(size_t)l * (ms->z + 1)]);
  SC_HIP(hipStreamSynchronize(ms->stream));
  SC_HIP(hipMemcpy(ms->d_thr, t.data(), t.size() * sizeof(Threshold), hipMemcpyHostToDevice));
  SC_HIP(hipMemcpy(ms->d_beta, ms->beta_by_label.data(), ms->n * sizeof(double), hipMemcpyHostToDevice));
  ms->tables_ready = true;
  return SCAMPER_OK;
}

int ensure_ms_series(scamper_multispin *ms, uint64_t extra) {
  const uint64_t need = ms->series_len + extra;
  if (need <= ms->series_cap) return SCAMPER_OK;
  const uint64_t cap = std::max<uint64_t>({need, ms->series_cap * 2, 256});
  const size_t slot = 2ull * ms->n * sizeof(double);
  double *fresh = nullptr;
  if (hipMalloc(&fresh, cap * slot) != hipSuccess) return fail(SCAMPER_ENOMEM, "multispin series");
  if (ms->series_len) SC_HIP(hipMemcpyAsync(fresh, ms->d_series, ms->series_len * slot, hipMemcpyDeviceToDevice, ms->stream));
  SC_HIP(hipStreamSynchronize(ms->stream));
  hipFree(ms->d_series);
  ms->d_series = fresh;
  ms->series_cap = cap;
  return SCAMPER_OK;
}

// site -> (colour, index) of the colour-split geometry
inline uint64_t split_index(const StencilGeom &g, uint64_t site, uint32_t *colour) {
  const uint64_t lx = 2ull * g.lx2;
  const uint64_t x = site % lx, row = site / lx;
  const uint64_t y = row % g.ly, zz = row / g.ly;
  *colour = (uint32_t)((x + y + zz) & 1);
  return row * g.lx2 + (x >> 1);
}

}  // namespace

extern "C" {

int scamper_multispin_create(const scamper_lattice *lh, double j, uint32_t n_bundles, uint64_t seed,
                             scamper_multispin **out) {
  if (!lh || !out) return fail(SCAMPER_EINVAL, "null argument");
  if (n_bundles == 0 || n_bundles > 16) return fail(SCAMPER_EINVAL, "n_bundles must be in 1..16");
  if (!std::isfinite(j)) return fail(SCAMPER_EINVAL, "coupling must be finite");
  const Lattice &l = lh->lat;
  // MultiSpinIsing::new requires uniform degree and unit, loop-free bonds
  // (multi_spin.rs:50-75); on the device the lattice is a periodic
  // hypercube with even sides.
  if (!l.stencil)
    return fail(SCAMPER_EUNSUPPORTED, "multi-spin coding runs on periodic hypercubic lattices with even sides");
  auto ms = std::unique_ptr<scamper_multispin>(new (std::nothrow) scamper_multispin);
  if (!ms) return fail(SCAMPER_ENOMEM, "multispin handle");
  ms->geom = StencilGeom{l.nd, (uint32_t)(l.dims[0] / 2), (uint32_t)l.dims[1], (uint32_t)l.dims[2], l.n_sites / 2};
  set_divisors(ms->geom);
  ms->z = 2 * l.nd;
  ms->n_sites = l.n_sites;
  ms->n_bundles = n_bundles;
  ms->n = 64 * n_bundles;
  ms->j = j;
  ms->seed = seed;
  ms->beta_by_label.assign(ms->n, 1.0);
  SC_HIP(hipStreamCreateWithFlags(&ms->stream, hipStreamNonBlocking));
  const size_t words = (size_t)n_bundles * l.n_sites;
  if (hipMalloc(&ms->d_words, words * 8) != hipSuccess) return fail(SCAMPER_ENOMEM, "multispin words");
  SC_HIP(hipMalloc(&ms->d_keys, ms->n * 8));
  SC_HIP(hipMalloc(&ms->d_thr, (size_t)ms->n * (ms->z + 1) * sizeof(Threshold)));
  SC_HIP(hipMalloc(&ms->d_labels, ms->n * 4));
  SC_HIP(hipMalloc(&ms->d_beta, ms->n * 8));
  SC_HIP(hipMalloc(&ms->d_counts, (size_t)n_bundles * kMsCountSlots * 128 * 8));
  SC_HIP(hipMalloc(&ms->d_stats, 4 * 8));
  std::vector<uint64_t> keys(ms->n);
  std::vector<int32_t> labels(ms->n);
  for (uint32_t r = 0; r < ms->n; ++r) {
    keys[r] = stream_seed(seed, r, 0);
    labels[r] = (int32_t)r;
  }
  SC_HIP(hipMemcpy(ms->d_keys, keys.data(), ms->n * 8, hipMemcpyHostToDevice));
  SC_HIP(hipMemcpy(ms->d_labels, labels.data(), ms->n * 4, hipMemcpyHostToDevice));
  SC_HIP(hipMemset(ms->d_counts, 0, (size_t)n_bundles * kMsCountSlots * 128 * 8));
  SC_HIP(hipMemset(ms->d_stats, 0, 4 * 8));
  for (uint32_t b = 0; b < n_bundles; ++b) {
    hipError_t e = launch_ms_init(ms->d_words, ms->d_keys, ms->geom, b, 0, ms->stream);
    if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "multispin init: %s", hipGetErrorString(e));
  }
  SC_HIP(hipStreamSynchronize(ms->stream));
  *out = ms.release();
  return SCAMPER_OK;
}

void scamper_multispin_free(scamper_multispin *ms) { delete ms; }

int scamper_multispin_set_betas(scamper_multispin *ms, const double *beta_by_label) {
  if (!ms || !beta_by_label) return fail(SCAMPER_EINVAL, "null argument");
  for (uint32_t l = 0; l < ms->n; ++l)
    if (!std::isfinite(beta_by_label[l]) || beta_by_label[l] < 0.0)
      return fail(SCAMPER_EINVAL, "beta must be finite and non-negative");
  ms->beta_by_label.assign(beta_by_label, beta_by_label + ms->n);
  return upload_thresholds(ms);
}

int scamper_multispin_init(scamper_multispin *ms, int random) {
  if (!ms) return fail(SCAMPER_EINVAL, "null argument");
  for (uint32_t b = 0; b < ms->n_bundles; ++b) {
    hipError_t e = launch_ms_init(ms->d_words, ms->d_keys, ms->geom, b, random != 0, ms->stream);
    if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "multispin init: %s", hipGetErrorString(e));
  }
  return SCAMPER_OK;
}

int scamper_multispin_upload(scamper_multispin *ms, const uint64_t *words) {
  if (!ms || !words) return fail(SCAMPER_EINVAL, "null argument");
  std::vector<uint64_t> split((size_t)ms->n_bundles * ms->n_sites);
  for (uint32_t b = 0; b < ms->n_bundles; ++b)
    for (uint64_t s = 0; s < ms->n_sites; ++s) {
      uint32_t c;
      const uint64_t i = split_index(ms->geom, s, &c);
      split[(2ull * b + c) * ms->geom.nh + i] = words[(uint64_t)b * ms->n_sites + s];
    }
  SC_HIP(hipStreamSynchronize(ms->stream));
  SC_HIP(hipMemcpy(ms->d_words, split.data(), split.size() * 8, hipMemcpyHostToDevice));
  return SCAMPER_OK;
}

int scamper_multispin_download(scamper_multispin *ms, uint64_t *words) {
  if (!ms || !words) return fail(SCAMPER_EINVAL, "null argument");
  std::vector<uint64_t> split((size_t)ms->n_bundles * ms->n_sites);
  SC_HIP(hipStreamSynchronize(ms->stream));
  SC_HIP(hipMemcpy(split.data(), ms->d_words, split.size() * 8, hipMemcpyDeviceToHost));
  for (uint32_t b = 0; b < ms->n_bundles; ++b)
    for (uint64_t s = 0; s < ms->n_sites; ++s) {
      uint32_t c;
      const uint64_t i = split_index(ms->geom, s, &c);
      words[(uint64_t)b * ms->n_sites + s] = split[(2ull * b + c) * ms->geom.nh + i];
    }
  return SCAMPER_OK;
}

namespace {

constexpr uint64_t kGraphSteps = 16;  // steps per captured chunk

// One step on `st`, its sweep number, exchange round and series slot taken
// from the device clock: the colour-0 and colour-1 sweeps, then the
// exchange kernel (which also turns the counters into energies, records
// them when measuring, and advances the clock), or the clock's tick alone.
int enqueue_ms_step(scamper_multispin *ms, int exchange, int measure, hipStream_t st) {
  const bool record = measure || exchange;
  const double n_bonds = (double)ms->z * (double)ms->n_sites / 2.0;
  const uint64_t ex_key = stream_seed(ms->seed, 0, 3);
  for (uint32_t c = 0; c < 2; ++c) {
    hipError_t e = launch_ms_sweep(ms->d_words, ms->d_thr, ms->d_labels, ms->d_keys, ms->geom, ms->n_bundles, c, 0,
                                   (record && c == 1) ? ms->d_counts : nullptr, st, ms->d_clock);
    if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "multispin sweep: %s", hipGetErrorString(e));
  }
  if (record) {
    hipError_t e = launch_ms_exchange(ms->d_labels, ms->d_counts, ms->d_beta, ms->n, n_bonds, (double)ms->n_sites, ms->j,
                                      ex_key, 0, exchange != 0, nullptr, nullptr, ms->d_stats, st, ms->d_clock,
                                      measure != 0, 1);
    if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "multispin exchange: %s", hipGetErrorString(e));
    return SCAMPER_OK;
  }
  SC_HIP(launch_clock_tick(ms->d_clock, 0, 0, st));
  return SCAMPER_OK;
}

// The captured chunk of kGraphSteps steps for (exchange, measure).
int ms_graph(scamper_multispin *ms, int exchange, int measure, hipGraphExec_t *out) {
  hipGraphExec_t &x = ms->exec[exchange != 0][measure != 0];
  if (!x) {
    if (!ms->cap_stream) SC_HIP(hipStreamCreateWithFlags(&ms->cap_stream, hipStreamNonBlocking));
    SC_HIP(hipStreamBeginCapture(ms->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = SCAMPER_OK;
    for (uint64_t t = 0; t < kGraphSteps && rc == SCAMPER_OK; ++t) rc = enqueue_ms_step(ms, exchange, measure, ms->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ms->cap_stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(SCAMPER_EDEVICE, "graph capture: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      hipGraphDestroy(g);
      return fail(SCAMPER_EDEVICE, "graph instantiation: %s", hipGetErrorString(ei));
    }
    ms->graph[exchange != 0][measure != 0] = g;
  }
  *out = x;
  return SCAMPER_OK;
}

}  // namespace

int scamper_multispin_step(scamper_multispin *ms, uint64_t n_steps, int exchange, int measure) {
  if (!ms) return fail(SCAMPER_EINVAL, "null argument");
  if (!ms->tables_ready)
    if (int rc = upload_thresholds(ms)) return rc;
  if (measure)
    if (int rc = ensure_ms_series(ms, n_steps)) return rc;
  if (n_steps == 0) return SCAMPER_OK;
  if (!ms->d_clock) SC_HIP(hipMalloc(&ms->d_clock, sizeof(StepClock)));
  // the clock starts where the host's counters are; whole chunks replay the
  // captured graph, the rest are launched one by one
  SC_HIP(launch_clock_set(ms->d_clock, ms->sweeps, exchange ? ms->exchange_steps + 1 : 0, ms->series_len,
                          ms->d_series, ms->stream));
  const uint64_t chunks = n_steps / kGraphSteps;
  if (chunks) {
    hipGraphExec_t x = nullptr;
    if (int rc = ms_graph(ms, exchange, measure, &x)) return rc;
    for (uint64_t k = 0; k < chunks; ++k) SC_HIP(hipGraphLaunch(x, ms->stream));
    ms->graph_launches += chunks;
  }
  for (uint64_t t = chunks * kGraphSteps; t < n_steps; ++t)
    if (int rc = enqueue_ms_step(ms, exchange, measure, ms->stream)) return rc;
  ms->sweeps += n_steps;
  if (exchange) ms->exchange_steps += n_steps;
  if (measure) ms->series_len += n_steps;
  return SCAMPER_OK;
}

int scamper_multispin_step_timed(scamper_multispin *ms, uint64_t n_steps, int exchange, int measure, float *ms_out) {
  if (!ms || !ms_out) return fail(SCAMPER_EINVAL, "null argument");
  if (!ms->tables_ready)
    if (int rc = upload_thresholds(ms)) return rc;
  if (measure)
    if (int rc = ensure_ms_series(ms, n_steps)) return rc;
  hipEvent_t a, b;
  SC_HIP(hipEventCreate(&a));
  SC_HIP(hipEventCreate(&b));
  SC_HIP(hipEventRecord(a, ms->stream));
  int rc = scamper_multispin_step(ms, n_steps, exchange, measure);
  if (rc == SCAMPER_OK) {
    SC_HIP(hipEventRecord(b, ms->stream));
    SC_HIP(hipEventSynchronize(b));
    SC_HIP(hipEventElapsedTime(ms_out, a, b));
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  return rc;
}

int scamper_multispin_labels(scamper_multispin *ms, int32_t *labels) {
  if (!ms || !labels) return fail(SCAMPER_EINVAL, "null argument");
  SC_HIP(hipStreamSynchronize(ms->stream));
  SC_HIP(hipMemcpy(labels, ms->d_labels, ms->n * 4, hipMemcpyDeviceToHost));
  return SCAMPER_OK;
}

int scamper_multispin_stats(scamper_multispin *ms, uint64_t *accepted, uint64_t *near_ties, uint64_t *exchange_steps) {
  if (!ms) return fail(SCAMPER_EINVAL, "null argument");
  unsigned long long s[4];
  SC_HIP(hipStreamSynchronize(ms->stream));
  SC_HIP(hipMemcpy(s, ms->d_stats, sizeof s, hipMemcpyDeviceToHost));
  if (accepted) *accepted = s[0];
  if (near_ties) *near_ties = s[1];
  if (exchange_steps) *exchange_steps = ms->exchange_steps;
  return SCAMPER_OK;
}

int scamper_multispin_series(scamper_multispin *ms, double *energy, double *magnetization, uint64_t max_steps,
                             uint64_t *n_steps) {
  if (!ms) return fail(SCAMPER_EINVAL, "null argument");
  const uint64_t k = std::min(max_steps, ms->series_len);
  if (n_steps) *n_steps = ms->series_len;
  if (k == 0) return SCAMPER_OK;
  std::vector<double> buf(k * 2ull * ms->n);
  SC_HIP(hipStreamSynchronize(ms->stream));
  SC_HIP(hipMemcpy(buf.data(), ms->d_series, buf.size() * 8, hipMemcpyDeviceToHost));
  for (uint64_t t = 0; t < k; ++t) {
    if (energy) std::memcpy(energy + t * ms->n, buf.data() + t * 2ull * ms->n, ms->n * 8);
    if (magnetization) std::memcpy(magnetization + t * ms->n, buf.data() + t * 2ull * ms->n + ms->n, ms->n * 8);
  }
  return SCAMPER_OK;
}

}  // extern "C"
