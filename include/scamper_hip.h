/*
 * scamper_hip.h -- C ABI of libscamper_hip.so, the MI355X (gfx950) classical
 * lattice sweep.
 *
 * The entry points replace the reference's CPU sweep surface in
 * UynajGI/scamper (CMC.rs / Carlo.rs).  Each declaration cites the reference
 * interface it stands in for.  Conventions:
 *   - every function returns an int status (SCAMPER_OK = 0) and never
 *     unwinds; scamper_last_error() returns the message of the calling
 *     thread's last failure (the reference's panics and CarloError values
 *     become status codes);
 *   - handles are not thread-safe: one handle per host thread, or external
 *     locking (the reference's Algorithm/System are Send, not Sync);
 *   - host pointers are borrowed for the duration of the call only; device
 *     buffers belong to the handle;
 *   - spins cross the ABI in the reference's layout (flat, site-major f64:
 *     lattice/state.rs:19-28) or as site-major int8 / f32 for speed.
 */
#ifndef SCAMPER_HIP_H
#define SCAMPER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ statuses */
enum {
  SCAMPER_OK = 0,
  SCAMPER_EINVAL = 1,       /* invalid argument (CarloError::InvalidConfig) */
  SCAMPER_ENOMEM = 2,       /* device or host allocation failed */
  SCAMPER_EDEVICE = 3,      /* HIP runtime error */
  SCAMPER_EINVARIANT = 4,   /* invariant violation (the reference panics) */
  SCAMPER_EUNSUPPORTED = 5, /* valid in the reference, not on this path */
  SCAMPER_EBUDGET = 6,      /* admission refused: HBM budget exhausted */
};

/* Message of the calling thread's last failure; returns its full length. */
size_t scamper_last_error(char *buf, size_t len);
/* Library version "major.minor.patch" and build target. */
const char *scamper_version(void);

/* ------------------------------------------------------------- lattices */
typedef struct scamper_lattice scamper_lattice;

/* build_hypercubic (CMC.rs/src/lattice/graph.rs:417-451); build_chain and
 * build_square (graph.rs:402-414) are the 1-D and 2-D cases.  Site index
 * x + Lx*(y + Ly*z); bonds to +1 along each axis, wrapped when pbc. */
int scamper_lattice_hypercubic(const uint64_t *dims, int n_dims, int pbc, scamper_lattice **out);
/* build_triangular / build_honeycomb / build_kagome (graph.rs:454-543). */
int scamper_lattice_triangular(uint64_t lx, uint64_t ly, scamper_lattice **out);
int scamper_lattice_honeycomb(uint64_t lx, uint64_t ly, scamper_lattice **out);
int scamper_lattice_kagome(uint64_t lx, uint64_t ly, scamper_lattice **out);
/* CsrLattice::try_from_edges (graph.rs:131-184): physical bonds stored once;
 * parallel bonds and self-loops allowed; weight may be NULL (all 1.0). */
int scamper_lattice_from_edges(uint64_t n_sites, const uint64_t *src, const uint64_t *dst,
                               const double *weight, uint64_t n_edges, scamper_lattice **out);
/* n_sites, n_edges (CsrLattice::n_edges), n_bonds (incidences, graph.rs:
 * 112-124) and the number of colour classes of the greedy colouring. */
int scamper_lattice_info(const scamper_lattice *lat, uint64_t *n_sites, uint64_t *n_edges,
                         uint64_t *n_bonds, uint32_t *n_colours);
/* CSR arrays in the reference's order (offsets[n+1], neighbors/edge_ids[n_bonds]). */
int scamper_lattice_csr(const scamper_lattice *lat, uint64_t *offsets, uint64_t *neighbors,
                        uint64_t *edge_ids);
/* Physical edges in CsrLattice::edges order (graph.rs:112-184): source,
 * target, BondType code and weight, n_edges of each; any output may be NULL.
 * Codes follow BondType's declaration order (graph.rs:12-28): 0 generic,
 * 1 chain_x, 2 square_x, 3 square_y, 4 square_z, 5 cubic_x, 6 cubic_y,
 * 7 cubic_z, 8 tri_x, 9 tri_y, 10 tri_diag, 11 honey_x, 12 honey_y,
 * 13 kagome; the builders label as build_chain/build_square, the 3-D
 * lattices of build_lattice_from_params (classical_mc.rs:361-383),
 * build_triangular/honeycomb/kagome and try_from_edges (generic) do. */
int scamper_lattice_edges(const scamper_lattice *lat, uint64_t *source, uint64_t *target, uint8_t *kind,
                          double *weight);
/* Colour of each site and its rank inside its colour class. */
int scamper_lattice_colours(const scamper_lattice *lat, uint32_t *colour, uint64_t *class_index);
void scamper_lattice_free(scamper_lattice *lat);

/* --------------------------------------------------------------- models */
enum { SCAMPER_ISING = 0, SCAMPER_POTTS = 1, SCAMPER_ON = 2 };

/* IsingModel{j} (models.rs:29-33), PottsModel{j,q} (models.rs:190-194),
 * ONModel<D>{j} (models.rs:384-388; spin_dim = D: D = 2, the XY model, and
 * D = 3, the Heisenberg model, on any lattice: periodic hypercubic ones with
 * even sides as stencils, every other CSR lattice, with any finite bond
 * weights, class by class). */
typedef struct {
  int32_t kind;
  int32_t q;
  int32_t spin_dim;
  int32_t reserved;
  double j;
} scamper_model;

/* ---------------------------------------------------------------- state */
typedef struct scamper_state scamper_state;

/* A batch of n_replicas independent System values (lattice/state.rs:19-28)
 * of one lattice and model, device-resident.  Replica r draws from the
 * Philox key RngStreamKey(seed).with_replica(r) (Carlo.rs/src/rng_stream.rs:
 * 126-154).  Every replica starts cold (ordered_spin) at beta = 1.0
 * (classical_mc.rs:447-449 default). */
int scamper_state_create(const scamper_lattice *lat, const scamper_model *model,
                         uint32_t n_replicas, uint64_t seed, scamper_state **out);
void scamper_state_free(scamper_state *st);
/* System::set_beta (state.rs:67-73). */
int scamper_state_set_beta(scamper_state *st, uint32_t replica, double beta);
int scamper_state_get_beta(const scamper_state *st, uint32_t replica, double *beta);
/* initial_state "cold" / "hot" (classical_mc.rs:452-466): ordered_spin, or
 * a random start keyed by (seed, replica, site). */
int scamper_state_init(scamper_state *st, uint32_t replica, int random);
/* Upload / download one replica in the reference's layout (site-major f64,
 * spin_dim values per site); uploads are validated like
 * Hamiltonian::validate_configuration (interaction.rs:40-58). */
int scamper_state_upload_f64(scamper_state *st, uint32_t replica, const double *spins);
int scamper_state_download_f64(scamper_state *st, uint32_t replica, double *spins);
/* Site-major int8 (Ising +-1, Potts 0..q-1) or f32 (O(N), spin_dim per site). */
int scamper_state_upload_i8(scamper_state *st, uint32_t replica, const int8_t *spins);
int scamper_state_download_i8(scamper_state *st, uint32_t replica, int8_t *spins);
int scamper_state_upload_f32(scamper_state *st, uint32_t replica, const float *spins);
int scamper_state_download_f32(scamper_state *st, uint32_t replica, float *spins);
/* Parallel tempering over the state's replicas (ParallelTemperingCompatible
 * for ClassicalMC, CMC.rs/src/classical_mc.rs:203-229, driven by
 * PtExchange::try_exchange, Carlo.rs/src/parallel_tempering.rs:362-419).
 * scamper_state_set_ladder gives temperature label l the inverse
 * temperature beta_by_label[l] (n_replicas values, at most 1024 replicas)
 * and replica r label r.  scamper_state_exchange queues one exchange round
 * on the state's stream: the exact energy of every replica, then pairs of
 * adjacent labels with offset (step-1)&1, w = (beta_self - beta_partner)
 * E_self on each side, the lower label deciding with
 * accept_log_probability(w_lower + w_upper, u) on a counter-based draw of
 * (seed, lower label, step); accepted pairs swap labels and each replica
 * takes its new label's beta (change_parameter).  Labels stay on the device
 * between rounds, so rounds and sweeps queue without host synchronisation.
 * scamper_state_set_beta ends tempering (every replica keeps the beta of
 * the label it holds). */
int scamper_state_set_ladder(scamper_state *st, const double *beta_by_label);
int scamper_state_exchange(scamper_state *st, uint64_t step);
/* n_steps tempering steps: one sweep of every replica with `algorithm`,
 * then one exchange round numbered by the sweep count (the reference's
 * ParallelTemperingMC::step with tempering_interval 1,
 * Carlo.rs/src/parallel_tempering.rs:307-355).  With measure != 0 each step's exact
 * energy and magnetisation of every replica join the series.  The steps'
 * counters live on the device, so for O(N) on stencils and the Ising
 * Metropolis kernel on Lx % 32 == 0 lattices chunks of 16 steps replay one
 * captured HIP graph; other states launch step by step with identical
 * results. */
int scamper_state_pt_step(scamper_state *st, int algorithm, uint64_t n_steps, int measure);
/* Captured-graph launches made by scamper_state_pt_step so far. */
int scamper_state_pt_graph_launches(const scamper_state *st, uint64_t *n);
/* Label held by each replica (n_replicas values). */
int scamper_state_labels(scamper_state *st, int32_t *labels);
/* Energies of every replica used by the last exchange round. */
int scamper_state_exchange_energies(scamper_state *st, double *energy);
/* Accepted exchanges, decisions whose ln(u) fell within 1e-13 relative of
 * the log ratio (where the device and host libm could disagree), and
 * exchange rounds so far (since the ladder was set). */
int scamper_state_exchange_stats(scamper_state *st, uint64_t *accepted, uint64_t *near_ties, uint64_t *rounds);

/* Sweeps done so far (the counter in every Philox draw). */
int scamper_state_sweeps(const scamper_state *st, uint64_t *sweeps);
/* Set that counter: the whole random-number state of a replica is (seed,
 * replica, sweeps), so a state restored from a snapshot continues the
 * original trajectory exactly (the role of the Context/RNG checkpoint,
 * Carlo.rs/src/context.rs:348-420, rng_checkpoint.rs:15-60). */
int scamper_state_set_sweeps(scamper_state *st, uint64_t sweeps);

/* ------------------------------------------------------------- sweeping */
enum {
  SCAMPER_METROPOLIS = 0,        /* MetropolisCore (algorithms/metropolis.rs:16-111) */
  SCAMPER_HEAT_BATH = 1,         /* HeatBathCore (algorithms/heat_bath.rs:13-72) */
  SCAMPER_CONT_HEAT_BATH = 2,    /* ContinuousHeatBathCore (heat_bath.rs:75-134) */
  SCAMPER_OVER_RELAXATION = 3,   /* MicrocanonicalCore (algorithms/microcanonical.rs:13-85) */
  SCAMPER_HYBRID_HB_OR = 4,      /* HybridCore<ContinuousHeatBath, Microcanonical> (hybrid.rs:10-64) */
  SCAMPER_SWENDSEN_WANG = 5,     /* SWCore (algorithms/swendsen_wang.rs:12-144), Ising, Potts, O(2), O(3) */
  SCAMPER_WOLFF = 6,             /* WolffCore (algorithms/wolff.rs:12-141), Ising, Potts, O(2), O(3) */
};

/* Algorithm::sweep_with_phase (algorithms/common.rs:35-55) for every
 * replica, n_sweeps times, visiting colour classes in ascending order (one
 * cluster update per "sweep" for SCAMPER_SWENDSEN_WANG / SCAMPER_WOLFF:
 * bonds, cluster targets and the Wolff seed come from counter-based draws of
 * their own streams; J * bond weight must be >= 0, as ClusterModel asserts
 * in models.rs:139-142, 321-324, 549-553, else SCAMPER_EUNSUPPORTED; O(3)
 * spins take Wolff's embedding (models.rs:516-577): each update reflects
 * clusters about one random direction, in f64, stored back as FP32).  With
 * measure != 0 each sweep's energy and magnetisation of every replica are
 * appended to the device-resident series (ClassicalMC::measure,
 * classical_mc.rs:187-190).  Asynchronous on the state's stream. */
int scamper_sweep(scamper_state *st, int algorithm, uint64_t n_sweeps, int measure);
/* Algorithm::sweep_with_phase with its phase argument (SimulationPhase,
 * algorithms/common.rs:9-55): 0 = Thermalization, 1 = Measurement.  Only
 * adaptive proposals act on the phase (allows_adaptation); every algorithm
 * above is a fixed kernel, so both phases sweep as scamper_sweep does. */
int scamper_sweep_with_phase(scamper_state *st, int algorithm, uint64_t n_sweeps, int measure, int phase);
/* Block until the state's queued work is done. */
int scamper_synchronize(scamper_state *st);
/* scamper_sweep bracketed by HIP events on the state's stream: *ms is the
 * device time of the n_sweeps (the reference's Run::step sweep timer,
 * Carlo.rs/src/run.rs:197-211). */
int scamper_sweep_timed(scamper_state *st, int algorithm, uint64_t n_sweeps, int measure, float *ms);

/* Exact energy (compute_total_energy, interaction.rs:183-212) and
 * magnetisation (Measurable::magnetization, models.rs:103-110, 276-292,
 * 500-514) of one replica, recomputed from its spins. */
int scamper_state_observables(scamper_state *st, uint32_t replica, double *energy,
                              double *magnetization);
/* Number of measured sweeps in the series, and the per-sweep "Energy" and
 * "Magnetization" samples of one replica (observables/energy.rs:7-25,
 * magnetization.rs:7-31).  Either output may be NULL. */
int scamper_series_length(const scamper_state *st, uint64_t *n);
int scamper_series_get(scamper_state *st, uint32_t replica, double *energy,
                       double *magnetization, uint64_t max_n);
int scamper_series_clear(scamper_state *st);

/* ----------------------------------------------- multi-spin + tempering */
typedef struct scamper_multispin scamper_multispin;

/* MultiSpinIsing::new (CMC.rs/src/multi_spin.rs:50-101) for 64 * n_bundles
 * Ising replicas packed one bit per replica in u64 site words (bit set =
 * spin +1), with one temperature label per bit instead of the reference's
 * single shared beta.  Replica 64 b + r draws from the key of replica
 * 64 b + r, so each bit reproduces the single-replica state's trajectory.
 * The lattice must be a periodic hypercube with even sides (the reference
 * requires uniform degree and unit, loop-free bonds).  Starts cold. */
int scamper_multispin_create(const scamper_lattice *lat, double j, uint32_t n_bundles, uint64_t seed,
                             scamper_multispin **out);
void scamper_multispin_free(scamper_multispin *ms);
/* beta of every temperature label (64 * n_bundles values); replica r holds
 * label r until exchanges move labels (change_parameter, multi_spin.rs:360-375). */
int scamper_multispin_set_betas(scamper_multispin *ms, const double *beta_by_label);
/* Cold (all +1) or random start, bit r of site s = start of replica r. */
int scamper_multispin_init(scamper_multispin *ms, int random);
/* Site words, [bundle][site] site-major. */
int scamper_multispin_upload(scamper_multispin *ms, const uint64_t *words);
int scamper_multispin_download(scamper_multispin *ms, uint64_t *words);
/* n_steps of: one colour-ordered Metropolis sweep of every bit, then, when
 * exchange != 0, one replica-exchange round over the label ladder with
 * pairing offset (step-1)&1 and accept_log_probability(w_lower + w_upper)
 * (Carlo.rs/src/parallel_tempering.rs:362-419; log_weight_ratio
 * (beta_self - beta_partner) E_self, multi_spin.rs:352-358 per bit).  With
 * measure != 0 each step's per-replica energy and |m| are recorded. */
int scamper_multispin_step(scamper_multispin *ms, uint64_t n_steps, int exchange, int measure);
int scamper_multispin_step_timed(scamper_multispin *ms, uint64_t n_steps, int exchange, int measure,
                                 float *ms_out);
/* Temperature label held by each replica. */
int scamper_multispin_labels(scamper_multispin *ms, int32_t *labels);
/* Accepted exchanges, decisions whose ln(u) fell within 1e-13 relative of
 * the log ratio (where the device and host libm could disagree), and
 * exchange rounds so far. */
int scamper_multispin_stats(scamper_multispin *ms, uint64_t *accepted, uint64_t *near_ties,
                            uint64_t *exchange_steps);
/* Recorded steps: energy[step][replica], magnetization[step][replica]
 * (multi_spin.rs:220-286 "Energy_replica", "Magnetization_replica"). */
int scamper_multispin_series(scamper_multispin *ms, double *energy, double *magnetization,
                             uint64_t max_steps, uint64_t *n_steps);

/* ------------------------------------------------- statistics and driver */
/* Estimate (Carlo.rs/src/estimate.rs:9-21). */
typedef struct {
  double mean;
  double stderr_;
  double autocorr_time;
  uint64_t n_bins;
} scamper_estimate;

/* Accumulator::add over a whole series, then Accumulator::finalize
 * (Carlo.rs/src/measurements.rs:107-123,162-182,244-280): bins of
 * `binsize` samples, the partial bin included in the mean, stderr over bins,
 * tau from one rebinning by two. */
int scamper_binned_estimate(const double *samples, uint64_t n, uint64_t binsize,
                            scamper_estimate *out);

/* RunConfig (Carlo.rs/src/scheduler.rs:48-79). */
typedef struct {
  uint64_t thermalization_sweeps;
  uint64_t measurement_sweeps;
  uint64_t binsize;
  uint64_t base_seed;
} scamper_run_config;

/* Scheduler::run_one::<ClassicalMC<model, algorithm>> (scheduler.rs:94-155,
 * classical_mc.rs:427-490): one replica seeded with base_seed, cold or
 * random start, thermalisation then measured sweeps on the device, and the
 * DefaultObservableSet estimates out[0..4] = Energy, E2, Magnetization, M2,
 * M4 (observables/common.rs:81-96).  The per-sweep series may be copied to
 * energy / magnetization (each measurement_sweeps long) when not NULL. */
int scamper_run_one(const scamper_lattice *lat, const scamper_model *model, int algorithm,
                    double beta, int random_start, const scamper_run_config *cfg,
                    scamper_estimate out[5], double *energy, double *magnetization);

/* ----------------------------------------------------------- job pool
 * Many independent run_one jobs on one GPU at once: the device counterpart
 * of Scheduler::run_parallel -> Backend::spawn_tasks, where every task owns
 * its own MonteCarlo and Context (Carlo.rs/src/scheduler.rs:345-405,
 * Carlo.rs/src/backend/rayon.rs:36-49).  Jobs are Ising or Potts Metropolis
 * runs on periodic hypercubic lattices with even sides (1-3 dimensions).
 *
 * Jobs are admitted in submission order as device memory (the pool's spin
 * budget) and job rows free up; admitted jobs advance together, one sweep
 * of every running job per step, in graph-launched chunks of
 * `chunk_sweeps` steps.  A job's result depends only on its own spec: its
 * draws are keyed by (seed, sweep, colour, site), so it is bit-identical to
 * scamper_run_one on the same spec and to itself in any mix. */
typedef struct scamper_pool scamper_pool;

typedef struct {
  uint64_t spin_budget_bytes; /* device bytes for spins (one byte per site) */
  uint32_t max_jobs;          /* running jobs at once (1..4096) */
  uint32_t chunk_sweeps;      /* steps per graph launch (1..1024); 0 = 32 */
  uint32_t grid;              /* blocks per launch, multiple of 8; 0 = autotune: the lane-width
                                 crossover (1, 2 or 4 Mi resident colour-class sites for 16
                                 sites per lane, else 4), then the grid, then the chunk length
                                 (chunk_sweeps / 4, / 2 or chunk_sweeps), each by the median
                                 rate of three real chunks of one set of resident jobs, once
                                 the set holds 1 Mi colour-class sites (8 Mi for the grid) */
  uint32_t use_graphs;        /* 1 = hipGraph per chunk, 0 = plain launches */
  uint32_t int8_only;         /* 0 = Ising and Potts q <= 4 jobs run on bit planes (packed from
                                 and back to their int8 spins at every chunk; needs another half
                                 of the spin budget), 1 = every job on the int8 kernel; both give
                                 the same results */
  uint32_t reserved;          /* 0 */
} scamper_pool_config;

typedef struct {
  uint32_t n_dims;          /* 1, 2 or 3 */
  uint32_t reserved;
  uint64_t dims[3];         /* even sides, periodic */
  scamper_model model;      /* SCAMPER_ISING or SCAMPER_POTTS (q <= 6) */
  double beta;
  int32_t algorithm;        /* SCAMPER_METROPOLIS */
  int32_t random_start;
  uint64_t thermalization_sweeps;
  uint64_t measurement_sweeps;
  uint64_t binsize;
  uint64_t seed;            /* base seed, as scamper_run_config.base_seed */
  uint32_t keep_series;     /* keep the per-sweep E and M for scamper_pool_result */
  uint32_t keep_spins;      /* keep the final spins (site-major int8) */
} scamper_job_spec;

enum { SCAMPER_JOB_QUEUED = 0, SCAMPER_JOB_RUNNING = 1, SCAMPER_JOB_DONE = 2 };

typedef struct {
  uint64_t submitted, queued, running, done; /* running = jobs holding a row */
  uint64_t spin_bytes_in_use;
  uint64_t site_updates;  /* attempted single-site updates enqueued so far */
  uint64_t chunks;
  uint32_t grid;          /* blocks per launch in use (after autotuning) */
  uint32_t tuned;         /* 1 once the autotuner has settled */
  uint32_t lane16_from_mi; /* lane-width crossover it settled on: launches over at least this
                              many Mi colour-class sites run 16 sites per lane, smaller ones 4 */
  uint32_t chunk_sweeps;  /* steps per chunk it settled on */
} scamper_pool_stats;

int scamper_pool_create(const scamper_pool_config *cfg, scamper_pool **out);
void scamper_pool_free(scamper_pool *pool);
/* Queue a job.  SCAMPER_EBUDGET when its spins alone exceed the budget;
 * otherwise it waits for memory and a row. */
int scamper_pool_submit(scamper_pool *pool, const scamper_job_spec *spec, uint64_t *job_id);
/* Enqueue one chunk for the running jobs (after admitting queued ones) and
 * collect the chunk before the previous one; *n_completed (may be NULL)
 * gets the number of jobs finished by this call. */
int scamper_pool_step(scamper_pool *pool, uint32_t *n_completed);
/* Step until every submitted job is done. */
int scamper_pool_run(scamper_pool *pool);
/* Ids of jobs finished since the last call (up to cap; *n = count written). */
int scamper_pool_completed(scamper_pool *pool, uint64_t *ids, uint64_t cap, uint64_t *n);
int scamper_pool_job_state(const scamper_pool *pool, uint64_t job_id, int32_t *state, uint64_t *sweeps_done);
/* Estimates (Energy, E2, Magnetization, M2, M4) of a finished job, its
 * measured series if kept (each measurement_sweeps long) and its final
 * spins if kept (n_sites bytes).  NULL outputs are skipped. */
int scamper_pool_result(const scamper_pool *pool, uint64_t job_id, scamper_estimate out[5], double *energy,
                        double *magnetization, int8_t *spins);
/* Forget a finished job's results. */
int scamper_pool_release(scamper_pool *pool, uint64_t job_id);
int scamper_pool_stats_get(const scamper_pool *pool, scamper_pool_stats *out);

/* ------------------------------------------------- no device allocation
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): scamper_sweep's Ising Metropolis sweep (periodic 2-D or
 * 3-D lattice, Lx a multiple of 32, even sides) of n_rep <= 64 replicas held
 * in a caller-owned device buffer, d_spins: replica r's colour 0 then colour
 * 1 at d_spins + r * N, each colour N/2 bytes of +1/-1 in rows of Lx/2
 * (row y + Ly z, element i at x = 2i + ((c + y + z) & 1)), the layout a
 * state keeps.  Replica r draws from the key of replica replica0 + r of
 * `seed` and sweeps number sweep0 .. sweep0 + n_sweeps - 1, so the result
 * equals scamper_sweep on a state of that seed, bit for bit.  When energy or
 * magnetization is not NULL the last sweep is measured (n_rep values each).
 * The decision tables and counters live in one fixed library-owned device
 * array (about 44 KB per GPU, kept until the process ends) instead of a
 * state's buffers: calls that overlap on one GPU race.  Runs on `stream` (a
 * hipStream_t, NULL for the default stream) and returns when it is done. */
int scamper_ising_sweep_noalloc(int8_t *d_spins, const uint64_t *dims, int n_dims, double j, const double *beta,
                                uint32_t n_rep, uint64_t seed, uint32_t replica0, uint64_t sweep0,
                                uint64_t n_sweeps, double *energy, double *magnetization, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SCAMPER_HIP_H */
