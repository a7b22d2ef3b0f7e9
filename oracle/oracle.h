/*
 * oracle.h -- CPU restatement of the reference's classical lattice sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker.
 *
 * Every function restates one piece of UynajGI/scamper (CMC.rs / Carlo.rs)
 * and cites the file:line it follows.  The random-number layer (Philox
 * draws keyed by seed, replica, sweep, colour and site) is this project's own
 * definition of the "injected uniforms"; the per-site rules that consume the
 * uniforms are the reference's.
 */
#ifndef SCAMPER_ORACLE_H
#define SCAMPER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- RNG keys */

/* RngStreamKey::seed (Carlo.rs/src/rng_stream.rs:126-154). */
uint64_t orc_stream_seed(uint64_t base, uint64_t task, uint64_t run, uint64_t chain,
                         uint64_t replica, uint64_t thread, uint64_t phase,
                         uint64_t substream);
/* Philox4x32-10 (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Key of replica `replica` of a job seeded with `seed`. */
uint64_t orc_replica_key(uint64_t seed, uint64_t replica);
/* Key of the replica-exchange stream of a job seeded with `seed`. */
uint64_t orc_exchange_key(uint64_t seed);

/* Top (part 0) or mid (part 1) byte and low 16 bits of a site draw (see
 * DRAW LAYOUT in oracle.c). */
uint32_t orc_byte(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t stream,
                  uint32_t part);
uint32_t orc_low(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t stream);
/* 32-bit site draw = top << 24 | mid << 16 | low. */
uint32_t orc_word(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t stream);
/* Two 32-bit words of the continuous-spin stream for site j. */
void orc_cont_words(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t w[2]);
/* Four 32-bit words for the isotropic random start of continuous spins. */
void orc_cont_init_words(uint64_t key, uint64_t site, uint32_t w[4]);
/* 32-bit exchange draw for the pair whose lower label is `lower`. */
uint32_t orc_exchange_word(uint64_t key, uint64_t lower, uint64_t step);

/* -------------------------------------------------------------- acceptance */

/* accept_log_probability with an injected uniform u in [0,1)
 * (Carlo.rs/src/acceptance.rs:11-22). */
int orc_accept_log_probability(double log_acceptance, double u);
/* canonical_energy_term + MetropolisHastingsAcceptance
 * (CMC.rs/src/core/ensemble.rs:90-95,196-204; core/acceptance.rs:25-35). */
double orc_log_acceptance(double beta, double delta_energy);

/* ------------------------------------------------------------------- graph */

typedef struct {
  uint64_t source;
  uint64_t target;
  uint8_t kind;
  double weight;
} orc_bond; /* graph.rs:76-82 (32 bytes) */

typedef struct {
  uint64_t n_sites;
  uint64_t n_edges;
  uint64_t n_bonds; /* incidences */
  uint64_t *offsets;
  uint64_t *neighbors;
  uint64_t *edge_ids;
  orc_bond *edges;
  /* colouring (greedy, ascending site order) */
  uint32_t n_colours;
  uint32_t *colour;      /* [n_sites] */
  uint64_t *class_index; /* [n_sites] rank of the site inside its colour class */
  uint64_t *class_start; /* [n_colours+1] */
  uint64_t *class_sites; /* [n_sites] sites grouped by colour, ascending */
} orc_lattice;

/* CsrLattice::try_from_edges (graph.rs:131-184); returns NULL on error. */
orc_lattice *orc_lattice_from_edges(uint64_t n_sites, const uint64_t *src, const uint64_t *dst,
                                    const double *weight, const uint8_t *kind, uint64_t n_edges);
/* build_hypercubic (graph.rs:417-451). */
orc_lattice *orc_lattice_hypercubic(const uint64_t *dims, int n_dims, int pbc);
orc_lattice *orc_lattice_triangular(uint64_t lx, uint64_t ly);
orc_lattice *orc_lattice_honeycomb(uint64_t lx, uint64_t ly);
orc_lattice *orc_lattice_kagome(uint64_t lx, uint64_t ly);
void orc_lattice_free(orc_lattice *lat);
int orc_lattice_validate(const orc_lattice *lat);
void orc_lattice_get(const orc_lattice *lat, uint64_t *offsets, uint64_t *neighbors,
                     uint64_t *edge_ids, uint64_t *src, uint64_t *dst, double *weight);
uint32_t orc_lattice_n_colours(const orc_lattice *lat);
void orc_lattice_colours(const orc_lattice *lat, uint32_t *colour, uint64_t *class_index);

/* ------------------------------------------------------------------ models */

enum { ORC_ISING = 0, ORC_POTTS = 1, ORC_ON = 2 };

typedef struct {
  int kind;
  double j;
  int q;        /* Potts */
  int spin_dim; /* 1 for Ising/Potts, D for O(N) */
} orc_model;

/* local_energy / delta_energy / compute_total_energy
 * (interaction.rs:73-84,149-212). */
double orc_local_energy(const orc_model *m, const orc_lattice *lat, const double *spins,
                        uint64_t site, const double *proposed);
double orc_delta_energy(const orc_model *m, const orc_lattice *lat, const double *spins,
                        uint64_t site, const double *proposed);
double orc_total_energy(const orc_model *m, const orc_lattice *lat, const double *spins);
/* Measurable::magnetization (models.rs:103-110, 276-292, 500-514). */
double orc_magnetization(const orc_model *m, const double *spins, uint64_t n_sites);

/* ---------------------------------------------------- colour-ordered sweeps */

/* One Metropolis sweep over colour classes 0..n_colours-1, sites ascending
 * inside a class, with the draws above (metropolis.rs:67-106 per site rule).
 * Returns the number of accepted moves; *energy is updated by ΔE. */
uint64_t orc_sweep_metropolis(const orc_model *m, const orc_lattice *lat, double *spins,
                              double *energy, double beta, uint64_t key, uint64_t sweep);
/* Discrete heat bath (heat_bath.rs:40-72; models.rs:158-186, 340-380). */
void orc_sweep_heat_bath(const orc_model *m, const orc_lattice *lat, double *spins,
                         double *energy, double beta, uint64_t key, uint64_t sweep);
/* O(3) continuous heat bath (heat_bath.rs:102-134; models.rs:617-662). */
void orc_sweep_on3_heat_bath(const orc_model *m, const orc_lattice *lat, double *spins,
                             double *energy, double beta, uint64_t key, uint64_t sweep);
/* Continuous heat bath of ONModel<2> (models.rs:596-614) with the von Mises
 * angle from orc_von_mises (Best-Fisher, models.rs:666-689, on the words of
 * orc_vm_block: try k of site j uses block (j, k) of stream 14). */
void orc_sweep_on2_heat_bath(const orc_model *m, const orc_lattice *lat, double *spins, double *energy,
                             double beta, uint64_t key, uint64_t sweep);
double orc_von_mises(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, double kappa);
void orc_vm_block(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t k, uint32_t w[4]);
/* Over-relaxation (microcanonical.rs:42-80). */
void orc_sweep_over_relaxation(const orc_model *m, const orc_lattice *lat, double *spins,
                               double *energy);

/* Random starts keyed by (seed, replica, site). */
void orc_random_start(const orc_model *m, const orc_lattice *lat, uint64_t key, double *spins);

/* Fast int8 stencil restatement of the Ising Metropolis sweep on an even
 * Lx x Ly periodic square lattice (identical decisions to
 * orc_sweep_metropolis; parallel over rows). */
uint64_t orc_ising2d_i8_sweeps(int8_t *spins, uint64_t lx, uint64_t ly, double beta, double j,
                               uint64_t key, uint64_t sweep0, uint64_t n_sweeps, int threads);
/* Same for the 3-D periodic cube with even sides and for q-state Potts
 * (states 0..q-1) in 2-D/3-D. */
uint64_t orc_stencil_i8_sweeps(int8_t *spins, const uint64_t *dims, int n_dims, int model_kind,
                               int q, double beta, double j, int heat_bath, uint64_t key,
                               uint64_t sweep0, uint64_t n_sweeps, int threads);

/* ---------------------------------------------------------- multi-spin */

/* Per-bit-β multi-spin Metropolis sweep (multi_spin.rs:107-151 with one β
 * per bit): bit r of words[site] is replica `replica0 + r`.  beta_bits[r] is
 * the β of bit r. */
void orc_multispin_sweep(const orc_lattice *lat, uint64_t *words, const double *beta_bits,
                         double j, uint64_t seed, uint64_t replica0, uint64_t sweep);
/* Physical energy of bit r (multi_spin.rs:184-197 restricted to one bit). */
double orc_multispin_energy(const orc_lattice *lat, const uint64_t *words, int bit, double j);

/* ------------------------------------------------------ replica exchange */

/* One in-process exchange round (parallel_tempering.rs:362-419): `label[c]`
 * is the temperature index held by configuration c, `energy[c]` its
 * energy.  step is the 1-based exchange step.  Returns accepted swaps. */
int orc_pt_exchange(int n, int *label, const double *energy, const double *beta_ladder,
                    uint64_t seed, uint64_t step);

/* ------------------------------------------------------------- binning */

typedef struct {
  double mean;
  double stderr_;
  double autocorr_time;
  uint64_t n_bins;
} orc_estimate;

/* Accumulator::add + finalize + autocorr_time (measurements.rs:107-123,
 * 162-182, 244-280) and Estimate::from_bins_with_autocorr (estimate.rs:30-58)
 * over a whole sample series. */
orc_estimate orc_binned_estimate(const double *samples, uint64_t n, uint64_t binsize);

/* ----------------------------------------- faithful plain path (config 1) */

/* Scheduler::run_one::<ClassicalMC<IsingModel, MetropolisCore>>
 * (scheduler.rs:94-155) with the reference's own layout: f64 spins, usize
 * CSR, 32-byte Bond, Fisher-Yates visit order and xoshiro256++.
 * initial_state: 0 cold, 1 random.  out[5*?]: estimates of Energy, E2,
 * Magnetization, M2, M4 (in that order).  Returns the attempted flips. */
uint64_t orc_run_one_ising(const orc_lattice *lat, double j, double beta, int initial_state,
                           uint64_t base_seed, uint64_t therm, uint64_t meas, uint64_t binsize,
                           orc_estimate out[5], double *series_e, double *series_m);
/* Time-only variant for the CPU baseline: n sweeps of the same kernel. */
uint64_t orc_plain_sweeps(const orc_lattice *lat, double j, double beta, uint64_t base_seed,
                          uint64_t n_sweeps);
/* Independent plain chains on `threads` threads (Scheduler::run_parallel,
 * Carlo.rs/src/scheduler.rs:345-405): chain c as orc_plain_sweeps with base
 * seed base_seed + c, over one shared lattice; returns all attempted updates. */
uint64_t orc_plain_chains(const orc_lattice *lat, double j, double beta, uint64_t base_seed, uint64_t n_sweeps,
                          int n_chains, int threads);

/* xoshiro256++ seeded like rand_xoshiro's seed_from_u64 (for tests). */
void orc_xoshiro_seq(uint64_t seed, uint64_t n, uint64_t *out);

/* ----------------------------------------------------- cluster updates */

/* Counter-based uniform of cluster update `step` for index idx in stream s
 * (11 bond activation by physical edge index, 12 Swendsen-Wang cluster
 * target by the cluster's smallest site, 13 Wolff seed (idx 0) and Potts
 * target (idx 1)): word idx & 3 of Philox({idx >> 2, step, x(s), idx >> 34}).
 * u = word * 2^-32. */
uint32_t orc_cluster_word(uint64_t key, uint64_t idx, uint64_t step, uint32_t stream);
/* K with (u < p) <=> (word < K) for the bond probability of
 * cluster_bond_probability (models.rs:130-147 Ising: 1 - exp(-2 beta J w);
 * 312-329 Potts: 1 - exp(-beta J w)), after clamp_probability (models.rs:
 * 16-24): K = ceil(p 2^32) in [0, 2^32]. */
uint64_t orc_cluster_threshold(int kind, double j, double weight, double beta);
/* n Swendsen-Wang (wolff = 0; SWCore::sweep_with_phase, algorithms/
 * swendsen_wang.rs:73-135) or Wolff (wolff = 1; WolffCore, wolff.rs:72-134)
 * updates of Ising (+-1) or Potts (0..q-1) f64 site-major spins, steps step0
 * .. step0+n-1, with the draws above:
 *   bond e (source s, target t) is active iff s and t hold the same state
 *     and word(11, e) < K(weight e);
 *   SW: every cluster takes the state its smallest site draws: Ising +1 if
 *     word(12, root) < 2^31 else -1; Potts floor(u q);
 *   Wolff: the seed is floor(u N) of word(13, 0); its cluster (connected by
 *     active bonds) takes -seed (Ising) or, for Potts, the other state
 *     r = floor(u' (q-1)), r += (r >= seed state), u' from word(13, 1).
 * O(N) (spin_dim 2 or 3, D components per site): every update draws a
 *   direction r (orc_on_direction); a bond is active when word(11, e) 2^-32
 *   < orc_on_bond_probability; SW reflects a cluster about r when
 *   word(12, root) < 2^31, Wolff reflects the seed's; reflected spins are
 *   renormalised in f64 and rounded to FP32.
 * Returns the number of SW clusters, or the size of the last Wolff cluster,
 * of the last step. */
uint64_t orc_cluster_steps(const orc_model *m, const orc_lattice *lat, double *spins, double beta,
                           uint64_t key, uint64_t step0, uint64_t n, int wolff);
/* O(N) embedding clusters (D = 2 or 3): the update's direction (D doubles)
 * and the bond probability of two spins along it; the on3 forms are D = 3. */
void orc_on_direction(int d, uint64_t key, uint64_t step, double *r);
double orc_on_bond_probability(int d, double j, double weight, double beta, const double *left, const double *right,
                               const double *r);
void orc_on3_direction(uint64_t key, uint64_t step, double *r);
double orc_on3_bond_probability(double j, double weight, double beta, const double *left, const double *right,
                                const double *r);

#ifdef __cplusplus
}
#endif
#endif
