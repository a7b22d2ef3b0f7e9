/*
 * oracle.c -- CPU restatement of the reference's classical lattice sweep.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off so
 * that every f64 expression rounds exactly as the reference's Rust source
 * does on x86-64 (no FMA contraction); ln/exp/sin/cos come from glibc, which
 * is what Rust's f64 methods call on linux-gnu.
 *
 * DRAW LAYOUT (this project's injected uniforms; the reference draws from a
 * sequential xoshiro256++ stream instead, which no test can replay):
 *   key    = RngStreamKey(seed).with_replica(replica).seed()   (64 bit)
 *   site   = (colour c, rank j of the site inside colour class c)
 *   x(t)   = (sweep>>32 & 0xffff) | c<<16 | t<<24
 *   top    = byte (j & 15) of Philox4x32-10({ j>>4, sweep lo32, x(s),   j>>36 }, key)
 *   mid    = byte (j & 15) of Philox4x32-10({ j>>4, sweep lo32, x(s+1), j>>36 }, key)
 *   low    = half (j & 7)  of Philox4x32-10({ j>>3, sweep lo32, x(s+2), j>>35 }, key)
 *   word   = top << 24 | mid << 16 | low,   uniform u = word * 2^-32
 * (byte k of a 32-bit word is bits 8k..8k+7, half k bits 16k..16k+15.)
 * Streams s: 0 acceptance (Metropolis, Ising and Potts heat bath), 3 Potts
 * proposal, 6 discrete random start (j = site, c = 0, sweep = 0), 7
 * continuous-spin words (ctr[0] = j>>1, two words per site), 8 continuous
 * fallback (isotropic redraw), 9 replica exchange, 10 continuous random
 * start (ctr[0] = site).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_STREAM_ACCEPT 0u
#define ORC_STREAM_PROPOSE 3u
#define ORC_STREAM_INIT 6u
#define ORC_STREAM_CONT 7u
#define ORC_STREAM_CONT_FALLBACK 8u
#define ORC_STREAM_EXCHANGE 9u
#define ORC_STREAM_CONT_INIT 10u
#define ORC_STREAM_CLUSTER_BOND 11u
#define ORC_STREAM_CLUSTER_TARGET 12u
#define ORC_STREAM_CLUSTER_SEED 13u
#define ORC_STREAM_VON_MISES 14u

static const double TWO_M32 = 2.3283064365386963e-10; /* 2^-32 */
static const double TAU = 6.283185307179586;          /* std::f64::consts::TAU */
static const double PI = 3.141592653589793;           /* std::f64::consts::PI */

/* ------------------------------------------------------------ SplitMix64 */

/* mix64 (Carlo.rs/src/rng_stream.rs:146-152). */
static uint64_t mix64(uint64_t v) {
  v += 0x9E3779B97F4A7C15ull;
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
  return v ^ (v >> 31);
}

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* fold (Carlo.rs/src/rng_stream.rs:141-144). */
static uint64_t fold(uint64_t state, uint64_t domain, uint64_t value) {
  return mix64(state ^ domain ^ mix64(value + rotl64(domain, 17)));
}

/* RngStreamKey::seed (Carlo.rs/src/rng_stream.rs:126-135). */
uint64_t orc_stream_seed(uint64_t base, uint64_t task, uint64_t run, uint64_t chain,
                         uint64_t replica, uint64_t thread, uint64_t phase,
                         uint64_t substream) {
  uint64_t s = mix64(base ^ 0x53435554544C4552ull);
  s = fold(s, 0x5441534Bull, task);
  s = fold(s, 0x52554E00ull, run);
  s = fold(s, 0x434841494E00ull, chain);
  s = fold(s, 0x5245504C49434100ull, replica);
  s = fold(s, 0x544852454144ull, thread);
  s = fold(s, 0x504841534500ull, phase);
  return fold(s, 0x535542535452ull, substream);
}

uint64_t orc_replica_key(uint64_t seed, uint64_t replica) {
  return orc_stream_seed(seed, 0, 0, 0, replica, 0, 0, 0);
}

/* RngPhase::Exchange = 3 (Carlo.rs/src/rng_stream.rs:15-31). */
uint64_t orc_exchange_key(uint64_t seed) { return orc_stream_seed(seed, 0, 0, 0, 0, 0, 3, 0); }

/* --------------------------------------------------------------- Philox */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * x0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
    uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
    uint32_t y1 = (uint32_t)p1;
    uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
    uint32_t y3 = (uint32_t)p0;
    x0 = y0;
    x1 = y1;
    x2 = y2;
    x3 = y3;
  }
  out[0] = x0;
  out[1] = x1;
  out[2] = x2;
  out[3] = x3;
}

static void block(uint64_t key, uint32_t c0, uint64_t sweep, uint32_t colour, uint32_t stream,
                  uint32_t c3, uint32_t out[4]) {
  uint32_t ctr[4] = {c0, (uint32_t)sweep,
                     ((uint32_t)(sweep >> 32) & 0xFFFFu) | ((colour & 0xFFu) << 16) |
                         ((stream & 0xFFu) << 24),
                     c3};
  uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
  orc_philox4x32_10(ctr, k, out);
}

uint32_t orc_byte(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t stream,
                  uint32_t part) {
  uint32_t out[4];
  block(key, (uint32_t)(j >> 4), sweep, colour, stream + part, (uint32_t)(j >> 36), out);
  return (out[(j >> 2) & 3] >> (8 * (j & 3))) & 0xFFu;
}

uint32_t orc_low(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t stream) {
  uint32_t out[4];
  block(key, (uint32_t)(j >> 3), sweep, colour, stream + 2, (uint32_t)(j >> 35), out);
  return (out[(j >> 1) & 3] >> (16 * (j & 1))) & 0xFFFFu;
}

uint32_t orc_word(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t stream) {
  return (orc_byte(key, j, sweep, colour, stream, 0) << 24) |
         (orc_byte(key, j, sweep, colour, stream, 1) << 16) | orc_low(key, j, sweep, colour, stream);
}

void orc_cont_words(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t w[2]) {
  uint32_t out[4];
  block(key, (uint32_t)(j >> 1), sweep, colour, ORC_STREAM_CONT, (uint32_t)(j >> 33), out);
  w[0] = out[2 * (j & 1)];
  w[1] = out[2 * (j & 1) + 1];
}

static void cont_fallback_words(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour,
                                uint32_t w[4]) {
  block(key, (uint32_t)j, sweep, colour, ORC_STREAM_CONT_FALLBACK, (uint32_t)(j >> 32), w);
}

void orc_cont_init_words(uint64_t key, uint64_t site, uint32_t w[4]) {
  block(key, (uint32_t)site, 0, 0, ORC_STREAM_CONT_INIT, (uint32_t)(site >> 32), w);
}

uint32_t orc_exchange_word(uint64_t key, uint64_t lower, uint64_t step) {
  uint32_t out[4];
  block(key, (uint32_t)lower, step, 0, ORC_STREAM_EXCHANGE, (uint32_t)(lower >> 32), out);
  return out[0];
}

/* ----------------------------------------------------------- acceptance */

/* accept_log_probability (Carlo.rs/src/acceptance.rs:11-22) with the
 * uniform injected instead of drawn. */
int orc_accept_log_probability(double log_acceptance, double u) {
  if (isnan(log_acceptance)) return 0;
  if (log_acceptance >= 0.0) return 1;
  if (log_acceptance == -INFINITY) return 0;
  return log(fmax(u, DBL_MIN)) < log_acceptance;
}

/* canonical_energy_term (core/ensemble.rs:196-204), plus log_jacobian 0.0
 * (ensemble.rs:23-28,93) and the symmetric proposal's log ratio 0.0
 * (core/acceptance.rs:33). */
double orc_log_acceptance(double beta, double delta_energy) {
  double term;
  if (delta_energy == INFINITY)
    term = -INFINITY;
  else if (delta_energy == -INFINITY)
    term = INFINITY;
  else
    term = -beta * delta_energy;
  return (term + 0.0) + 0.0;
}

/* ------------------------------------------------------------------ graph */

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) abort();
  return p;
}

/* Greedy colouring in ascending site order: each site takes the smallest
 * colour none of its already-coloured neighbours holds.  On bipartite
 * hypercubic lattices with even sides this is the checkerboard parity. */
static void colour_lattice(orc_lattice *lat) {
  uint64_t n = lat->n_sites;
  lat->colour = xcalloc(n, sizeof(uint32_t));
  uint32_t maxc = 0;
  uint32_t cap = 64;
  uint8_t *used = xcalloc(cap, 1);
  for (uint64_t s = 0; s < n; ++s) {
    uint64_t deg = lat->offsets[s + 1] - lat->offsets[s];
    if (deg + 2 > cap) {
      free(used);
      cap = (uint32_t)(deg + 2);
      used = xcalloc(cap, 1);
    }
    memset(used, 0, cap);
    for (uint64_t e = lat->offsets[s]; e < lat->offsets[s + 1]; ++e) {
      uint64_t nb = lat->neighbors[e];
      if (nb < s && lat->colour[nb] < cap) used[lat->colour[nb]] = 1;
    }
    uint32_t c = 0;
    while (used[c]) ++c;
    lat->colour[s] = c;
    if (c + 1 > maxc) maxc = c + 1;
  }
  free(used);
  lat->n_colours = maxc;
  lat->class_start = xcalloc(maxc + 1, sizeof(uint64_t));
  for (uint64_t s = 0; s < n; ++s) lat->class_start[lat->colour[s] + 1]++;
  for (uint32_t c = 0; c < maxc; ++c) lat->class_start[c + 1] += lat->class_start[c];
  lat->class_sites = xcalloc(n, sizeof(uint64_t));
  lat->class_index = xcalloc(n, sizeof(uint64_t));
  uint64_t *cursor = xcalloc(maxc, sizeof(uint64_t));
  for (uint64_t s = 0; s < n; ++s) {
    uint32_t c = lat->colour[s];
    lat->class_index[s] = cursor[c];
    lat->class_sites[lat->class_start[c] + cursor[c]] = s;
    cursor[c]++;
  }
  free(cursor);
}

/* CsrLattice::try_from_edges (CMC.rs/src/lattice/graph.rs:131-184). */
orc_lattice *orc_lattice_from_edges(uint64_t n_sites, const uint64_t *src, const uint64_t *dst,
                                    const double *weight, const uint8_t *kind, uint64_t n_edges) {
  if (n_sites == 0) return NULL;
  uint64_t *degree = xcalloc(n_sites, sizeof(uint64_t));
  for (uint64_t e = 0; e < n_edges; ++e) {
    if (src[e] >= n_sites || dst[e] >= n_sites) {
      free(degree);
      return NULL;
    }
    if (weight && !isfinite(weight[e])) {
      free(degree);
      return NULL;
    }
    degree[src[e]]++;
    degree[dst[e]]++;
  }
  orc_lattice *lat = xcalloc(1, sizeof(orc_lattice));
  lat->n_sites = n_sites;
  lat->n_edges = n_edges;
  lat->offsets = xcalloc(n_sites + 1, sizeof(uint64_t));
  for (uint64_t s = 0; s < n_sites; ++s) lat->offsets[s + 1] = lat->offsets[s] + degree[s];
  lat->n_bonds = lat->offsets[n_sites];
  lat->neighbors = xcalloc(lat->n_bonds, sizeof(uint64_t));
  lat->edge_ids = xcalloc(lat->n_bonds, sizeof(uint64_t));
  lat->edges = xcalloc(n_edges, sizeof(orc_bond));
  uint64_t *cursor = degree; /* reuse */
  for (uint64_t s = 0; s < n_sites; ++s) cursor[s] = lat->offsets[s];
  for (uint64_t e = 0; e < n_edges; ++e) {
    lat->edges[e].source = src[e];
    lat->edges[e].target = dst[e];
    lat->edges[e].kind = kind ? kind[e] : 0;
    lat->edges[e].weight = weight ? weight[e] : 1.0;
    uint64_t left = cursor[src[e]]++;
    lat->neighbors[left] = dst[e];
    lat->edge_ids[left] = e;
    uint64_t right = cursor[dst[e]]++;
    lat->neighbors[right] = src[e];
    lat->edge_ids[right] = e;
  }
  free(degree);
  if (!orc_lattice_validate(lat)) {
    orc_lattice_free(lat);
    return NULL;
  }
  colour_lattice(lat);
  return lat;
}

/* build_hypercubic (CMC.rs/src/lattice/graph.rs:417-451). */
orc_lattice *orc_lattice_hypercubic(const uint64_t *dims, int n_dims, int pbc) {
  if (n_dims <= 0) return NULL;
  uint64_t n = 1, strides[8];
  for (int a = 0; a < n_dims; ++a) {
    if (dims[a] == 0) return NULL;
    strides[a] = n;
    n *= dims[a];
  }
  uint64_t cap = n * (uint64_t)n_dims, m = 0;
  uint64_t *src = xcalloc(cap, sizeof(uint64_t)), *dst = xcalloc(cap, sizeof(uint64_t));
  uint8_t *kind = xcalloc(cap, 1);
  for (uint64_t site = 0; site < n; ++site) {
    uint64_t coords[8], rem = site;
    for (int a = n_dims - 1; a >= 0; --a) {
      coords[a] = rem / strides[a];
      rem %= strides[a];
    }
    for (int a = 0; a < n_dims; ++a) {
      uint64_t c = coords[a], next;
      if (c + 1 < dims[a])
        next = c + 1;
      else if (pbc && dims[a] > 1)
        next = 0;
      else
        continue;
      src[m] = site;
      dst[m] = site - c * strides[a] + next * strides[a];
      kind[m] = (uint8_t)a;
      ++m;
    }
  }
  orc_lattice *lat = orc_lattice_from_edges(n, src, dst, NULL, kind, m);
  free(src);
  free(dst);
  free(kind);
  return lat;
}

/* build_triangular (graph.rs:454-472). */
orc_lattice *orc_lattice_triangular(uint64_t lx, uint64_t ly) {
  if (lx < 2 || ly < 2) return NULL;
  uint64_t n = lx * ly, m = 0;
  uint64_t *src = xcalloc(3 * n, sizeof(uint64_t)), *dst = xcalloc(3 * n, sizeof(uint64_t));
  for (uint64_t y = 0; y < ly; ++y)
    for (uint64_t x = 0; x < lx; ++x) {
      uint64_t s = y * lx + x;
      src[m] = s;
      dst[m++] = y * lx + (x + 1) % lx;
      src[m] = s;
      dst[m++] = ((y + 1) % ly) * lx + x;
      src[m] = s;
      dst[m++] = ((y + 1) % ly) * lx + (x + 1) % lx;
    }
  orc_lattice *lat = orc_lattice_from_edges(n, src, dst, NULL, NULL, m);
  free(src);
  free(dst);
  return lat;
}

/* CsrLattice::try_from_adjacency (graph.rs:194-249): pairs (i<j) in
 * ascending order, multiplicity preserved; self-loops appended last. */
static int cmp_pair(const void *a, const void *b) {
  const uint64_t *x = a, *y = b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}

static orc_lattice *from_adjacency(uint64_t n, const uint64_t *adj, int deg) {
  /* every row has exactly `deg` entries; no self loops in the builders */
  uint64_t total = n * (uint64_t)deg, m = 0;
  uint64_t *pairs = xcalloc(2 * total, sizeof(uint64_t));
  for (uint64_t s = 0; s < n; ++s)
    for (int k = 0; k < deg; ++k) {
      uint64_t t = adj[s * deg + k];
      if (s < t) {
        pairs[2 * m] = s;
        pairs[2 * m + 1] = t;
        ++m;
      }
    }
  qsort(pairs, m, 2 * sizeof(uint64_t), cmp_pair);
  uint64_t *src = xcalloc(m, sizeof(uint64_t)), *dst = xcalloc(m, sizeof(uint64_t));
  for (uint64_t e = 0; e < m; ++e) {
    src[e] = pairs[2 * e];
    dst[e] = pairs[2 * e + 1];
  }
  orc_lattice *lat = orc_lattice_from_edges(n, src, dst, NULL, NULL, m);
  free(pairs);
  free(src);
  free(dst);
  return lat;
}

/* build_honeycomb (graph.rs:475-505). */
orc_lattice *orc_lattice_honeycomb(uint64_t lx, uint64_t ly) {
  if (lx < 2 || ly < 2 || lx % 2) return NULL;
  uint64_t n = lx * ly;
  uint64_t *adj = xcalloc(3 * n, sizeof(uint64_t));
  for (uint64_t y = 0; y < ly; ++y)
    for (uint64_t x = 0; x < lx; ++x) {
      uint64_t s = y * lx + x;
      uint64_t matched = (x % 2 == 0) ? ((y + 1) % ly) * lx + (x + 1) % lx
                                      : ((y + ly - 1) % ly) * lx + (x + lx - 1) % lx;
      adj[3 * s] = y * lx + (x + 1) % lx;
      adj[3 * s + 1] = y * lx + (x + lx - 1) % lx;
      adj[3 * s + 2] = matched;
    }
  orc_lattice *lat = from_adjacency(n, adj, 3);
  free(adj);
  return lat;
}

/* build_kagome (graph.rs:508-543). */
orc_lattice *orc_lattice_kagome(uint64_t lx, uint64_t ly) {
  if (lx < 2 || ly < 2) return NULL;
  uint64_t n = 3 * lx * ly;
  uint64_t *adj = xcalloc(4 * n, sizeof(uint64_t));
#define KI(sub, x, y) ((sub) + 3 * ((x) + (y) * lx))
  for (uint64_t y = 0; y < ly; ++y)
    for (uint64_t x = 0; x < lx; ++x) {
      uint64_t *a0 = adj + 4 * KI(0, x, y), *a1 = adj + 4 * KI(1, x, y),
               *a2 = adj + 4 * KI(2, x, y);
      a0[0] = KI(1, x, y);
      a0[1] = KI(2, x, y);
      a0[2] = KI(1, (x + lx - 1) % lx, y);
      a0[3] = KI(2, x, (y + ly - 1) % ly);
      a1[0] = KI(0, x, y);
      a1[1] = KI(2, x, y);
      a1[2] = KI(0, (x + 1) % lx, y);
      a1[3] = KI(2, (x + lx - 1) % lx, y);
      a2[0] = KI(0, x, y);
      a2[1] = KI(1, x, y);
      a2[2] = KI(0, x, (y + 1) % ly);
      a2[3] = KI(1, (x + 1) % lx, y);
    }
#undef KI
  orc_lattice *lat = from_adjacency(n, adj, 4);
  free(adj);
  return lat;
}

void orc_lattice_free(orc_lattice *lat) {
  if (!lat) return;
  free(lat->offsets);
  free(lat->neighbors);
  free(lat->edge_ids);
  free(lat->edges);
  free(lat->colour);
  free(lat->class_index);
  free(lat->class_start);
  free(lat->class_sites);
  free(lat);
}

/* CsrLattice::validate (graph.rs:259-326). */
int orc_lattice_validate(const orc_lattice *lat) {
  if (lat->n_sites == 0 || lat->offsets[0] != 0) return 0;
  for (uint64_t s = 0; s < lat->n_sites; ++s)
    if (lat->offsets[s] > lat->offsets[s + 1]) return 0;
  if (lat->offsets[lat->n_sites] != lat->n_bonds) return 0;
  uint64_t *inc = xcalloc(lat->n_edges, sizeof(uint64_t));
  uint64_t *ep = xcalloc(2 * lat->n_edges, sizeof(uint64_t));
  int ok = 1;
  for (uint64_t s = 0; s < lat->n_sites && ok; ++s)
    for (uint64_t i = lat->offsets[s]; i < lat->offsets[s + 1]; ++i) {
      uint64_t nb = lat->neighbors[i], e = lat->edge_ids[i];
      if (nb >= lat->n_sites || e >= lat->n_edges) {
        ok = 0;
        break;
      }
      orc_bond b = lat->edges[e];
      uint64_t other = (s == b.source) ? b.target : (s == b.target ? b.source : UINT64_MAX);
      if (other != nb) {
        ok = 0;
        break;
      }
      inc[e]++;
      if (b.source == b.target || s == b.source)
        ep[2 * e]++;
      else
        ep[2 * e + 1]++;
    }
  for (uint64_t e = 0; e < lat->n_edges && ok; ++e) {
    if (inc[e] != 2) ok = 0;
    orc_bond b = lat->edges[e];
    if (b.source == b.target) {
      if (ep[2 * e] != 2 || ep[2 * e + 1] != 0) ok = 0;
    } else if (ep[2 * e] != 1 || ep[2 * e + 1] != 1)
      ok = 0;
  }
  free(inc);
  free(ep);
  return ok;
}

void orc_lattice_get(const orc_lattice *lat, uint64_t *offsets, uint64_t *neighbors,
                     uint64_t *edge_ids, uint64_t *src, uint64_t *dst, double *weight) {
  if (offsets) memcpy(offsets, lat->offsets, (lat->n_sites + 1) * sizeof(uint64_t));
  if (neighbors) memcpy(neighbors, lat->neighbors, lat->n_bonds * sizeof(uint64_t));
  if (edge_ids) memcpy(edge_ids, lat->edge_ids, lat->n_bonds * sizeof(uint64_t));
  for (uint64_t e = 0; e < lat->n_edges; ++e) {
    if (src) src[e] = lat->edges[e].source;
    if (dst) dst[e] = lat->edges[e].target;
    if (weight) weight[e] = lat->edges[e].weight;
  }
}

uint32_t orc_lattice_n_colours(const orc_lattice *lat) { return lat->n_colours; }

void orc_lattice_colours(const orc_lattice *lat, uint32_t *colour, uint64_t *class_index) {
  memcpy(colour, lat->colour, lat->n_sites * sizeof(uint32_t));
  memcpy(class_index, lat->class_index, lat->n_sites * sizeof(uint64_t));
}

/* ----------------------------------------------------------------- models */

/* bond_energy of IsingModel / PottsModel / ONModel
 * (models.rs:58-60, 230-236, 450-452; dot 427-429). */
static double bond_energy(const orc_model *m, const double *l, const double *r, double w) {
  switch (m->kind) {
    case ORC_ISING:
      return -m->j * w * l[0] * r[0];
    case ORC_POTTS:
      return ((uint64_t)l[0] == (uint64_t)r[0]) ? -m->j * w : 0.0;
    default: {
      double d = -0.0;
      for (int c = 0; c < m->spin_dim; ++c) d += l[c] * r[c];
      return -m->j * w * d;
    }
  }
}

/* PairInteraction blanket local_energy (interaction.rs:149-181). */
double orc_local_energy(const orc_model *m, const orc_lattice *lat, const double *spins,
                        uint64_t site, const double *proposed) {
  int d = m->spin_dim;
  double energy = 0.0; /* onsite_energy default (interaction.rs:131-133) */
  for (uint64_t i = lat->offsets[site]; i < lat->offsets[site + 1]; ++i) {
    orc_bond e = lat->edges[lat->edge_ids[i]];
    if (e.source == site && e.target == site) {
      energy += 0.5 * bond_energy(m, proposed, proposed, e.weight);
      continue;
    }
    if (e.source == site)
      energy += bond_energy(m, proposed, spins + e.target * d, e.weight);
    else
      energy += bond_energy(m, spins + e.source * d, proposed, e.weight);
  }
  return energy;
}

/* Hamiltonian::delta_energy default (interaction.rs:73-84). */
double orc_delta_energy(const orc_model *m, const orc_lattice *lat, const double *spins,
                        uint64_t site, const double *proposed) {
  const double *old = spins + site * m->spin_dim;
  return orc_local_energy(m, lat, spins, site, proposed) -
         orc_local_energy(m, lat, spins, site, old);
}

/* compute_total_energy (interaction.rs:183-212). */
double orc_total_energy(const orc_model *m, const orc_lattice *lat, const double *spins) {
  int d = m->spin_dim;
  double onsite = -0.0;
  for (uint64_t s = 0; s < lat->n_sites; ++s) onsite += 0.0;
  double bonds = -0.0;
  for (uint64_t e = 0; e < lat->n_edges; ++e) {
    orc_bond b = lat->edges[e];
    bonds += bond_energy(m, spins + b.source * d, spins + b.target * d, b.weight);
  }
  return onsite + bonds;
}

double orc_magnetization(const orc_model *m, const double *spins, uint64_t n) {
  if (n == 0) return 0.0;
  if (m->kind == ORC_ISING) { /* models.rs:103-110 */
    double s = -0.0;
    for (uint64_t i = 0; i < n; ++i) s += spins[i];
    return fabs(s) / (double)n;
  }
  if (m->kind == ORC_POTTS) { /* models.rs:276-292 */
    uint64_t *counts = xcalloc((size_t)m->q, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t st = (uint64_t)spins[i];
      if (st < (uint64_t)m->q) counts[st]++;
    }
    uint64_t largest = 0;
    for (int k = 0; k < m->q; ++k)
      if (counts[k] > largest) largest = counts[k];
    free(counts);
    double nd = (double)n;
    return ((double)m->q * (double)largest - nd) / (nd * (double)(m->q - 1));
  }
  /* O(N) models.rs:500-514 */
  double sum[8] = {0};
  for (uint64_t i = 0; i < n; ++i)
    for (int c = 0; c < m->spin_dim; ++c) sum[c] += spins[i * m->spin_dim + c];
  double sq = -0.0;
  for (int c = 0; c < m->spin_dim; ++c) sq += sum[c] * sum[c];
  return sqrt(sq) / (double)n;
}

/* ------------------------------------------------------------- samplers */

/* random_other_state (models.rs:212-218) with the proposal index drawn as
 * floor(word * (q-1) / 2^32). */
static uint64_t potts_propose(uint64_t current, int q, uint32_t word) {
  uint64_t state = ((uint64_t)word * (uint64_t)(q - 1)) >> 32;
  if (state >= current) state += 1;
  return state;
}

/* IsingModel::heat_bath_sample_site (models.rs:158-186). */
static double ising_heat_bath(const orc_model *m, const orc_lattice *lat, const double *spins,
                              uint64_t site, double beta, double u) {
  double field = -0.0;
  for (uint64_t i = lat->offsets[site]; i < lat->offsets[site + 1]; ++i) {
    orc_bond e = lat->edges[lat->edge_ids[i]];
    if (e.source == e.target) continue;
    field += m->j * e.weight * spins[lat->neighbors[i]];
  }
  double x = 2.0 * beta * field;
  double p_plus = (x >= 0.0) ? 1.0 / (1.0 + exp(-x)) : exp(x) / (1.0 + exp(x));
  return (u < p_plus) ? 1.0 : -1.0;
}

/* PottsModel::heat_bath_sample_site (models.rs:340-380). */
static double potts_heat_bath(const orc_model *m, const orc_lattice *lat, const double *spins,
                              uint64_t site, double beta, double u) {
  double lw[256];
  int q = m->q;
  for (int k = 0; k < q; ++k) lw[k] = 0.0;
  for (uint64_t i = lat->offsets[site]; i < lat->offsets[site + 1]; ++i) {
    orc_bond e = lat->edges[lat->edge_ids[i]];
    if (e.source == e.target) continue;
    uint64_t st = (uint64_t)spins[lat->neighbors[i]];
    if (st < (uint64_t)q) lw[st] += beta * m->j * e.weight;
  }
  double max_log = -INFINITY;
  for (int k = 0; k < q; ++k) max_log = fmax(max_log, lw[k]);
  double total = 0.0;
  for (int k = 0; k < q; ++k) {
    lw[k] = exp(lw[k] - max_log);
    total += lw[k];
  }
  double threshold = u * total;
  for (int k = 0; k < q; ++k) {
    threshold -= lw[k];
    if (threshold <= 0.0) return (double)k;
  }
  return (double)(q - 1);
}

static void normalize(double *v, int d) {
  double sq = -0.0;
  for (int c = 0; c < d; ++c) sq += v[c] * v[c];
  double norm = sqrt(sq);
  if (norm > 1e-15)
    for (int c = 0; c < d; ++c) v[c] /= norm;
}

/* ONModel::random_unit_vector (models.rs:405-422) from four words. */
static void random_unit_vector(int d, const uint32_t w[4], double *out) {
  int len = 0, k = 0;
  while (len < d && k < 4) {
    double u1 = fmax((double)w[k] * TWO_M32, DBL_MIN);
    double u2 = (double)w[k + 1] * TWO_M32;
    k += 2;
    double radius = sqrt(-2.0 * log(u1));
    double angle = TAU * u2;
    out[len++] = radius * cos(angle);
    if (len < d) out[len++] = radius * sin(angle);
  }
  normalize(out, d);
}

/* local_field (models.rs:579-595). */
static void local_field(const orc_model *m, const orc_lattice *lat, const double *spins,
                        uint64_t site, double *out) {
  int d = m->spin_dim;
  for (int c = 0; c < d; ++c) out[c] = 0.0;
  for (uint64_t i = lat->offsets[site]; i < lat->offsets[site + 1]; ++i) {
    orc_bond e = lat->edges[lat->edge_ids[i]];
    if (e.source == e.target) continue;
    double scale = m->j * e.weight;
    uint64_t base = lat->neighbors[i] * d;
    for (int c = 0; c < d; ++c) out[c] += scale * spins[base + c];
  }
}

/* ContinuousHeatBathable for ONModel<3> (models.rs:617-662). */
static void on3_heat_bath(const orc_model *m, const orc_lattice *lat, const double *spins,
                          uint64_t site, double beta, const uint32_t w[2],
                          const uint32_t fallback[4], double *out) {
  double f[3];
  local_field(m, lat, spins, site, f);
  double norm = sqrt((-0.0 + f[0] * f[0]) + f[1] * f[1] + f[2] * f[2]);
  double kappa = beta * norm;
  if (kappa < 1e-12) {
    random_unit_vector(3, fallback, out);
    return;
  }
  double u = fmax((double)w[0] * TWO_M32, DBL_MIN);
  double cos_theta = 1.0 + log(u + (1.0 - u) * exp(-2.0 * kappa)) / kappa;
  cos_theta = fmin(fmax(cos_theta, -1.0), 1.0);
  double sin_theta = sqrt(1.0 - cos_theta * cos_theta);
  double phi = TAU * ((double)w[1] * TWO_M32);
  double lx = sin_theta * cos(phi);
  double ly = sin_theta * sin(phi);
  double ux = f[0] / norm, uy = f[1] / norm, uz = f[2] / norm;
  double transverse = sqrt(ux * ux + uy * uy);
  if (transverse < 1e-12) {
    out[0] = lx;
    out[1] = ly;
    out[2] = (uz >= 0.0) ? cos_theta : -cos_theta;
  } else {
    double inverse = 1.0 / transverse;
    out[0] = -uy * inverse * lx - ux * uz * inverse * ly + ux * cos_theta;
    out[1] = ux * inverse * lx - uy * uz * inverse * ly + uy * cos_theta;
    out[2] = transverse * ly + uz * cos_theta;
  }
}

/* Best-Fisher von Mises sampler (sample_von_mises, models.rs:666-689) on
 * counter-based draws: try k of site j takes block (j, k) of the von Mises
 * stream, word 0 for the angle pi u, word 1 for the acceptance uniform; the
 * accepted try's word 2 gives the sign (+acos below 2^31, else -acos).  With
 * kappa < 1e-8 the angle is 2 pi u from word 0 of try 0.  A 64th rejection in
 * a row (probability below 0.35^64) keeps that try's value. */
#define ORC_VM_TRIES 64

void orc_vm_block(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, uint32_t k, uint32_t w[4]) {
  block(key, (uint32_t)j, sweep, colour, ORC_STREAM_VON_MISES, (uint32_t)(j >> 32) | (k << 20), w);
}

double orc_von_mises(uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, double kappa) {
  uint32_t w[4];
  if (kappa < 1e-8) {
    orc_vm_block(key, j, sweep, colour, 0, w);
    return (double)w[0] * TWO_M32 * TAU;
  }
  const double tau = 1.0 + sqrt(1.0 + 4.0 * kappa * kappa);
  const double rho = (tau - sqrt(2.0 * tau)) / (2.0 * kappa);
  const double r = (1.0 + rho * rho) / (2.0 * rho);
  double value = 0.0;
  for (uint32_t k = 0; k < ORC_VM_TRIES; ++k) {
    orc_vm_block(key, j, sweep, colour, k, w);
    const double z = cos(PI * ((double)w[0] * TWO_M32));
    value = (1.0 + r * z) / (r + z);
    const double c = kappa * (r - value);
    const double u = fmax((double)w[1] * TWO_M32, DBL_MIN);
    if (c * (2.0 - c) > u || log(c / u) + 1.0 >= c) break;
  }
  const double a = acos(fmin(fmax(value, -1.0), 1.0));
  return w[2] < 0x80000000u ? a : -a;
}

/* ContinuousHeatBathable for ONModel<2> (models.rs:596-614): the angle about
 * the local field from the von Mises draw; a vanishing field (|h| < 1e-15)
 * draws a uniform direction from the fallback words. */
static void on2_heat_bath(const orc_model *m, const orc_lattice *lat, const double *spins, uint64_t site,
                          double beta, uint64_t key, uint64_t j, uint64_t sweep, uint32_t colour, double *out) {
  double f[2];
  local_field(m, lat, spins, site, f);
  const double norm = sqrt((-0.0 + f[0] * f[0]) + f[1] * f[1]);
  if (norm < 1e-15) {
    uint32_t fb[4];
    cont_fallback_words(key, j, sweep, colour, fb);
    random_unit_vector(2, fb, out);
    return;
  }
  const double theta = orc_von_mises(key, j, sweep, colour, beta * norm) + atan2(f[1], f[0]);
  out[0] = cos(theta);
  out[1] = sin(theta);
}

/* ------------------------------------------------------ colour-ordered sweeps */

/* MetropolisCore::sweep_with_phase (algorithms/metropolis.rs:67-106) with the
 * visit order "colour classes ascending, sites ascending inside a class";
 * per site: propose (proposal.rs:53-62; models.rs:98-100, 212-218,
 * 267-269), metropolis_hastings_step (core/trial.rs:68-101), commit
 * (lattice/state.rs:127-136). */
uint64_t orc_sweep_metropolis(const orc_model *m, const orc_lattice *lat, double *spins,
                              double *energy, double beta, uint64_t key, uint64_t sweep) {
  uint64_t accepted = 0;
  for (uint32_t c = 0; c < lat->n_colours; ++c) {
    for (uint64_t idx = lat->class_start[c]; idx < lat->class_start[c + 1]; ++idx) {
      uint64_t site = lat->class_sites[idx], j = idx - lat->class_start[c];
      double proposed;
      if (m->kind == ORC_ISING)
        proposed = -spins[site];
      else
        proposed = (double)potts_propose((uint64_t)spins[site], m->q,
                                         orc_word(key, j, sweep, c, ORC_STREAM_PROPOSE));
      double de = orc_delta_energy(m, lat, spins, site, &proposed);
      double log_a = orc_log_acceptance(beta, de);
      double u = (double)orc_word(key, j, sweep, c, ORC_STREAM_ACCEPT) * TWO_M32;
      if (orc_accept_log_probability(log_a, u)) {
        spins[site] = proposed;
        *energy += de;
        ++accepted;
      }
    }
  }
  return accepted;
}

/* HeatBathCore::sweep_with_phase (algorithms/heat_bath.rs:40-72). */
void orc_sweep_heat_bath(const orc_model *m, const orc_lattice *lat, double *spins,
                         double *energy, double beta, uint64_t key, uint64_t sweep) {
  for (uint32_t c = 0; c < lat->n_colours; ++c) {
    for (uint64_t idx = lat->class_start[c]; idx < lat->class_start[c + 1]; ++idx) {
      uint64_t site = lat->class_sites[idx], j = idx - lat->class_start[c];
      double u = (double)orc_word(key, j, sweep, c, ORC_STREAM_ACCEPT) * TWO_M32;
      double proposed = (m->kind == ORC_ISING) ? ising_heat_bath(m, lat, spins, site, beta, u)
                                               : potts_heat_bath(m, lat, spins, site, beta, u);
      double de = orc_delta_energy(m, lat, spins, site, &proposed);
      spins[site] = proposed;
      *energy += de;
    }
  }
}

/* ContinuousHeatBathCore::sweep_with_phase (algorithms/heat_bath.rs:102-134)
 * for ONModel<3>. */
void orc_sweep_on3_heat_bath(const orc_model *m, const orc_lattice *lat, double *spins,
                             double *energy, double beta, uint64_t key, uint64_t sweep) {
  for (uint32_t c = 0; c < lat->n_colours; ++c) {
    for (uint64_t idx = lat->class_start[c]; idx < lat->class_start[c + 1]; ++idx) {
      uint64_t site = lat->class_sites[idx], j = idx - lat->class_start[c];
      uint32_t w[2], fb[4];
      orc_cont_words(key, j, sweep, c, w);
      cont_fallback_words(key, j, sweep, c, fb);
      double proposed[3];
      on3_heat_bath(m, lat, spins, site, beta, w, fb, proposed);
      double de = orc_delta_energy(m, lat, spins, site, proposed);
      memcpy(spins + 3 * site, proposed, sizeof proposed);
      *energy += de;
    }
  }
}

/* ContinuousHeatBathCore::sweep_with_phase (algorithms/heat_bath.rs:102-134)
 * for ONModel<2>. */
void orc_sweep_on2_heat_bath(const orc_model *m, const orc_lattice *lat, double *spins, double *energy,
                             double beta, uint64_t key, uint64_t sweep) {
  for (uint32_t c = 0; c < lat->n_colours; ++c) {
    for (uint64_t idx = lat->class_start[c]; idx < lat->class_start[c + 1]; ++idx) {
      uint64_t site = lat->class_sites[idx], j = idx - lat->class_start[c];
      double proposed[2];
      on2_heat_bath(m, lat, spins, site, beta, key, j, sweep, c, proposed);
      double de = orc_delta_energy(m, lat, spins, site, proposed);
      memcpy(spins + 2 * site, proposed, sizeof proposed);
      *energy += de;
    }
  }
}

/* MicrocanonicalCore::sweep_with_phase (algorithms/microcanonical.rs:42-80). */
void orc_sweep_over_relaxation(const orc_model *m, const orc_lattice *lat, double *spins,
                               double *energy) {
  int d = m->spin_dim;
  for (uint32_t c = 0; c < lat->n_colours; ++c) {
    for (uint64_t idx = lat->class_start[c]; idx < lat->class_start[c + 1]; ++idx) {
      uint64_t site = lat->class_sites[idx];
      double f[8], old[8], refl[8];
      local_field(m, lat, spins, site, f);
      double nsq = -0.0;
      for (int k = 0; k < d; ++k) nsq += f[k] * f[k];
      if (nsq < 1e-28) continue;
      double proj = -0.0;
      for (int k = 0; k < d; ++k) {
        old[k] = spins[site * d + k];
        proj += old[k] * f[k];
      }
      proj = proj / nsq;
      for (int k = 0; k < d; ++k) refl[k] = 2.0 * proj * f[k] - old[k];
      double de = orc_delta_energy(m, lat, spins, site, refl);
      memcpy(spins + site * d, refl, d * sizeof(double));
      *energy += de;
    }
  }
}

void orc_random_start(const orc_model *m, const orc_lattice *lat, uint64_t key, double *spins) {
  for (uint64_t s = 0; s < lat->n_sites; ++s) {
    if (m->kind == ORC_ISING) {
      spins[s] = (orc_byte(key, s, 0, 0, ORC_STREAM_INIT, 0) & 1u) ? 1.0 : -1.0;
    } else if (m->kind == ORC_POTTS) {
      spins[s] = (double)(((uint64_t)orc_word(key, s, 0, 0, ORC_STREAM_INIT) * (uint64_t)m->q) >> 32);
    } else {
      uint32_t w[4];
      orc_cont_init_words(key, s, w);
      random_unit_vector(m->spin_dim, w, spins + s * m->spin_dim);
    }
  }
}

/* ------------------------------------------- fast int8 stencil restatement */

/* Per-thread cache of the top and mid blocks of 16 sites. */
typedef struct {
  uint64_t group;
  uint32_t stream;
  uint32_t top[4], mid[4];
  int valid;
} top_cache;

static uint32_t cached_word(top_cache *tc, uint64_t key, uint64_t j, uint64_t sweep, uint32_t c,
                            uint32_t stream) {
  uint64_t g = j >> 4;
  if (!tc->valid || tc->group != g || tc->stream != stream) {
    block(key, (uint32_t)g, sweep, c, stream, (uint32_t)(j >> 36), tc->top);
    block(key, (uint32_t)g, sweep, c, stream + 1, (uint32_t)(j >> 36), tc->mid);
    tc->group = g;
    tc->stream = stream;
    tc->valid = 1;
  }
  uint32_t top = (tc->top[(j >> 2) & 3] >> (8 * (j & 3))) & 0xFFu;
  uint32_t mid = (tc->mid[(j >> 2) & 3] >> (8 * (j & 3))) & 0xFFu;
  return (top << 24) | (mid << 16) | orc_low(key, j, sweep, c, stream);
}

/* Colour-ordered sweeps on an int8 hypercubic periodic lattice with even
 * sides.  Site = x + Lx*(y + Ly*z); colour = parity of x+y+z; rank inside
 * the colour class = site >> 1.  Ising spins are +-1, Potts states 0..q-1.
 * With J = +-1 (or any J whose multiples are exact) the f64 ΔE here equals
 * the CSR restatement's bit for bit; decisions use the same draws. */
uint64_t orc_stencil_i8_sweeps(int8_t *spins, const uint64_t *dims, int n_dims, int model_kind,
                               int q, double beta, double j, int heat_bath, uint64_t key,
                               uint64_t sweep0, uint64_t n_sweeps, int threads) {
  uint64_t lx = dims[0], ly = n_dims > 1 ? dims[1] : 1, lz = n_dims > 2 ? dims[2] : 1;
  uint64_t rows = ly * lz;
  uint64_t accepted = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  for (uint64_t t = 0; t < n_sweeps; ++t) {
    uint64_t sweep = sweep0 + t;
    for (uint32_t c = 0; c < 2; ++c) {
#pragma omp parallel for schedule(static) reduction(+ : accepted)
      for (uint64_t row = 0; row < rows; ++row) {
        top_cache ha = {0}, hp = {0};
        uint64_t y = row % ly, z = row / ly;
        uint64_t ym = (y + ly - 1) % ly, yp = (y + 1) % ly;
        uint64_t zm = (z + lz - 1) % lz, zp = (z + 1) % lz;
        for (uint64_t x = (c + y + z) & 1; x < lx; x += 2) {
          uint64_t site = x + lx * row;
          uint64_t jj = site >> 1;
          uint64_t xm = (x + lx - 1) % lx, xp = (x + 1) % lx;
          int8_t nb[6];
          int nn = 0;
          nb[nn++] = spins[xm + lx * row];
          nb[nn++] = spins[xp + lx * row];
          if (n_dims > 1) {
            nb[nn++] = spins[x + lx * (ym + ly * z)];
            nb[nn++] = spins[x + lx * (yp + ly * z)];
          }
          if (n_dims > 2) {
            nb[nn++] = spins[x + lx * (y + ly * zm)];
            nb[nn++] = spins[x + lx * (y + ly * zp)];
          }
          int8_t s = spins[site];
          if (model_kind == ORC_ISING) {
            int sum = 0;
            for (int k = 0; k < nn; ++k) sum += nb[k];
            if (heat_bath) {
              double field = j * (double)sum;
              double x2 = 2.0 * beta * field;
              double pp = (x2 >= 0.0) ? 1.0 / (1.0 + exp(-x2)) : exp(x2) / (1.0 + exp(x2));
              uint32_t w = cached_word(&ha, key, jj, sweep, c, ORC_STREAM_ACCEPT);
              spins[site] = ((double)w * TWO_M32 < pp) ? 1 : -1;
            } else {
              double de = 2.0 * j * (double)s * (double)sum;
              double la = orc_log_acceptance(beta, de);
              int acc;
              if (la >= 0.0)
                acc = 1;
              else {
                uint32_t w = cached_word(&ha, key, jj, sweep, c, ORC_STREAM_ACCEPT);
                acc = orc_accept_log_probability(la, (double)w * TWO_M32);
              }
              if (acc) {
                spins[site] = (int8_t)-s;
                accepted++;
              }
            }
          } else {
            int cnt[256];
            for (int k = 0; k < q; ++k) cnt[k] = 0;
            for (int k = 0; k < nn; ++k) cnt[(uint8_t)nb[k]]++;
            if (heat_bath) {
              double lw[256];
              for (int k = 0; k < q; ++k) {
                lw[k] = 0.0;
                for (int r = 0; r < cnt[k]; ++r) lw[k] += beta * j * 1.0;
              }
              double mx = -INFINITY, total = 0.0;
              for (int k = 0; k < q; ++k) mx = fmax(mx, lw[k]);
              for (int k = 0; k < q; ++k) {
                lw[k] = exp(lw[k] - mx);
                total += lw[k];
              }
              uint32_t w = cached_word(&ha, key, jj, sweep, c, ORC_STREAM_ACCEPT);
              double thr = (double)w * TWO_M32 * total;
              int ns = q - 1;
              for (int k = 0; k < q; ++k) {
                thr -= lw[k];
                if (thr <= 0.0) {
                  ns = k;
                  break;
                }
              }
              spins[site] = (int8_t)ns;
            } else {
              uint32_t pw = cached_word(&hp, key, jj, sweep, c, ORC_STREAM_PROPOSE);
              uint64_t ns = potts_propose((uint64_t)(uint8_t)s, q, pw);
              double de = -j * (double)cnt[ns] - (-j * (double)cnt[(uint8_t)s]);
              double la = orc_log_acceptance(beta, de);
              int acc;
              if (la >= 0.0)
                acc = 1;
              else {
                uint32_t w = cached_word(&ha, key, jj, sweep, c, ORC_STREAM_ACCEPT);
                acc = orc_accept_log_probability(la, (double)w * TWO_M32);
              }
              if (acc) {
                spins[site] = (int8_t)ns;
                accepted++;
              }
            }
          }
        }
      }
    }
  }
  return accepted;
}

uint64_t orc_ising2d_i8_sweeps(int8_t *spins, uint64_t lx, uint64_t ly, double beta, double j,
                               uint64_t key, uint64_t sweep0, uint64_t n_sweeps, int threads) {
  uint64_t dims[2] = {lx, ly};
  return orc_stencil_i8_sweeps(spins, dims, 2, ORC_ISING, 2, beta, j, 0, key, sweep0, n_sweeps,
                               threads);
}

/* ------------------------------------------------------------- multi-spin */

/* MultiSpinIsing::sweep (multi_spin.rs:107-151) with one β per bit, the
 * colour-ordered visit and per-(replica, site) draws; the acceptance LUT is
 * rebuilt per bit as in rebuild_accept_lut (multi_spin.rs:200-208). */
void orc_multispin_sweep(const orc_lattice *lat, uint64_t *words, const double *beta_bits,
                         double j, uint64_t seed, uint64_t replica0, uint64_t sweep) {
  uint64_t z = lat->offsets[1] - lat->offsets[0];
  double *lut = xcalloc(64 * (z + 1), sizeof(double));
  uint64_t keys[64];
  for (int r = 0; r < 64; ++r) {
    keys[r] = orc_replica_key(seed, replica0 + (uint64_t)r);
    for (uint64_t k = 0; k <= z; ++k) {
      double de = 2.0 * j * (2.0 * (double)k - (double)z);
      double v = -beta_bits[r] * de;
      lut[r * (z + 1) + k] = v < 0.0 ? v : 0.0;
    }
  }
  for (uint32_t c = 0; c < lat->n_colours; ++c) {
    for (uint64_t idx = lat->class_start[c]; idx < lat->class_start[c + 1]; ++idx) {
      uint64_t site = lat->class_sites[idx], jj = idx - lat->class_start[c];
      uint64_t word = words[site];
      uint8_t anti[64] = {0};
      for (uint64_t i = lat->offsets[site]; i < lat->offsets[site + 1]; ++i) {
        uint64_t x = word ^ words[lat->neighbors[i]];
        for (int r = 0; r < 64; ++r) anti[r] += (uint8_t)((x >> r) & 1u);
      }
      uint64_t flip = 0;
      for (int r = 0; r < 64; ++r) {
        uint64_t k = z - anti[r];
        double la = lut[r * (z + 1) + k];
        int acc;
        if (la >= 0.0)
          acc = 1;
        else
          acc = orc_accept_log_probability(
              la, (double)orc_word(keys[r], jj, sweep, c, ORC_STREAM_ACCEPT) * TWO_M32);
        if (acc) flip |= 1ull << r;
      }
      words[site] = word ^ flip;
    }
  }
  free(lut);
}

double orc_multispin_energy(const orc_lattice *lat, const uint64_t *words, int bit, double j) {
  double e = -0.0;
  for (uint64_t k = 0; k < lat->n_edges; ++k) {
    orc_bond b = lat->edges[k];
    double l = ((words[b.source] >> bit) & 1u) ? 1.0 : -1.0;
    double r = ((words[b.target] >> bit) & 1u) ? 1.0 : -1.0;
    e += -j * b.weight * l * r;
  }
  return e;
}

/* ------------------------------------------------------ replica exchange */

/* PtExchange::try_exchange (Carlo.rs/src/parallel_tempering.rs:362-419)
 * run for every pair at once: pairing offset (step-1)&1, weights
 * w = (β_self − β_partner)·E_self (classical_mc.rs:209-214), decision by the
 * lower label on w_lower + w_upper, accepted pairs swap labels. */
int orc_pt_exchange(int n, int *label, const double *energy, const double *beta_ladder,
                    uint64_t seed, uint64_t step) {
  if (n < 2) return 0;
  int *config_of = xcalloc((size_t)n, sizeof(int));
  for (int c = 0; c < n; ++c) config_of[label[c]] = c;
  uint64_t offset = (step > 0 ? step - 1 : 0) & 1u;
  uint64_t key = orc_exchange_key(seed);
  int accepted = 0;
  for (int lower = (int)offset; lower + 1 < n; lower += 2) {
    int a = config_of[lower], b = config_of[lower + 1];
    double wa = (beta_ladder[lower] - beta_ladder[lower + 1]) * energy[a];
    double wb = (beta_ladder[lower + 1] - beta_ladder[lower]) * energy[b];
    double u = (double)orc_exchange_word(key, (uint64_t)lower, step) * TWO_M32;
    if (orc_accept_log_probability(wa + wb, u)) {
      label[a] = lower + 1;
      label[b] = lower;
      ++accepted;
    }
  }
  free(config_of);
  return accepted;
}

/* ---------------------------------------------------------------- binning */

static orc_estimate estimate_from_bins(const double *bins, uint64_t n, double tau) {
  orc_estimate e = {0.0, 0.0, 1.0, 0};
  if (n == 0) return e;
  double nd = (double)n, sum = -0.0;
  for (uint64_t i = 0; i < n; ++i) sum += bins[i];
  double mean = sum / nd;
  double var = 0.0;
  if (nd > 1.0) {
    double ss = -0.0;
    for (uint64_t i = 0; i < n; ++i) ss += (bins[i] - mean) * (bins[i] - mean);
    var = ss / (nd - 1.0);
  }
  e.mean = mean;
  e.stderr_ = sqrt(var) / sqrt(nd);
  e.autocorr_time = tau > 1.0 ? tau : 1.0;
  e.n_bins = n;
  return e;
}

/* Accumulator::autocorr_time (measurements.rs:244-280). */
static double autocorr_time(const double *means, uint64_t nb) {
  if (nb < 2) return 1.0;
  double n = (double)nb, mu = -0.0;
  for (uint64_t i = 0; i < nb; ++i) mu += means[i];
  mu /= n;
  double v = -0.0;
  for (uint64_t i = 0; i < nb; ++i) v += (means[i] - mu) * (means[i] - mu);
  v /= (n - 1.0);
  uint64_t rn = nb / 2;
  if (rn < 2) return 1.0;
  double *rb = xcalloc(rn, sizeof(double));
  double rmu = -0.0;
  for (uint64_t i = 0; i < rn; ++i) {
    rb[i] = (means[2 * i] + means[2 * i + 1]) / 2.0;
    rmu += rb[i];
  }
  rmu /= (double)rn;
  double vr = -0.0;
  for (uint64_t i = 0; i < rn; ++i) vr += (rb[i] - rmu) * (rb[i] - rmu);
  vr /= ((double)rn - 1.0);
  free(rb);
  double expected = v / 2.0;
  if (expected <= 0.0) return 1.0;
  double r = vr / expected - 1.0;
  return 0.5 * (r > 0.0 ? r : 0.0);
}

/* Accumulator::add / finalize (measurements.rs:107-123,162-182). */
orc_estimate orc_binned_estimate(const double *samples, uint64_t n, uint64_t binsize) {
  uint64_t cap = binsize ? n / binsize + 2 : n + 2;
  double *bins = xcalloc(cap, sizeof(double));
  uint64_t nb = 0, filling = 0;
  double current = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    current += samples[i];
    filling++;
    if (filling >= binsize) {
      bins[nb++] = current / (double)binsize;
      current = 0.0;
      filling = 0;
    }
  }
  double tau = autocorr_time(bins, nb);
  uint64_t total = nb;
  if (filling > 0) bins[total++] = current / ((double)filling * 1.0);
  orc_estimate e = estimate_from_bins(bins, total, tau);
  free(bins);
  return e;
}

/* ------------------------------------------ faithful plain path (config 1) */

typedef struct {
  uint64_t s[4];
} xoshiro;

/* Xoshiro256PlusPlus::seed_from_u64 via SplitMix64 (rand_xoshiro 0.8). */
static void xo_seed(xoshiro *x, uint64_t state) {
  for (int i = 0; i < 4; ++i) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    x->s[i] = z ^ (z >> 31);
  }
}

static uint64_t xo_next(xoshiro *x) {
  uint64_t *s = x->s;
  uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return result;
}

void orc_xoshiro_seq(uint64_t seed, uint64_t n, uint64_t *out) {
  xoshiro x;
  if (seed == UINT64_MAX) { /* raw state {1,2,3,4} for the known-answer test */
    x.s[0] = 1;
    x.s[1] = 2;
    x.s[2] = 3;
    x.s[3] = 4;
  } else
    xo_seed(&x, seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = xo_next(&x);
}

/* random::<f64>(): 53 high bits scaled by 2^-53. */
static double xo_f64(xoshiro *x) { return (double)(xo_next(x) >> 11) * 1.1102230246251565e-16; }

/* random_range(0..=hi) by Canon's widening-multiply method. */
static uint64_t xo_range_inclusive(xoshiro *x, uint64_t hi) {
  uint64_t range = hi + 1;
  if (range == 0) return xo_next(x);
  __uint128_t m = (__uint128_t)xo_next(x) * range;
  uint64_t result = (uint64_t)(m >> 64), lo = (uint64_t)m;
  if (lo > (uint64_t)(-range)) {
    __uint128_t m2 = (__uint128_t)xo_next(x) * range;
    uint64_t new_hi = (uint64_t)(m2 >> 64);
    if (lo + new_hi < lo) result += 1;
  }
  return result;
}

static void plain_sweep(const orc_model *m, const orc_lattice *lat, double *spins,
                        double *energy, double beta, uint64_t *order, xoshiro *rng) {
  uint64_t n = lat->n_sites;
  /* SiteOrder::prepare RandomPermutation (core/visit.rs:30-59) */
  for (uint64_t i = n - 1; i >= 1; --i) {
    uint64_t sw = xo_range_inclusive(rng, i);
    uint64_t tmp = order[i];
    order[i] = order[sw];
    order[sw] = tmp;
  }
  for (uint64_t k = 0; k < n; ++k) {
    uint64_t site = order[k];
    double proposed = -spins[site];
    double de = orc_delta_energy(m, lat, spins, site, &proposed);
    double la = orc_log_acceptance(beta, de);
    int acc;
    if (isnan(la))
      acc = 0;
    else if (la >= 0.0)
      acc = 1;
    else if (la == -INFINITY)
      acc = 0;
    else
      acc = log(fmax(xo_f64(rng), DBL_MIN)) < la;
    if (acc) {
      spins[site] = proposed;
      *energy += de;
    }
  }
}

/* Scheduler::run_one (Carlo.rs/src/scheduler.rs:94-155) for
 * ClassicalMC<IsingModel, MetropolisCore> with DefaultObservableSet
 * (observables/common.rs:81-96; energy.rs:7-25; magnetization.rs:7-31). */
uint64_t orc_run_one_ising(const orc_lattice *lat, double j, double beta, int initial_state,
                           uint64_t base_seed, uint64_t therm, uint64_t meas, uint64_t binsize,
                           orc_estimate out[5], double *series_e, double *series_m) {
  orc_model m = {ORC_ISING, j, 2, 1};
  uint64_t n = lat->n_sites;
  xoshiro rng;
  xo_seed(&rng, orc_stream_seed(base_seed, 0, 0, 0, 0, 0, 0, 0));
  double *spins = xcalloc(n, sizeof(double));
  for (uint64_t s = 0; s < n; ++s)
    spins[s] = initial_state ? (((uint32_t)(xo_next(&rng) >> 32) >> 31) ? 1.0 : -1.0) : 1.0;
  double energy = orc_total_energy(&m, lat, spins);
  uint64_t *order = xcalloc(n, sizeof(uint64_t));
  for (uint64_t s = 0; s < n; ++s) order[s] = s;
  for (uint64_t t = 0; t < therm; ++t) plain_sweep(&m, lat, spins, &energy, beta, order, &rng);
  double *se = series_e ? series_e : xcalloc(meas, sizeof(double));
  double *sm = series_m ? series_m : xcalloc(meas, sizeof(double));
  for (uint64_t t = 0; t < meas; ++t) {
    plain_sweep(&m, lat, spins, &energy, beta, order, &rng);
    se[t] = energy;
    sm[t] = orc_magnetization(&m, spins, n);
  }
  double *tmp = xcalloc(meas ? meas : 1, sizeof(double));
  out[0] = orc_binned_estimate(se, meas, binsize);
  for (uint64_t t = 0; t < meas; ++t) tmp[t] = se[t] * se[t];
  out[1] = orc_binned_estimate(tmp, meas, binsize);
  out[2] = orc_binned_estimate(sm, meas, binsize);
  for (uint64_t t = 0; t < meas; ++t) tmp[t] = sm[t] * sm[t];
  out[3] = orc_binned_estimate(tmp, meas, binsize);
  for (uint64_t t = 0; t < meas; ++t) tmp[t] = (sm[t] * sm[t]) * (sm[t] * sm[t]);
  out[4] = orc_binned_estimate(tmp, meas, binsize);
  free(tmp);
  if (!series_e) free(se);
  if (!series_m) free(sm);
  free(order);
  free(spins);
  return n * (therm + meas);
}

uint64_t orc_plain_sweeps(const orc_lattice *lat, double j, double beta, uint64_t base_seed,
                          uint64_t n_sweeps) {
  orc_model m = {ORC_ISING, j, 2, 1};
  uint64_t n = lat->n_sites;
  xoshiro rng;
  xo_seed(&rng, orc_stream_seed(base_seed, 0, 0, 0, 0, 0, 0, 0));
  double *spins = xcalloc(n, sizeof(double));
  for (uint64_t s = 0; s < n; ++s)
    spins[s] = (((uint32_t)(xo_next(&rng) >> 32)) >> 31) ? 1.0 : -1.0;
  double energy = orc_total_energy(&m, lat, spins);
  uint64_t *order = xcalloc(n, sizeof(uint64_t));
  for (uint64_t s = 0; s < n; ++s) order[s] = s;
  for (uint64_t t = 0; t < n_sweeps; ++t) plain_sweep(&m, lat, spins, &energy, beta, order, &rng);
  free(order);
  free(spins);
  return n * n_sweeps;
}

/* Scheduler::run_parallel (Carlo.rs/src/scheduler.rs:345-405): independent
 * plain chains, one per task on its own thread, each as orc_plain_sweeps
 * with base seed base_seed + chain, over one shared read-only lattice.
 * Returns the attempted updates of all chains. */
uint64_t orc_plain_chains(const orc_lattice *lat, double j, double beta, uint64_t base_seed, uint64_t n_sweeps,
                          int n_chains, int threads) {
  uint64_t total = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel for schedule(static, 1) reduction(+ : total)
  for (int c = 0; c < n_chains; ++c) total += orc_plain_sweeps(lat, j, beta, base_seed + (uint64_t)c, n_sweeps);
  return total;
}

/* ----------------------------------------------------- cluster updates */

uint32_t orc_cluster_word(uint64_t key, uint64_t idx, uint64_t step, uint32_t stream) {
  uint32_t out[4];
  block(key, (uint32_t)(idx >> 2), step, 0, stream, (uint32_t)(idx >> 34), out);
  return out[idx & 3];
}

/* clamp_probability (models.rs:16-24). */
static double clamp_probability(double v) {
  if (isfinite(v)) return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
  return v > 0.0 ? 1.0 : 0.0;
}

uint64_t orc_cluster_threshold(int kind, double j, double weight, double beta) {
  const double coupling = j * weight;
  if (coupling == 0.0) return 0;
  /* Ising 1 - exp(-2 beta c) (models.rs:146), Potts 1 - exp(-beta c) (models.rs:328) */
  const double arg = kind == ORC_ISING ? -2.0 * beta * coupling : -beta * coupling;
  const double p = clamp_probability(1.0 - exp(arg));
  return (uint64_t)ceil(p * 4294967296.0); /* p 2^32 is exact; k < K <=> k 2^-32 < p */
}

/* Union-find with the smaller site as the surviving root, so every root is
 * its component's smallest site (SWCore's union by rank picks another root;
 * the partition, the only thing the update depends on, is the same). */
static uint64_t uf_find(uint64_t *parent, uint64_t x) {
  uint64_t r = x;
  while (parent[r] != r) r = parent[r];
  while (parent[x] != r) {
    const uint64_t next = parent[x];
    parent[x] = r;
    x = next;
  }
  return r;
}

static void uf_union(uint64_t *parent, uint64_t a, uint64_t b) {
  a = uf_find(parent, a);
  b = uf_find(parent, b);
  if (a == b) return;
  if (a < b) parent[b] = a;
  else parent[a] = b;
}

/* O(N) embedding clusters, N = 2 or 3 (ClusterModel for ONModel<D>,
 * models.rs:516-577).  The embedding direction r of an update is uniform on
 * the circle or sphere, drawn by rejection from the square or cube: try k
 * takes words 4 + Dk .. 3 + D(k + 1) of the seed stream, x = 2 u - 1 per
 * component, and accepts 2^-40 < |x|^2 <= 1; r is x normalised as in
 * normalize (models.rs:432-439).  The 64th rejection in a row (probability
 * 0.215^64 or 0.476^64) leaves r = the last unit axis.  Sums run left to
 * right; only +, *, / and sqrt are used, all correctly rounded, so the device
 * draws the same bits. */
#define ORC_ON_TRIES 64

static double dotn(int d, const double *a, const double *b) {
  double acc = a[0] * b[0];
  for (int c = 1; c < d; ++c) acc = acc + a[c] * b[c];
  return acc;
}

void orc_on_direction(int d, uint64_t key, uint64_t step, double *r) {
  for (uint64_t k = 0; k < ORC_ON_TRIES; ++k) {
    double x[3];
    for (int c = 0; c < d; ++c)
      x[c] = 2.0 * ((double)orc_cluster_word(key, 4 + (uint64_t)d * k + c, step, ORC_STREAM_CLUSTER_SEED) * 0x1p-32) -
             1.0;
    const double n2 = dotn(d, x, x);
    if (n2 > 0x1p-40 && n2 <= 1.0) {
      const double norm = sqrt(n2);
      for (int c = 0; c < d; ++c) r[c] = x[c] / norm;
      return;
    }
  }
  for (int c = 0; c < d; ++c) r[c] = c == d - 1 ? 1.0 : 0.0;
}

void orc_on3_direction(uint64_t key, uint64_t step, double *r) { orc_on_direction(3, key, step, r); }

/* cluster_bond_probability (models.rs:537-558): p = 1 - exp(-2 beta J w
 * (s_i . r)(s_j . r)) when the product is positive, else 0. */
double orc_on_bond_probability(int d, double j, double weight, double beta, const double *left, const double *right,
                               const double *r) {
  const double coupling = j * weight;
  const double product = dotn(d, left, r) * dotn(d, right, r);
  if (coupling == 0.0 || product <= 0.0) return 0.0;
  return clamp_probability(1.0 - exp(-2.0 * beta * coupling * product));
}

double orc_on3_bond_probability(double j, double weight, double beta, const double *left, const double *right,
                                const double *r) {
  return orc_on_bond_probability(3, j, weight, beta, left, right, r);
}

/* transform_cluster_spin (models.rs:560-575): s - 2 (s . r) r, normalised;
 * then rounded to FP32, the device's storage. */
static void on_reflect(int d, double *s, const double *r) {
  const double proj = dotn(d, s, r);
  double v[3];
  for (int c = 0; c < d; ++c) v[c] = s[c] - 2.0 * proj * r[c];
  const double norm = sqrt(dotn(d, v, v));
  if (norm > 1e-15)
    for (int c = 0; c < d; ++c) v[c] /= norm;
  for (int c = 0; c < d; ++c) s[c] = (double)(float)v[c];
}

/* n Swendsen-Wang or Wolff updates of O(N) spins (D components per site,
 * site-major).  Bonds: word e of the bond stream below p 2^32.  Swendsen-Wang
 * reflects a cluster when its root's target word is below 2^31 (the
 * reference's coin, sw_cluster_auxiliary); Wolff reflects the seed's. */
static uint64_t on_cluster_steps(const orc_model *m, const orc_lattice *lat, double *spins, double beta,
                                 uint64_t key, uint64_t step0, uint64_t n, int wolff) {
  const int d = m->spin_dim;
  const uint64_t ns = lat->n_sites, ne = lat->n_edges;
  uint64_t *parent = (uint64_t *)malloc((ns ? ns : 1) * sizeof(uint64_t));
  uint64_t result = 0;
  for (uint64_t k = 0; k < n; ++k) {
    const uint64_t step = step0 + k;
    double r[3];
    orc_on_direction(d, key, step, r);
    for (uint64_t s = 0; s < ns; ++s) parent[s] = s;
    for (uint64_t e = 0; e < ne; ++e) {
      const orc_bond *b = &lat->edges[e];
      const double p =
          orc_on_bond_probability(d, m->j, b->weight, beta, spins + d * b->source, spins + d * b->target, r);
      if ((double)orc_cluster_word(key, e, step, ORC_STREAM_CLUSTER_BOND) * 0x1p-32 < p)
        uf_union(parent, b->source, b->target);
    }
    result = 0;
    if (!wolff) {
      for (uint64_t s = 0; s < ns; ++s) {
        const uint64_t root = uf_find(parent, s);
        if (root == s) ++result;
        if (orc_cluster_word(key, root, step, ORC_STREAM_CLUSTER_TARGET) < 0x80000000u) on_reflect(d, spins + d * s, r);
      }
    } else if (ns) {
      const uint64_t seed = ((uint64_t)orc_cluster_word(key, 0, step, ORC_STREAM_CLUSTER_SEED) * ns) >> 32;
      const uint64_t root = uf_find(parent, seed);
      for (uint64_t s = 0; s < ns; ++s)
        if (uf_find(parent, s) == root) {
          on_reflect(d, spins + d * s, r);
          ++result;
        }
    }
  }
  free(parent);
  return result;
}

uint64_t orc_cluster_steps(const orc_model *m, const orc_lattice *lat, double *spins, double beta,
                           uint64_t key, uint64_t step0, uint64_t n, int wolff) {
  if (m->kind == ORC_ON)
    return m->spin_dim == 2 || m->spin_dim == 3 ? on_cluster_steps(m, lat, spins, beta, key, step0, n, wolff) : 0;
  const uint64_t ns = lat->n_sites, ne = lat->n_edges;
  uint64_t *parent = (uint64_t *)malloc((ns ? ns : 1) * sizeof(uint64_t));
  uint64_t result = 0;
  for (uint64_t k = 0; k < n; ++k) {
    const uint64_t step = step0 + k;
    for (uint64_t s = 0; s < ns; ++s) parent[s] = s;
    /* every physical edge once, in CsrLattice::edges order (swendsen_wang.rs:85-103) */
    for (uint64_t e = 0; e < ne; ++e) {
      const orc_bond *b = &lat->edges[e];
      if (spins[b->source] != spins[b->target]) continue;
      const uint64_t kk = orc_cluster_threshold(m->kind, m->j, b->weight, beta);
      if ((uint64_t)orc_cluster_word(key, e, step, ORC_STREAM_CLUSTER_BOND) < kk)
        uf_union(parent, b->source, b->target);
    }
    if (!wolff) {
      result = 0;
      for (uint64_t s = 0; s < ns; ++s) {
        const uint64_t r = uf_find(parent, s);
        if (r == s) ++result;
        const uint32_t w = orc_cluster_word(key, r, step, ORC_STREAM_CLUSTER_TARGET);
        spins[s] = m->kind == ORC_ISING ? (w < 0x80000000u ? 1.0 : -1.0)
                                        : (double)(((uint64_t)w * (uint64_t)m->q) >> 32);
      }
    } else if (ns) {
      const uint64_t seed = ((uint64_t)orc_cluster_word(key, 0, step, ORC_STREAM_CLUSTER_SEED) * ns) >> 32;
      double target;
      if (m->kind == ORC_ISING) {
        target = -spins[seed];
      } else {
        const uint64_t s0 = (uint64_t)spins[seed];
        uint64_t r = ((uint64_t)orc_cluster_word(key, 1, step, ORC_STREAM_CLUSTER_SEED) * (uint64_t)(m->q - 1)) >> 32;
        r += r >= s0;
        target = (double)r;
      }
      const uint64_t root = uf_find(parent, seed);
      result = 0;
      for (uint64_t s = 0; s < ns; ++s)
        if (uf_find(parent, s) == root) {
          spins[s] = target;
          ++result;
        }
    }
  }
  free(parent);
  return result;
}
