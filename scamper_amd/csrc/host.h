// Host-side structures of libscamper_hip.so (no device code here).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/scamper_hip.h"
#include "internal.h"

namespace scamper {

// One physical undirected bond (CMC.rs/src/lattice/graph.rs:76-82).
// Bond labels in the order of the reference's BondType (CMC.rs/src/lattice/
// graph.rs:12-28); the serialised labels are BondType::as_label (graph.rs:33-50).
enum BondKind : uint8_t {
  kBondGeneric = 0,
  kBondChainX,
  kBondSquareX,
  kBondSquareY,
  kBondSquareZ,
  kBondCubicX,
  kBondCubicY,
  kBondCubicZ,
  kBondTriX,
  kBondTriY,
  kBondTriDiag,
  kBondHoneyX,
  kBondHoneyY,
  kBondKagome,
};

struct Bond {
  uint64_t source;
  uint64_t target;
  uint8_t kind;  // BondKind
  double weight;
};

// CSR multigraph with the greedy colouring used by every sweep.
struct Lattice {
  uint64_t n_sites = 0;
  std::vector<uint64_t> offsets, neighbors, edge_ids;
  std::vector<Bond> edges;
  uint32_t n_colours = 0;
  std::vector<uint32_t> colour;
  std::vector<uint64_t> class_index, class_start, class_sites;
  // Set when the lattice is a periodic hypercube with even sides (<= 3-D):
  // the sweep then runs on the implicit checkerboard stencil.
  // The CSR arrays of a stencil lattice are built only on request
  // (materialize): at 16384^2 they would take ~28 GB of host memory.
  bool stencil = false;
  bool csr_ready = false;
  int nd = 0;
  uint64_t dims[3] = {1, 1, 1};
  uint64_t n_edge_count = 0;
  uint64_t n_bond_count = 0;

  uint64_t degree(uint64_t s) const { return offsets[s + 1] - offsets[s]; }
};

// Record the calling thread's last error (scamper_last_error) and return code.
int fail(int code, const char *fmt, ...);

#define SC_HIP(call)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) return ::scamper::fail(SCAMPER_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// RngStreamKey(base).with_replica(replica).with_phase(phase).seed()
// (Carlo.rs/src/rng_stream.rs:126-154).
uint64_t stream_seed(uint64_t base, uint64_t replica, uint64_t phase);

// Builders; on failure return false and set `err`.
bool lattice_from_edges(uint64_t n_sites, std::vector<Bond> edges, Lattice &out, std::string &err);
bool lattice_hypercubic(const uint64_t *dims, int n_dims, bool pbc, Lattice &out, std::string &err);
bool lattice_triangular(uint64_t lx, uint64_t ly, Lattice &out, std::string &err);
bool lattice_honeycomb(uint64_t lx, uint64_t ly, Lattice &out, std::string &err);
bool lattice_kagome(uint64_t lx, uint64_t ly, Lattice &out, std::string &err);
// Build the CSR arrays and colouring of a stencil lattice.
bool materialize(Lattice &l, std::string &err);

// Decision thresholds (exact against glibc's log/exp on the host):
//   accept_threshold(a): number of k in [0, 2^32) for which
//     accept_log_probability(a, k * 2^-32) holds (Carlo.rs/src/acceptance.rs:11-22);
//   below_threshold(p):  number of k with k * 2^-32 < p.
uint64_t accept_threshold(double log_acceptance);
uint64_t below_threshold(double p);
double log_acceptance(double beta, double delta_energy);
// Bond threshold of a cluster update: an aligned bond of weight w is active
// when its draw is below it (cluster_bond_probability, models.rs:130-147,
// 312-329, clamped as clamp_probability, models.rs:16-24).
uint64_t cluster_bond_threshold(int kind, double beta, double j, double weight);

// Tables for one replica; `z` is the coordination of the stencil.
void ising_metropolis_table(double beta, double j, int z, Threshold *thr);
void ising_heat_bath_table(double beta, double j, int z, Threshold *thr);
void potts_metropolis_table(double beta, double j, int z, Threshold *thr);
void potts_heat_bath_exp(double beta, double j, int z, double *table /* (z+1)^2 */);
void csr_ising_metropolis_table(double beta, double j, int w_max, Threshold *thr);
void csr_ising_heat_bath_table(double beta, double j, int w_max, Threshold *thr);

}  // namespace scamper

struct scamper_lattice {
  scamper::Lattice lat;
};
