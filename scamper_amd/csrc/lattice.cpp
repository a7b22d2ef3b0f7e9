// Lattice construction on the host: CSR multigraphs as the reference builds
// them (CMC.rs/src/lattice/graph.rs), plus the greedy colouring that orders
// every sweep.
#include <algorithm>
#include <cmath>
#include <map>

#include "host.h"

namespace scamper {
namespace {

bool validate(const Lattice &l, std::string &err) {
  // CsrLattice::validate (graph.rs:259-326)
  std::vector<uint32_t> inc(l.edges.size(), 0), first(l.edges.size(), 0), second(l.edges.size(), 0);
  for (uint64_t s = 0; s < l.n_sites; ++s)
    for (uint64_t i = l.offsets[s]; i < l.offsets[s + 1]; ++i) {
      const uint64_t nb = l.neighbors[i], e = l.edge_ids[i];
      if (nb >= l.n_sites || e >= l.edges.size()) {
        err = "invalid incidence " + std::to_string(i) + " at site " + std::to_string(s);
        return false;
      }
      const Bond &b = l.edges[e];
      const uint64_t other = s == b.source ? b.target : (s == b.target ? b.source : UINT64_MAX);
      if (other != nb) {
        err = "incidence " + std::to_string(i) + " does not match physical edge " + std::to_string(e);
        return false;
      }
      inc[e]++;
      if (b.source == b.target || s == b.source) first[e]++; else second[e]++;
    }
  for (size_t e = 0; e < l.edges.size(); ++e) {
    const bool loop = l.edges[e].source == l.edges[e].target;
    if (inc[e] != 2 || (loop ? (first[e] != 2 || second[e] != 0) : (first[e] != 1 || second[e] != 1))) {
      err = "physical edge " + std::to_string(e) + " has invalid incidences";
      return false;
    }
  }
  return true;
}

// Greedy colouring in ascending site order: every site takes the smallest
// colour that no lower-numbered neighbour holds.  Same-colour sites share
// no bond, so a colour class can be updated in parallel; on periodic
// hypercubes with even sides it reproduces the checkerboard parity.
void colour(Lattice &l) {
  const uint64_t n = l.n_sites;
  l.colour.assign(n, 0);
  std::vector<uint64_t> mark;
  uint32_t nc = 0;
  for (uint64_t s = 0; s < n; ++s) {
    mark.assign(l.degree(s) + 2, UINT64_MAX);
    for (uint64_t i = l.offsets[s]; i < l.offsets[s + 1]; ++i) {
      const uint64_t nb = l.neighbors[i];
      if (nb < s && l.colour[nb] < mark.size()) mark[l.colour[nb]] = s;
    }
    uint32_t c = 0;
    while (mark[c] == s) ++c;
    l.colour[s] = c;
    nc = std::max(nc, c + 1);
  }
  l.n_colours = nc;
  l.class_start.assign(nc + 1, 0);
  for (uint64_t s = 0; s < n; ++s) l.class_start[l.colour[s] + 1]++;
  for (uint32_t c = 0; c < nc; ++c) l.class_start[c + 1] += l.class_start[c];
  l.class_sites.assign(n, 0);
  l.class_index.assign(n, 0);
  std::vector<uint64_t> cur(nc, 0);
  for (uint64_t s = 0; s < n; ++s) {
    const uint32_t c = l.colour[s];
    l.class_index[s] = cur[c];
    l.class_sites[l.class_start[c] + cur[c]++] = s;
  }
}

}  // namespace

// CsrLattice::try_from_edges (graph.rs:131-184).
bool lattice_from_edges(uint64_t n_sites, std::vector<Bond> edges, Lattice &out, std::string &err) {
  if (n_sites == 0) {
    err = "lattice must contain at least one site";
    return false;
  }
  std::vector<uint64_t> degree(n_sites, 0);
  for (size_t e = 0; e < edges.size(); ++e) {
    const Bond &b = edges[e];
    if (b.source >= n_sites || b.target >= n_sites) {
      err = "edge " + std::to_string(e) + " endpoint out of range";
      return false;
    }
    if (!std::isfinite(b.weight)) {
      err = "edge " + std::to_string(e) + " has non-finite weight";
      return false;
    }
    degree[b.source]++;
    degree[b.target]++;
  }
  Lattice l;
  l.n_sites = n_sites;
  l.offsets.assign(n_sites + 1, 0);
  for (uint64_t s = 0; s < n_sites; ++s) l.offsets[s + 1] = l.offsets[s] + degree[s];
  const uint64_t nb = l.offsets[n_sites];
  l.neighbors.assign(nb, 0);
  l.edge_ids.assign(nb, 0);
  std::vector<uint64_t> cursor(l.offsets.begin(), l.offsets.end() - 1);
  for (size_t e = 0; e < edges.size(); ++e) {
    const Bond &b = edges[e];
    const uint64_t left = cursor[b.source]++;
    l.neighbors[left] = b.target;
    l.edge_ids[left] = e;
    const uint64_t right = cursor[b.target]++;
    l.neighbors[right] = b.source;
    l.edge_ids[right] = e;
  }
  l.edges = std::move(edges);
  if (!validate(l, err)) return false;
  colour(l);
  l.csr_ready = true;
  l.n_edge_count = l.edges.size();
  l.n_bond_count = l.neighbors.size();
  out = std::move(l);
  return true;
}

namespace {
// build_hypercubic (graph.rs:417-451): site = x + Lx*(y + Ly*z), one bond
// to +1 along every axis, wrapped when periodic and the side exceeds one.
// Bond labels as build_chain / build_square (graph.rs:402-414) and the 3-D
// lattices of build_lattice_from_params (classical_mc.rs:361-383) give them;
// generic beyond three dimensions.
bool build_hypercubic_csr(const uint64_t *dims, int n_dims, bool pbc, Lattice &out, std::string &err) {
  static const uint8_t kAxisKind[3][3] = {{kBondChainX, 0, 0},
                                          {kBondSquareX, kBondSquareY, 0},
                                          {kBondCubicX, kBondCubicY, kBondCubicZ}};
  uint64_t strides[8], n = 1;
  for (int a = 0; a < n_dims; ++a) {
    strides[a] = n;
    n *= dims[a];
  }
  std::vector<Bond> edges;
  edges.reserve(n * n_dims);
  for (uint64_t site = 0; site < n; ++site) {
    uint64_t rem = site, coords[8];
    for (int a = n_dims - 1; a >= 0; --a) {
      coords[a] = rem / strides[a];
      rem %= strides[a];
    }
    for (int a = 0; a < n_dims; ++a) {
      uint64_t next;
      if (coords[a] + 1 < dims[a]) next = coords[a] + 1;
      else if (pbc && dims[a] > 1) next = 0;
      else continue;
      const uint8_t kind = n_dims <= 3 ? kAxisKind[n_dims - 1][a] : (uint8_t)kBondGeneric;
      edges.push_back(Bond{site, site - coords[a] * strides[a] + next * strides[a], kind, 1.0});
    }
  }
  return lattice_from_edges(n, std::move(edges), out, err);
}
}  // namespace

bool lattice_hypercubic(const uint64_t *dims, int n_dims, bool pbc, Lattice &out, std::string &err) {
  if (n_dims <= 0 || n_dims > 8) {
    err = "between one and eight lattice dimensions are supported";
    return false;
  }
  uint64_t n = 1;
  for (int a = 0; a < n_dims; ++a) {
    if (dims[a] == 0) {
      err = "lattice dimensions must be positive";
      return false;
    }
    n *= dims[a];
  }
  bool even = pbc && n_dims <= 3;
  for (int a = 0; a < n_dims; ++a) even = even && dims[a] >= 2 && dims[a] % 2 == 0;
  if (!even) return build_hypercubic_csr(dims, n_dims, pbc, out, err);
  // Periodic, even sides: the checkerboard stencil needs no CSR.  The greedy
  // colouring of this graph is the parity of x+y+z and a site's rank inside
  // its class is site >> 1 (every lower-numbered neighbour has the other
  // parity, so by induction each site takes colour = parity).
  Lattice l;
  l.n_sites = n;
  l.stencil = true;
  l.nd = n_dims;
  for (int a = 0; a < n_dims; ++a) l.dims[a] = dims[a];
  l.n_colours = 2;
  l.n_edge_count = n * (uint64_t)n_dims;
  l.n_bond_count = 2 * l.n_edge_count;
  out = std::move(l);
  return true;
}

bool materialize(Lattice &l, std::string &err) {
  if (l.csr_ready) return true;
  Lattice full;
  if (!build_hypercubic_csr(l.dims, l.nd, true, full, err)) return false;
  full.stencil = l.stencil;
  full.nd = l.nd;
  for (int a = 0; a < 3; ++a) full.dims[a] = l.dims[a];
  l = std::move(full);
  return true;
}

// build_triangular (graph.rs:454-472).
bool lattice_triangular(uint64_t lx, uint64_t ly, Lattice &out, std::string &err) {
  if (lx < 2 || ly < 2) {
    err = "triangular lattice needs Lx,Ly >= 2";
    return false;
  }
  std::vector<Bond> e;
  for (uint64_t y = 0; y < ly; ++y)
    for (uint64_t x = 0; x < lx; ++x) {
      const uint64_t s = y * lx + x;
      e.push_back(Bond{s, y * lx + (x + 1) % lx, kBondTriX, 1.0});
      e.push_back(Bond{s, ((y + 1) % ly) * lx + x, kBondTriY, 1.0});
      e.push_back(Bond{s, ((y + 1) % ly) * lx + (x + 1) % lx, kBondTriDiag, 1.0});
    }
  return lattice_from_edges(lx * ly, std::move(e), out, err);
}

namespace {
// CsrLattice::try_from_adjacency (graph.rs:194-249) for loop-free rows.
bool from_adjacency(const std::vector<std::vector<uint64_t>> &rows, Lattice &out, std::string &err) {
  std::map<std::pair<uint64_t, uint64_t>, std::pair<uint64_t, uint64_t>> counts;
  for (uint64_t s = 0; s < rows.size(); ++s)
    for (uint64_t t : rows[s]) {
      if (s == t) {
        err = "self-loop adjacency is not produced by the builders";
        return false;
      }
      auto &c = counts[{std::min(s, t), std::max(s, t)}];
      (s < t ? c.first : c.second)++;
    }
  std::vector<Bond> e;
  for (const auto &kv : counts) {
    if (kv.second.first != kv.second.second) {
      err = "adjacency is not symmetric";
      return false;
    }
    for (uint64_t k = 0; k < kv.second.first; ++k)
      e.push_back(Bond{kv.first.first, kv.first.second, kBondGeneric, 1.0});
  }
  return lattice_from_edges(rows.size(), std::move(e), out, err);
}
}  // namespace

// build_honeycomb (graph.rs:475-505).
bool lattice_honeycomb(uint64_t lx, uint64_t ly, Lattice &out, std::string &err) {
  if (lx < 2 || ly < 2 || lx % 2) {
    err = "honeycomb lattice needs even Lx >= 2 and Ly >= 2";
    return false;
  }
  std::vector<std::vector<uint64_t>> adj(lx * ly);
  for (uint64_t y = 0; y < ly; ++y)
    for (uint64_t x = 0; x < lx; ++x) {
      const uint64_t matched = x % 2 == 0 ? ((y + 1) % ly) * lx + (x + 1) % lx
                                          : ((y + ly - 1) % ly) * lx + (x + lx - 1) % lx;
      adj[y * lx + x] = {y * lx + (x + 1) % lx, y * lx + (x + lx - 1) % lx, matched};
    }
  if (!from_adjacency(adj, out, err)) return false;
  // labels (graph.rs:497-504): honey_x within a row, honey_y across rows
  for (Bond &b : out.edges) b.kind = b.source / lx == b.target / lx ? kBondHoneyX : kBondHoneyY;
  return true;
}

// build_kagome (graph.rs:508-543).
bool lattice_kagome(uint64_t lx, uint64_t ly, Lattice &out, std::string &err) {
  if (lx < 2 || ly < 2) {
    err = "kagome lattice needs Lx,Ly >= 2";
    return false;
  }
  auto idx = [&](uint64_t sub, uint64_t x, uint64_t y) { return sub + 3 * (x + y * lx); };
  std::vector<std::vector<uint64_t>> adj(3 * lx * ly);
  for (uint64_t y = 0; y < ly; ++y)
    for (uint64_t x = 0; x < lx; ++x) {
      adj[idx(0, x, y)] = {idx(1, x, y), idx(2, x, y), idx(1, (x + lx - 1) % lx, y), idx(2, x, (y + ly - 1) % ly)};
      adj[idx(1, x, y)] = {idx(0, x, y), idx(2, x, y), idx(0, (x + 1) % lx, y), idx(2, (x + lx - 1) % lx, y)};
      adj[idx(2, x, y)] = {idx(0, x, y), idx(1, x, y), idx(0, x, (y + 1) % ly), idx(1, (x + 1) % lx, y)};
    }
  if (!from_adjacency(adj, out, err)) return false;
  for (Bond &b : out.edges) b.kind = kBondKagome;  // graph.rs:539-541
  return true;
}

}  // namespace scamper
