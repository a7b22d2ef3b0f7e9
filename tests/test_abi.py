"""CPU checks of the C ABI: the library loads, exports every declared entry
point, and its host-side lattice construction and colouring match the
oracle's restatement of graph.rs (no device calls here)."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(scamper_\w+)\s*\(", text))
    return names


def test_exports_every_declared_symbol(sc):
    lib_path = sc.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\b(scamper_\w+)$", out, flags=re.M))
    missing = _declared() - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    assert set(sc.SIGNATURES) == _declared()


def test_version(sc):
    assert sc.lib().scamper_version().decode().endswith("gfx950")


@pytest.mark.parametrize("builder,args", [
    ("hypercubic", ([4, 4], True)), ("hypercubic", ([5, 3], True)), ("hypercubic", ([4, 4], False)),
    ("hypercubic", ([2, 2], True)), ("hypercubic", ([3, 4, 2], True)), ("hypercubic", ([6], True)),
    ("triangular", (3, 3)), ("honeycomb", (4, 4)), ("kagome", (3, 3)),
])
def test_lattice_matches_oracle(sc, orc, builder, args):
    lat = getattr(sc.Lattice, builder)(*args)
    ol = getattr(orc.Lattice, builder)(*args)
    assert (lat.n_sites, lat.n_edges, lat.n_bonds) == (ol.n_sites, ol.n_edges, ol.n_bonds)
    off, nb, eid = lat.csr()
    ocsr = ol.csr()
    assert np.array_equal(off, ocsr["offsets"])
    assert np.array_equal(nb, ocsr["neighbors"])
    assert np.array_equal(eid, ocsr["edge_ids"])
    col, idx = lat.colours()
    ocol, oidx = ol.colours()
    assert np.array_equal(col, ocol) and np.array_equal(idx, oidx)
    assert lat.n_colours == ol.n_colours


def test_reference_graph_counts(sc):
    # graph.rs:549-575
    assert sc.Lattice.chain(4, False).n_bonds == 6
    assert sc.Lattice.chain(4, True).n_bonds == 8
    assert sc.Lattice.chain(1, True).n_bonds == 0
    assert sc.Lattice.square(2, 2, False).n_bonds == 8
    assert sc.Lattice.square(2, 2, True).n_bonds == 16
    assert sc.Lattice.square(4, 4, True).n_edges == 32
    tri = sc.Lattice.triangular(3, 3)
    off, _, _ = tri.csr()
    assert all(np.diff(off) == 6)
    for lx, ly in [(2, 3), (4, 4), (6, 5)]:
        off, _, _ = sc.Lattice.honeycomb(lx, ly).csr()
        assert all(np.diff(off) == 3)
    off, _, _ = sc.Lattice.kagome(3, 3).csr()
    assert all(np.diff(off) == 4)


def test_checkerboard_colouring(sc):
    for dims in [(8, 8), (16, 4), (4, 6, 8), (10,)]:
        lat = sc.Lattice.hypercubic(list(dims))
        col, idx = lat.colours()
        sites = np.arange(lat.n_sites)
        coords = np.unravel_index(sites, dims[::-1])
        parity = sum(coords) & 1
        assert np.array_equal(col, parity.astype(np.uint32))
        assert np.array_equal(idx, sites >> 1)
        assert lat.n_colours == 2


def test_odd_periodic_lattice_is_properly_coloured(sc):
    # an odd periodic ring is not bipartite, so at least three classes
    lat = sc.Lattice.square(5, 5)
    assert lat.n_colours >= 3
    off, nb, _ = lat.csr()
    col, _ = lat.colours()
    for s in range(lat.n_sites):
        assert all(col[s] != col[int(t)] for t in nb[off[s]:off[s + 1]])


def test_invalid_edges_rejected(sc):
    with pytest.raises(sc.ScamperError):
        sc.Lattice.from_edges(2, [0], [5])
    with pytest.raises(sc.ScamperError):
        sc.Lattice.from_edges(2, [0], [1], [float("inf")])
    with pytest.raises(sc.ScamperError):
        sc.Lattice.from_edges(0, [], [])


def test_short_arrays_never_reach_the_library(sc):
    """Uploads hand the library a pointer it reads a fixed count from; the
    bindings refuse arrays of any other length first."""
    a = np.zeros(7, np.int8)
    assert sc._sized(a, 7, "spins") is a
    for n in (6, 8):
        with pytest.raises(ValueError, match="expected"):
            sc._sized(a, n, "spins")
