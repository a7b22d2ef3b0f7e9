"""Model check of the lock-free union-find the cluster kernels use.

The kernels (scamper_amd/csrc/k_cluster.hip: find_root / unite on global
memory, find_lds / unite_lds on LDS) link the larger root under the smaller
one with a compare-and-swap, halve paths with plain relaxed stores, and read
parents with relaxed loads.  Relaxed atomics guarantee per-location coherence
and no more: a load may return any value of that location's modification
order that is not older than what the same thread saw there before, while a
compare-and-swap always acts on the latest value.  Lanes of one wave running
in lockstep are one of the interleavings modelled here.

The model runs the kernels' operations as threads stepping one memory
operation at a time under random schedules and random stale reads, then
checks what the kernels rely on:
  * the partition equals a sequential union-find's, every set's root is its
    smallest site and parent[x] <= x throughout;
  * with the tile phase's depth-1 forest as the start (each site pointing at
    its tile-local root, the smallest site of its local cluster) and the face
    merges starting from the loaded local roots as k_bonds_faces does, the
    entry of a site that is not a local root is never rewritten (the tile
    transform reads its local root from that entry);
  * root walks with halving but no links, as the transforms run them after a
    kernel boundary, all return the true root.
A mutant that links by load-compare-store instead of a compare-and-swap must
fail the same check, so that the model can tell a racy link from a safe one.
"""
import random

import pytest


class Memory:
    """Per-location modification orders; each thread keeps, per location,
    the index of the newest write it has seen."""

    def __init__(self, values):
        self.hist = [[v] for v in values]

    def latest(self, loc):
        return self.hist[loc][-1]

    def snapshot(self):
        return [h[-1] for h in self.hist]


def run(mem, threads, rng, stale=0.5):
    """Step the threads' generators to completion in a random order.  A load
    returns the latest value with probability 1 - stale, else a uniformly
    chosen value no older than the thread last saw there."""
    seen = [dict() for _ in threads]
    pending = [None] * len(threads)
    live = []
    for i, g in enumerate(threads):
        try:
            pending[i] = next(g)
            live.append(i)
        except StopIteration:
            pass
    steps = 0
    while live:
        steps += 1
        assert steps < 200_000, "no progress"
        k = rng.randrange(len(live))
        i = live[k]
        op = pending[i]
        h = mem.hist[op[1]]
        if op[0] == "ld":
            lo = seen[i].get(op[1], 0)
            idx = len(h) - 1 if rng.random() >= stale else rng.randrange(lo, len(h))
            seen[i][op[1]] = idx
            res = h[idx]
        elif op[0] == "st":
            h.append(op[2])
            seen[i][op[1]] = len(h) - 1
            res = None
        else:  # compare-and-swap on the latest value
            res = h[-1]
            if res == op[2]:
                h.append(op[3])
            seen[i][op[1]] = len(h) - 1
        try:
            pending[i] = threads[i].send(res)
        except StopIteration:
            live.pop(k)


# The kernels' algorithm, one memory operation per yield
# (k_cluster.hip find_root / unite and find_lds / unite_lds).

def find_root(x):
    while True:
        p = yield ("ld", x)
        if p == x:
            return x
        gp = yield ("ld", p)
        if gp == p:
            return p
        yield ("st", x, gp)
        x = gp


def unite(a, b, cas_link=True):
    while True:
        a = yield from find_root(a)
        b = yield from find_root(b)
        if a == b:
            return
        if a < b:
            a, b = b, a
        if cas_link:
            old = yield ("cas", a, a, b)
            if old == a:
                return
            a = old
        else:  # the mutant: check, then store
            cur = yield ("ld", a)
            if cur == a:
                yield ("st", a, b)
                return
            a = cur


def unite_from_loaded(s, t):
    """k_bonds_faces: unite the entries loaded for both ends."""
    a = yield ("ld", s)
    b = yield ("ld", t)
    yield from unite(a, b)


def walk_into(out, x):
    out[x] = yield from find_root(x)


def edges_thread(edges, cas_link=True):
    for a, b in edges:
        yield from unite(a, b, cas_link)


def faces_thread(edges):
    for a, b in edges:
        yield from unite_from_loaded(a, b)


# ------------------------------------------------------------------ checks

def components(n, edges, start=None):
    par = list(range(n)) if start is None else list(start)

    def f(x):
        while par[x] != x:
            x = par[x]
        return x

    for a, b in edges:
        ra, rb = f(a), f(b)
        if ra != rb:
            par[max(ra, rb)] = min(ra, rb)
    return [f(x) for x in range(n)]


def roots_of(values):
    out = []
    for x in range(len(values)):
        seen = 0
        while values[x] != x:
            assert values[x] < x, "parent above its site"
            x = values[x]
            seen += 1
            assert seen <= len(values), "cycle"
        out.append(x)
    return out


def split(edges, n_threads, rng):
    edges = list(edges)
    rng.shuffle(edges)
    return [edges[i::n_threads] for i in range(n_threads)]


def random_graph(rng, n, p):
    return [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < p]


def check_partition(values, truth):
    got = roots_of(values)
    assert got == truth, "partition differs from the sequential union-find"
    for x, r in enumerate(got):
        assert r == min(y for y in range(len(got)) if got[y] == r), "root is not the smallest site"


@pytest.mark.parametrize("trial", range(6))
def test_unite_under_interleavings_and_stale_reads(trial):
    rng = random.Random(1000 + trial)
    for _ in range(60):
        n = rng.randrange(6, 28)
        edges = random_graph(rng, n, rng.uniform(0.05, 0.3))
        threads = [edges_thread(part) for part in split(edges, rng.randrange(2, 9), rng)]
        mem = Memory(range(n))
        run(mem, threads, rng, stale=rng.choice([0.0, 0.3, 0.7]))
        for h in mem.hist:
            assert all(v <= h[0] for v in h), "parent above its site"
        check_partition(mem.snapshot(), components(n, edges))


@pytest.mark.parametrize("trial", range(4))
def test_face_merges_keep_non_local_root_entries(trial):
    """Tiles of 8 sites with a depth-1 forest (the x-run / LDS phase's
    output), then face merges from the loaded local roots, then concurrent
    root walks: every non-local-root entry is unchanged and every walk finds
    the true root."""
    rng = random.Random(2000 + trial)
    for _ in range(40):
        tile, n_tiles = 8, rng.randrange(2, 6)
        n = tile * n_tiles
        start, inner = [], []
        for t in range(n_tiles):
            base = t * tile
            local = [(base + a, base + b) for a, b in random_graph(rng, tile, 0.25)]
            inner += local
            start += components(n, local)[base:base + tile]
        local_roots = {x for x in range(n) if start[x] == x}
        faces = [(a, b) for a in range(n) for b in range(n)
                 if a // tile < b // tile and rng.random() < 0.06]
        mem = Memory(start)
        run(mem, [faces_thread(p) for p in split(faces, rng.randrange(2, 7), rng)], rng,
            stale=rng.choice([0.0, 0.5]))
        after = mem.snapshot()
        for x in range(n):
            if x not in local_roots:
                assert len(mem.hist[x]) == 1, "a non-local-root entry was rewritten"
        truth = components(n, inner + faces)
        check_partition(after, truth)
        # the transform: a kernel boundary, then root walks with halving only
        mem = Memory(after)
        out = {}
        walks = [walk_into(out, x) for x in sorted(local_roots) for _ in range(2)]
        run(mem, walks, rng, stale=0.5)
        assert all(out[x] == truth[x] for x in local_roots)
        check_partition(mem.snapshot(), truth)


def test_model_catches_a_racy_link():
    """A link by load, compare and plain store loses unions when two threads
    link the same root: the model must find such a schedule."""
    rng = random.Random(7)
    failures = 0
    for _ in range(300):
        n = rng.randrange(6, 16)
        edges = random_graph(rng, n, 0.3)
        threads = [edges_thread(part, cas_link=False) for part in split(edges, 4, rng)]
        mem = Memory(range(n))
        run(mem, threads, rng, stale=0.3)
        try:
            check_partition(mem.snapshot(), components(n, edges))
        except AssertionError:
            failures += 1
    assert failures > 0
