"""A CPU model of the device union-find of the cluster updates
(scamper_amd/csrc/k_cluster.hip: find_root with path halving, unite linking
the larger root under the smaller by compare-and-swap), run under adversarial
interleavings and a weak memory model.

Each device thread is a generator that yields at every memory operation of
the kernel's code.  A random scheduler picks which thread runs next.  Memory
keeps every location's history of values in modification order.  A relaxed
load may return any value not older than the one the loading thread last saw
at that location: per-location coherence and nothing more, which is what
relaxed agent-scope atomics promise.  Compare-and-swap acts on the newest
value.  After all threads finish (the kernel boundary), every site's root
must be the smallest site of its component in the graph of the united edges,
and each parent must be an ancestor no larger than its site.  That is the
invariant the labelling relies on: each cluster's root is its smallest site
whatever order the unions run in.
"""
import random

import numpy as np


class Memory:
    def __init__(self, n):
        self.hist = [[i] for i in range(n)]  # parent[i] starts at i

    def latest(self, x):
        return self.hist[x][-1]


class Thread:
    def __init__(self, mem, rng):
        self.mem, self.rng = mem, rng
        self.seen = {}

    def load(self, x):
        h = self.mem.hist[x]
        lo = self.seen.get(x, 0)
        k = self.rng.randint(lo, len(h) - 1)  # any version not older than the last seen
        self.seen[x] = k
        return h[k]

    def store(self, x, v):
        self.mem.hist[x].append(v)
        self.seen[x] = len(self.mem.hist[x]) - 1

    def cas(self, x, expect, new):
        h = self.mem.hist[x]
        old = h[-1]
        if old == expect:
            h.append(new)
        self.seen[x] = len(h) - 1
        return old

    # find_root (k_cluster.hip:39-48): one yield per memory operation
    def find_root(self, x):
        while True:
            yield
            p = self.load(x)
            if p == x:
                return x
            yield
            gp = self.load(p)
            if gp == p:
                return p
            yield
            self.store(x, gp)
            x = gp

    # unite (k_cluster.hip:50-66)
    def unite(self, a, b):
        while True:
            a = yield from self.find_root(a)
            b = yield from self.find_root(b)
            if a == b:
                return
            if a < b:
                a, b = b, a
            yield
            old = self.cas(a, a, b)
            if old == a:
                return
            a = old

    def run(self, edges):
        for a, b in edges:
            yield from self.unite(a, b)


def components_min(n, edges):
    parent = list(range(n))

    def f(x):
        while parent[x] != x:
            x = parent[x]
        return x
    for a, b in edges:
        ra, rb = f(a), f(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    return [f(x) for x in range(n)]


def simulate(n, edges, n_threads, seed):
    rng = random.Random(seed)
    mem = Memory(n)
    work = [edges[t::n_threads] for t in range(n_threads)]
    threads = [Thread(mem, rng).run(w) for w in work]
    live = list(range(n_threads))
    while live:
        t = rng.choice(live)
        try:
            next(threads[t])
        except StopIteration:
            live.remove(t)
    final = [mem.latest(x) for x in range(n)]

    def root(x):
        while final[x] != x:
            assert final[x] < x, "a parent larger than its site"
            x = final[x]
        return x
    return [root(x) for x in range(n)]


def test_model_matches_sequential_components():
    rng = np.random.default_rng(3)
    for trial in range(300):
        n = int(rng.integers(4, 24))
        m = int(rng.integers(1, 3 * n))
        edges = [tuple(int(v) for v in rng.integers(0, n, 2)) for _ in range(m)]
        want = components_min(n, edges)
        for sched in range(4):
            got = simulate(n, edges, n_threads=int(rng.integers(2, 9)), seed=1000 * trial + sched)
            assert got == want, (trial, sched, edges)


def test_model_on_a_grid_torus():
    """A 6 x 6 torus with every bond active, edges in the kernel's order
    (+x, +y of each site), one thread per edge: one cluster rooted at 0."""
    L = 6
    edges = []
    for y in range(L):
        for x in range(L):
            s = x + L * y
            edges += [(s, (x + 1) % L + L * y), (s, x + L * ((y + 1) % L))]
    for seed in range(20):
        assert simulate(L * L, edges, n_threads=len(edges), seed=seed) == [0] * (L * L)
