"""scamper_ising_sweep_noalloc against scamper_sweep, bit for bit.

The no-allocation entry point sweeps replicas held in a device buffer that
the caller owns (allocated here through the HIP runtime's own hipMalloc, as
such a caller would), keeping its tables and counters in the library's fixed
device array.  Two calls in a row, the second on a smaller lattice with
fewer replicas and a later sweep counter, must each give the spins, energies
and magnetisations that a state of the same seed gives through
scamper_sweep.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _split(site_major, dims):
    """Site-major spins -> the colour-split layout (colour 0, then colour 1;
    rows of Lx/2 bytes, element i of row (y, z) at x = 2i + ((c + y + z) & 1))."""
    lx = dims[0]
    rest = int(np.prod(dims[1:]))
    a = np.asarray(site_major, np.int8).reshape(rest, lx)
    ly = dims[1]
    out = np.empty(lx * rest, np.int8)
    nh = lx * rest // 2
    rows = np.arange(rest)
    par = ((rows % ly) + (rows // ly)) & 1  # (y + z) & 1 per row
    for c in (0, 1):
        start = (c + par) & 1  # x of element 0 in each row
        idx = start[:, None] + 2 * np.arange(lx // 2)[None, :]
        out[c * nh:(c + 1) * nh] = a[rows[:, None], idx].ravel()
    return out


def _run_case(sc, hip, dims, betas, seed, sweep0, n_sweeps):
    n_rep = len(betas)
    lat = sc.Lattice.hypercubic(list(dims))
    n = lat.n_sites
    st = sc.State(lat, sc.model(sc.ISING), n_rep, seed)
    starts = []
    for r, b in enumerate(betas):
        st.init(True, r)
        st.set_beta(float(b), r)
        starts.append(st.download_i8(r))
    st.set_sweeps(sweep0)
    if n_sweeps > 1:
        st.sweep(sc.METROPOLIS, n_sweeps - 1)
    st.sweep(sc.METROPOLIS, 1, measure=True)

    host = np.concatenate([_split(s, dims) for s in starts])
    ptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(ptr), host.nbytes) == 0
    try:
        assert hip.hipMemcpy(ptr, host.ctypes.data, host.nbytes, H2D) == 0
        d = np.asarray(dims, np.uint64)
        b = np.asarray(betas, np.float64)
        e = np.zeros(n_rep)
        m = np.zeros(n_rep)
        rc = sc.lib().scamper_ising_sweep_noalloc(ptr, d.ctypes.data, len(dims), 1.0, b.ctypes.data, n_rep, seed, 0,
                                                  sweep0, n_sweeps, e.ctypes.data, m.ctypes.data, None)
        assert rc == 0, sc.last_error()
        got = np.empty_like(host)
        assert hip.hipMemcpy(got.ctypes.data, ptr, host.nbytes, D2H) == 0
    finally:
        hip.hipFree(ptr)
    for r in range(n_rep):
        want = _split(st.download_i8(r), dims)
        assert np.array_equal(got[r * n:(r + 1) * n], want), f"replica {r}: spins"
        e_r, m_r = st.series(r)
        assert e[r] == e_r[-1] and m[r] == m_r[-1], f"replica {r}: energy / magnetisation"
    st.close()


def test_noalloc_equals_state_sweep(sc, hip):
    print("case noalloc 256x256, 4 replicas, then 64x32, 2 replicas", flush=True)
    _run_case(sc, hip, (256, 256), [0.3, 0.4406868, 0.6, 1.0], seed=9, sweep0=0, n_sweeps=5)
    _run_case(sc, hip, (64, 32), [0.25, 0.5], seed=10, sweep0=3, n_sweeps=3)


def test_noalloc_3d(sc, hip):
    print("case noalloc 32x16x8, 3 replicas", flush=True)
    _run_case(sc, hip, (32, 16, 8), [0.2, 0.2216544, 0.3], seed=4, sweep0=7, n_sweeps=4)


def test_noalloc_rejects_sizes_beyond_the_array(sc, hip):
    d = np.asarray([64, 64], np.uint64)
    b = np.full(65, 0.4)
    ptr = C.c_void_p(1)  # never dereferenced: the call fails on n_rep first
    rc = sc.lib().scamper_ising_sweep_noalloc(ptr, d.ctypes.data, 2, 1.0, b.ctypes.data, 65, 1, 0, 0, 1, None, None,
                                              None)
    assert rc == 1 and "1..64" in sc.last_error()
    d2 = np.asarray([48, 48], np.uint64)  # Lx/2 = 24, not a multiple of 16
    rc = sc.lib().scamper_ising_sweep_noalloc(ptr, d2.ctypes.data, 2, 1.0, b.ctypes.data, 1, 1, 0, 0, 1, None, None,
                                              None)
    assert rc == 5
