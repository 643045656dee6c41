"""Host-side inputs of the path: the matrix generators against the reference's committed fixture and
structure rules, and the right-hand-side RNG against the C++ standard's mt19937_64 (CPU only)."""
import hashlib
import os

import numpy as np
import pytest

from amp_amd import hostsetup as hs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bubbles3d_24_check.npz")


def _digest_matches(A, rhs, ref):
    A = A.tocsr()
    A.sort_indices()
    h = hashlib.sha256()
    h.update(A.indptr.astype("<i8").tobytes())
    h.update(A.indices.astype("<i8").tobytes())
    if A.shape[0] != int(ref["n"]) or A.nnz != int(ref["nnz"]) or h.hexdigest() != ref["pattern_sha256"].item().decode():
        return False
    amax = np.abs(A.data).max()
    w = 1.0 + (np.arange(A.shape[1]) % 7) / 7.0
    ok = np.abs(A @ w - ref["wsum"]).max() <= 1e-13 * amax
    ok &= np.all(np.abs(A.data[::16] - ref["data16"]) <= 1e-14 * np.abs(ref["data16"]))
    ok &= np.abs(A.diagonal() - ref["diag"]).max() <= 1e-14 * amax
    ok &= np.abs(rhs[::4] - ref["rhs4"]).max() <= 1e-13 * np.abs(ref["rhs4"]).max()
    return bool(ok)


BUBBLES_CHECK = dict(nx=24, nbubbles=150, radius=0.04, rho_ratio=1000.0, cfl=0.5, time=3.0, seed=2024)


def test_bubbles3d_matches_reference_fixture():
    """bubbles3d 24^3 as the reference's matgen wrote it (test/gamg/run_gamg_23038420/check_s2l.mtx,
    log.check_matgen:1): identical pattern; values equal to rounding (the reference was built with
    -O3 -march=native, so FMA contraction moves a few last bits: <= 1e-14 relative per value)."""
    ref = np.load(GOLDEN)
    A, rhs, _, sym, _ = hs.generate("bubbles3d", **BUBBLES_CHECK)
    assert sym
    assert _digest_matches(A, rhs, ref)


def test_fixture_check_rejects_a_different_field():
    """The digest check is sensitive: another bubble seed, or one more bubble, fails it."""
    ref = np.load(GOLDEN)
    for change in (dict(seed=2025), dict(nbubbles=151)):
        A, rhs, _, _, _ = hs.generate("bubbles3d", **{**BUBBLES_CHECK, **change})
        assert not _digest_matches(A, rhs, ref)


@pytest.mark.parametrize("time", [0.0, 11.0])
def test_bubbles3d_structure(time):
    """Symmetric; zero row sums below the top layer; face coefficients = mean of beta = 1/rho times
    area / distance with beta from alpha = rhs / cell volume (test/test_bubbles.cpp:58-95)."""
    n, rho = 16, 1000.0
    A, rhs, _, _, _ = hs.generate("bubbles3d", nx=n, nbubbles=12, radius=0.12, cfl=1.0, time=time, rho_ratio=rho)
    h = 1.0 / n
    alpha = rhs / h ** 3
    assert alpha.min() >= -1e-15 and alpha.max() <= 1.0 + 1e-12 and alpha.max() > 0.9
    beta = 1.0 / (alpha / rho + (1.0 - alpha))
    assert abs(A - A.T).max() <= 1e-15 * abs(A).max()
    C = A.tocoo()
    off = C.row != C.col
    expect = 0.5 * (beta[C.row[off]] + beta[C.col[off]]) * h
    assert np.abs(-C.data[off] - expect).max() / expect.min() < 1e-13
    rows = np.arange(A.shape[0])
    below_top = rows // (n * n) != n - 1
    s = np.asarray(A.sum(axis=1)).ravel()
    assert np.abs(s[below_top] / A.diagonal()[below_top]).max() < 1e-12


@pytest.mark.parametrize("name,nx,n,nnz", [("laplace2d", 256, 65536, 326656), ("laplace3d", 64, 262144, 1810432),
                                            ("poisson2d", 256, 65536, 326656), ("poisson3d", 64, 262144, 1810432)])
def test_stencil_sizes(name, nx, n, nnz):
    """5- and 7-point stencils: n and nnz(A) of the 256^2 and 64^3 grids; symmetric, diagonally
    dominant with positive diagonal."""
    A, _, _, sym, _ = hs.generate(name, nx=nx)
    assert A.shape == (n, n) and A.nnz == nnz and sym
    d = A.diagonal()
    assert d.min() > 0
    offsum = np.asarray(abs(A).sum(axis=1)).ravel() - d
    assert np.all(offsum <= d * (1 + 1e-14))
    assert abs(A - A.T).max() == 0.0


def test_aniso_and_high_contrast():
    A, _, _, sym, _ = hs.generate("aniso2d", nx=32, aniso=1e-3)
    assert sym and A.shape == (1024, 1024)
    ratio = sorted(set(np.round(-A.data[A.data < 0], 12)))
    assert ratio[0] / ratio[-1] == pytest.approx(1e-3, rel=1e-9)
    B, _, _, sym, _ = hs.generate("highcontrast3d", nx=16, seed=3, contrast=1e6, channel_fraction=0.1)
    assert sym and B.shape == (4096, 4096)
    assert np.abs(B.data).max() / np.abs(B.data).min() > 1e5


# ---------------------------------------------------------------------- RNG of the right-hand side
class MT19937_64:
    """std::mt19937_64 from the C++ standard's parameters ([rand.predef]); its 10000th output for the
    default seed 5489 is required to be 9981545732273789042."""

    M64 = (1 << 64) - 1

    def __init__(self, seed=5489):
        self.mt = [0] * 312
        self.mt[0] = seed & self.M64
        for i in range(1, 312):
            p = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (p ^ (p >> 62)) + i) & self.M64
        self.i = 312

    def _twist(self):
        mt = self.mt
        for k in range(312):
            x = (mt[k] & ~((1 << 31) - 1) & self.M64) | (mt[(k + 1) % 312] & ((1 << 31) - 1))
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[k] = mt[(k + 156) % 312] ^ xa
        self.i = 0

    def __call__(self):
        if self.i >= 312:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self.M64


def test_mt19937_64_known_answer():
    g = MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


@pytest.mark.parametrize("seed", [12345, 7])
def test_random_rhs_is_libstdcxx_uniform(seed):
    """uniform_real_distribution<double>(-1, 1) over mt19937_64 in libstdc++: one 64-bit draw x per value,
    u = double(x) / 2^64, value = 2 u - 1 (domain_decomposer.cpp:158-164)."""
    n = 600
    g = MT19937_64(seed)
    expect = np.array([float(g()) / 18446744073709551616.0 * 2.0 - 1.0 for _ in range(n)])
    assert np.array_equal(hs.random_rhs(n, seed), expect)


def test_random_rhs_seq_is_deterministic_and_distinct():
    a = hs.random_rhs_seq(500, 12345, 3, 1)
    assert np.array_equal(a, hs.random_rhs_seq(500, 12345, 3, 1))
    assert not np.array_equal(a, hs.random_rhs_seq(500, 12345, 3, 2))
    assert np.all(np.abs(a) <= 1.0)
