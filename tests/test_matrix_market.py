"""Matrix Market input/output, step-file sequences and the pattern hash of the host side
(src/partition/matrix_market_io.cpp, src/partition/domain_decomposer.cpp:80-175,345-352), CPU only.

What pins them: the reference's committed generator fixture (test/gamg/run_gamg_23038420/
check_s2l*.mtx) read with this reader matches the digest in tests/golden (when the reference is
present), and a write -> read round trip of the generated matrix matches it too; scipy.io.mmread,
an independent parser, on files of every supported header; Python's printf formatting for the
step names; the symmetry-detection cases of the reference's test/test_subdomain.cpp:105-147. The
OpenFOAM LDU adapter is pinned by lduMatrix::Amul's face loop, by the OpenFOAM solver's own export
format read back, and by the reference's check_of.mtx split into cells and faces."""
import os

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from amp_amd import hostsetup as hs
from test_generators import GOLDEN, _digest_matches

REF_DIR = "/root/reference/test/gamg/run_gamg_23038420"


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


def _same_matrix(A, B):
    A = sp.csr_matrix(A)
    B = sp.csr_matrix(B)
    A.sort_indices()
    B.sort_indices()
    return (A.shape == B.shape and np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
            and np.array_equal(A.data, B.data))


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "check_s2l.mtx")), reason="reference fixture absent")
def test_reads_reference_fixture_to_the_golden_digest():
    """check_s2l.mtx (bubbles3d 24^3, 93,312 entries) and its array rhs, read by this reader."""
    A, sym = hs.read_matrix(os.path.join(REF_DIR, "check_s2l.mtx"))
    b = hs.read_vector(os.path.join(REF_DIR, "check_s2l_rhs.mtx"))
    assert not sym  # the header says general
    assert _digest_matches(A, b, np.load(GOLDEN))


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "check_of.mtx")), reason="reference fixture absent")
def test_reads_reference_openfoam_matrix_like_scipy():
    """check_of.mtx (the OpenFOAM pressure matrix of the same case) against scipy's parser."""
    A, _ = hs.read_matrix(os.path.join(REF_DIR, "check_of.mtx"))
    S = sp.csr_matrix(sio.mmread(os.path.join(REF_DIR, "check_of.mtx")))
    S.eliminate_zeros()
    assert _same_matrix(A, S)
    b = hs.read_vector(os.path.join(REF_DIR, "check_of_rhs.mtx"))
    assert np.array_equal(b, np.asarray(sio.mmread(os.path.join(REF_DIR, "check_of_rhs.mtx"))).ravel())


def test_generated_round_trip_matches_golden_digest(tmp_path):
    """bubbles3d 24^3 with the fixture's parameters, written and read back: bitwise the generated
    matrix (17 significant digits round-trip every double), and the fixture's digest."""
    A, rhs, _, _, _ = hs.generate("bubbles3d", nx=24, nbubbles=150, radius=0.04, rho_ratio=1000.0, cfl=0.5,
                                  time=3.0, seed=2024)
    hs.write_matrix(tmp_path / "a.mtx", A)
    hs.write_vector(tmp_path / "b.mtx", rhs)
    B, _ = hs.read_matrix(tmp_path / "a.mtx")
    c = hs.read_vector(tmp_path / "b.mtx")
    assert _same_matrix(A, B)
    assert np.array_equal(rhs, c)
    assert _digest_matches(B, c, np.load(GOLDEN))
    # scipy reads the written file to the same matrix
    assert _same_matrix(sp.csr_matrix(sio.mmread(str(tmp_path / "a.mtx"))), B)


@pytest.mark.parametrize("field,symmetry", [("real", "general"), ("real", "symmetric"), ("real", "skew-symmetric"),
                                            ("integer", "general"), ("integer", "symmetric"),
                                            ("pattern", "general"), ("pattern", "symmetric")])
def test_headers_like_scipy(tmp_path, field, symmetry):
    """Every supported header on a random 40 x 40 matrix: symmetric storage expanded, the mirror
    negated for skew-symmetric, against scipy.io.mmread of the same file."""
    rng = np.random.default_rng(7)
    n = 40
    entries = set()
    while len(entries) < 120:
        i, j = (int(x) for x in rng.integers(0, n, 2))
        if symmetry != "general" and i < j:
            i, j = j, i
        if symmetry == "skew-symmetric" and i == j:
            continue
        entries.add((i, j))
    lines = [f"%%MatrixMarket matrix coordinate {field} {symmetry}", "% a comment", "", f"{n} {n} {len(entries)}"]
    for i, j in sorted(entries):
        v = "" if field == "pattern" else (f" {int(rng.integers(1, 9))}" if field == "integer" else f" {rng.normal():.17g}")
        lines.append(f"{i + 1} {j + 1}{v}")
    path = _write(tmp_path / "m.mtx", "\n".join(lines) + "\n")
    A, sym = hs.read_matrix(path)
    S = sp.csr_matrix(sio.mmread(path)).astype(np.float64)
    assert sym == (symmetry == "symmetric")
    assert _same_matrix(A, S)


def test_zeros_duplicates_case_and_comments(tmp_path):
    """Explicit zeros are dropped, duplicates summed, the banner is read case-insensitively and
    comment / empty lines after it are skipped (matrix_market_io.cpp:62-81,118-121)."""
    path = _write(tmp_path / "m.mtx", "%%MatrixMarket MATRIX Coordinate REAL General\n%c1\n\n%c2\n3 3 6\n"
                  "1 1 2.0\n1 2 0.0\n2 2 1.5\n2 2 2.5\n3 1 -1\n3 3 1e-3\n")
    A, sym = hs.read_matrix(path)
    assert not sym
    assert A.nnz == 4
    assert np.array_equal(A.toarray(), np.array([[2.0, 0, 0], [0, 4.0, 0], [-1.0, 0, 1e-3]]))


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1 0\n", "complex"),
    ("%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n1 1 1\n", "complex"),
    ("%%MatrixMarket matrix array real general\n2 1\n1\n2\n", "only coordinate"),
    ("%%MatrixMarket vector coordinate real general\n2 2 1\n1 1 1\n", "not a Matrix Market"),
    ("just text\n", "not a Matrix Market"),
    ("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1\n", "square"),
    ("%%MatrixMarket matrix coordinate real general\n0 0 0\n", "invalid size"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n", "end of file"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", "out of range"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1\n", "out of range"),
])
def test_matrix_errors(tmp_path, text, msg):
    path = _write(tmp_path / "bad.mtx", text)
    with pytest.raises(RuntimeError, match=msg):
        hs.read_matrix(path)


def test_missing_file(tmp_path):
    with pytest.raises(RuntimeError, match="cannot open"):
        hs.read_matrix(tmp_path / "none.mtx")


def test_vectors(tmp_path):
    """Array with one column, coordinate (entries added into their rows), errors."""
    v = np.random.default_rng(3).normal(size=17)
    hs.write_vector(tmp_path / "v.mtx", v)
    assert np.array_equal(hs.read_vector(tmp_path / "v.mtx"), v)
    path = _write(tmp_path / "c.mtx", "%%MatrixMarket matrix coordinate real general\n4 1 3\n2 1 1.5\n4 1 -2\n2 1 0.5\n")
    assert np.array_equal(hs.read_vector(path), [0.0, 2.0, 0.0, -2.0])
    path = _write(tmp_path / "w.mtx", "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(RuntimeError, match="one column"):
        hs.read_vector(path)
    path = _write(tmp_path / "s.mtx", "%%MatrixMarket matrix array real general\n3 1\n1\n2\n")
    with pytest.raises(RuntimeError, match="end of file"):
        hs.read_vector(path)
    assert hs.read_vector(_write(tmp_path / "e.mtx", "%%MatrixMarket matrix array real general\n0 1\n")).size == 0


@pytest.mark.parametrize("pattern", ["p_%04d.mtx", "p_%d.mtx", "p_%4d.mtx", "m%02d", "%d.mtx", "seq/%06d/A.mtx"])
def test_step_file_names_like_printf(pattern):
    for step in (0, 7, 123, 99999):
        assert hs.step_file_name(pattern, step) == pattern % step


def test_step_file_name_rules():
    assert hs.step_file_name("A.mtx", 5) == "A.mtx"
    for bad in ("p_%s.mtx", "p_%d_%d.mtx", "p_%", "p_%04x.mtx", "p_%%d.mtx"):
        with pytest.raises(ValueError, match="exactly one"):
            hs.step_file_name(bad, 1)


def _fnv_bytes(words):
    # FNV-1a over a byte string, with the reference's start value
    h = 1469598103934665603
    for byte in np.asarray(words, "<i8").tobytes():
        h = ((h ^ byte) * 1099511628211) % (1 << 64)
    return h


def test_pattern_hash():
    """FNV-1a steps over the bytes of n, rowptr and col (domain_decomposer.cpp:116-130): the same
    pattern with other values hashes equal; a moved column, a different n or an extra entry do not."""
    A = sp.csr_matrix(hs.generate("poisson2d", nx=6)[0])
    B = A.copy()
    B.data = B.data * 3.0 + 1.0
    assert hs.hash_pattern(A) == hs.hash_pattern(B)
    assert hs.hash_pattern(A) == _fnv_bytes(np.concatenate([[A.shape[0]], A.indptr, A.indices]))
    C = A.copy()
    C.indices = C.indices.copy()
    C.indices[3] += 1 if C.indices[3] + 1 < A.shape[0] else -1
    assert hs.hash_pattern(C) != hs.hash_pattern(A)
    D = sp.csr_matrix(hs.generate("poisson2d", nx=7)[0])
    assert hs.hash_pattern(D) != hs.hash_pattern(A)
    E = A.tolil()
    E[0, A.shape[0] - 1] = 1.0
    assert hs.hash_pattern(sp.csr_matrix(E)) != hs.hash_pattern(A)
    # FNV-1a with its published offset basis would give another value: the start value matters
    h = 14695981039346656037
    for byte in np.concatenate([[A.shape[0]], A.indptr, A.indices]).astype("<i8").tobytes():
        h = ((h ^ byte) * 1099511628211) % (1 << 64)
    assert h != hs.hash_pattern(A)


@pytest.mark.parametrize("lower,upper,symmetric", [(-1.0, 0.0, False), (0.0, -1.0, False), (-1.0, -1.0, True),
                                                   (-1.0, -0.5, False)])
def test_symmetry_detection_cases(tmp_path, lower, upper, symmetric):
    """The cases of test/test_subdomain.cpp:105-147: a file with a general header is symmetric
    when its values are, whatever it stores."""
    n = 64
    rows, cols, vals = [], [], []
    for i in range(n):
        rows.append(i), cols.append(i), vals.append(4.0)
        if i > 0 and lower != 0.0:
            rows.append(i), cols.append(i - 1), vals.append(lower)
        if i > 0 and upper != 0.0:
            rows.append(i - 1), cols.append(i), vals.append(upper)
    hs.write_matrix(tmp_path / "s.mtx", sp.csr_matrix((vals, (rows, cols)), shape=(n, n)))
    p = hs.load_problem(matrix=str(tmp_path / "s.mtx"), rhs="ones")
    assert p.symmetric == symmetric
    assert np.array_equal(p.b, np.ones(n))


def test_load_problem(tmp_path):
    """loadProblem (domain_decomposer.cpp:133-175): step patterns, the right-hand-side kinds, the
    generator path and the errors."""
    A = hs.generate("poisson2d", nx=8)[0]
    for s in range(2):
        hs.write_matrix(tmp_path / f"p_{s:02d}.mtx", A * (1.0 + s))
    p = hs.load_problem(matrix=str(tmp_path / "p_%02d.mtx"), step=1, seed=99)
    assert _same_matrix(p.A, A * 2.0) and p.symmetric and p.block_size == 1
    assert np.array_equal(p.b, hs.random_rhs(A.shape[0], 99))
    assert "p_01.mtx" in p.description
    hs.write_vector(tmp_path / "b.mtx", np.arange(A.shape[0], dtype=float))
    assert np.array_equal(hs.load_problem(matrix=str(tmp_path / "p_00.mtx"), rhs=str(tmp_path / "b.mtx")).b,
                          np.arange(A.shape[0]))
    hs.write_vector(tmp_path / "short.mtx", np.ones(3))
    with pytest.raises(RuntimeError, match="size does not match"):
        hs.load_problem(matrix=str(tmp_path / "p_00.mtx"), rhs=str(tmp_path / "short.mtx"))
    with pytest.raises(RuntimeError, match="requires --generate"):
        hs.load_problem(matrix=str(tmp_path / "p_00.mtx"), rhs="generated")
    with pytest.raises(RuntimeError, match="block size"):
        hs.load_problem(matrix=str(tmp_path / "p_00.mtx"), block_size=5)
    g = hs.load_problem(generate_family="bubbles2d", rhs="generated", step=2, gen_kw=dict(nx=12, nbubbles=3))
    A2, rhs2, _, sym2, _ = hs.generate("bubbles2d", nx=12, nbubbles=3, time=2.0)
    assert _same_matrix(g.A, A2) and np.array_equal(g.b, rhs2) and g.symmetric == sym2
    with pytest.raises(ValueError):
        hs.load_problem(rhs="ones")
    with pytest.raises(ValueError):
        hs.load_problem(matrix="a.mtx", generate_family="poisson2d")


def test_refactorize_refuses_a_new_pattern():
    """DomainDecomposer::updateValues refuses a step whose size, pattern or symmetry changed
    (domain_decomposer.cpp:345-352); refactorize does the same."""
    A, rhs, xyz, sym, bs = hs.generate("poisson2d", nx=12)
    setup = hs.build_setup(A, rhs, 4, xyz=xyz, symmetric=sym, nev=4)
    hs.refactorize(setup, A * 2.0, rhs)  # same pattern: accepted
    E = A.tolil()
    E[0, A.shape[0] - 1] = E[A.shape[0] - 1, 0] = -1e-3
    with pytest.raises(ValueError, match="pattern"):
        hs.refactorize(setup, sp.csr_matrix(E), rhs)
    with pytest.raises(ValueError, match="size"):
        hs.refactorize(setup, hs.generate("poisson2d", nx=13)[0])
    with pytest.raises(ValueError, match="symmetry"):
        hs.refactorize(setup, A, rhs, symmetric=False)


def _random_ldu(n=300, nfaces=900, seed=5):
    rng = np.random.default_rng(seed)
    faces = set()
    while len(faces) < nfaces:
        a, b = (int(x) for x in rng.integers(0, n, 2))
        if a != b:
            faces.add((min(a, b), max(a, b)))
    faces = sorted(faces, key=lambda f: (f[0], f[1]))
    la = np.array([f[0] for f in faces], np.int64)
    ua = np.array([f[1] for f in faces], np.int64)
    return (rng.uniform(1.0, 2.0, n), la, ua, rng.normal(size=nfaces), rng.normal(size=nfaces),
            rng.permutation(n).astype(np.int64), rng.normal(size=n))


def test_ldu_product_is_openfoams_amul():
    """A x of the built matrix equals lduMatrix::Amul's face loop: y_c = D_c x_c, then for every
    face y_l += upper_f x_u and y_u += lower_f x_l (with the cells renumbered by gid)."""
    diag, la, ua, up, lo, gid, src = _random_ldu()
    A, b, sym = hs.from_ldu(diag, la, ua, up, lo, gid, src)
    assert not sym
    x_cells = np.random.default_rng(9).normal(size=diag.size)
    y = diag * x_cells
    np.add.at(y, la, up * x_cells[ua])
    np.add.at(y, ua, lo * x_cells[la])
    x = np.empty_like(x_cells)
    x[gid] = x_cells  # global order
    assert np.allclose((A @ x)[gid], y, rtol=1e-14, atol=1e-14)
    bc = np.empty_like(src)
    bc[gid] = src
    assert np.array_equal(b, bc)
    # the orientation matters: with lower and upper swapped the product differs
    A2, _, _ = hs.from_ldu(diag, la, ua, lo, up, gid, src)
    assert not np.allclose((A2 @ x)[gid], y)


def test_ldu_equals_the_file_route(tmp_path):
    """The OpenFOAM solver's Matrix Market export (bubblyPressureFoam.C:120-170: the diagonal lines,
    then each face's upper and lower lines, 1-based global cells, 17 digits), read back, is
    bitwise the in-memory matrix; exact zeros are dropped on both routes."""
    diag, la, ua, up, lo, gid, src = _random_ldu(seed=11)
    up[::17] = 0.0
    lines = ["%%MatrixMarket matrix coordinate real general", f"{diag.size} {diag.size} {diag.size + 2 * la.size}"]
    lines += [f"{gid[c] + 1} {gid[c] + 1} {diag[c]:.17g}" for c in range(diag.size)]
    for f in range(la.size):
        lines.append(f"{gid[la[f]] + 1} {gid[ua[f]] + 1} {up[f]:.17g}")
        lines.append(f"{gid[ua[f]] + 1} {gid[la[f]] + 1} {lo[f]:.17g}")
    path = _write(tmp_path / "of.mtx", "\n".join(lines) + "\n")
    A, _, _ = hs.from_ldu(diag, la, ua, up, lo, gid, src)
    B, _ = hs.read_matrix(path)
    assert _same_matrix(A, B)
    S, _, sym = hs.from_ldu(diag, la, ua, up)  # symmetric lduMatrix: no lower coefficients
    assert sym and (S != S.T).nnz == 0


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "check_of.mtx")), reason="reference fixture absent")
def test_ldu_rebuilds_the_reference_openfoam_matrix():
    """check_of.mtx (written by the OpenFOAM solver from its lduMatrix) split back into cells and
    faces gives the same matrix through the adapter."""
    A, _ = hs.read_matrix(os.path.join(REF_DIR, "check_of.mtx"))
    U = sp.triu(A, k=1).tocoo()
    Lt = sp.csr_matrix(sp.tril(A, k=-1).T)
    lower = np.asarray(Lt[U.row, U.col]).ravel()
    B, _, _ = hs.from_ldu(A.diagonal(), U.row, U.col, U.data, lower)
    assert _same_matrix(A, B)


def test_ldu_errors():
    d = np.ones(3)
    with pytest.raises(ValueError, match="invalid cells"):
        hs.from_ldu(d, [0], [0], [1.0])
    with pytest.raises(ValueError, match="invalid cells"):
        hs.from_ldu(d, [0], [3], [1.0])
    with pytest.raises(ValueError, match="permutation"):
        hs.from_ldu(d, [0], [1], [1.0], gid=[0, 0, 1])
    with pytest.raises(ValueError, match="sizes"):
        hs.from_ldu(d, [0, 1], [1], [1.0])
