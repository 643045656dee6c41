"""Host setup stand-in: everything the apply path consumes, built without the reference.

The reference builds these objects with its own host code, which does not ship with this
library (its setup stays on the host and is uploaded once): the generated matrix and
right-hand side (src/partition/domain_decomposer.cpp:133-175), the partition and the
renumbering to contiguous ownership (:254-269), the overlapping subdomains
(src/partition/subdomain.cpp:34-302), the local factorizations
(src/linalg/local_solver.cpp:436-451), the local eigenproblems and the coarse basis
(src/preconditioner/local_eigensolver.cpp:215-421, src/preconditioner/coarse_space.cpp:51-75)
and the coarse matrix A_00 = W^T A W (src/preconditioner/coarse_space.cpp:77-206).

The index maps follow the reference's rules exactly (they are the contract of the on-device
gather): owned unknowns [0, n_I) in increasing global order, then the ghosts sorted by global
id; the overlap is the set of vertices at distance <= ``overlap`` from the owned block in
G(A + A^T) (distance 1 is the reference's only setting). The numerical parts (factor,
eigenvectors) are this library's own and are not meant to be bit-identical with Eigen/Spectra.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes
import os
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import _lib
from ._lib import ptr


# --------------------------------------------------------------------------- generators
def generate(name: str, **kw):
    """Generates a matrix family (include/s2l_setup.h). Returns (A csr int64, rhs, xyz, symmetric, block_size)."""
    L = _lib.setup_lib()
    o = _lib.GenOpts()
    L.s2l_gen_defaults(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise ValueError(f"unknown generator option {k}")
        setattr(o, k, v)
    h = L.s2l_generate(name.encode(), ctypes.byref(o))
    if not h:
        raise ValueError(L.s2l_setup_last_error().decode())
    try:
        n = L.s2l_problem_n(h)
        nnz = L.s2l_problem_nnz(h)
        rp = np.empty(n + 1, np.int64)
        col = np.empty(nnz, np.int64)
        val = np.empty(nnz, np.float64)
        rhs = np.empty(n, np.float64)
        xyz = np.empty(3 * n, np.float64)
        L.s2l_problem_copy(h, ptr(rp, ctypes.c_int64), ptr(col, ctypes.c_int64), ptr(val, ctypes.c_double),
                           ptr(rhs, ctypes.c_double), ptr(xyz, ctypes.c_double))
        sym = bool(L.s2l_problem_symmetric(h))
        bs = L.s2l_problem_block_size(h)
    finally:
        L.s2l_problem_free(h)
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    return A, rhs, xyz.reshape(n, 3), sym, bs


def random_rhs(n: int, seed: int = 12345) -> np.ndarray:
    """uniform[-1, 1] with std::mt19937_64(seed) in the original order (domain_decomposer.cpp:158-164)."""
    out = np.empty(n, np.float64)
    _lib.setup_lib().s2l_random_rhs(n, seed, ptr(out, ctypes.c_double))
    return out


def random_rhs_seq(n: int, seed: int, step: int, corrector: int) -> np.ndarray:
    """Right-hand side of one solve of a sequence (src/solver/time_loop.cpp:31-43)."""
    out = np.empty(n, np.float64)
    _lib.setup_lib().s2l_random_rhs_seq(n, seed, step, corrector, ptr(out, ctypes.c_double))
    return out


# --------------------------------------------------------------------------- Matrix Market files
def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def _problem_csr(L, h) -> sp.csr_matrix:
    n = L.s2l_problem_n(h)
    nnz = L.s2l_problem_nnz(h)
    rp = np.empty(n + 1, np.int64)
    col = np.empty(nnz, np.int64)
    val = np.empty(nnz, np.float64)
    L.s2l_problem_copy(h, ptr(rp, ctypes.c_int64), ptr(col, ctypes.c_int64), ptr(val, ctypes.c_double), None, None)
    return sp.csr_matrix((val, col, rp), shape=(n, n))


def read_matrix(path) -> tuple[sp.csr_matrix, bool]:
    """Sparse matrix of a Matrix Market file and its header's symmetric flag
    (MatrixMarketIO::readMatrix, src/partition/matrix_market_io.cpp:92-140): coordinate real,
    integer or pattern; general, symmetric (expanded) or skew-symmetric (expanded, mirror negated);
    explicit zeros dropped, duplicates summed, rows sorted by column. RuntimeError on complex,
    Hermitian or array files, a non-square matrix, indices out of range or a short file."""
    L = _lib.setup_lib()
    h = L.s2l_read_matrix(_path(path))
    if not h:
        raise RuntimeError(L.s2l_setup_last_error().decode())
    try:
        return _problem_csr(L, h), bool(L.s2l_problem_symmetric(h))
    finally:
        L.s2l_problem_free(h)


def read_vector(path) -> np.ndarray:
    """Dense vector of a Matrix Market file: array with one column, or coordinate
    (MatrixMarketIO::readVector, matrix_market_io.cpp:142-172)."""
    L = _lib.setup_lib()
    n = ctypes.c_int64(0)
    if L.s2l_read_vector(_path(path), None, 0, ctypes.byref(n)) != 0:
        raise RuntimeError(L.s2l_setup_last_error().decode())
    out = np.empty(n.value, np.float64)
    if L.s2l_read_vector(_path(path), ptr(out, ctypes.c_double), out.size, ctypes.byref(n)) != 0:
        raise RuntimeError(L.s2l_setup_last_error().decode())
    return out


def write_matrix(path, A) -> None:
    """Writes A as `coordinate real general`, 17 significant digits (matrix_market_io.cpp:174-186)."""
    L = _lib.setup_lib()
    A = sp.csr_matrix(A)
    if A.shape[0] != A.shape[1]:
        raise ValueError("write_matrix: the matrix must be square")
    rp = np.ascontiguousarray(A.indptr, np.int64)
    col = np.ascontiguousarray(A.indices, np.int64)
    val = np.ascontiguousarray(A.data, np.float64)
    if L.s2l_write_matrix(_path(path), A.shape[0], ptr(rp, ctypes.c_int64), ptr(col, ctypes.c_int64),
                          ptr(val, ctypes.c_double)) != 0:
        raise RuntimeError(L.s2l_setup_last_error().decode())


def write_vector(path, v) -> None:
    """Writes v as `array real general` with one column (matrix_market_io.cpp:188-194)."""
    L = _lib.setup_lib()
    v = np.ascontiguousarray(v, np.float64).reshape(-1)
    if L.s2l_write_vector(_path(path), v.size, ptr(v, ctypes.c_double)) != 0:
        raise RuntimeError(L.s2l_setup_last_error().decode())


def step_file_name(pattern: str, step: int) -> str:
    """Name of step ``step`` of a file sequence: one %d / %Nd / %0Nd conversion takes the step, a
    name without '%' is returned as is (domain_decomposer.cpp:80-97). ValueError on other patterns."""
    L = _lib.setup_lib()
    pat = pattern.encode()
    buf = ctypes.create_string_buffer(len(pat) + 32)
    n = L.s2l_step_file_name(pat, step, buf, len(buf))
    if n < 0:
        raise ValueError(L.s2l_setup_last_error().decode())
    return buf.value.decode()


def hash_pattern(A) -> int:
    """Hash of the sparsity pattern (size, row pointers, columns) of a CSR matrix
    (domain_decomposer.cpp:116-130); two steps of a sequence must have the same one."""
    A = sp.csr_matrix(A)
    rp = np.ascontiguousarray(A.indptr, np.int64)
    col = np.ascontiguousarray(A.indices, np.int64)
    return int(_lib.setup_lib().s2l_hash_pattern(A.shape[0], ptr(rp, ctypes.c_int64), ptr(col, ctypes.c_int64)))


def from_ldu(diag, lower_addr, upper_addr, upper, lower=None, gid=None, source=None):
    """Global matrix (and right-hand side) of an OpenFOAM lduMatrix, as the reference's OpenFOAM
    solver writes it to Matrix Market (openfoam/bubblyPressureFoam/bubblyPressureFoam.C:120-170):
    diag[c] at (gid[c], gid[c]); face f gives upper[f] at (gid[l_f], gid[u_f]) and lower[f] at
    (gid[u_f], gid[l_f]); lower None = symmetric (lower = upper); b[gid[c]] = source[c].
    Returns (A csr, b or None, symmetric)."""
    L = _lib.setup_lib()
    d = np.ascontiguousarray(diag, np.float64)
    la = np.ascontiguousarray(lower_addr, np.int64)
    ua = np.ascontiguousarray(upper_addr, np.int64)
    up = np.ascontiguousarray(upper, np.float64)
    lo = None if lower is None else np.ascontiguousarray(lower, np.float64)
    g = None if gid is None else np.ascontiguousarray(gid, np.int64)
    src = None if source is None else np.ascontiguousarray(source, np.float64)
    n, nf = d.size, la.size
    if ua.size != nf or up.size != nf or (lo is not None and lo.size != nf) or (g is not None and g.size != n) \
            or (src is not None and src.size != n):
        raise ValueError("from_ldu: array sizes do not match the cells and faces")
    h = L.s2l_from_ldu(n, nf, ptr(la, ctypes.c_int64), ptr(ua, ctypes.c_int64), ptr(d, ctypes.c_double),
                       ptr(up, ctypes.c_double), ptr(lo, ctypes.c_double), ptr(g, ctypes.c_int64),
                       ptr(src, ctypes.c_double))
    if not h:
        raise ValueError(L.s2l_setup_last_error().decode())
    try:
        A = _problem_csr(L, h)
        b = None
        if L.s2l_problem_rhs_size(h):
            b = np.empty(n, np.float64)
            L.s2l_problem_copy(h, None, None, None, ptr(b, ctypes.c_double), None)
        return A, b, bool(L.s2l_problem_symmetric(h))
    finally:
        L.s2l_problem_free(h)


@dataclass
class Problem:
    """A read or generated global problem (GlobalProblem, domain_decomposer.cpp:66-74)."""

    A: sp.csr_matrix
    b: np.ndarray
    symmetric: bool
    block_size: int = 1
    xyz: np.ndarray | None = None
    description: str = ""


def load_problem(*, matrix: str | None = None, generate_family: str | None = None, rhs: str = "random",
                 seed: int = 12345, step: int = 0, block_size: int = 0, gen_kw: dict | None = None) -> Problem:
    """The matrix of step ``step`` and its right-hand side (loadProblem, domain_decomposer.cpp:133-175):
    ``matrix`` is a Matrix Market file or a step pattern (``p_%04d.mtx``), ``generate_family`` a
    generator family (time = step); rhs is "random" (mt19937_64(seed), uniform[-1, 1] in the
    original order), "ones", "generated" (the generator's) or a Matrix Market vector file. The
    symmetric flag is the header's or the generator's, else detected numerically."""
    if (matrix is None) == (generate_family is None):
        raise ValueError("either matrix or generate_family must be given")
    xyz = None
    gen_rhs = None
    if generate_family is not None:
        A, gen_rhs, xyz, sym, bs = generate(generate_family, **{**(gen_kw or {}), "time": float(step)})
        desc = generate_family
    else:
        name = step_file_name(matrix, step)
        A, sym = read_matrix(name)
        bs = 1
        desc = f"{name} (n={A.shape[0]}, nnz={A.nnz})"
    if block_size > 0:
        bs = block_size
    n = A.shape[0]
    if n % bs != 0:
        raise RuntimeError("n is not a multiple of the block size")
    if not sym:
        sym = is_numerically_symmetric(A)
    if rhs == "random":
        b = random_rhs(n, seed)
    elif rhs == "ones":
        b = np.ones(n)
    elif rhs == "generated":
        if gen_rhs is None:
            raise RuntimeError("--rhs generated requires --generate")
        b = gen_rhs
    else:
        b = read_vector(rhs)
        if b.size != n:
            raise RuntimeError("right-hand side size does not match the matrix")
    return Problem(A, b, sym, bs, xyz, desc)


def is_numerically_symmetric(A: sp.csr_matrix) -> bool:
    """|a_ij - a_ji| <= 1e-14 max|a| on both triangles (domain_decomposer.cpp:49-66)."""
    amax = np.abs(A.data).max() if A.nnz else 0.0
    D = (A - A.T).tocsr()
    return bool(D.nnz == 0 or np.abs(D.data).max() <= 1e-14 * amax)


# --------------------------------------------------------------------------- partition
def partition(A: sp.csr_matrix, nparts: int, xyz=None, block_size: int = 1) -> np.ndarray:
    """Partition vector of G(A + A^T) (stand-in for METIS, graph_partitioner.cpp:17-94)."""
    L = _lib.setup_lib()
    A = A.tocsr()
    rp = np.ascontiguousarray(A.indptr, np.int64)
    col = np.ascontiguousarray(A.indices, np.int64)
    part = np.empty(A.shape[0], np.int32)
    xy = None if xyz is None else np.ascontiguousarray(xyz, np.float64).reshape(-1)
    rc = L.s2l_partition(A.shape[0], ptr(rp, ctypes.c_int64), ptr(col, ctypes.c_int64), nparts,
                         ptr(xy, ctypes.c_double), block_size, ptr(part, ctypes.c_int32))
    if rc != 0:
        raise RuntimeError(L.s2l_setup_last_error().decode())
    return part


@dataclass
class Decomposition:
    """Renumbered global system: part p owns [ownership[p], ownership[p+1]) (domain_decomposer.cpp:254-269)."""

    A: sp.csr_matrix            # renumbered matrix, rows sorted by column
    b: np.ndarray               # renumbered rhs
    ownership: np.ndarray       # int64, nparts + 1
    new_to_old: np.ndarray
    old_to_new: np.ndarray
    part: np.ndarray            # partition vector in the original order
    symmetric: bool
    block_size: int = 1


def decompose(A: sp.csr_matrix, b: np.ndarray, part: np.ndarray, symmetric: bool, block_size: int = 1) -> Decomposition:
    """Contiguous ownership; inside a part the original order is kept (domain_decomposer.cpp:254-269)."""
    n = A.shape[0]
    nparts = int(part.max()) + 1
    count = np.bincount(part, minlength=nparts).astype(np.int64)
    if np.any(count == 0):
        raise RuntimeError("the partition produced an empty subdomain")
    ownership = np.zeros(nparts + 1, np.int64)
    np.cumsum(count, out=ownership[1:])
    # stable sort by part = original order kept inside a part
    new_to_old = np.argsort(part, kind="stable").astype(np.int64)
    old_to_new = np.empty(n, np.int64)
    old_to_new[new_to_old] = np.arange(n, dtype=np.int64)
    Ar = A[new_to_old][:, new_to_old].tocsr()
    Ar.sort_indices()
    return Decomposition(Ar, np.ascontiguousarray(b[new_to_old]), ownership, new_to_old, old_to_new,
                         np.asarray(part), symmetric, block_size)


# --------------------------------------------------------------------------- subdomains
@dataclass
class Subdomain:
    """What subdomain i knows (src/partition/subdomain.cpp): Omega_i = [owned | ghosts]."""

    index: int
    begin: int                  # first owned global row
    n_owned: int
    ghosts: np.ndarray          # sorted global ids (int64)
    Aii: sp.csr_matrix          # A(Omega_i, Omega_i), local numbering
    Aowned: sp.csr_matrix       # A(Omega_I, Omega_i), local numbering
    lump: np.ndarray            # s_i(j) = sum_{k outside Omega_i} |A(j, k)| for the ghost rows
    D: np.ndarray               # partition of unity on Omega_i
    boolean: bool

    @property
    def n_local(self) -> int:
        return self.n_owned + len(self.ghosts)

    def global_ids(self) -> np.ndarray:
        """globalId(l) for every local l (subdomain.cpp:150-153)."""
        return np.concatenate([np.arange(self.begin, self.begin + self.n_owned, dtype=np.int64), self.ghosts])


def _neighbors(A: sp.csr_matrix, AT: sp.csr_matrix, rows: np.ndarray) -> np.ndarray:
    """Columns of A(rows, :) and of A^T(rows, :) (the graph G(A + A^T)), unique."""
    ca = A[rows].indices if len(rows) else np.empty(0, np.int32)
    cb = AT[rows].indices if len(rows) else np.empty(0, np.int32)
    return np.unique(np.concatenate([ca, cb]))


def build_subdomains(dec: Decomposition, overlap: int = 1, pou: str = "boolean") -> list[Subdomain]:
    """Overlapping subdomains from the renumbered matrix (subdomain.cpp:34-302).

    overlap = 1 is the reference's distance-one rule (subdomain.cpp:169-198); overlap > 1 repeats
    it from the previous layer (no reference counterpart)."""
    if pou not in ("boolean", "multiplicity"):
        raise ValueError(f"unknown partition of unity '{pou}'")
    if overlap < 1:
        raise ValueError("overlap must be >= 1")
    A = dec.A
    AT = A.T.tocsr()
    own = dec.ownership
    N = len(own) - 1
    subs = []
    ghost_lists = []
    for i in range(N):
        b, e = int(own[i]), int(own[i + 1])
        inside = np.arange(b, e, dtype=np.int64)
        frontier = inside
        members = inside
        for _ in range(overlap):
            nb = _neighbors(A, AT, frontier)
            new = np.setdiff1d(nb, members, assume_unique=False)
            members = np.union1d(members, new)
            frontier = new
        ghosts = np.setdiff1d(members, inside).astype(np.int64)  # sorted, unique
        ghost_lists.append(ghosts)
    mult = np.ones(A.shape[0]) if pou == "multiplicity" else None
    if mult is not None:
        for g in ghost_lists:
            mult[g] += 1.0
    for i in range(N):
        b, e = int(own[i]), int(own[i + 1])
        ghosts = ghost_lists[i]
        gid = np.concatenate([np.arange(b, e, dtype=np.int64), ghosts])
        nl = len(gid)
        loc = np.full(A.shape[0], -1, np.int64) if nl > 0 else None
        loc[gid] = np.arange(nl)
        rows = A[gid]                       # rows of Omega_i, global columns
        rr = np.repeat(np.arange(nl), np.diff(rows.indptr))
        cc = loc[rows.indices]
        inside = cc >= 0
        Aii = sp.csr_matrix((rows.data[inside], (rr[inside], cc[inside])), shape=(nl, nl))
        Aii.sum_duplicates()
        Aii.sort_indices()
        nI = e - b
        lump = np.zeros(nl - nI)
        outside = ~inside
        np.add.at(lump, rr[outside] - nI, np.abs(rows.data[outside]))
        if np.any(rr[outside] < nI):
            raise RuntimeError("Subdomain: owned-row column outside the overlap")
        Aown = Aii[:nI].tocsr()
        if pou == "boolean":
            D = np.zeros(nl)
            D[:nI] = 1.0
        else:
            D = 1.0 / mult[gid]
        loc[gid] = -1
        subs.append(Subdomain(i, b, nI, ghosts, Aii, Aown, lump, D, pou == "boolean"))
    return subs


# --------------------------------------------------------------------------- local factors
@dataclass
class LocalFactor:
    """A_ii(perm, perm) = L U (Cholesky: U = L^T). L and U^T in CSC, rows sorted, diagonal first."""

    n: int
    kind: int                   # 0 Cholesky, 1 LU (L unit lower)
    perm: np.ndarray            # int32, factor position -> local index
    l_colptr: np.ndarray
    l_row: np.ndarray
    l_val: np.ndarray
    ut_colptr: np.ndarray
    ut_row: np.ndarray
    ut_val: np.ndarray
    sn_begin: np.ndarray
    flops: float

    @property
    def nnz_l(self) -> int:
        return len(self.l_val)

    @property
    def nnz_u(self) -> int:
        return len(self.ut_val)


def factorize(Aii: sp.csr_matrix, kind: int = 0, leaf: int = 0) -> LocalFactor:
    """Nested-dissection supernodal Cholesky (kind 0, LU fallback) or LU without pivoting (kind 1)."""
    L = _lib.setup_lib()
    A = Aii.tocsr()
    A.sort_indices()
    n = A.shape[0]
    rp = np.ascontiguousarray(A.indptr, np.int32)
    col = np.ascontiguousarray(A.indices, np.int32)
    val = np.ascontiguousarray(A.data, np.float64)
    h = L.s2l_factorize(n, ptr(rp, ctypes.c_int32), ptr(col, ctypes.c_int32), ptr(val, ctypes.c_double), kind, leaf)
    if not h:
        raise RuntimeError(L.s2l_setup_last_error().decode())
    try:
        nn = ctypes.c_int32()
        nl = ctypes.c_int64()
        nu = ctypes.c_int64()
        ns = ctypes.c_int32()
        kd = ctypes.c_int()
        fl = ctypes.c_double()
        L.s2l_factor_info(h, ctypes.byref(nn), ctypes.byref(nl), ctypes.byref(nu), ctypes.byref(ns),
                          ctypes.byref(kd), ctypes.byref(fl))
        perm = np.empty(n, np.int32)
        lcp = np.empty(n + 1, np.int64)
        lr = np.empty(nl.value, np.int32)
        lv = np.empty(nl.value, np.float64)
        ucp = np.empty(n + 1, np.int64)
        ur = np.empty(nu.value, np.int32)
        uv = np.empty(nu.value, np.float64)
        L.s2l_factor_copy(h, ptr(perm, ctypes.c_int32), ptr(lcp, ctypes.c_int64), ptr(lr, ctypes.c_int32),
                          ptr(lv, ctypes.c_double), ptr(ucp, ctypes.c_int64), ptr(ur, ctypes.c_int32),
                          ptr(uv, ctypes.c_double))
        sn = np.empty(ns.value + 1, np.int32)
        L.s2l_factor_supernodes(h, ptr(sn, ctypes.c_int32))
    finally:
        L.s2l_factor_free(h)
    return LocalFactor(n, kd.value, perm, lcp, lr, lv, ucp, ur, uv, sn, fl.value)


def factorize_all(subs: list[Subdomain], kind: int = 0, threads: int | None = None, leaf: int = 0) -> list[LocalFactor]:
    """Factorizes every A_ii (ctypes releases the GIL: one thread per subdomain)."""
    threads = threads or min(len(subs), os.cpu_count() or 1, 16)
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda s: factorize(s.Aii, kind, leaf), subs))


# --------------------------------------------------------------------------- coarse space
def lumped_matrix(sd: Subdomain) -> sp.csr_matrix:
    """A~_ii = A_ii - diag(0, s_i) on the ghost rows (local_block_splitting.cpp:9-21)."""
    At = sd.Aii.tolil(copy=True)
    for j, s in enumerate(sd.lump):
        l = sd.n_owned + j
        At[l, l] = At[l, l] - s
    return At.tocsr()


def _norm_estimate(M: sp.csr_matrix) -> float:
    """sqrt(||M||_1 ||M||_inf) (local_eigensolver.cpp:33-43)."""
    if M.shape[0] == 0:
        return 0.0
    a = abs(M)
    return float(np.sqrt(a.sum(axis=0).max() * a.sum(axis=1).max()))


def _select(lam, U, nev, tau):
    """Eigenvectors with |lambda| >= 1/tau among the first nev by decreasing modulus; complex pairs
    contribute Re and Im (selectEigenvectors, local_eigensolver.cpp:127-160)."""
    order = np.argsort(-np.abs(lam), kind="stable")
    lam = lam[order][:nev]
    U = U[:, order][:, :nev]
    thr = 1.0 / tau
    out = []
    for j, l in enumerate(lam):
        if abs(l) < thr:
            continue
        if abs(np.imag(l)) <= 1e-10 * abs(l):
            out.append(np.real(U[:, j]))
            continue
        if np.imag(l) < 0 and any(abs(lam[q] - np.conj(l)) <= 1e-8 * abs(l) for q in range(len(lam)) if q != j):
            continue
        out.append(np.real(U[:, j]))
        out.append(np.imag(U[:, j]))
    return out


def local_eigenvectors(sd: Subdomain, nev: int, tau: float = 0.6, symmetric: bool = True,
                       shift_rel: float = 1e-10, dense_threshold: int = 400, tol: float = 1e-8) -> np.ndarray:
    """Z_i: eigenvectors of (A~_ii + sigma I)^{-1} B_i, B_i = D_i A_ii D_i, with |lambda| >= 1/tau.

    Follows the pencil of LocalEigensolver::compute (local_eigensolver.cpp:215-226); the kernel of
    A~_ii is handled by the shift (its directions with B q != 0 get eigenvalues ~ 1/sigma and are
    selected first), instead of the reference's explicit kernel detection."""
    n = sd.n_local
    if nev <= 0:
        return np.zeros((n, 0))
    Dd = sp.diags(sd.D)
    B = (Dd @ sd.Aii @ Dd).tocsr()
    At = lumped_matrix(sd)
    normA = _norm_estimate(At) or 1.0
    sigma = shift_rel * normA
    ncv = max(2 * nev + 1, nev + 20)
    for attempt in range(4):
        As = (At + sigma * sp.identity(n, format="csr")).tocsc()
        try:
            if n <= dense_threshold or ncv >= n:
                if symmetric:
                    lam, U = sla.eigh(B.toarray(), As.toarray())
                else:
                    lam, U = sla.eig(np.linalg.solve(As.toarray(), B.toarray()))
            else:
                lu = spla.splu(As)
                if symmetric:
                    Minv = spla.LinearOperator((n, n), matvec=lu.solve, dtype=np.float64)
                    # a fixed start vector per subdomain makes the setup reproducible (ARPACK draws
                    # a random one otherwise); pseudo-random, so no symmetry of the subdomain hides
                    # an eigenvector from it as a constant vector could
                    v0 = np.random.default_rng(0x5EED + sd.index).uniform(-1.0, 1.0, n)
                    lam, U = spla.eigsh(B, k=nev, M=As, Minv=Minv, which="LA", tol=tol, ncv=min(ncv, n), v0=v0)
                else:
                    T = spla.LinearOperator((n, n), matvec=lambda x: lu.solve(B @ x), dtype=np.float64)
                    v0 = np.ones(n)
                    lam, U = spla.eigs(T, k=nev, which="LM", tol=tol, ncv=min(ncv, n), v0=v0)
            break
        except (np.linalg.LinAlgError, RuntimeError, spla.ArpackError):
            sigma *= 1e3
    else:
        raise RuntimeError(f"local eigensolver failed on subdomain {sd.index}")
    vecs = _select(np.asarray(lam), np.asarray(U), nev, tau)
    if not vecs:
        return np.zeros((n, 0))
    Z = np.stack([v / (np.linalg.norm(v) or 1.0) for v in vecs], axis=1)
    return Z


def coarse_basis(sd: Subdomain, Z: np.ndarray, orth_tol: float = 1e-10) -> np.ndarray:
    """V_i = orth(D_i Z_i) by column-pivoted QR with a relative rank threshold; exact zeros on
    the ghost rows with the Boolean PoU (coarse_space.cpp:55-65)."""
    n = sd.n_local
    if Z.shape[1] == 0:
        return np.zeros((n, 0))
    DZ = sd.D[:, None] * Z
    Q, R, _ = sla.qr(DZ, mode="economic", pivoting=True)
    d = np.abs(np.diag(R))
    r = int(np.sum(d > orth_tol * d[0])) if d.size and d[0] > 0 else 0
    V = np.ascontiguousarray(Q[:, :r])
    if sd.boolean:
        V[sd.n_owned:, :] = 0.0
    return V


def _eig_task(args):
    sd, nev, tau, sym, tol = args
    return local_eigenvectors(sd, nev, tau, sym, tol=tol)


@dataclass
class CoarseSpace:
    """V_i blocks, coarse offsets and A_00 = W^T A W (dense, FP64)."""

    V: list                      # n_local x m_i per subdomain
    offsets: np.ndarray          # int64, N + 1
    E: np.ndarray                # n0 x n0

    @property
    def n0(self) -> int:
        return int(self.offsets[-1])


def assemble_W(dec: Decomposition, subs: list[Subdomain], V: list) -> sp.csr_matrix:
    """W = R_0^T (n x n0): V_i placed at the rows Omega_i and columns [off_i, off_i + m_i)."""
    rows, cols, vals = [], [], []
    off = 0
    for sd, Vi in zip(subs, V):
        m = Vi.shape[1]
        gid = sd.global_ids()
        rr, cc = np.nonzero(Vi)
        rows.append(gid[rr])
        cols.append(off + cc)
        vals.append(Vi[rr, cc])
        off += m
    n = dec.A.shape[0]
    if off == 0:
        return sp.csr_matrix((n, 0))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, off))


def build_coarse(dec: Decomposition, subs: list[Subdomain], nev: int, tau: float = 0.6, workers: int = 0,
                 tol: float = 1e-8, orth_tol: float = 1e-10, Z: list | None = None) -> CoarseSpace:
    """Local eigenproblems -> V_i -> A_00 (coarse_space.cpp:51-206)."""
    sym = dec.symmetric
    if Z is None:
        tasks = [(sd, nev, tau, sym, tol) for sd in subs]
        if workers and workers > 1 and len(subs) > 1 and _lib.fork_is_safe():
            import multiprocessing as mp
            ctx = mp.get_context("fork")
            with ctx.Pool(min(workers, len(subs))) as pool:
                Z = pool.map(_eig_task, tasks)
        else:
            Z = [_eig_task(t) for t in tasks]
    V = [coarse_basis(sd, z, orth_tol) for sd, z in zip(subs, Z)]
    offsets = np.zeros(len(subs) + 1, np.int64)
    np.cumsum([v.shape[1] for v in V], out=offsets[1:])
    W = assemble_W(dec, subs, V)
    E = (W.T @ (dec.A @ W)).toarray() if W.shape[1] else np.zeros((0, 0))
    return CoarseSpace(V, offsets, E)


# --------------------------------------------------------------------------- whole setup
@dataclass
class SchwarzSetup:
    """Everything uploaded to the device once per setup."""

    dec: Decomposition
    subs: list
    factors: list
    coarse: CoarseSpace
    one_level: str = "ras"
    correction: str = "deflated"
    times: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.dec.A.shape[0]


def device_eigenvectors(dec: Decomposition, subs: list, factors: list, nev: int, tau: float = 0.6, tol: float = 1e-8,
                        shift_rel: float = 1e-10, dense_threshold: int = 400, max_restarts: int = 100,
                        device_index: int = 0) -> tuple[list, dict]:
    """Z_i of every subdomain from the device eigensolver (s2l_local_eigensolve: the operator
    (A~_ii + sigma_i I)^{-1} B_i of ProjectedOperator, local_eigensolver.cpp:174-211, applied by
    batched device solves inside one thick-restart Lanczos process per subdomain), selected as
    local_eigenvectors does (|lambda| >= 1/tau among the nev largest, unit 2-norm). Subdomains of at
    most ``dense_threshold`` rows take the dense host path, as in the reference
    (local_eigensolver.cpp:223-225); a subdomain whose Lanczos process did not converge within
    ``max_restarts`` is solved again on the host. Needs symmetric A (Cholesky local factors)."""
    from . import device
    if not dec.symmetric or any(f.kind != 0 for f in factors):
        raise ValueError("the device eigensolver needs a symmetric matrix (Cholesky local factors)")
    n = len(subs)
    if nev <= 0:
        return [np.zeros((sd.n_local, 0)) for sd in subs], {}
    ncv = max(2 * nev + 1, nev + 20)
    small = np.array([sd.n_local <= dense_threshold or ncv >= sd.n_local for sd in subs], np.int32)
    stats = {"device_subdomains": int(n - small.sum())}
    Zs = [None] * n
    if small.sum() < n:
        t0 = time.perf_counter()
        ctx = device.HipContext(device_index)
        try:
            ctx.upload_setup(SchwarzSetup(dec, subs, factors, None), "double", coarse=False)
            t1 = time.perf_counter()
            lam, Zall, conv, info = ctx.local_eigensolve(nev, ncv, shift_rel, tol, max_restarts, skip=small)
            stats.update(upload_s=t1 - t0, eigensolve_s=time.perf_counter() - t1, restarts=int(info[0]),
                         lanczos_steps=int(info[1]), shift_attempts=int(info[2]), converged=int(info[3]))
            lptr = ctx._keep["local_ptr"]
        finally:
            ctx.close()
        for i, sd in enumerate(subs):
            if small[i] or not conv[i]:
                continue
            U = Zall[:, lptr[i]:lptr[i + 1]].T
            vecs = _select(lam[i], U, nev, tau)
            Zs[i] = (np.stack([v / (np.linalg.norm(v) or 1.0) for v in vecs], axis=1) if vecs
                     else np.zeros((sd.n_local, 0)))
    stats["host_subdomains"] = sum(z is None for z in Zs)
    for i, sd in enumerate(subs):
        if Zs[i] is None:
            Zs[i] = local_eigenvectors(sd, nev, tau, True, tol=tol)
    return Zs, stats


def build_setup(A, b, nparts, *, xyz=None, symmetric=None, block_size=1, overlap=1, pou="boolean",
                nev=20, tau=0.6, factorization="auto", one_level="ras", correction="deflated",
                eig_workers=0, eig_tol=1e-8, part=None, leaf=0, eig_backend: str = "host") -> SchwarzSetup:
    """Full host setup of the two-level preconditioner on ``nparts`` subdomains. ``eig_backend``
    "device" solves the local eigenproblems on the GPU (symmetric matrices; device_eigenvectors)."""
    t = {}
    t0 = time.perf_counter()
    A = sp.csr_matrix(A)
    A.sort_indices()
    if symmetric is None:
        symmetric = is_numerically_symmetric(A)
    if part is None:
        part = partition(A, nparts, xyz, block_size) if nparts > 1 else np.zeros(A.shape[0], np.int32)
    dec = decompose(A, b, part, symmetric, block_size)
    t["partition"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    subs = build_subdomains(dec, overlap, pou)
    t["subdomains"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    kind = 1 if (factorization == "lu" or not symmetric) else 0
    factors = factorize_all(subs, kind, leaf=leaf)
    t["factorization"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    if correction == "none":
        coarse = CoarseSpace([np.zeros((s.n_local, 0)) for s in subs], np.zeros(len(subs) + 1, np.int64), np.zeros((0, 0)))
    elif eig_backend == "device":
        Z, st = device_eigenvectors(dec, subs, factors, nev, tau, eig_tol)
        t["eigensolver_device"] = st
        coarse = build_coarse(dec, subs, nev, tau, eig_workers, eig_tol, Z=Z)
    elif eig_backend == "host":
        coarse = build_coarse(dec, subs, nev, tau, eig_workers, eig_tol)
    else:
        raise ValueError(f"eig_backend must be 'host' or 'device', not {eig_backend!r}")
    t["coarse"] = time.perf_counter() - t0
    return SchwarzSetup(dec, subs, factors, coarse, one_level, correction if correction != "none" else "additive", t)


def refactorize(setup: SchwarzSetup, A_new, b_new=None, *, overlap: int = 1, pou: str = "boolean",
                leaf: int = 0, symmetric: bool | None = None) -> SchwarzSetup:
    """Refactorize level for new values of A with the same pattern (two_level_schwarz.cpp:35-41,
    coarse_space.cpp:95-105,186-206, subdomain.cpp:107-148): the partition, the subdomains and the
    coarse basis V are kept; the local matrices, their factors and A_00 = W^T A W are recomputed.
    overlap and pou must be those of the setup. ValueError, as the reference's updateValues, when
    the size, the sparsity pattern or (if ``symmetric`` is given) the symmetry changed, or when
    the overlap sets changed: all need a rebuild."""
    t = {}
    t0 = time.perf_counter()
    dec0 = setup.dec
    A = sp.csr_matrix(A_new)
    A.sort_indices()
    # the checks of DomainDecomposer::updateValues (domain_decomposer.cpp:345-352)
    if A.shape != dec0.A.shape:
        raise ValueError("refactorize: the size changed; rebuild instead")
    if symmetric is not None and bool(symmetric) != dec0.symmetric:
        raise ValueError("refactorize: the symmetry of the matrix changed; rebuild instead")
    b = np.asarray(b_new, np.float64) if b_new is not None else dec0.b[dec0.old_to_new]
    dec = decompose(A, b, dec0.part, dec0.symmetric, dec0.block_size)
    if hash_pattern(dec.A) != hash_pattern(dec0.A):
        raise ValueError("refactorize: the sparsity pattern changed; rebuild instead")
    subs = build_subdomains(dec, overlap, pou)
    for s0, s1 in zip(setup.subs, subs):
        if s0.n_owned != s1.n_owned or not np.array_equal(s0.ghosts, s1.ghosts):
            raise ValueError(f"subdomain {s0.index}: the overlap changed with the new matrix; rebuild instead")
    t["subdomains"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    kind = 0 if setup.factors[0].ut_colptr is None else 1
    factors = factorize_all(subs, kind, leaf=leaf)
    t["factorization"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    V = setup.coarse.V
    W = assemble_W(dec, subs, V)
    E = (W.T @ (dec.A @ W)).toarray() if W.shape[1] else np.zeros((0, 0))
    t["coarse"] = time.perf_counter() - t0
    return SchwarzSetup(dec, subs, factors, CoarseSpace(V, setup.coarse.offsets.copy(), E), setup.one_level,
                        setup.correction, t)

