"""Shared problem builders and independent dense references for the tests (CPU only)."""
import functools
import json
import os
import pickle
import subprocess
import sys
import tempfile

import numpy as np

from amp_amd import hostsetup as hs


@functools.lru_cache(maxsize=32)
def problem(name, kw_items, N, seed=12345, rhs="random", **_):
    """(A, b, xyz, symmetric, block_size) of a generated problem; rhs 'random' (seeded uniform[-1, 1],
    domain_decomposer.cpp:158-164) or 'generated' (the generator's own right-hand side)."""
    A, g, xyz, sym, bs = hs.generate(name, **dict(kw_items))
    b = hs.random_rhs(A.shape[0], seed) if rhs == "random" else g
    return A, b, xyz, sym, bs


@functools.lru_cache(maxsize=64)
def setup(name, kw_items, N, opts_items=(), rhs="random"):
    """Host setup of a generated problem on N subdomains (cached across tests)."""
    A, b, xyz, sym, bs = problem(name, kw_items, N, rhs=rhs)
    return hs.build_setup(A, b, N, xyz=xyz, symmetric=sym, block_size=bs, **dict(opts_items))


@functools.lru_cache(maxsize=4)
def setup_large(name, kw_items, N, opts_items=(), rhs="random", workers=16):
    """Host setup of a large problem built in a child process (tests/setup_child.py), whose
    eigensolver pool may fork: this process may have initialised the HIP runtime. Cached."""
    args = dict(name=name, kw=dict(kw_items), N=N, opts=dict(opts_items), rhs=rhs, workers=workers)
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "setup.pkl")
        here = os.path.dirname(os.path.abspath(__file__))
        subprocess.run([sys.executable, os.path.join(here, "setup_child.py"), out, json.dumps(args)], check=True,
                       timeout=600)
        with open(out, "rb") as f:   # written by setup_child.py just now
            return pickle.load(f)


def items(d):
    return tuple(sorted(d.items()))


def dense_apply_reference(st, r, one_level, correction):
    """z = M^{-1} r assembled densely from the global matrix, the index sets Omega_i, the partitions
    of unity D_i and the local bases V_i, with dense LU solves of every A_ii and of
    A_00 = R_0 A R_0^T (the sequential reference of test/test_preconditioner.cpp:104-172)."""
    A = st.dec.A.toarray()
    n = A.shape[0]

    def one(v):
        out = np.zeros(n)
        for sd in st.subs:
            gid = sd.global_ids()
            yi = np.linalg.solve(A[np.ix_(gid, gid)], v[gid])
            d = sd.D if one_level == "ras" else np.ones(len(gid))
            np.add.at(out, gid, d * yi)
        return out

    n0 = st.coarse.n0
    if n0 == 0:
        return one(r)
    R0T = np.zeros((n, n0))
    off = 0
    for sd, V in zip(st.subs, st.coarse.V):
        gid = sd.global_ids()
        m = V.shape[1]
        R0T[gid, off:off + m] += V
        off += m
    E = R0T.T @ A @ R0T

    def Q(v):
        return R0T @ np.linalg.solve(E, R0T.T @ v)

    if correction == "additive":
        return Q(r) + one(r)
    x0 = Q(r)
    return x0 + one(r - A @ x0)


def brute_force_ghosts(A, begin, end, dist):
    """Unknowns at graph distance 1..dist from [begin, end) in G(A + A^T), by breadth-first search
    over explicit adjacency sets (test/test_subdomain.cpp:44-53 for dist = 1)."""
    C = A.tocoo()
    adj = [set() for _ in range(A.shape[0])]
    for i, j in zip(C.row.tolist(), C.col.tolist()):
        if i != j:
            adj[i].add(j)
            adj[j].add(i)
    seen = set(range(begin, end))
    frontier = set(seen)
    for _ in range(dist):
        nxt = set()
        for u in frontier:
            nxt |= adj[u]
        nxt -= seen
        seen |= nxt
        frontier = nxt
    return sorted(seen - set(range(begin, end)))
