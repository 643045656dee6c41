"""GPU tests of the device local eigensolver (s2l_local_eigensolve, the operator of the reference's
ProjectedOperator, local_eigensolver.cpp:174-211, inside a thick-restart Lanczos process per
subdomain) through the C-ABI.

The eigenpairs are pinned three ways: by their own residuals ||B z - lambda (A~ + sigma I) z||
computed here on the host from the matrices of hostsetup, by ARPACK (scipy eigsh in shift-invert
form, tol 1e-12) on the same pencils, and end to end by the coarse space they give: the same coarse
dimension and the same GMRES iterations as the host setup."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from amp_amd import hostsetup as hs

from cases import items, problem

pytestmark = pytest.mark.gpu

CASES = [("poisson3d", dict(nx=24), 4, 1, 8),
         ("bubbles3d", dict(nx=24, nbubbles=4, radius=0.15), 4, 1, 8),
         ("poisson3d", dict(nx=20), 8, 2, 12)]


def _pieces(name, kw, N, overlap):
    print(f"case {name} {kw} N={N} overlap={overlap}", flush=True)
    A, b, xyz, sym, bs = problem(name, items(kw), N)
    part = hs.partition(A, N, xyz)
    dec = hs.decompose(A, b, part, sym)
    subs = hs.build_subdomains(dec, overlap)
    factors = hs.factorize_all(subs, 0)
    return dec, subs, factors


def _pencil(sd, shift_rel=1e-10):
    Dd = sp.diags(sd.D)
    B = (Dd @ sd.Aii @ Dd).tocsr()
    At = hs.lumped_matrix(sd)
    sigma = shift_rel * (hs._norm_estimate(At) or 1.0)
    As = (At + sigma * sp.identity(sd.n_local, format="csr")).tocsc()
    return B, As


def _context(dec, subs, factors):
    from amp_amd import device
    ctx = device.HipContext()
    ctx.upload_setup(hs.SchwarzSetup(dec, subs, factors, None), "double", coarse=False)
    return ctx


@pytest.mark.parametrize("name,kw,N,overlap,nev", CASES)
def test_eigenpairs_residuals_and_arpack(name, kw, N, overlap, nev):
    dec, subs, factors = _pieces(name, kw, N, overlap)
    ctx = _context(dec, subs, factors)
    ncv = max(2 * nev + 1, nev + 20)
    lam, Z, conv, info = ctx.local_eigensolve(nev, ncv, 1e-10, 1e-8, 100)
    lptr = ctx._keep["local_ptr"]
    print(f"restarts {info[0]} steps {info[1]} attempts {info[2]} converged {info[3]}", flush=True)
    assert conv.all() and info[3] == N and info[2] == 1
    for i, sd in enumerate(subs):
        B, As = _pencil(sd)
        U = Z[:, lptr[i]:lptr[i + 1]].T
        assert np.all(np.diff(lam[i]) <= 0)  # decreasing
        for c in range(nev):
            z = U[:, c]
            Az = As @ z
            res = np.linalg.norm(B @ z - lam[i, c] * Az) / (abs(lam[i, c]) * np.linalg.norm(Az))
            assert res <= 1e-6, (i, c, res)
            assert abs(z @ Az - 1.0) < 1e-8  # unit length in the A~ + sigma I norm
        G = U.T @ (As @ U)
        assert np.abs(G - np.eye(nev)).max() < 1e-8  # orthonormal in that inner product
        lu = spla.splu(As)
        Minv = spla.LinearOperator(As.shape, matvec=lu.solve, dtype=np.float64)
        v0 = np.random.default_rng(7 + i).uniform(-1, 1, sd.n_local)
        la, _ = spla.eigsh(B, k=nev, M=As, Minv=Minv, which="LA", tol=1e-12, ncv=ncv, v0=v0)
        la = np.sort(la)[::-1]
        assert np.allclose(lam[i], la, rtol=1e-7, atol=0), (i, lam[i], la)


def test_skipped_subdomains_untouched():
    dec, subs, factors = _pieces(*CASES[0][:4])
    ctx = _context(dec, subs, factors)
    skip = np.array([1, 0, 1, 0], np.int32)
    lam, Z, conv, info = ctx.local_eigensolve(8, 28, 1e-10, 1e-8, 100, skip=skip)
    lptr = ctx._keep["local_ptr"]
    lam_all, Z_all, conv_all, _ = ctx.local_eigensolve(8, 28, 1e-10, 1e-8, 100)
    assert list(conv) == [0, 1, 0, 1] and info[3] == 2
    for i in range(4):
        sl = slice(lptr[i], lptr[i + 1])
        if skip[i]:
            assert not Z[:, sl].any() and not lam[i].any()
        else:
            # one subdomain's process does not depend on the others being solved
            assert np.array_equal(Z[:, sl], Z_all[:, sl]) and np.array_equal(lam[i], lam_all[i])


def test_refused_configurations():
    dec, subs, factors = _pieces(*CASES[0][:4])
    ctx = _context(dec, subs, factors)
    with pytest.raises(RuntimeError, match="ncv"):
        ctx.local_eigensolve(8, 9, 1e-10, 1e-8, 10)
    with pytest.raises(RuntimeError, match="ncv"):
        ctx.local_eigensolve(8, 200, 1e-10, 1e-8, 10)
    from amp_amd import device
    ctx32 = device.HipContext()
    ctx32.upload_setup(hs.SchwarzSetup(dec, subs, factors, None), "single", coarse=False)
    with pytest.raises(RuntimeError, match="FP64"):
        ctx32.local_eigensolve(8, 28, 1e-10, 1e-8, 10)
    A, b, xyz, sym, bs = problem("convdiff2d", items(dict(nx=40)), 4)
    dn = hs.decompose(A, b, hs.partition(A, 4, xyz), False)
    sn = hs.build_subdomains(dn, 1)
    with pytest.raises(ValueError, match="symmetric"):
        hs.device_eigenvectors(dn, sn, hs.factorize_all(sn, 1), 8)


@pytest.mark.parametrize("name,kw,N,overlap,nev", CASES[:2])
def test_device_coarse_space_gives_host_iterations(name, kw, N, overlap, nev):
    """The whole setup with the device eigensolver: the same coarse dimension, the same GMRES(50)
    iterations to 1e-8 and the same subspaces V_i as with ARPACK on the host."""
    from amp_amd import device
    A, b, xyz, sym, bs = problem(name, items(kw), N)
    sth = hs.build_setup(A, b, N, xyz=xyz, symmetric=sym, overlap=overlap, nev=nev, eig_tol=1e-8)
    std = hs.build_setup(A, b, N, xyz=xyz, symmetric=sym, overlap=overlap, nev=nev, eig_tol=1e-8, eig_backend="device")
    ed = std.times["eigensolver_device"]
    print(f"device eigensolver {ed}", flush=True)
    assert ed["device_subdomains"] == N and ed["host_subdomains"] == 0
    assert sth.coarse.n0 == std.coarse.n0
    for Vh, Vd in zip(sth.coarse.V, std.coarse.V):
        if Vh.shape[1] == 0:
            continue
        s = np.linalg.svd(Vh.T @ Vd, compute_uv=False)  # cosines of the principal angles
        assert s.min() > 1 - 1e-6
    res = []
    for st in (sth, std):
        M = device.HipTwoLevelSchwarz(st, "double")
        x, r = M.ctx.gmres(st.dec.b, None, 50, 1e-8, 1000)
        res.append(r)
    assert res[0].converged and res[1].converged
    assert abs(res[0].iterations - res[1].iterations) <= 1
