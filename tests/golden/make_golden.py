"""Writes tests/golden/bubbles3d_24_check.npz from the reference's committed generator fixture
test/gamg/run_gamg_23038420/check_s2l.mtx and check_s2l_rhs.mtx (bubbles3d 24^3, 150 bubbles,
R = 0.04, rho ratio 1000, cfl 0.5, t = 3, bubble seed 2024; log.check_matgen:1 in that folder).

Only a digest of the data is kept (the full matrix is 3 MB of text):
  * sha256 of the CSR pattern (row pointers and sorted column indices as int64),
  * the row sums weighted by w_j = 1 + (j mod 7) / 7 (every entry of a row enters with its own
    weight, so a wrong coefficient changes its row's sum),
  * every 16th stored value (CSR order), every right-hand-side entry whose index is a multiple of 4,
  * the diagonal.
Run in a container that holds the reference: python tests/golden/make_golden.py
"""
import hashlib
import os

import numpy as np
import scipy.io as sio

SRC = "/root/reference/test/gamg/run_gamg_23038420"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bubbles3d_24_check.npz")


def digest(A, rhs):
    A = A.tocsr()
    A.sort_indices()
    h = hashlib.sha256()
    h.update(A.indptr.astype("<i8").tobytes())
    h.update(A.indices.astype("<i8").tobytes())
    w = 1.0 + (np.arange(A.shape[1]) % 7) / 7.0
    return dict(n=np.int64(A.shape[0]), nnz=np.int64(A.nnz), pattern_sha256=np.bytes_(h.hexdigest()),
                wsum=A @ w, data16=A.data[::16].copy(), rhs4=rhs[::4].copy(), diag=A.diagonal())


def main():
    A = sio.mmread(os.path.join(SRC, "check_s2l.mtx"))
    rhs = np.asarray(sio.mmread(os.path.join(SRC, "check_s2l_rhs.mtx"))).ravel()
    np.savez_compressed(OUT, **digest(A, rhs))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
