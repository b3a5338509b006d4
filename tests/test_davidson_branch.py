"""Why davidson_clean runs its dense branch at every size: the reference's
iterative branch (src/davidson_clean.F90:120-323, taken for nbasis >= 200
and never called by the reference's driver) restated step by step in numpy
breaks down on a small noisy ERI matrix.

The restatement keeps the reference's block size m = nbasis, its NMAX =
nbasis iterations per restart, the un-orthonormalized basis U (random
normalized start block, then normalized residuals E u - W u), the
generalized eigenproblem H v = E S v with S = U^T U, H = U^T W U read from
their upper triangles (dsygv 'U'), the convergence test
(||r_{m/2+1}|| < 1e-8 or E_m < 0), the re-contraction after NMAX - 1
iterations, and the pass loop that subtracts the negative Ritz pairs from
the canonical list entries only, moy(kcp) -= sum_k E_k U(ii,k) U(jj,k),
before the eight permutations rebuild W for the next pass.

On the 7-AO water matrix with 2e-3 noise (13 negative eigenvalues):
* for most start blocks S stops being positive definite (the reference
  stops with 'DSYGV Diagonalization failed');
* a run that finishes needs many passes, removes far more Ritz pairs than W
  has negative eigenvalues (the canonical update followed by the eight
  permutations is not a projection of W, so every pass creates new negative
  directions) and ends at a start-block-dependent list that differs from
  the dense branch by more than 1e-4.
The dense branch (rocSOLVER syevd on the GPU) is exact and deterministic;
it is what zv_davidson_clean computes.
"""
import numpy as np
import scipy.linalg as sl

from test_gpu_cleanup import gather, negative_removed, noisy_water, product_matrix


def canonical_rows(n, q):
    ijkl = np.stack(q, axis=1) - 1
    return ijkl[:, 0] + ijkl[:, 1] * n, ijkl[:, 2] + ijkl[:, 3] * n


def reference_iterative_branch(n, q, moy, seed, max_pass=200):
    """src/davidson_clean.F90:125-321 on the list moy; returns
    (moy, passes, removed) or raises scipy.linalg.LinAlgError where dsygv
    would fail.  Random start blocks come from numpy (the reference's
    random_number stream is compiler-defined)."""
    N = n * n
    mmax = nmax = n
    rng = np.random.default_rng(seed)
    ii, jj = canonical_rows(n, q)
    moy = moy.copy()
    m, passes, removed = mmax, 0, 0
    while m > 0 and passes < max_pass:
        m = mmax
        W = product_matrix(n, q, moy)
        U = np.zeros((N, nmax * m))
        Wb = np.zeros((N, nmax * m))
        U[:, :m] = rng.standard_normal((N, m))
        U[:, :m] /= np.linalg.norm(U[:, :m], axis=0)
        converged = False
        while not converged:
            for it in range(1, nmax):
                sh = m * (it - 1)
                K = sh + m
                Wb[:, sh:K] = W @ U[:, sh:K]
                S = U[:, :K].T @ U[:, :K]
                H = U[:, :K].T @ Wb[:, :K]
                up = np.triu(np.ones((K, K), bool))
                Hs = np.where(up, H, H.T)
                Ss = np.where(up, S, S.T)
                _, V = sl.eigh(Hs, Ss)
                E = np.diag(V.T @ (H @ V))[:m]
                U[:, K:K + m] = U[:, :K] @ V[:, :m]
                Wb[:, K:K + m] = Wb[:, :K] @ V[:, :m]
                R = E * U[:, K:K + m] - Wb[:, K:K + m]
                rn = np.linalg.norm(R, axis=0)
                U[:, K:K + m] = R / rn
                converged = rn[m // 2] < 1e-8 or E[m - 1] < 0
                if converged:
                    break
            U[:, :m] = U[:, :K] @ V[:, :m]
        passes += 1
        kk = 0
        while kk < m and E[kk] < 0:
            kk += 1
        m = kk
        moy -= (U[ii, :m] * E[:m] * U[jj, :m]).sum(axis=1)
        removed += m
    return moy, passes, removed


def test_reference_iterative_branch_breaks_down_on_noisy_water():
    bas, q, vals = noisy_water()
    n = bas.n
    Wc, nneg = negative_removed(product_matrix(n, q, vals))
    dense = gather(n, q, Wc, False)
    assert nneg == 13
    failed, finished = 0, []
    for seed in range(8):
        try:
            finished.append(reference_iterative_branch(n, q, vals, seed))
        except sl.LinAlgError:
            failed += 1
    assert failed >= 5, failed
    for moy, passes, removed in finished:
        assert passes > 4 and removed > 2 * nneg
        assert np.abs(moy - dense).max() > 1e-4
        # the pass loop's own fixed point: the rebuilt W has no negative
        # eigenvalue beyond the 1e-8 residual tolerance
        assert np.linalg.eigvalsh(product_matrix(n, q, moy)).min() > -1e-6
