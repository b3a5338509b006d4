"""GPU cleanups of the pair-packed ERI supermatrix V[P][Q] = (ik|jl)
(zv_rsvd_clean, zv_davidson_lowest), pinned by the mathematics: numpy's
dense symmetric eigendecomposition of the same V on small N.

* block Davidson, plain and with the shift-invert preconditioner: the nev
  lowest eigenvalues of a noisy 30-AO ERI matrix (N = 465; the noise leaves
  half the spectrum negative, the wanted values a tight cluster at its lower
  edge) within the residual bound |theta - lambda| <= ||r|| <= tol, and the
  cleaned list equal to the dense removal of those eigenpairs; on the
  analytic (positive semidefinite) ERIs the change stays at rounding level;
* randomized Rayleigh-Ritz: on a matrix of exact rank 40 (20 negative
  eigenvalues) a rank-48 range finder captures the range, so the Ritz values
  are the eigenvalues and the cleanup is the dense projection onto the
  positive part (1e-12 with the FP64 sketch; 1e-7 with the FP32 sketch,
  whose range carries FP32 rounding that the FP64 Rayleigh-Ritz step only
  partly removes).
"""
import numpy as np
import pytest

from slater_stoch_amd import basis as B
from slater_stoch_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from slater_stoch_amd.zvmc import Engine
    e = Engine(0)
    yield e
    e.close()


def _pq(bas):
    is_, js, ks, ls = B.quartet_mapping(bas.n)
    return (is_ * (is_ - 1)) // 2 + ks - 1, (js * (js - 1)) // 2 + ls - 1


def pair_matrix(bas, eri):
    P, Q = _pq(bas)
    V = np.zeros((bas.npair, bas.npair))
    V[P, Q] = eri
    V[Q, P] = eri
    return V


def pair_list(bas, V):
    P, Q = _pq(bas)
    return 0.5 * (V[P, Q] + V[Q, P])


@pytest.fixture(scope="module")
def mol30(engine):
    bas = configs.random_molecule(seed=7, n_heavy=1, n_h=3, box=3.0, ng0=4)
    assert bas.n == 30 and bas.npair == 465
    engine.set_basis(bas)
    eri, _, _ = engine.gauss_eri()
    return bas, eri


# (noise, tol, preconditioner, iteration bound): noise 1e-6 leaves the ten
# lowest eigenvalues (-4e-5) 1e-7..1e-6 apart; noise 1e-8 leaves them
# (-4e-7) 1e-9..1e-8 apart, 6.5e-10 between the 10th and the 11th, under a
# largest eigenvalue of 16.  A numpy restatement of the same iteration needs
# 169 / 891 / 1411 plain iterations and 19 / 21 / 24 with the shift-invert
# preconditioner for the three cases.
DAVIDSON_CASES = [(1e-6, 1e-10, "none", 400), (1e-6, 1e-10, "shift-invert", 60),
                  (1e-8, 1e-10, "none", 2000), (1e-8, 1e-10, "shift-invert", 60),
                  (1e-8, 1e-11, "shift-invert", 60)]


@pytest.mark.parametrize("noise,tol,precond,bound", DAVIDSON_CASES)
def test_davidson_lowest_matches_dense(engine, mol30, noise, tol, precond, bound):
    bas, eri = mol30
    engine.set_basis(bas)
    noisy = eri + noise * np.random.default_rng(3).standard_normal(eri.size)
    got, ev, res, nrem, iters = engine.davidson_lowest(noisy, nev=10, tol=tol, max_iter=3000,
                                                       precond=precond)
    assert res.max() <= tol
    V = pair_matrix(bas, noisy)
    w, U = np.linalg.eigh(V)
    # the Ritz values lie within their residuals of the ten lowest eigenvalues
    np.testing.assert_allclose(ev, w[:10], rtol=0, atol=2 * tol)
    assert nrem == int((w[:10] < 0).sum()) == 10
    lo = U[:, :10]
    ref = pair_list(bas, V - (lo * np.minimum(w[:10], 0.0)) @ lo.T)
    # the removed part (a projection onto the span of the ten) is known to
    # about ||r|| / (lambda_11 - lambda_10) of its size (Davis-Kahan)
    gap = w[10] - w[9]
    bound_removed = 3.0 * res.max() / gap * np.abs(w[:10]).max() + 1e-13
    np.testing.assert_allclose(got, ref, rtol=0, atol=bound_removed)
    assert 0 < iters <= bound


def test_davidson_on_positive_semidefinite_input(engine, mol30):
    """Analytic ERIs: a near-continuum of eigenvalues below 1e-12 under a
    largest one of 16, where residual-driven expansion converges slowly.
    Whatever it reaches in 300 iterations: the Ritz values interlace with the
    spectrum (theta_j >= lambda_j) and lie within their residuals of it, and
    only converged negative pairs, of rounding size, are removed."""
    bas, eri = mol30
    engine.set_basis(bas)
    got, ev, res, nrem, iters = engine.davidson_lowest(eri, nev=4, tol=1e-10, max_iter=300)
    w = np.linalg.eigvalsh(pair_matrix(bas, eri))
    assert np.all(ev >= w[:4] - 1e-13)
    for t, r in zip(ev, res):
        assert np.min(np.abs(w - t)) <= r + 1e-13
    assert nrem <= int(np.sum((ev < 0) & (res <= 1e-10)))
    np.testing.assert_allclose(got, eri, rtol=0, atol=1e-12)


def _low_rank(bas, seed):
    rng = np.random.default_rng(seed)
    U = np.linalg.qr(rng.standard_normal((bas.npair, 40)))[0]
    lam = np.concatenate([rng.uniform(0.1, 1.0, 20), -rng.uniform(0.01, 0.5, 20)])
    V = (U * lam) @ U.T
    Vp = (U[:, lam > 0] * lam[lam > 0]) @ U[:, lam > 0].T
    return pair_list(bas, V), pair_list(bas, Vp), np.sort(lam)


@pytest.mark.parametrize("sketch,tol", [("fp64", 1e-12), ("fp32", 1e-7)])
def test_rsvd_clean_exact_low_rank(engine, mol30, sketch, tol):
    bas, _ = mol30
    engine.set_basis(bas)
    eri, ref, lam = _low_rank(bas, 11)
    got, ritz, nrem = engine.rsvd_clean(eri, rank=48, q=2, seed=4, sketch=sketch)
    big = np.abs(ritz) > 1e-10
    assert big.sum() == 40
    np.testing.assert_allclose(ritz[big], lam, rtol=0, atol=1e-12)
    # the 20 negative pairs are converged (exact range); the near-zero Ritz
    # values of the 8 surplus directions may be either sign
    assert int((ritz < -1e-10).sum()) == 20 and nrem >= 20
    np.testing.assert_allclose(got, ref, rtol=0, atol=tol)


def test_rsvd_clean_on_positive_semidefinite_input(engine, mol30):
    bas, eri = mol30
    engine.set_basis(bas)
    got, ritz, _ = engine.rsvd_clean(eri, rank=64, q=2, seed=2, sketch="fp32")
    assert ritz.min() > -1e-12 * ritz.max()
    np.testing.assert_allclose(got, eri, rtol=0, atol=1e-12 * np.abs(eri).max())
