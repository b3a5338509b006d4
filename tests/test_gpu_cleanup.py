"""GPU checks of the PSD cleanup (src/svd_clean.F90, src/davidson_clean.F90).

The reference's cleanup calls LAPACK through BLAS libraries this image does
not have, so it is not part of the reference build; both cleanups are pinned
by what the mathematics fixes:
* the dense branch of davidson_clean equals W - sum_{e<0} e u u^T computed
  with numpy.linalg.eigh on the same product-basis matrix;
* svd_clean's randomized passes converge to the same projection, averaged
  over the eight permutations;
* a positive semidefinite input comes back unchanged.
The final average over the eight permutations (src/svd_clean.F90:372-405)
mixes row and column indices (a partial transpose of W), so the cleaned list
itself need not be positive semidefinite; the tests compare against the same
averaged projection instead.
"""
import numpy as np
import pytest

from oracle_lib import golden
from slater_stoch_amd import basis as B
from slater_stoch_amd import configs

pytestmark = pytest.mark.gpu

PERMS = [(0, 1, 2, 3), (2, 1, 0, 3), (2, 3, 0, 1), (0, 3, 2, 1),
         (1, 0, 3, 2), (1, 2, 3, 0), (3, 2, 1, 0), (3, 0, 1, 2)]


@pytest.fixture(scope="module")
def engine():
    from slater_stoch_amd.zvmc import Engine
    e = Engine(0)
    yield e
    e.close()


def product_matrix(n, q, vals):
    """W((i,j),(k,l)) = <ij|kl> from the unique list (src/svd_clean.F90:97-130)."""
    N = n * n
    W = np.zeros((N, N))
    ijkl = np.stack(q, axis=1) - 1  # columns i, j, k, l (0-based)
    for p in PERMS:
        a, b, c, d = (ijkl[:, p[0]], ijkl[:, p[1]], ijkl[:, p[2]], ijkl[:, p[3]])
        W[a + b * n, c + d * n] = vals
    return W


def gather(n, q, W, avg8):
    ijkl = np.stack(q, axis=1) - 1
    out = np.zeros(ijkl.shape[0])
    for p in (PERMS if avg8 else PERMS[:1]):
        a, b, c, d = (ijkl[:, p[0]], ijkl[:, p[1]], ijkl[:, p[2]], ijkl[:, p[3]])
        out += W[a + b * n, c + d * n] * (0.125 if avg8 else 1.0)
    return out


def negative_removed(W):
    e, U = np.linalg.eigh(W)
    neg = e < 0
    return W - (U[:, neg] * e[neg]) @ U[:, neg].T, int(neg.sum())


def noisy_water():
    bas = configs.water_minimal(ng0=20)
    g = golden("water_ng20_eri.npz")
    q = [g["is_"], g["js"], g["ks"], g["ls"]]
    rng = np.random.default_rng(4)
    vals = g["eri"] + 2e-3 * rng.standard_normal(g["eri"].size)
    return bas, q, vals


def test_davidson_clean_equals_dense_projection(engine):
    bas, q, vals = noisy_water()
    engine.set_basis(bas)
    got, nrem = engine.davidson_clean(vals)
    Wc, nneg = negative_removed(product_matrix(bas.n, q, vals))
    assert nrem == nneg > 0
    np.testing.assert_allclose(got, gather(bas.n, q, Wc, False), rtol=0, atol=1e-12)


def test_svd_clean_equals_projection_full_rank(engine):
    bas, q, vals = noisy_water()
    engine.set_basis(bas)
    got, nrem = engine.svd_clean(vals, rank=bas.n ** 2, q=12, seed=3)
    Wc, nneg = negative_removed(product_matrix(bas.n, q, vals))
    assert nrem == nneg > 0
    ref = gather(bas.n, q, Wc, True)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * np.abs(ref).max())


def test_svd_clean_leaves_psd_input(engine):
    bas = configs.water_minimal(ng0=20)
    g = golden("water_ng20_eri.npz")
    engine.set_basis(bas)
    got, _ = engine.svd_clean(g["eri"], rank=49, q=12, seed=1)
    np.testing.assert_allclose(got, g["eri"], rtol=0, atol=1e-12)


def test_svd_clean_partial_rank_passes(engine):
    """rank < N: several randomized passes, as the reference runs for nbasis^2 > 512."""
    geom = B.geometry_from_angstrom(["O", "H", "H"], [[0, 0, 0.1173], [0, 0.7572, -0.4692],
                                                      [0, -0.7572, -0.4692]])
    bas = B.expand_basis(geom, {"H": [("1S", 1.24), ("2P", 1.5)],
                                "O": [("1S", 7.66), ("2S", 2.25), ("2P", 2.25), ("3D", 2.0)]},
                         ng0=6)
    engine.set_basis(bas)
    eri, _, _ = engine.gauss_eri()
    q = B.quartet_mapping(bas.n)
    rng = np.random.default_rng(8)
    # noise 1e-5 on the unique values turns three of the smallest eigenvalues
    # of W (~3e-3 apart from those) negative, down to -7.4e-5; the passes must
    # find those (at noise 1e-8 W stays positive definite).  A later pass may
    # also remove a negative direction of rounding size that the
    # eight-permutation average of an earlier pass leaves behind (which one
    # depends on the last bits of the ERIs), so the count is at least the
    # dense one and the result equals the dense projection.
    vals = eri + 1e-5 * rng.standard_normal(eri.size)
    got, nrem = engine.svd_clean(vals, rank=96, q=12, seed=5)
    Wc, nneg = negative_removed(product_matrix(bas.n, q, vals))
    ref = gather(bas.n, q, Wc, True)
    assert nrem >= nneg > 0
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * scale)


def test_svd_clean_finds_small_negative_eigenvalues_110_ao(engine):
    """The 110-AO molecule's analytic ERIs plus a short Monte Carlo run give a
    W (N = 12100) with six negative eigenvalues that change W by 5e-11, far
    below its largest (~3).  The randomized passes reach them only in the last,
    clustered passes, where the singular vectors of the r x r factor must stay
    accurate.  svd_clean must remove what the dense eigendecomposition removes:
    its distance to the dense projection stays at the eight-permutation average
    (1.1e-11), well below the size of the removed part (a solver that misses
    them leaves the full 5e-11)."""
    bas = configs.random_molecule()
    engine.set_basis(bas)
    eri, _, _ = engine.gauss_eri()
    mean, _ = engine.mc_blocks(10, 400, "fp64", seed=1)
    vals = eri + mean
    dense, nneg = engine.davidson_clean(vals)
    assert nneg > 0
    removed_part = np.abs(dense - vals).max()
    got, nrem = engine.svd_clean(vals, rank=512, q=12, seed=1)
    assert nrem > 0
    assert np.abs(got - dense).max() < 0.5 * removed_part, (np.abs(got - dense).max(), removed_part)
