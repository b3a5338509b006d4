"""The 200-AO workload on the GPU: 40 C/H atoms in a 10 A box (5 C with
1S,2S,2S,2P,3D, 35 H with 1S,2P; the reference's example/VB1 exponents),
20100 pairs, 202,015,050 quartets, 20-Gaussian fits.

* analytic ERIs (shell kernels, Schwarz driver src/two_e_slater.F90:219-284)
  against the oracle's gauss_ijkl (src/gauss_ijkl.F90:51-154) on quartets
  stratified by the orders of the four AOs and the number of distinct
  centres: |d| <= 1e-10 |ref| + 1e-14;
* the screened count recomputed from precond, screened entries exactly 0,
  diagonal quartets equal to precond;
* the fixed-sample estimator (unsplit instances, as timed) against the
  exact (extended-precision) evaluation on 5,000 quartets per estimator
  class pair;
* at full size, two free-running runs on independent Philox streams (FP64
  and FP32 estimators) agree within their standard errors on all quartets;
* the shift-invert Davidson's 10 lowest Ritz values of the 20100 x 20100
  pair supermatrix against a dense eigensolve of the same matrix.
"""
import numpy as np
import pytest

from estimator_ref import fp64_estimate
from slater_stoch_amd import basis as B
from slater_stoch_amd import configs
from test_gpu_estimator import _check, stratified_quartets

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from slater_stoch_amd.zvmc import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def c5(engine):
    bas = configs.large_ch()
    assert bas.n == 200 and bas.nint == 202015050
    engine.set_basis(bas)
    eri, pre, nscr = engine.gauss_eri()
    q = B.quartet_mapping(bas.n)
    return bas, eri, pre, nscr, q


def test_c5_screening_and_diagonal(c5):
    bas, eri, pre, nscr, (is_, js, ks, ls) = c5
    scr = np.abs(pre[ks - 1, is_ - 1] * pre[ls - 1, js - 1]) < 1e-30
    diag = (is_ == js) & (ks == ls)
    assert nscr == int(scr.sum())
    assert np.all(eri[scr & ~diag] == 0.0)
    np.testing.assert_array_equal(eri[diag], pre[ks[diag] - 1, is_[diag] - 1])
    assert np.all(np.isfinite(eri))


def test_c5_analytic_stratified_vs_oracle(c5, oracle):
    bas, eri, pre, nscr, (is_, js, ks, ls) = c5
    oracle.set_basis(bas)
    lao = bas.npower.sum(1) + 2 * (bas.i_type == 2)
    nuc = bas.nucleus
    rng = np.random.default_rng(29)
    # a random million live quartets, then up to 3 per stratum
    live = np.flatnonzero(np.abs(pre[ks - 1, is_ - 1] * pre[ls - 1, js - 1]) >= 1e-30)
    cand = rng.choice(live, 1_000_000, replace=False)
    a, b, c, d = (x[cand] - 1 for x in (is_, ks, js, ls))
    orders = np.sort(np.stack([lao[a], lao[b], lao[c], lao[d]]), axis=0)
    cents = np.sort(np.stack([nuc[a], nuc[b], nuc[c], nuc[d]]), axis=0)
    ncent = 1 + (np.diff(cents, axis=0) != 0).sum(0)
    key = ((orders[0] * 8 + orders[1]) * 8 + orders[2]) * 8 + orders[3]
    key = key * 8 + ncent
    pick = []
    for k in np.unique(key):
        idx = cand[key == k]
        pick.extend(idx[:3])
    pick = np.array(sorted(pick))
    assert pick.size >= 60
    ref, _ = oracle.gauss_eri([is_[pick], js[pick], ks[pick], ls[pick]], None)
    np.testing.assert_allclose(eri[pick], ref, rtol=1e-10, atol=1e-14)


@pytest.fixture(scope="module")
def c5_fixed(c5, oracle):
    """Fixed samples (4 blocks of 40) and the extended-precision estimator on
    5,000 quartets of each estimator class pair."""
    bas = c5[0]
    oracle.set_basis(bas)
    nbl, npts = 4, 40
    rng = np.random.default_rng(31)
    r1, r2 = rng.standard_normal((nbl * npts, 3)), rng.standard_normal((nbl * npts, 3))
    pi0 = oracle.pi0(r1, r2)
    pick, nclass = stratified_quartets(bas, 5000, 7)
    assert nclass == 9
    ref, scale = fp64_estimate(oracle, bas, r1, r2, pi0, nbl, npts, pick=pick)
    return nbl, npts, (r1, r2, pi0), pick, ref, scale


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_c5_fixed_sample_estimator_subset(engine, c5, c5_fixed, prec):
    from slater_stoch_amd.zvmc import TUNE_EST_SPLIT
    bas = c5[0]
    engine.set_basis(bas)
    nbl, npts, (r1, r2, pi0), pick, ref, scale = c5_fixed
    got, _ = engine.mc_blocks(nbl, npts, prec, samples=(r1, r2, pi0))
    try:
        engine.set_tuning(TUNE_EST_SPLIT, 1)
        one, _ = engine.mc_blocks(nbl, npts, prec, samples=(r1, r2, pi0))
    finally:
        engine.set_tuning(TUNE_EST_SPLIT, 0)
    assert np.array_equal(got, one)  # the automatic choice is the unsplit kernel
    _check(got[pick], ref, scale, prec, (prec, "200 AO"))


def test_c5_independent_streams_agree(engine, c5):
    bas = c5[0]
    engine.set_basis(bas)
    m1, e1 = engine.mc_blocks(10, 4000, "fp64", seed=5, stream=1)
    m2, e2 = engine.mc_blocks(10, 4000, "fp32", seed=5, stream=2)
    assert np.all(np.isfinite(m1)) and np.all(np.isfinite(m2))
    assert np.all(e1 >= 0) and np.all(e2 >= 0)
    z = (m1 - m2) / np.sqrt(e1 ** 2 + e2 ** 2 + 1e-300)
    # ten blocks: the block-variance stderr is itself noisy (t with 9 dof
    # and the reference's biased variance), hence the wide single bound
    assert np.mean(np.abs(z) > 4.0) < 0.02
    assert np.abs(z).max() < 40.0


def test_c5_davidson_shift_invert_vs_dense(engine, c5):
    """The 10 lowest eigenpairs of the 20100 x 20100 pair supermatrix of the
    final ERIs (analytic + a 10 x 4000 FP64 pilot mean), by block Davidson
    with the shift-invert preconditioner, against a dense FP64 eigensolve of
    the same matrix on the GPU (torch.linalg.eigvalsh in a child process):
    every Ritz value within twice the tolerance of its eigenvalue.  The MC
    noise leaves 745 eigenvalues below -1e-9 Ha, the lowest -5.1e-6,
    3e-8..3e-6 apart, under a largest one of 23: the plain iteration had not
    found a negative one after 300 iterations."""
    from dense_eig import lowest_eigenvalues
    bas, eri = c5[0], c5[1]
    engine.set_basis(bas)
    mean, _ = engine.mc_blocks(10, 4000, "fp64", seed=1, stream=0)
    vals = eri + mean
    tol = 1e-9
    _, ev, res, nrem, iters = engine.davidson_lowest(vals, nev=10, tol=tol, max_iter=400,
                                                     precond="shift-invert")
    w = lowest_eigenvalues(bas.n, vals, 11)
    assert res.max() <= tol, (iters, res)
    np.testing.assert_allclose(ev, w[:10], rtol=0, atol=2 * tol)
    assert nrem == int((w[:10] < 0).sum())
    print(f"davidson shift-invert on the 200-AO pair matrix: {iters} iterations, lowest "
          f"{ev[0]:.6e} (dense {w[0]:.6e})")
