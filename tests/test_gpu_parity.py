"""GPU parity: the HIP library (through its C-ABI) against the CPU oracle and
the reference's golden vectors.

Tolerances:
* fixed-sample MC, FP64 parity estimator: bit-identical to the oracle (both
  use the shared exp, the reference's operation order and FP32 casts); against
  the reference's own golden vectors 1e-15 Ha (libm exp differs by an ulp).
* every estimator against a cast-free FP64 numpy evaluation of the formula
  on the same samples, per quartet within tol * sum_m |t_m| / M: 1e-12 for the
  tiled FP64 kernel, 2e-5 for the tiled FP32 kernel, 2e-6 for the parity
  kernel (its FP32 casts are the reference's).
* analytic ERIs: McMurchie-Davidson on the GPU vs the reference's
  polynomial recursion: |d| <= 1e-10 |ref| + 1e-14.
"""
import numpy as np
import pytest
from scipy import special

from oracle_lib import golden
from slater_stoch_amd import basis as B
from slater_stoch_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from slater_stoch_amd.zvmc import Engine
    e = Engine(0)
    yield e
    e.close()


def _mixed_basis():
    geom = B.geometry_from_angstrom(["O", "H", "N"],
                                    [[0.1, -0.2, 0.05], [0.3, 1.05, -0.4], [-1.1, 0.2, 0.7]])
    basis = {"O": [("1S", 7.7), ("2S", 2.2), ("3S", 1.4), ("2P", 2.3), ("3P", 1.1), ("3D", 1.9)],
             "H": [("1S", 1.3), ("2P", 1.6)],
             "N": [("4F", 1.7), ("5G", 2.1)]}
    return B.expand_basis(geom, basis, ng0=4)


# ------------------------------------------------------------------ MC path
def test_pair_samples_bitwise(engine, oracle):
    bas = configs.water_minimal(ng0=20)
    engine.set_basis(bas)
    oracle.set_basis(bas)
    g = golden("water_ng20_pairs.npz")
    for i, k in list(g["pairs"]) + [(7, 1), (5, 5)]:
        got = engine.pair_samples(i, k, g["r1"], g["r2"])
        ref = oracle.pair_samples(i, k, g["r1"], g["r2"])
        for name in ("ut1", "ut2", "rho1", "rho2", "rhog1", "rhog2", "jac1", "jac2"):
            assert np.array_equal(got[name], ref[name]), (i, k, name)


@pytest.mark.parametrize("tag,nbl,npts", [("", 10, 40), ("b", 4, 200)])
def test_mc_fixed_parity_bitwise(engine, oracle, tag, nbl, npts):
    bas = configs.water_minimal(ng0=20)
    engine.set_basis(bas)
    oracle.set_basis(bas)
    g = golden("water_ng20_mc.npz")
    smp = (g["r1" + tag], g["r2" + tag], g["pi0" + tag])
    mean, err, blk = engine.mc_blocks(nbl, npts, "fp64_parity", samples=smp, block_means=True)
    q = B.quartet_mapping(bas.n)
    om, oe, es, eg = oracle.mc_fixed(nbl, npts, *smp, q, blocks=True)
    assert np.array_equal(mean, om)
    assert np.array_equal(err, oe)
    assert np.array_equal(blk, es - eg)
    np.testing.assert_allclose(mean, g["mean" + tag], rtol=0, atol=1e-15)
    np.testing.assert_allclose(err, g["err" + tag], rtol=0, atol=1e-15)


def test_mc_fixed_parity_bitwise_all_orbital_types(engine, oracle):
    bas = _mixed_basis()
    engine.set_basis(bas)
    oracle.set_basis(bas)
    rng = np.random.default_rng(21)
    M = 3 * 80
    r1, r2 = rng.standard_normal((M, 3)), rng.standard_normal((M, 3))
    pi0 = oracle.pi0(r1, r2)
    mean, err = engine.mc_blocks(3, 80, "fp64_parity", samples=(r1, r2, pi0))
    q = B.quartet_mapping(bas.n)
    om, oe = oracle.mc_fixed(3, 80, r1, r2, pi0, q)
    assert np.array_equal(mean, om)
    assert np.array_equal(err, oe)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5), ("fp64_parity", 2e-6),
                                      ("bf16", 8e-3)])
def test_estimators_match_fp64_reference(engine, oracle, prec, tol):
    """Every estimator against a cast-free FP64 numpy evaluation of the same
    formula on the same samples; the bound is tol * sum_m|t_m|/M per quartet
    (FP32 casts of the parity form: 2 x 2^-24 per sample; FP32 tiles: FP32
    coordinates and hardware rsqrt).  BF16 products: each factor of the two
    products K A1 A2 and K B1 B2 is rounded to BF16 (relative 2^-9), so each
    product is within (1 + 2^-9)^2 - 1 = 3.9e-3 of itself and the bound is
    taken over sum_m (|K A1 A2| + |K B1 B2|) / M, with a factor 2 for the
    FP32 parts."""
    from estimator_ref import fp64_estimate
    for bas, nbl, npts in ((configs.water_minimal(ng0=4), 4, 400), (_mixed_basis(), 2, 40)):
        engine.set_basis(bas)
        oracle.set_basis(bas)
        rng = np.random.default_rng(3)
        r1, r2 = rng.standard_normal((nbl * npts, 3)), rng.standard_normal((nbl * npts, 3))
        pi0 = oracle.pi0(r1, r2)
        ref, scale, scale2 = fp64_estimate(oracle, bas, r1, r2, pi0, nbl, npts, split=True)
        if prec == "bf16":
            scale = scale2
        got, _ = engine.mc_blocks(nbl, npts, prec, samples=(r1, r2, pi0))
        bad = np.abs(got - ref) > tol * scale + 1e-300
        assert not bad.any(), (prec, np.max(np.abs(got - ref) / scale))


def test_device_sampler_is_philox_plus_inverse_erfc(engine, oracle):
    """The device stream is Philox4x32-10 (counter = sample index, stream,
    block) mapped by x = 2 - 2u, r = sqrt(2) erfc^-1(x): recomputed here from
    the oracle's Philox and scipy's erfcinv."""
    seed, stream, m0, m = 0x5EED, 7, 1000, 512
    r1, r2, pi0 = engine.draw_samples(m, seed=seed, stream=stream, m0=m0)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    exp_r = np.zeros((m, 6))
    for s in range(m):
        g = m0 + s
        u = []
        for t in range(3):
            o = oracle.philox([g & 0xFFFFFFFF, g >> 32, stream & 0xFFFFFFFF,
                               ((stream >> 32) << 2) | t], key)
            for a, b in ((o[0], o[1]), (o[2], o[3])):
                v = (int(a) >> 5) * (1 << 26) + (int(b) >> 6)
                u.append((v + 0.5) * 2.0 ** -53)
        exp_r[s] = np.sqrt(2.0) * special.erfcinv(2.0 - 2.0 * np.array(u))
    np.testing.assert_allclose(np.hstack([r1, r2]), exp_r, rtol=1e-13, atol=1e-14)
    np.testing.assert_array_equal(pi0, oracle.pi0(r1, r2))


def test_device_sampler_statistics(engine):
    r1, r2, _ = engine.draw_samples(200000, seed=11, stream=3)
    x = np.concatenate([r1.ravel(), r2.ravel()])
    assert abs(x.mean()) < 5 * 1.0 / np.sqrt(x.size)
    assert abs(x.var() - 1.0) < 5 * np.sqrt(2.0 / x.size)
    from scipy import stats
    assert stats.kstest(x[:50000], "norm").pvalue > 1e-4


def test_free_running_equals_fixed_on_device_samples(engine, oracle):
    """Philox mode = fixed-sample mode fed with the device's own samples, and
    both equal the oracle on those samples, bit for bit."""
    bas = configs.water_minimal(ng0=20)
    engine.set_basis(bas)
    oracle.set_basis(bas)
    nbl, npts, seed, stream = 5, 120, 99, 2
    mean, err = engine.mc_blocks(nbl, npts, "fp64_parity", seed=seed, stream=stream)
    smp = engine.draw_samples(nbl * npts, seed=seed, stream=stream)
    mf, ef = engine.mc_blocks(nbl, npts, "fp64_parity", samples=smp)
    assert np.array_equal(mean, mf) and np.array_equal(err, ef)
    om, oe = oracle.mc_fixed(nbl, npts, *smp, B.quartet_mapping(bas.n))
    assert np.array_equal(mean, om) and np.array_equal(err, oe)


def test_independent_streams_agree_statistically(engine):
    """Two free-running runs (FP64 and FP32, different streams) agree within
    their standard errors: z-scores of 406 quartets."""
    bas = configs.water_minimal(ng0=4)
    engine.set_basis(bas)
    m1, e1 = engine.mc_blocks(20, 2000, "fp64", seed=1, stream=1)
    m2, e2 = engine.mc_blocks(20, 2000, "fp32", seed=1, stream=2)
    z = (m1 - m2) / np.sqrt(e1 ** 2 + e2 ** 2 + 1e-300)
    assert np.mean(np.abs(z) > 4.0) < 0.02
    assert np.abs(z).max() < 8.0


# ------------------------------------------------------------------ analytic
def test_gauss_eri_water_matches_reference(engine):
    bas = configs.water_minimal(ng0=20)
    engine.set_basis(bas)
    g = golden("water_ng20_eri.npz")
    eri, pre, nscr = engine.gauss_eri()
    np.testing.assert_allclose(eri, g["eri"], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(pre, g["precond"], rtol=1e-10, atol=1e-14)
    q = [g["is_"], g["js"], g["ks"], g["ls"]]
    ref_scr = np.sum(np.abs(g["precond"][q[2] - 1, q[0] - 1] * g["precond"][q[3] - 1, q[1] - 1])
                     < 1e-30)
    assert nscr == ref_scr


def test_gauss_ijkl_all_orbital_types(engine):
    bas = _mixed_basis()
    engine.set_basis(bas)
    g = golden("mixed_ng4_eri.npz")
    q = B.quartet_mapping(bas.n)
    pick = g["pick"]
    got = engine.gauss_ijkl(q[0][pick], q[2][pick], q[1][pick], q[3][pick])
    scale = np.abs(g["eri"]).max()
    # high angular momentum (f, g) near T = 1 is where the reference's upward
    # Boys recursion is least accurate: allow 1e-9 of the largest integral there
    np.testing.assert_allclose(got, g["eri"], rtol=1e-9, atol=1e-12 * scale)


def test_gauss_ijkl_rand12_subset(engine):
    bas = configs.random_molecule()
    engine.set_basis(bas)
    g = golden("rand12_ng20_eri.npz")
    qs = g["quartets"]
    got = engine.gauss_ijkl(qs[:, 0], qs[:, 1], qs[:, 2], qs[:, 3])
    np.testing.assert_allclose(got, g["eri"], rtol=1e-10, atol=1e-14)


def test_gauss_eri_full_list_equals_explicit_list(engine):
    """The class-batched full run and the explicit-list path agree, and the
    full run honours the 8-fold permutational symmetry."""
    bas = _mixed_basis()
    engine.set_basis(bas)
    eri, pre, _ = engine.gauss_eri()
    q = B.quartet_mapping(bas.n)
    rng = np.random.default_rng(5)
    pick = rng.choice(q[0].size, 400, replace=False)
    i, j, k, l = [a[pick] for a in q]
    a = engine.gauss_ijkl(i, k, j, l)
    b = engine.gauss_ijkl(j, l, i, k)
    c = engine.gauss_ijkl(k, i, l, j)
    keep = np.abs(pre[k - 1, i - 1] * pre[l - 1, j - 1]) >= 1e-30
    np.testing.assert_allclose(eri[pick][keep], a[keep], rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(c, a, rtol=1e-12, atol=1e-16)


def test_gauss_eri_every_quartet_of_all_orbital_types(engine):
    """Every quartet of the mixed basis (s, r^2-type s, p, r^2-type p, d, f, g
    AOs): the shell kernels, which skip the structural zeros of each lane
    pair's Hermite coefficients (a box of per-axis power sums, a union of
    boxes for r^2-type AOs), against the component kernels, which sum over
    every coefficient.  A pattern missing a nonzero coefficient would drop a
    term here."""
    bas = _mixed_basis()
    engine.set_basis(bas)
    eri, pre, _ = engine.gauss_eri()
    assert any(k[0] == "shell" for k in engine.eri_stats())
    i, j, k, l = B.quartet_mapping(bas.n)
    keep = np.abs(pre[k - 1, i - 1] * pre[l - 1, j - 1]) >= 1e-30
    a = engine.gauss_ijkl(i[keep], k[keep], j[keep], l[keep])
    scale = np.abs(a).max()
    np.testing.assert_allclose(eri[keep], a, rtol=1e-12, atol=1e-15 * scale)


def _kcp_lookup(is_, js, ks, ls, n):
    """Reference position of quartet (i k | j l) (any of its 8 equivalent forms)."""
    key = lambda a, b, c, d: ((a.astype(np.int64) * (n + 1) + b) * (n + 1) + c) * (n + 1) + d
    keys = key(is_, js, ks, ls)
    order = np.argsort(keys)
    sk = keys[order]

    def find(i, k, j, l):
        for a, b, c, d in ((i, k, j, l), (k, i, j, l), (i, k, l, j), (k, i, l, j),
                           (j, l, i, k), (l, j, i, k), (j, l, k, i), (l, j, k, i)):
            kk = key(np.array([a]), c, b, d)[0]  # is = e1a, js = e2a, ks = e1b, ls = e2b
            t = np.searchsorted(sk, kk)
            if t < sk.size and sk[t] == kk:
                return order[t]
        raise KeyError((i, k, j, l))
    return find


def test_gauss_eri_shells_on_110_ao_molecule(engine):
    """The full Schwarz run on the 110-AO molecule (shell kernels for its
    s, p and d shell pairs): the reference's values of 48 stratified quartets,
    screened quartets exactly zero (src/two_e_slater.F90:268-272), diagonal
    quartets equal to precond, and agreement with the component kernels."""
    bas = configs.random_molecule()
    engine.set_basis(bas)
    eri, pre, nscr = engine.gauss_eri()
    st = engine.eri_stats()
    assert any(k[0] == "shell" for k in st)
    n = bas.n
    is_, js, ks, ls = B.quartet_mapping(n)
    g = golden("rand12_ng20_eri.npz")
    find = _kcp_lookup(is_, js, ks, ls, n)
    idx = np.array([find(*map(int, qq)) for qq in g["quartets"]])
    np.testing.assert_allclose(eri[idx], g["eri"], rtol=1e-10, atol=1e-14)
    scr = np.abs(pre[ks - 1, is_ - 1] * pre[ls - 1, js - 1]) < 1e-30
    diag = (is_ == js) & (ks == ls)
    assert nscr == int(scr.sum())
    assert np.all(eri[scr & ~diag] == 0.0)
    np.testing.assert_array_equal(eri[diag], pre[ks[diag] - 1, is_[diag] - 1])
    rng = np.random.default_rng(11)
    live = np.flatnonzero(~scr)
    pick = rng.choice(live, 3000, replace=False)
    a = engine.gauss_ijkl(is_[pick], ks[pick], js[pick], ls[pick])
    scale = np.abs(eri).max()
    np.testing.assert_allclose(eri[pick], a, rtol=1e-11, atol=1e-15 * scale)


def test_errors_are_reported(engine):
    from slater_stoch_amd.zvmc import ZVError
    bas = configs.water_minimal(ng0=4)
    engine.set_basis(bas)
    with pytest.raises(ZVError, match="multiple of nw"):
        engine.mc_blocks(2, 50, "fp64")
    with pytest.raises(ZVError, match="outside"):
        engine.gauss_ijkl([8], [1], [1], [1])
