"""GPU checks of the exact one-centre Slater integrals
(zv_exact_monocenter; src/two_e_slater.F90:318-365, src/compute_int_slater.F90).

* every one-centre quartet of a basis with each shell type the reference's
  formulas accept equals the reference build's values (golden fixture) and
  the oracle's independent restatement; other entries stay untouched and the
  standard errors of the one-centre quartets become 0;
* an f function on any atom is an error, where the reference stops;
* the method's own consistency: for the one-centre quartets of water the
  Gaussian-fit ERIs plus the Monte Carlo correction (an unbiased estimate of
  the STO integral) agree with the exact values within their standard
  errors.
"""
import numpy as np
import pytest

from oracle_lib import Oracle, golden
from slater_stoch_amd import basis as B
from slater_stoch_amd import configs
from test_oracle import mono_basis

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from slater_stoch_amd.zvmc import Engine
    e = Engine(0)
    yield e
    e.close()


def test_exact_monocenter_matches_reference(engine):
    g = golden("mono_exact.npz")
    bas = mono_basis()
    engine.set_basis(bas)
    vals = np.full(bas.nint, 7.0)
    err = np.full(bas.nint, 3.0)
    nm = engine.exact_monocenter(vals, err)
    assert nm == g["kcp"].size == 9526
    ref = g["exact"]
    scale = np.abs(ref).max()
    np.testing.assert_allclose(vals[g["kcp"]], ref, rtol=1e-12, atol=1e-14 * scale)
    assert (err[g["kcp"]] == 0).all()
    rest = np.setdiff1d(np.arange(bas.nint), g["kcp"])
    assert (vals[rest] == 7.0).all() and (err[rest] == 3.0).all()
    # the oracle's restatement on a sample, including zero-by-symmetry quartets
    o = Oracle()
    o.set_basis(bas)
    pick = np.random.default_rng(5).choice(g["kcp"].size, 300, replace=False)
    want = [o.mono_exact(g["is_"][t], g["ks"][t], g["js"][t], g["ls"][t]) for t in pick]
    np.testing.assert_allclose(vals[g["kcp"][pick]], want, rtol=1e-12, atol=1e-14 * scale)


def test_exact_monocenter_rejects_f_functions(engine):
    from slater_stoch_amd.zvmc import ZVError
    geom = B.geometry_from_angstrom(["O", "N"], [[0, 0, 0], [0, 0, 1.2]])
    bas = B.expand_basis(geom, {"O": [("1S", 7.7), ("2P", 2.3)], "N": [("4F", 1.7)]}, ng0=4)
    engine.set_basis(bas)
    with pytest.raises(ZVError):
        engine.exact_monocenter(np.zeros(bas.nint))


def test_monte_carlo_is_unbiased_on_one_centre_quartets(engine):
    """Water, 20-Gaussian fits: analytic fit ERIs + ZVMC correction (40 blocks
    x 20000 samples, FP64) against the exact one-centre values.  With 40
    blocks the standard error's own noise is small enough (Student t, 39
    degrees of freedom) for |z| < 6 over the 122 one-centre quartets."""
    bas = configs.water_minimal(ng0=20)
    engine.set_basis(bas)
    eri, _, _ = engine.gauss_eri()
    mean, err = engine.mc_blocks(40, 20000, "fp64", seed=2024)
    total = eri + mean
    exact = total.copy()
    e0 = err.copy()
    nm = engine.exact_monocenter(exact, e0)
    one = np.flatnonzero(e0 == 0)
    assert nm == one.size == 122
    live = err[one] > 0
    z = (total[one][live] - exact[one][live]) / err[one][live]
    assert np.abs(z).max() < 6.0, np.abs(z).max()
    assert 0.3 < np.mean(z * z) < 3.0, np.mean(z * z)
    # exact zeros by symmetry: the estimate is zero too or within its error
    dead = ~live
    assert np.all(np.abs(total[one][dead] - exact[one][dead]) < 1e-12)
