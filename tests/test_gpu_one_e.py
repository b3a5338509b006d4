"""GPU checks of the one-electron integrals over the Gaussian fits
(zv_one_e; src/one_e_slater.F90:87-95, src/one_elect.F90).

* water (20 Gaussians) and a basis with 3S-5G (4 Gaussians): every S, V, K
  entry equals the reference build (V within the reference's own 1e-9
  coefficient cut, see test_oracle.ONE_E_TOL) and the oracle's
  McMurchie-Davidson restatement to 1e-12;
* the 110-AO molecule: the matrices are symmetric, S is positive definite
  (a Gram matrix) and K has a positive diagonal.
"""
import numpy as np
import pytest

from oracle_lib import Oracle, golden
from slater_stoch_amd import configs
from test_oracle import ONE_E_TOL, one_e_mixed_basis

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from slater_stoch_amd.zvmc import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("tag", ["water", "mixed"])
def test_one_e_matches_reference_and_oracle(engine, tag):
    g = golden("one_e.npz")
    bas = configs.water_minimal(ng0=20) if tag == "water" else one_e_mixed_basis()
    engine.set_basis(bas)
    got = dict(zip("SVK", engine.one_e()))
    o = Oracle()
    o.set_basis(bas)
    want = dict(zip("SVK", o.one_e()))
    for name in "SVK":
        ref = g[f"{tag}_{name}"]
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got[name], ref, rtol=0, atol=ONE_E_TOL[name] * scale,
                                   err_msg=f"{tag} {name} vs reference")
        np.testing.assert_allclose(got[name], want[name], rtol=0, atol=1e-12 * scale,
                                   err_msg=f"{tag} {name} vs oracle")


def test_one_e_properties_on_110_ao_molecule(engine):
    bas = configs.random_molecule()
    engine.set_basis(bas)
    S, V, K = engine.one_e()
    for M in (S, V, K):
        assert np.abs(M - M.T).max() <= 1e-12 * np.abs(M).max()
    assert np.linalg.eigvalsh(0.5 * (S + S.T)).min() > 0
    assert (np.diag(K) > 0).all() and (np.diag(V) < 0).all()


def test_one_e_rejects_missing_nuclei(engine):
    from slater_stoch_amd.zvmc import ZVError
    engine.set_basis(configs.water_minimal(ng0=4))
    with pytest.raises(ZVError):
        engine.one_e(charges=[], nuclei=np.zeros((0, 3)))
