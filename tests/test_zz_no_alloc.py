"""The allocation-free entry point zv_gauss_ijkl_list_noalloc against
zv_gauss_ijkl_list (same kernels and summation order: bit for bit), on two
calls in a row, the second with a smaller list, on water and on the 110-AO
molecule; its size limit and its precondition (primitive-pair tables built)
are reported as errors."""
import numpy as np
import pytest

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


def _lists(bas, sizes, seed):
    is_, js, ks, ls = B.quartet_mapping(bas.n)
    rng = np.random.default_rng(seed)
    out = []
    for m in sizes:
        pick = rng.choice(is_.size, m, replace=m > is_.size)
        out.append((is_[pick], ks[pick], js[pick], ls[pick]))
    return out


@pytest.mark.parametrize("name", ["water", "rand110"])
def test_noalloc_equals_list_entry_point(engine, name):
    bas = configs.water_minimal(ng0=20) if name == "water" else configs.random_molecule()
    engine.set_basis(bas)
    big, small = _lists(bas, (40000, 3000) if name == "rand110" else (5000, 300), 7)
    ref_big = engine.gauss_ijkl(*big)          # builds the primitive-pair tables
    ref_small = engine.gauss_ijkl(*small)
    got_big = engine.gauss_ijkl_noalloc(*big)
    got_small = engine.gauss_ijkl_noalloc(*small)
    np.testing.assert_array_equal(got_big, ref_big)
    np.testing.assert_array_equal(got_small, ref_small)
    # the second call leaves nothing of the first behind: again, reversed
    np.testing.assert_array_equal(engine.gauss_ijkl_noalloc(*big), ref_big)


def test_noalloc_matches_reference_golden(engine):
    bas = configs.random_molecule()
    engine.set_basis(bas)
    g = golden("rand12_ng20_eri.npz")
    qs = g["quartets"]
    engine.gauss_ijkl(qs[:1, 0], qs[:1, 1], qs[:1, 2], qs[:1, 3])
    got = engine.gauss_ijkl_noalloc(qs[:, 0], qs[:, 1], qs[:, 2], qs[:, 3])
    np.testing.assert_allclose(got, g["eri"], rtol=1e-10, atol=1e-14)


def test_noalloc_limits_and_state(engine):
    from slater_stoch_amd import zvmc as Z
    bas = configs.water_minimal(ng0=4)
    engine.set_basis(bas)
    one = [np.array([1])] * 4
    with pytest.raises(Z.ZVError, match="primitive-pair tables"):
        engine.gauss_ijkl_noalloc(*one)
    engine.gauss_ijkl(*one)
    too_many = [np.ones(Z.NOALLOC_MAXQ + 1, np.int32)] * 4
    with pytest.raises(Z.ZVError, match="at most 65536"):
        engine.gauss_ijkl_noalloc(*too_many)
    full = [np.ones(Z.NOALLOC_MAXQ, np.int32)] * 4
    got = engine.gauss_ijkl_noalloc(*full)
    np.testing.assert_array_equal(got, engine.gauss_ijkl(*full))
    # a one-quartet list splits the primitive loop 32 ways: other rounding
    np.testing.assert_allclose(got, engine.gauss_ijkl(*one)[0], rtol=1e-13)
