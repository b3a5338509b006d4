"""Targeted refinement (zv_mc_refine): quartets whose kept standard error is
above the target get a fresh run of the unit-pair estimator over the unit
pairs that hold them, on the samples a full zv_mc_blocks run with the same
seed, stream and blocks would draw, so

* every refined quartet equals that full run's value up to summation order:
  within 1e-3 of its standard error (reordering a sum of M terms changes it
  by ~eps sum|t|, about eps sqrt(M) standard errors: 1e-14 in FP64, 6e-6 in
  FP32 at M = 8000);
* every other quartet keeps its pilot mean and standard error, bit for bit;
* the count of refined quartets and the reported maximum match the arrays.
"""
import numpy as np
import pytest

from slater_stoch_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from slater_stoch_amd.zvmc import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["water", "mol30"])
def test_refine_equals_full_run_on_selected_quartets(engine, prec, name):
    from slater_stoch_amd import zvmc as Z
    bas = (configs.water_minimal(ng0=4) if name == "water" else
           configs.random_molecule(seed=7, n_heavy=1, n_h=3, box=3.0, ng0=4))
    engine.set_basis(bas)
    engine.set_tuning(Z.TUNE_EST_SPLIT, 1)
    try:
        pm, pe = engine.mc_blocks(10, 400, prec, seed=3, stream=0)
        target = float(np.percentile(pe, 90))
        sel = pe > target
        nref, emax = engine.mc_refine(10, 800, prec, target=target, seed=3, stream=5)
        got_m = engine.get_result(Z.RESULT_MEAN)
        got_e = engine.get_result(Z.RESULT_STDERR)
        fm, fe = engine.mc_blocks(10, 800, prec, seed=3, stream=5)
    finally:
        engine.set_tuning(Z.TUNE_EST_SPLIT, 0)
    assert nref == int(sel.sum()) > 0
    np.testing.assert_array_less(np.abs(got_m[sel] - fm[sel]), 1e-3 * fe[sel] + 1e-300)
    np.testing.assert_array_less(np.abs(got_e[sel] - fe[sel]), 1e-3 * fe[sel] + 1e-300)
    np.testing.assert_array_equal(got_m[~sel], pm[~sel])
    np.testing.assert_array_equal(got_e[~sel], pe[~sel])
    assert emax == got_e.max()


def test_refine_with_nothing_above_target(engine):
    from slater_stoch_amd import zvmc as Z
    engine.set_basis(configs.water_minimal(ng0=4))
    _, pe = engine.mc_blocks(10, 400, "fp32", seed=1, stream=0)
    nref, emax = engine.mc_refine(10, 400, "fp32", target=2 * pe.max(), seed=1, stream=1)
    assert nref == 0 and emax == pe.max()
    with pytest.raises(Z.ZVError, match="group estimator"):
        engine.mc_refine(10, 400, "fp64_parity", target=1e-6)
