"""GPU checks of the device-resident results (zv_get_result, zv_result_max).

A NULL host output keeps the per-quartet array on the device; fetched
afterwards it carries the same bits as the copy the call would have made,
and the on-device maximum equals numpy's over that array.  A new basis
invalidates the kept results.  Kept results also show that the analytic
ERIs repeat bit for bit (concurrent shell kernels, threaded planner)."""
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


def test_device_results_equal_host_outputs(engine):
    from slater_stoch_amd import zvmc as Z
    engine.set_basis(configs.water_minimal(ng0=20))
    eri, pre, nscr = engine.gauss_eri()
    none, pre2, nscr2 = engine.gauss_eri(device_only=True)
    assert none is None and nscr2 == nscr
    np.testing.assert_array_equal(pre2, pre)
    kept = engine.get_result(Z.RESULT_ERI)
    np.testing.assert_array_equal(kept, eri)
    assert engine.result_max(Z.RESULT_ERI) == eri.max()

    mean, err = engine.mc_blocks(10, 400, "fp64", seed=7, stream=3)
    m2, e2 = engine.mc_blocks(10, 400, "fp64", seed=7, stream=3, device_only=True)
    assert m2 is None and e2 is None
    timings = engine.timings()
    assert timings["estimator"][1] > 0
    np.testing.assert_array_equal(engine.get_result(Z.RESULT_MEAN), mean)
    np.testing.assert_array_equal(engine.get_result(Z.RESULT_STDERR), err)
    assert engine.result_max(Z.RESULT_STDERR) == err.max()
    assert engine.result_max(Z.RESULT_MEAN) == mean.max()
    # the queries leave the timings of the last computing call alone
    assert engine.timings() == timings


def test_device_results_on_110_ao_molecule(engine):
    """Sizes above one reduction pass of the maximum (18.6e6 quartets)."""
    from slater_stoch_amd import zvmc as Z
    engine.set_basis(configs.random_molecule())
    engine.mc_blocks(10, 40, "fp32", seed=3, device_only=True)
    err = engine.get_result(Z.RESULT_STDERR)
    assert err.size == 18_638_565
    assert engine.result_max(Z.RESULT_STDERR) == err.max()


def test_new_basis_invalidates_results(engine):
    from slater_stoch_amd import zvmc as Z
    engine.set_basis(configs.water_minimal(ng0=4))
    engine.gauss_eri(device_only=True)
    engine.get_result(Z.RESULT_ERI)
    engine.set_basis(configs.water_minimal(ng0=4))
    for which in (Z.RESULT_ERI, Z.RESULT_MEAN, Z.RESULT_STDERR):
        with pytest.raises(Z.ZVError):
            engine.get_result(which)
        with pytest.raises(Z.ZVError):
            engine.result_max(which)


def test_analytic_eris_repeat_bit_for_bit(engine):
    """The shell kernels of a Schwarz phase run side by side on several
    streams and their work plan is built on host threads; partial sums are
    added in a fixed order, so two calls give the same bits."""
    from slater_stoch_amd import zvmc as Z
    engine.set_basis(configs.random_molecule())
    engine.gauss_eri(device_only=True)
    first = engine.get_result(Z.RESULT_ERI)
    engine.gauss_eri(device_only=True)
    np.testing.assert_array_equal(engine.get_result(Z.RESULT_ERI), first)


def _with_pair_order_too_high(bas):
    """A copy of the basis whose first AO is an i_type 2 AO with powers
    (2, 2, 1): every axis passes the per-axis check, but its pair with
    itself has order 2 (5 + 2) = 14, above the supported 8."""
    import copy
    b = copy.deepcopy(bas)
    b.npower = np.array(b.npower, copy=True)
    b.i_type = np.array(b.i_type, copy=True)
    b.npower[0] = (2, 2, 1)
    b.i_type[0] = 2
    return b


def test_rejected_basis_keeps_the_previous_one(engine):
    """A basis rejected by validation is refused before anything is freed:
    the previous basis and its kept results stay usable.  No kernel runs on
    the rejected input."""
    from slater_stoch_amd import zvmc as Z
    good = configs.water_minimal(ng0=4)
    engine.set_basis(good)
    eri, _, _ = engine.gauss_eri()
    with pytest.raises(Z.ZVError, match="angular momentum 14 > 8"):
        engine.set_basis(_with_pair_order_too_high(good))
    engine.basis = good
    np.testing.assert_array_equal(engine.get_result(Z.RESULT_ERI), eri)
    again, _, _ = engine.gauss_eri()
    np.testing.assert_array_equal(again, eri)


def test_rejected_basis_on_a_fresh_context_leaves_no_basis():
    from slater_stoch_amd import zvmc as Z
    good = configs.water_minimal(ng0=4)
    with Z.Engine(0) as eng:
        with pytest.raises(Z.ZVError, match="angular momentum"):
            eng.set_basis(_with_pair_order_too_high(good))
        eng.basis = good
        with pytest.raises(Z.ZVError, match="zv_set_basis has not been called"):
            eng.get_result(Z.RESULT_ERI)
        with pytest.raises(Z.ZVError, match="zv_set_basis has not been called"):
            eng.gauss_eri()
