"""GPU parity of the group estimator instances, split and unsplit.

The group estimator (k_est_group<T, NR, NC, TR, TC, SPLIT, BF>) has two
instances per class pair: SPLIT = false when the tiles alone fill the device
(large molecules: these are the ones timed on the 110-AO and 200-AO
workloads) and SPLIT = true when each tile's samples are spread over several
workgroups (small molecules).  The library picks the split; ZV_TUNE_EST_SPLIT
forces it, so both instances of all nine class-pair kernels are compared
here with an exact (extended-precision) evaluation of the reference's estimator
(src/two_e_slater.F90:422-454) on the same fixed samples.

Bounds per quartet, tol * sum_m |t_m| / M (tests/estimator_ref.py): 1e-12
for FP64 tiles, 2e-5 for FP32 tiles (FP32 coordinates, hardware rsqrt),
8e-3 of the split-product scale for the BF16 products.  The single-precision
forms also get an absolute floor of 1e-20 Ha (1e-14 of the 1e-6 Ha target
error): on large molecules the fields of distant pairs are so small that
their products underflow FP32 (normal range ends at 1.2e-38), so quartets
whose whole correction is of that size come out as 0 or with few digits.
"""
import numpy as np
import pytest

from estimator_ref import fp64_estimate
from slater_stoch_amd import basis as B
from slater_stoch_amd import configs
from slater_stoch_amd.shells import pair_unit_class

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-5, "bf16": 8e-3}
FLOOR = {"fp64": 0.0, "fp32": 1e-20, "bf16": 1e-20}


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


def _samples(oracle, m, seed):
    rng = np.random.default_rng(seed)
    r1, r2 = rng.standard_normal((m, 3)), rng.standard_normal((m, 3))
    return r1, r2, oracle.pi0(r1, r2)


def _check(got, ref, scale, prec, what):
    bad = np.abs(got - ref) > TOL[prec] * scale + FLOOR[prec] + 1e-300
    assert not bad.any(), (what, int(bad.sum()),
                           float(np.max(np.abs(got - ref) / (scale + 1e-300))),
                           float(np.max(np.abs(ref[bad]))), float(np.max(scale[bad])))


def stratified_quartets(bas, per_class, seed):
    """Reference positions of up to per_class random quartets for each of the
    nine pairs of estimator unit classes (row = electron-1 pair P, column =
    electron-2 pair Q), so that every class-pair kernel is exercised."""
    is_, js, ks, ls = B.quartet_mapping(bas.n)
    P = (is_ * (is_ - 1)) // 2 + ks - 1
    Q = (js * (js - 1)) // 2 + ls - 1
    cls = pair_unit_class(bas)
    key = 3 * cls[P] + cls[Q]
    rng = np.random.default_rng(seed)
    out = []
    for k in range(9):
        idx = np.flatnonzero(key == k)
        if idx.size:
            out.append(rng.choice(idx, min(per_class, idx.size), replace=False))
    return np.sort(np.concatenate(out)), len(out)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("split", [1, 4])
@pytest.mark.parametrize("tiles", [0, 1])
def test_split_and_unsplit_instances_small_bases(engine, oracle, prec, split, tiles):
    """Both tile-shape sets (ZV_TUNE_EST_TILES), each split and unsplit."""
    from slater_stoch_amd.zvmc import TUNE_EST_SPLIT, TUNE_EST_TILES
    try:
        engine.set_tuning(TUNE_EST_SPLIT, split)
        engine.set_tuning(TUNE_EST_TILES, tiles)
        for bas, nbl, npts in ((configs.water_minimal(ng0=4), 4, 400), (_mixed_basis(), 2, 40)):
            engine.set_basis(bas)
            oracle.set_basis(bas)
            r1, r2, pi0 = _samples(oracle, nbl * npts, 3)
            ref, scale, scale2 = fp64_estimate(oracle, bas, r1, r2, pi0, nbl, npts, split=True)
            got, _ = engine.mc_blocks(nbl, npts, prec, samples=(r1, r2, pi0))
            _check(got, ref, scale2 if prec == "bf16" else scale, prec,
                   (prec, split, tiles, bas.n))
    finally:
        engine.set_tuning(TUNE_EST_SPLIT, 0)
        engine.set_tuning(TUNE_EST_TILES, 0)


@pytest.fixture(scope="module")
def c2_fixed(oracle):
    """The 110-AO molecule, 10 blocks of 40 fixed samples and the
    extended-precision estimator on 20,000 quartets of each class pair."""
    bas = configs.random_molecule()
    oracle.set_basis(bas)
    nbl, npts = 10, 40
    r1, r2, pi0 = _samples(oracle, nbl * npts, 17)
    pick, nclass = stratified_quartets(bas, 20000, 5)
    assert nclass == 9
    ref, scale = fp64_estimate(oracle, bas, r1, r2, pi0, nbl, npts, pick=pick)
    return bas, nbl, npts, (r1, r2, pi0), pick, ref, scale


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_timed_estimator_on_110_ao_molecule(engine, c2_fixed, prec):
    """The instances bench.py times: on the 110-AO molecule the library runs
    the estimator unsplit (its automatic result equals the forced unsplit one
    bit for bit), and 10 blocks of 40 fixed samples agree with the exact
    evaluation on 20,000 quartets of each of the nine class pairs."""
    from slater_stoch_amd.zvmc import TUNE_EST_SPLIT
    bas, nbl, npts, (r1, r2, pi0), pick, ref, scale = c2_fixed
    engine.set_basis(bas)
    auto, auto_err = engine.mc_blocks(nbl, npts, prec, samples=(r1, r2, pi0))
    try:
        engine.set_tuning(TUNE_EST_SPLIT, 1)
        one, one_err = engine.mc_blocks(nbl, npts, prec, samples=(r1, r2, pi0))
    finally:
        engine.set_tuning(TUNE_EST_SPLIT, 0)
    assert np.array_equal(auto, one) and np.array_equal(auto_err, one_err)
    _check(auto[pick], ref, scale, prec, (prec, "110 AO"))
