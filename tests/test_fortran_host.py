"""The Fortran host (two_e_slater_hip): the reference's input deck in, the
reference's bielec_ao format out, through the iso_c_binding C-ABI."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import golden
from slater_stoch_amd import configs

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "slater_stoch_amd" / "two_e_slater_hip"


def _run(tmp_path, npts=400, env=None, ng0=20):
    if not EXE.exists():
        pytest.skip("two_e_slater_hip not built (no Fortran compiler here)")
    bas = configs.water_minimal(ng0=ng0)
    configs.write_deck(bas, tmp_path, name="h2o", npts=npts, ng0=ng0)
    e = dict(os.environ)
    e.update(env or {})
    return bas, subprocess.run([str(EXE)], cwd=tmp_path, env=e, capture_output=True, text=True,
                               timeout=600)


def test_deck_is_parsed_and_missing_device_stops(tmp_path):
    """Host part of the driver on the CPU: ENUCL and the basis size as the
    reference prints them, then a loud stop when no gfx950 device exists."""
    import ctypes
    from slater_stoch_amd import zvmc
    h = ctypes.c_void_p()
    if zvmc.load_library().zv_create(0, ctypes.byref(h)) == 0:
        zvmc.load_library().zv_destroy(h)
        pytest.skip("a device is present")
    bas, p = _run(tmp_path)
    assert p.returncode != 0
    ref = golden("water_ng20_basis.npz")
    enucl = float(p.stdout.split("ENUCL=")[1].split()[0])
    assert enucl == pytest.approx(float(ref["enucl"]), rel=1e-14)
    assert "Number of basis functions  7" in p.stdout
    assert "zv_create failed" in p.stdout


@pytest.mark.gpu
def test_bielec_ao_matches_the_c_abi(tmp_path):
    bas, p = _run(tmp_path, env={"ZVMC_PRECISION": "fp64_parity", "ZVMC_SVD_CLEAN": "0"})
    assert p.returncode == 0, p.stdout + p.stderr
    lines = (tmp_path / "bielec_ao").read_text().splitlines()
    assert len(lines) == bas.nint
    # format of src/two_e_slater.F90:611: '(4(I5,X),2D22.15)'
    assert all(len(ln) == 4 * 6 + 2 * 22 for ln in lines)
    idx = np.array([[int(ln[6 * t:6 * t + 5]) for t in range(4)] for ln in lines])
    val = np.array([float(ln[24:46].replace("D", "E")) for ln in lines])
    err = np.array([float(ln[46:68].replace("D", "E")) for ln in lines])
    g = golden("water_ng20_eri.npz")
    assert np.array_equal(idx, np.stack([g["is_"], g["js"], g["ks"], g["ls"]], axis=1))
    from slater_stoch_amd.zvmc import Engine
    with Engine(0) as eng:
        eng.set_basis(bas)
        eri, _, _ = eng.gauss_eri()
        mean, stderr = eng.mc_blocks(10, 400, "fp64_parity", seed=1, stream=0)
    np.testing.assert_allclose(val, eri + mean, rtol=1e-14, atol=1e-300)  # 15 printed digits
    np.testing.assert_allclose(err, stderr, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(val, g["eri"] + mean, rtol=1e-10, atol=1e-14)


@pytest.mark.gpu
def test_driver_with_cleanup_runs(tmp_path):
    bas, p = _run(tmp_path, npts=80, ng0=4)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Removed" in p.stdout
    assert len((tmp_path / "bielec_ao").read_text().splitlines()) == bas.nint


@pytest.mark.gpu
def test_exact_monocenter_branch(tmp_path):
    """ZVMC_EXACT_MONOCENTER=1 (the reference's do_exact_monocenter): the
    one-centre quartets carry the exact Slater integral and error 0, the
    others the Gaussian-fit ERI plus the Monte Carlo mean as before."""
    bas, p = _run(tmp_path, env={"ZVMC_PRECISION": "fp64_parity", "ZVMC_SVD_CLEAN": "0",
                                 "ZVMC_EXACT_MONOCENTER": "1"})
    assert p.returncode == 0, p.stdout + p.stderr
    assert int(p.stdout.split("One-center integrals computed exactly:")[1].split()[0]) == 122
    lines = (tmp_path / "bielec_ao").read_text().splitlines()
    val = np.array([float(ln[24:46].replace("D", "E")) for ln in lines])
    err = np.array([float(ln[46:68].replace("D", "E")) for ln in lines])
    from slater_stoch_amd.zvmc import Engine
    with Engine(0) as eng:
        eng.set_basis(bas)
        eri, _, _ = eng.gauss_eri()
        mean, stderr = eng.mc_blocks(10, 400, "fp64_parity", seed=1, stream=0)
        want, werr = eri + mean, stderr.copy()
        nm = eng.exact_monocenter(want, werr)
    assert nm == 122 and (err == 0).sum() >= 122
    np.testing.assert_allclose(val, want, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(err, werr, rtol=1e-14, atol=1e-300)


@pytest.mark.gpu
def test_one_e_files(tmp_path):
    """ZVMC_ONE_E=1 writes the reference's one_e_slater files
    (src/one_e_slater.F90:123-136) with the C-ABI's S, V, K."""
    bas, p = _run(tmp_path, npts=80, env={"ZVMC_SVD_CLEAN": "0", "ZVMC_ONE_E": "1"})
    assert p.returncode == 0, p.stdout + p.stderr
    from slater_stoch_amd.zvmc import Engine
    with Engine(0) as eng:
        eng.set_basis(bas)
        want = eng.one_e()
    for name, M in zip(("overlap_ao_ex", "nuclear_ao_ex", "kinetic_ao_ex"), want):
        lines = (tmp_path / name).read_text().splitlines()
        assert len(lines) == bas.n ** 2 and all(len(ln) == 2 * 6 + 22 for ln in lines)
        ik = np.array([[int(ln[0:5]), int(ln[6:11])] for ln in lines])
        val = np.array([float(ln[12:34].replace("D", "E")) for ln in lines])
        assert np.array_equal(ik[:, 0], np.repeat(np.arange(1, bas.n + 1), bas.n))
        np.testing.assert_allclose(val, M[ik[:, 0] - 1, ik[:, 1] - 1], rtol=1e-14, atol=1e-300)
