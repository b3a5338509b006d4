"""CPU-side checks of the HIP library: it loads, exports every entry point of
include/zvmc.h, and its host logic (quartet indexing) matches the reference."""
import re
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import golden
from slater_stoch_amd import basis as B
from slater_stoch_amd import zvmc

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "zvmc.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t)\s+(zv_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert _declared() == sorted(zvmc.SIGNATURES)


def test_library_loads_and_exports_every_symbol():
    lib = zvmc.load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_library_mapping_matches_reference():
    ref = golden("water_ng20_eri.npz")
    q = zvmc.build_mapping(7)
    for a, b in zip(q, [ref["is_"], ref["js"], ref["ks"], ref["ls"]]):
        assert np.array_equal(a, b)
    for n in (1, 2, 5, 19):
        for a, b in zip(zvmc.build_mapping(n), B.quartet_mapping(n)):
            assert np.array_equal(a, b)
    assert zvmc.load_library().zv_nint(200) == 202015050


def test_no_device_is_reported_not_faked():
    """Without a gfx950 device the engine refuses to start (no CPU fallback)."""
    import ctypes
    lib = zvmc.load_library()
    h = ctypes.c_void_p()
    rc = lib.zv_create(1000, ctypes.byref(h))
    assert rc == zvmc.ZV_ERR_NODEV


def test_missing_library_raises(tmp_path):
    with pytest.raises(zvmc.ZVError):
        zvmc.load_library(tmp_path / "libzvmc.so")
