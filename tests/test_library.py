"""CPU-side checks of the HIP library: it loads, exports every entry point of
include/zvmc.h, links into a C program through that header, and its host
logic (quartet indexing) matches the reference."""
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


C_CALLER = r"""
#include <stdio.h>
#include "zvmc.h"
int main(void)
{
    zv_context *ctx = NULL;
    int rc = zv_create(0, &ctx);
    if (rc == ZV_ERR_NODEV) return 0;              /* no gfx950 device here */
    if (rc != 0) return 2;
    char buf[64];
    rc = zv_gauss_eri(ctx, NULL, NULL, NULL);      /* no basis yet: an error */
    zv_last_error(ctx, buf, (int)sizeof buf);
    zv_destroy(ctx);
    return rc == ZV_ERR_STATE ? 0 : 3;
}
"""


def test_header_is_c_and_the_library_links_from_c(tmp_path):
    """include/zvmc.h compiles as C11 and a C program linked against
    libzvmc.so runs: without a device zv_create reports ZV_ERR_NODEV, with
    one a call before zv_set_basis reports ZV_ERR_STATE."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    zvmc.load_library()  # the library is built
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    libdir = ROOT / "slater_stoch_amd"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src),
                    f"-L{libdir}", "-lzvmc", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                   check=True, capture_output=True)
    assert subprocess.run([str(exe)], capture_output=True, timeout=120).returncode == 0
