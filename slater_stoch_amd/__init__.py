"""MI355X-native ZVMC two-electron-integral engine (slater_stoch hot path).

The compute path is the HIP library ``libzvmc.so`` (built from ``csrc/``)
behind the C-ABI declared in ``include/zvmc.h``; ``zvmc`` is its ctypes
binding, ``basis``/``configs`` the host-side input layer.
"""
