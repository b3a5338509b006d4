"""Host mirror of the library's shell and estimator-unit bookkeeping.

A shell is a run of consecutive AOs with equal centre, Gaussian fit, type
and order (at most 15): the Cartesian components of one STO shell
(csrc/zvmc_api.hip, zv_set_basis).  A shell pair's AO pairs are split
greedily into estimator units of 6, 3 and 1 components (csrc/k_mc.hip,
group_ready); the estimator launches one kernel per pair of unit classes.
"""
from __future__ import annotations

import numpy as np

UNIT_SIZES = (6, 3, 1)


def shell_firsts(bas):
    """First AO (0-based) of every shell, then n."""
    n = bas.n
    lao = bas.npower.sum(1) + 2 * (bas.i_type == 2)
    first = []
    for i in range(n):
        same = (i > 0 and i - first[-1] < 15 and bas.i_type[i] == bas.i_type[i - 1]
                and bas.n_gauss[i] == bas.n_gauss[i - 1] and lao[i] == lao[i - 1]
                and np.array_equal(bas.center[i], bas.center[i - 1])
                and np.array_equal(bas.g_gauss[i], bas.g_gauss[i - 1])
                and np.array_equal(bas.c_gauss[:, i], bas.c_gauss[:, i - 1]))
        if not same:
            first.append(i)
    first.append(n)
    return first


def shell_units(bas):
    """Per shell pair (I, K), I >= K: the first shell I and its units as
    arrays of 0-based pair indices p = i(i+1)/2 + k."""
    first = shell_firsts(bas)
    out = []
    for I in range(len(first) - 1):
        for K in range(I + 1):
            comps = [i * (i + 1) // 2 + k for i in range(first[I], first[I + 1])
                     for k in range(first[K], first[K + 1]) if k <= i]
            units, at = [], 0
            for size in UNIT_SIZES:
                while len(comps) - at >= size:
                    units.append(np.array(comps[at:at + size]))
                    at += size
            out.append((I, units))
    return out


def pair_unit_class(bas):
    """Unit class (0: 1 component, 1: 3, 2: 6) of every AO pair."""
    cls = np.zeros(bas.npair, np.int64)
    code = {1: 0, 3: 1, 6: 2}
    for _, units in shell_units(bas):
        for u in units:
            cls[u] = code[u.size]
    return cls


def kernel_unit_pairs(bas):
    """Unit pairs (row unit U, column unit V) holding at least one quartet
    with row pair P >= column pair Q: the 1/r12 evaluations per sample."""
    sps = shell_units(bas)
    nsh = max(I for I, _ in sps) + 1
    per_I = np.zeros(nsh)
    for I, units in sps:
        per_I[I] += len(units)
    below = np.cumsum(per_I) - per_I           # units with first shell < I
    total = float(np.sum(per_I * below))
    by_I = {}
    for I, units in sps:
        by_I.setdefault(I, []).extend(units)
    for units in by_I.values():                # same first shell: per unit pair
        hi = np.array([u.max() for u in units])
        lo = np.array([u.min() for u in units])
        total += float(np.sum(hi[:, None] >= lo[None, :]))
    return total


def mfma_tile_use(bas, tile=16):
    """How well the estimator's per-shell-pair-pair products would fill
    matrix-core tiles.  For a row shell pair s and a column shell pair t the
    estimator's sum over samples, sum_m K_m(s,t) A_P,m B_Q,m for P in s and
    Q in t, is an |s| x |t| product with the samples as the inner dimension;
    on MFMA it runs as ceil(|s|/tile) x ceil(|t|/tile) tiles of tile x tile.
    Returns (share of the canonical quartets whose shell pairs both have at
    least ``tile`` components, useful fraction of the tile area summed over
    all shell-pair pairs weighted by their quartets)."""
    first = shell_firsts(bas)
    sizes = []
    for I in range(len(first) - 1):
        for K in range(I + 1):
            ni, nk = first[I + 1] - first[I], first[K + 1] - first[K]
            sizes.append(ni * (ni + 1) // 2 if I == K else ni * nk)
    w = np.array(sizes, dtype=np.float64)
    pad = np.ceil(w / tile) * tile
    big = w >= tile
    # all ordered (s, t) pairs: quartets ~ |s||t| (the canonical half is the
    # same share for every class up to the diagonal)
    quart = np.outer(w, w)
    share = float(quart[np.ix_(big, big)].sum() / quart.sum())
    used = float(quart.sum() / np.outer(pad, pad).sum())
    return share, used
