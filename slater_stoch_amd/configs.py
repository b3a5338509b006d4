"""Seeded synthetic workloads for the ZVMC ERI engine.

* ``water_minimal``   -- water, minimal STO basis (7 AOs, 406 quartets), ng0 = 20
* ``random_molecule`` -- random atoms in a box with min separation; C/N/O get
  1S,2S,2S,2P,2P,3D and H gets 1S,1S,2P (double zeta + polarisation);
  zeta ~ U[0.8, 8] per (element, shell) because the reference basis file is
  per element (src/read_basis.F90:28-36).  12 atoms / 110 AOs by default.
* ``large_ch``        -- 40 C/H atoms in a 10 Angstrom box, 200 AOs up to l = 2,
  with the C and H exponents of the reference's example basis
* ``batch``           -- 64 random molecules of uneven size

Every generator returns a ``Basis`` (see basis.py) and can write the
reference's input deck (``j_in``, ``<mol>.xyz``, basis file) with
``write_deck`` so the same case runs through the Fortran host.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .basis import Basis, expand_basis, geometry_from_angstrom

HEAVY_SHELLS = ["1S", "2S", "2S", "2P", "2P", "3D"]
H_SHELLS = ["1S", "1S", "2P"]


def water_minimal(ng0=20, a_zv=1.8) -> Basis:
    """Water: O (0,0,0.1173) A, H (0,+-0.7572,-0.4692) A; H 1S 1.24, O 1S 7.66,
    O 2S 2.25, O 2P 2.25; ng0 Gaussians per STO; a_ZV 1.8."""
    geom = geometry_from_angstrom(["O", "H", "H"],
                                  [[0.0, 0.0, 0.1173], [0.0, 0.7572, -0.4692],
                                   [0.0, -0.7572, -0.4692]])
    basis = {"H": [("1S", 1.24)], "O": [("1S", 7.66), ("2S", 2.25), ("2P", 2.25)]}
    return expand_basis(geom, basis, ng0=ng0, a_zv=a_zv)


def _place_atoms(rng, natom, box, dmin):
    pts = []
    while len(pts) < natom:
        p = rng.uniform(0.0, box, 3)
        if all(np.linalg.norm(p - q) >= dmin for q in pts):
            pts.append(p)
    return np.array(pts)


def random_molecule(seed=2024, n_heavy=5, n_h=7, box=6.0, dmin=0.9, zmin=0.8, zmax=8.0,
                    heavy=("C", "N", "O"), heavy_shells=None, h_shells=None, ng0=20,
                    a_zv=1.8) -> Basis:
    """Seeded random molecule; defaults give 12 atoms and 110 AOs."""
    rng = np.random.default_rng(seed)
    heavy_shells = HEAVY_SHELLS if heavy_shells is None else heavy_shells
    h_shells = H_SHELLS if h_shells is None else h_shells
    labels = list(rng.choice(list(heavy), n_heavy)) + ["H"] * n_h
    order = rng.permutation(len(labels))
    labels = [labels[t] for t in order]
    coords = _place_atoms(rng, len(labels), box, dmin)
    basis = {}
    for el in sorted(set(labels)):
        shells = h_shells if el == "H" else heavy_shells
        basis[el] = [(s, float(rng.uniform(zmin, zmax))) for s in shells]
    geom = geometry_from_angstrom(labels, coords)
    return expand_basis(geom, basis, ng0=ng0, a_zv=a_zv)


# C and H shells of the reference's own example basis (example/VB1:2,4,40,42-43,45,48):
# the valence-range exponents a production basis uses, so that the 200-AO
# workload is screened and sampled like a real molecule (compact exponents
# drawn at random would let the Schwarz test drop almost every quartet).
LARGE_CH_BASIS = {"C": [("1S", 5.206173), ("2S", 2.057774), ("2S", 1.329176), ("2P", 1.748226),
                        ("3D", 1.953933)],
                  "H": [("1S", 1.165255), ("2P", 1.481937)]}


def large_ch(seed=40, ng0=20, a_zv=1.8, n_c=5, n_h=35, box=10.0, dmin=0.9) -> Basis:
    """40 atoms (5 C, 35 H) at seeded positions in a 10 A box (min separation
    0.9 A): C 1S,2S,2S,2P,3D (12 AOs), H 1S,2P (4 AOs) -> 200 AOs, 20100
    pairs, with the exponents of ``LARGE_CH_BASIS``."""
    rng = np.random.default_rng(seed)
    labels = ["C"] * n_c + ["H"] * n_h
    labels = [labels[t] for t in rng.permutation(len(labels))]
    coords = _place_atoms(rng, len(labels), box, dmin)
    geom = geometry_from_angstrom(labels, coords)
    return expand_basis(geom, LARGE_CH_BASIS, ng0=ng0, a_zv=a_zv)


def batch(seed=64, count=64, ng0=20, a_zv=1.8, max_ao=150):
    """``count`` molecules of 2-20 atoms with seeded, uneven sizes, built as
    ``random_molecule`` (heavy atoms 15 AOs, H 5 AOs).  The number of heavy
    atoms is capped so that no molecule exceeds ``max_ao`` orbitals
    (10-150 AOs for the defaults)."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        natom = int(rng.integers(2, 21))
        max_heavy = max(0, min(natom, (max_ao - 5 * natom) // 10))
        nh = natom - int(rng.integers(0, max_heavy + 1))
        out.append(random_molecule(seed=int(rng.integers(1 << 30)), n_heavy=natom - nh, n_h=nh,
                                   box=3.0 + 0.35 * natom, ng0=ng0, a_zv=a_zv))
    return out


def write_deck(bas: Basis, directory, name="mol", npts=4000, num_simulation=1, ng0=20,
               i_add_2p=0, i_add_3d=0, g_thr=100.0, i_add_large_g=0):
    """Write j_in, <name>.xyz and <name>.basis in the reference's formats
    (example/j_in:1-13, src/read_geometry.F90:13-22, src/read_basis.F90:28-37)."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    geom = bas.geometry
    with open(d / f"{name}.xyz", "w") as f:
        f.write(f"{len(geom.labels)}\n{name}\n")
        for lab, c in zip(geom.labels, geom.coords):
            a = c / (1.0 / 0.52917721092)
            f.write(f"{lab:2s} {float(a[0])!r} {float(a[1])!r} {float(a[2])!r}\n")
    seen = {}
    for b in range(bas.n):
        el = geom.labels[bas.nucleus[b] - 1]
        key = (el, bas.names[b], float(bas.zeta[b]))
        seen.setdefault(el, [])
        if key not in seen[el]:
            seen[el].append(key)
    with open(d / f"{name}.basis", "w") as f:
        for el, rows in seen.items():
            for _, orb, z in rows:
                f.write(f"{el:2s} {orb} {z!r}\n")
    with open(d / "j_in", "w") as f:
        f.write("Simulation number\n")
        f.write(f"{num_simulation}\n")
        f.write("Molecule\n")
        f.write(f"{name}\n")
        f.write("Name of basis filename\n")
        f.write(f"{name}.basis\n")
        f.write("Ng for gaussian approximation of STO (and i_add_2p,i_add_3d,g_thr,i_add_large_g)\n")
        f.write(f" {ng0} {i_add_2p} {i_add_3d} {g_thr} {i_add_large_g}\n")
        f.write("ZVMC for TWO-electron integrals. Number of MC steps?\n")
        f.write(f" {npts}\n")
        f.write("************************************************************\n")
        f.write("ZV PARAMETER a coordinate transformation?\n")
        f.write(f"{bas.a_zv}\n")
    return d
