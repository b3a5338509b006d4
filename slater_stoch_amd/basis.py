"""Host-side basis model: geometry, STO basis file, Gaussian-fit expansion.

Mirrors the reference's input layer so the arrays handed to the C-ABI are
the ones the reference's module ``common_data`` would hold:

* ``read_xyz``          -- src/read_geometry.F90:1-41 (Angstrom -> bohr, E_nuc)
* ``read_basis_file``   -- src/read_basis.F90:27-71 ("ATOM ORB zeta" lines, coef 1)
* ``expand_basis``      -- src/build_gaussian_expansion_of_orbitals.F90:9-471
                           and src/build_c_g_gauss_STO.F90:1-31
* ``quartet_mapping``   -- src/build_mapping_ijkl.F90:9-44
* ``nuclei_class``      -- src/compare_nuclei.F90:1-139

Array layouts follow the Fortran column-major declarations, so
``center`` is (n, 3) C-ordered == Fortran ``center(3, n)`` and
``c_gauss`` is (2, n, ldg) C-ordered == Fortran ``c_gauss(ldg, n, 2)``.
"""
from __future__ import annotations

import dataclasses
from pathlib import Path

import numpy as np

ANG_TO_BOHR = 1.0 / 0.52917721092  # src/read_geometry.F90:5

# src/number_atom.F90:4-8
ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P",
            "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
            "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"]

_P = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
_D = [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]
_F = [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (1, 0, 2), (0, 3, 0), (0, 2, 1),
      (0, 1, 2), (0, 0, 3)]
_G = [(4, 0, 0), (3, 1, 0), (2, 2, 0), (1, 3, 0), (0, 4, 0), (3, 0, 1), (2, 1, 1), (1, 2, 1),
      (0, 3, 1), (2, 0, 2), (1, 1, 2), (0, 2, 2), (1, 0, 3), (0, 1, 3), (0, 0, 4)]

# name -> (n_STO, i_type, Cartesian components, which ng increment applies)
# src/build_gaussian_expansion_of_orbitals.F90:20-466
SHELLS = {
    "1S": (0, 1, [(0, 0, 0)], "s"),
    "2S": (1, 2, [(0, 0, 0)], "s"),
    "3S": (2, 2, [(0, 0, 0)], "s"),
    "2P": (0, 1, _P, "p"),
    "3P": (1, 2, _P, "p"),
    "3D": (0, 1, _D, "d"),
    "4F": (0, 1, _F, "d"),
    "5G": (0, 1, _G, "d"),
}

_FIT_CACHE = None


def fit_tables():
    """(c_fit, g_fit) as dicts ng -> array(ng), from data/sto_fit_tables.txt."""
    global _FIT_CACHE
    if _FIT_CACHE is None:
        c_fit, g_fit = {}, {}
        path = Path(__file__).with_name("data") / "sto_fit_tables.txt"
        for line in path.read_text().splitlines():
            if line.startswith("#") or not line.strip():
                continue
            ng, i, c, g = line.split()
            ng, i = int(ng), int(i)
            c_fit.setdefault(ng, [0.0] * ng)[i - 1] = float(c)
            g_fit.setdefault(ng, [0.0] * ng)[i - 1] = float(g)
        _FIT_CACHE = ({k: np.array(v) for k, v in c_fit.items()},
                      {k: np.array(v) for k, v in g_fit.items()})
    return _FIT_CACHE


class InputError(ValueError):
    """Raised where the reference would ``stop``."""


def number_atom(label: str) -> int:
    """src/number_atom.F90: 1-based atomic number of a 2-character label."""
    lab = label.strip()
    lab = lab[:1].upper() + lab[1:2].lower()
    if lab not in ELEMENTS:
        raise InputError(f"Atom not defined: :{label.strip()}:")
    return ELEMENTS.index(lab) + 1


@dataclasses.dataclass
class Geometry:
    labels: list
    coords: np.ndarray  # (natom, 3) bohr
    enucl: float


def geometry_from_angstrom(labels, coords_ang) -> Geometry:
    """Angstrom -> bohr and nuclear repulsion, src/read_geometry.F90:20-40."""
    coords = np.array([[ANG_TO_BOHR * float(v) for v in row] for row in coords_ang], dtype=np.float64)
    charge = [float(number_atom(l)) for l in labels]
    en = 0.0
    for i in range(len(labels)):
        for j in range(i + 1, len(labels)):
            d = coords[i] - coords[j]
            rij = np.sqrt((d[0] ** 2 + d[1] ** 2) + d[2] ** 2)
            en = en + charge[i] * charge[j] / rij
    return Geometry(list(labels), coords, float(en))


def read_xyz(path) -> Geometry:
    """src/read_geometry.F90:1-41 (first line count, second line comment)."""
    lines = Path(path).read_text().splitlines()
    nat = int(lines[0].split()[0])
    labels, coords = [], []
    for line in lines[2:2 + nat]:
        tok = line.split()
        labels.append(tok[0][:2])
        coords.append([float(t.replace("d", "e").replace("D", "E")) for t in tok[1:4]])
    return geometry_from_angstrom(labels, coords)


def read_basis_file(path) -> dict:
    """src/read_basis.F90:27-37: element -> list of (orbital, zeta), file order."""
    basis = {}
    for line in Path(path).read_text().splitlines():
        tok = line.split()
        if len(tok) < 3:
            continue
        el = ELEMENTS[number_atom(tok[0]) - 1]
        basis.setdefault(el, []).append((tok[1].upper(), float(tok[2])))
    return basis


@dataclasses.dataclass
class Basis:
    n: int
    center: np.ndarray    # (n,3) f64
    npower: np.ndarray    # (n,3) i32
    n_sto: np.ndarray     # (n,) i32
    i_type: np.ndarray    # (n,) i32
    zeta: np.ndarray      # (n,) f64  (g_slater)
    n_gauss: np.ndarray   # (n,) i32
    c_gauss: np.ndarray   # (2,n,ldg) f64
    g_gauss: np.ndarray   # (n,ldg) f64
    nucleus: np.ndarray   # (n,) i32, 1-based
    names: list
    a_zv: float
    geometry: Geometry = None

    @property
    def ldg(self) -> int:
        return self.g_gauss.shape[1]

    @property
    def npair(self) -> int:
        return self.n * (self.n + 1) // 2

    @property
    def nint(self) -> int:
        np_ = self.npair
        return np_ * (np_ + 1) // 2


def expand_basis(geom: Geometry, basis: dict, ng0=20, i_add_2p=0, i_add_3d=0, g_thr=100.0,
                 i_add_large_g=0, a_zv=1.8) -> Basis:
    """Shells -> Cartesian AOs with their Gaussian fits.

    src/build_gaussian_expansion_of_orbitals.F90:11-471 (order: nuclei in
    file order, basis lines of the element in file order, components in the
    order listed in SHELLS) and src/build_c_g_gauss_STO.F90:5-29:
      type 1: c1 = 1*c_fit,            g = (zeta*zeta)*g_fit
      type 2: c2 = 1*2*c_fit*g_fit*zeta, g = (zeta*zeta)*g_fit
    """
    c_fit, g_fit = fit_tables()
    rows = []
    for ia, lab in enumerate(geom.labels):
        el = ELEMENTS[number_atom(lab) - 1]
        for orb, zeta in basis.get(el, []):
            if orb not in SHELLS:
                raise InputError(f"orbital type {orb} is not supported")
            n_sto, itype, comps, kind = SHELLS[orb]
            ng = ng0 + {"s": 0, "p": i_add_2p, "d": i_add_3d}[kind]
            if zeta > g_thr:
                ng += i_add_large_g
            if ng not in c_fit:
                raise InputError(f"no Gaussian fit with {ng} functions")
            for comp in comps:
                rows.append((ia, orb, comp, n_sto, itype, zeta, ng))
    n = len(rows)
    if n == 0:
        raise InputError("empty basis")
    ldg = max(r[6] for r in rows)
    center = np.zeros((n, 3))
    npower = np.zeros((n, 3), np.int32)
    n_sto_a = np.zeros(n, np.int32)
    itype_a = np.zeros(n, np.int32)
    zeta_a = np.zeros(n)
    ngs = np.zeros(n, np.int32)
    c_gauss = np.zeros((2, n, ldg))
    g_gauss = np.zeros((n, ldg))
    nucleus = np.zeros(n, np.int32)
    names = []
    for b, (ia, orb, comp, n_sto, itype, zeta, ng) in enumerate(rows):
        center[b] = geom.coords[ia]
        npower[b] = comp
        n_sto_a[b] = n_sto
        itype_a[b] = itype
        zeta_a[b] = zeta
        ngs[b] = ng
        nucleus[b] = ia + 1
        names.append(orb)
        cf, gf = c_fit[ng], g_fit[ng]
        for m in range(ng):
            g_gauss[b, m] = (zeta * zeta) * gf[m]
            if itype == 1:
                c_gauss[0, b, m] = 1.0 * cf[m]
            else:
                c_gauss[1, b, m] = (((1.0 * 2.0) * cf[m]) * gf[m]) * zeta
    return Basis(n, center, npower, n_sto_a, itype_a, zeta_a, ngs, c_gauss, g_gauss, nucleus,
                 names, float(a_zv), geom)


def quartet_mapping(n: int):
    """Unique (ik|jl) in the reference order (src/build_mapping_ijkl.F90:16-43):
    l, k, j (>= l), i (>= k) nested, keeping i >= j and not (i == j and l > k).
    Returns 1-based int32 arrays is, js, ks, ls."""
    out = []
    kk, jj, ii = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), np.arange(1, n + 1),
                             indexing="ij")
    for l in range(1, n + 1):
        mask = (jj >= l) & (ii >= kk) & (ii >= jj) & ~((ii == jj) & (l > kk))
        k_, j_, i_ = kk[mask], jj[mask], ii[mask]
        out.append((i_, j_, k_, np.full(i_.shape, l)))
    is_ = np.concatenate([o[0] for o in out]).astype(np.int32)
    js_ = np.concatenate([o[1] for o in out]).astype(np.int32)
    ks_ = np.concatenate([o[2] for o in out]).astype(np.int32)
    ls_ = np.concatenate([o[3] for o in out]).astype(np.int32)
    np_ = n * (n + 1) // 2
    assert is_.size == np_ * (np_ + 1) // 2
    return is_, js_, ks_, ls_


def nuclei_class(bas: Basis, is_, js_, ks_, ls_) -> np.ndarray:
    """Number of distinct centres of each quartet (1..4), the classification
    of src/compare_nuclei.F90 used for error binning."""
    nuc = bas.nucleus
    v = np.stack([nuc[is_ - 1], nuc[js_ - 1], nuc[ks_ - 1], nuc[ls_ - 1]], axis=1)
    v.sort(axis=1)
    return (1 + (v[:, 1] != v[:, 0]) + (v[:, 2] != v[:, 1]) + (v[:, 3] != v[:, 2])).astype(np.int32)
