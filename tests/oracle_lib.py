"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP library against it and it
against the golden vectors written from the compiled reference.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "liboracle.so"
P = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
VP = ctypes.c_void_p


def _load():
    if not LIB.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    lib = ctypes.CDLL(str(LIB))
    lib.zvo_set_basis.restype = ctypes.c_int
    lib.zvo_set_basis.argtypes = [ctypes.c_int] + [VP] * 6 + [ctypes.c_int, VP, VP, ctypes.c_double]
    lib.zvo_exp.restype = ctypes.c_double
    lib.zvo_exp.argtypes = [ctypes.c_double]
    for f in ("zvo_u_orb", "zvo_u_gauss"):
        getattr(lib, f).restype = ctypes.c_double
        getattr(lib, f).argtypes = [ctypes.c_int, ctypes.c_double]
    lib.zvo_pair_samples.argtypes = [ctypes.c_int] * 3 + [VP] * 10
    lib.zvo_pi0.argtypes = [ctypes.c_int, VP, VP, VP]
    lib.zvo_nint.restype = ctypes.c_int64
    lib.zvo_nint.argtypes = [ctypes.c_int]
    lib.zvo_build_mapping.restype = ctypes.c_int64
    lib.zvo_build_mapping.argtypes = [ctypes.c_int, VP, VP, VP, VP]
    lib.zvo_compare_nuclei.restype = ctypes.c_int
    lib.zvo_compare_nuclei.argtypes = [ctypes.c_int] * 4
    lib.zvo_mc_fixed.restype = ctypes.c_int
    lib.zvo_mc_fixed.argtypes = [ctypes.c_int, ctypes.c_int, VP, VP, VP, ctypes.c_int64] + [VP] * 8
    lib.zvo_gauss_ijkl.restype = ctypes.c_double
    lib.zvo_gauss_ijkl.argtypes = [ctypes.c_int] * 4
    lib.zvo_precond.argtypes = [VP]
    lib.zvo_gauss_eri.argtypes = [ctypes.c_int64, VP, VP, VP, VP, VP, VP, VP]
    lib.zvo_rint.restype = ctypes.c_double
    lib.zvo_rint.argtypes = [ctypes.c_int, ctypes.c_double]
    lib.zvo_rint_sum.restype = ctypes.c_double
    lib.zvo_rint_sum.argtypes = [ctypes.c_int, ctypes.c_double, VP]
    lib.zvo_philox4x32_10.argtypes = [VP, VP, VP]
    lib.zvo_mono_exact.restype = ctypes.c_double
    lib.zvo_mono_exact.argtypes = [VP] * 5
    lib.zvo_gaunt.restype = ctypes.c_double
    lib.zvo_gaunt.argtypes = [ctypes.c_int] * 6
    lib.zvo_ylm_coef.argtypes = [ctypes.c_int] * 5 + [VP, VP]
    lib.zvo_one_e.argtypes = [ctypes.c_int] + [VP] * 5
    return lib


class Oracle:
    def __init__(self):
        self.lib = _load()
        self.basis = None

    def set_basis(self, bas):
        self.basis = bas
        rc = self.lib.zvo_set_basis(bas.n, P(bas.center), P(bas.npower), P(bas.n_sto),
                                    P(bas.i_type), P(bas.zeta), P(bas.n_gauss), bas.ldg,
                                    P(bas.c_gauss), P(bas.g_gauss), bas.a_zv)
        assert rc == 0

    def gauss_ijkl(self, i, k, j, l):
        return self.lib.zvo_gauss_ijkl(int(i), int(k), int(j), int(l))

    def pair_samples(self, i, k, r1, r2):
        ns = r1.shape[0]
        r1 = np.ascontiguousarray(r1, np.float64)
        r2 = np.ascontiguousarray(r2, np.float64)
        out = {name: np.zeros(ns) for name in ("rho1", "rho2", "rhog1", "rhog2", "jac1", "jac2")}
        ut1, ut2 = np.zeros((ns, 3)), np.zeros((ns, 3))
        self.lib.zvo_pair_samples(int(i), int(k), ns, P(r1), P(r2), P(ut1), P(ut2),
                                  P(out["rho1"]), P(out["rho2"]), P(out["rhog1"]),
                                  P(out["rhog2"]), P(out["jac1"]), P(out["jac2"]))
        out["ut1"], out["ut2"] = ut1, ut2
        return out

    def pi0(self, r1, r2):
        ns = r1.shape[0]
        out = np.zeros(ns)
        self.lib.zvo_pi0(ns, P(np.ascontiguousarray(r1)), P(np.ascontiguousarray(r2)), P(out))
        return out

    def mapping(self, n):
        nint = self.lib.zvo_nint(n)
        q = [np.zeros(nint, np.int32) for _ in range(4)]
        got = self.lib.zvo_build_mapping(n, *[P(a) for a in q])
        assert got == nint
        return q

    def mc_fixed(self, nbl, npts, r1, r2, pi0, quartets, blocks=False):
        is_, js_, ks_, ls_ = [np.ascontiguousarray(a, np.int32) for a in quartets]
        nq = is_.size
        mean, err = np.zeros(nq), np.zeros(nq)
        es = eg = None
        if blocks:
            es, eg = np.zeros((nbl, nq)), np.zeros((nbl, nq))
        rc = self.lib.zvo_mc_fixed(nbl, npts, P(np.ascontiguousarray(r1)),
                                   P(np.ascontiguousarray(r2)), P(np.ascontiguousarray(pi0)), nq,
                                   P(is_), P(js_), P(ks_), P(ls_), P(mean), P(err), P(es), P(eg))
        assert rc == 0
        return (mean, err, es, eg) if blocks else (mean, err)

    def precond(self):
        n = self.basis.n
        pre = np.zeros((n, n))
        self.lib.zvo_precond(P(pre))
        return pre

    def gauss_eri(self, quartets, precond):
        is_, js_, ks_, ls_ = [np.ascontiguousarray(a, np.int32) for a in quartets]
        out = np.zeros(is_.size)
        nscr = np.zeros(1, np.int64)
        self.lib.zvo_gauss_eri(is_.size, P(is_), P(js_), P(ks_), P(ls_), P(precond), P(out),
                               P(nscr))
        return out, int(nscr[0])

    def rint(self, n, rho):
        return self.lib.zvo_rint(int(n), float(rho))

    def mono_exact(self, i, k, j, l):
        """Exact one-centre Slater integral (ik|jl) of the loaded basis
        (1-based AO indices; the AOs are assumed to share one centre)."""
        b = self.basis
        idx = [int(i) - 1, int(k) - 1, int(j) - 1, int(l) - 1]
        nx = np.array([b.npower[a, 0] for a in idx], np.int32)
        ny = np.array([b.npower[a, 1] for a in idx], np.int32)
        nz = np.array([b.npower[a, 2] for a in idx], np.int32)
        ns = np.array([b.n_sto[a] for a in idx], np.int32)
        z = np.array([b.zeta[a] for a in idx], np.float64)
        return self.lib.zvo_mono_exact(P(nx), P(ny), P(nz), P(ns), P(z))

    def one_e(self):
        """(S, V, K) over the Gaussian fits of the loaded basis, nuclei of its
        geometry (charge = atomic number)."""
        from slater_stoch_amd import basis as B
        geo = self.basis.geometry
        q = np.array([float(B.number_atom(lab)) for lab in geo.labels])
        c = np.ascontiguousarray(np.asarray(geo.coords, np.float64))
        n = self.basis.n
        S, V, K = (np.zeros((n, n), order="F") for _ in range(3))
        self.lib.zvo_one_e(q.size, P(q), P(c), P(S), P(V), P(K))
        return S, V, K

    def philox(self, ctr, key):
        c = np.array(ctr, np.uint32)
        k = np.array(key, np.uint32)
        o = np.zeros(4, np.uint32)
        self.lib.zvo_philox4x32_10(P(c), P(k), P(o))
        return o


def golden(name):
    return np.load(ROOT / "tests" / "golden" / name, allow_pickle=False)


REF_LIB = ROOT / "oracle" / "_ref" / "libref.so"


class Reference:
    """The compiled reference (oracle/_ref/libref.so: the reference's own
    Fortran sources for the hot path plus the bind(C) harness
    oracle/ref_harness.F90), for timing its CPU path.  The deck is written in
    the reference's own formats and read by its own routines.  Raises
    FileNotFoundError when the library has not been built."""

    def __init__(self, bas, ng0=20):
        import tempfile
        from slater_stoch_amd import configs
        if not REF_LIB.exists():
            raise FileNotFoundError(REF_LIB)
        lib = ctypes.CDLL(str(REF_LIB))
        lib.ref_setup.restype = ctypes.c_int
        lib.ref_setup.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        lib.ref_gauss_ijkl.restype = ctypes.c_double
        lib.ref_gauss_ijkl.argtypes = [ctypes.c_int] * 4
        lib.ref_mc_fixed.argtypes = [ctypes.c_int, ctypes.c_int, VP, VP, VP,
                                     ctypes.c_int64] + [VP] * 6
        self.lib = lib
        d = Path(tempfile.mkdtemp())
        configs.write_deck(bas, d, name="mol", ng0=ng0)
        n = lib.ref_setup(str(d / "mol").encode(), str(d / "mol.basis").encode(), ng0, 0, 0,
                          100.0, 0, bas.a_zv)
        if n != bas.n:
            raise RuntimeError(f"reference read {n} orbitals, expected {bas.n}")

    def gauss_ijkl(self, i, k, j, l):
        return self.lib.ref_gauss_ijkl(int(i), int(k), int(j), int(l))

    def mc_fixed(self, nbl, npts, r1, r2, pi0, quartets):
        is_, js_, ks_, ls_ = [np.ascontiguousarray(a, np.int32) for a in quartets]
        mean, err = np.zeros(is_.size), np.zeros(is_.size)
        self.lib.ref_mc_fixed(nbl, npts, P(np.ascontiguousarray(r1)),
                              P(np.ascontiguousarray(r2)), P(np.ascontiguousarray(pi0)),
                              is_.size, P(is_), P(js_), P(ks_), P(ls_), P(mean), P(err))
        return mean, err
