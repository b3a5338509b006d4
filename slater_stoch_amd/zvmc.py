"""ctypes binding of libzvmc.so (the C-ABI declared in include/zvmc.h).

This is the Python-side mirror of the reference's hot-path entry points: the
analytic Gaussian-fit ERIs with their Schwarz driver (src/two_e_slater.F90:219-284),
gauss_ijkl for explicit quartets (src/gauss_ijkl.F90:51), the ZVMC correction
blocks (src/two_e_slater.F90:386-496) and the PSD cleanup
(src/svd_clean.F90).  Errors raise ``ZVError``, the counterpart of the
reference's ``stop``.  There is no CPU fallback: if the HIP library is
missing or no gfx950 device is present, constructing ``Engine`` raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libzvmc.so"

ZV_OK, ZV_ERR_ARG, ZV_ERR_HIP, ZV_ERR_STATE, ZV_ERR_NOMEM, ZV_ERR_NODEV = range(6)
SAMPLES_FIXED, SAMPLES_PHILOX = 0, 1
PREC_FP64_PARITY, PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2, 3
PRECISIONS = {"fp64_parity": PREC_FP64_PARITY, "fp64": PREC_FP64, "fp32": PREC_FP32,
              "bf16": PREC_BF16}
PHASES = ["sample", "densities", "estimator", "statistics", "prim_pairs", "analytic_eri",
          "copies", "cleanup"]

VP = ctypes.c_void_p
I32, I64, U64, F64 = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double

# name -> (restype, argtypes); every symbol declared in include/zvmc.h
SIGNATURES = {
    "zv_create": (I32, [I32, ctypes.POINTER(VP)]),
    "zv_destroy": (I32, [VP]),
    "zv_last_error": (I32, [VP, ctypes.c_char_p, I32]),
    "zv_set_basis": (I32, [VP, I32, VP, VP, VP, VP, VP, VP, I32, VP, VP, F64]),
    "zv_nint": (I64, [I32]),
    "zv_build_mapping": (I32, [I32, VP, VP, VP, VP]),
    "zv_gauss_eri": (I32, [VP, VP, VP, VP]),
    "zv_exact_monocenter": (I32, [VP, VP, VP, VP]),
    "zv_one_e": (I32, [VP, I32, VP, VP, VP, VP, VP]),
    "zv_gauss_ijkl_list": (I32, [VP, I64, VP, VP, VP, VP, VP]),
    "zv_mc_blocks": (I32, [VP, I32, I64, I32, U64, U64, I32, VP, VP, VP, VP, VP, VP]),
    "zv_draw_samples": (I32, [VP, U64, U64, I64, I64, VP, VP, VP]),
    "zv_pair_samples": (I32, [VP, I32, I32, I32, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    "zv_svd_clean": (I32, [VP, I32, I32, U64, VP, VP]),
    "zv_davidson_clean": (I32, [VP, VP, VP]),
    "zv_set_verbose": (I32, [VP, I32]),
    "zv_last_timings": (I32, [VP, VP, VP]),
    "zv_last_eri_stats": (I32, [VP, VP, VP, VP, VP]),
    "zv_get_result": (I32, [VP, I32, VP]),
    "zv_result_max": (I32, [VP, I32, VP]),
    "zv_set_tuning": (I32, [VP, I32, I64]),
    "zv_mc_refine": (I32, [VP, I32, I64, U64, U64, I32, F64, VP, VP]),
    "zv_rsvd_clean": (I32, [VP, I32, I32, U64, I32, VP, VP, VP]),
    "zv_davidson_lowest": (I32, [VP, I32, F64, I32, I32, VP, VP, VP, VP, VP]),
    "zv_gauss_ijkl_list_noalloc": (I32, [VP, I64, VP, VP, VP, VP, VP]),
}

NOALLOC_MAXQ = 65536

TUNE_EST_SPLIT, TUNE_MC_CHUNK, TUNE_ERI_ITEMS, TUNE_ERI_STREAMS, TUNE_EST_TILES = 0, 1, 2, 3, 4

RESULT_ERI, RESULT_MEAN, RESULT_STDERR = 0, 1, 2

_LIB = None


class ZVError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"zvmc error {code}: {msg}")
        self.code = code


def load_library(path=None):
    """dlopen libzvmc.so and declare every C-ABI signature (no device calls)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ZVError(ZV_ERR_STATE, f"{p} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(VP)


def build_mapping(n):
    """is, js, ks, ls (1-based) in the reference order, via the library."""
    lib = load_library()
    nint = lib.zv_nint(n)
    q = [np.zeros(nint, np.int32) for _ in range(4)]
    rc = lib.zv_build_mapping(n, *[_p(a) for a in q])
    if rc != ZV_OK:
        raise ZVError(rc, "zv_build_mapping failed")
    return q


class Engine:
    """One context on one gfx950 device (one HIP stream)."""

    def __init__(self, device=0, verbose=0):
        self.lib = load_library()
        h = VP()
        rc = self.lib.zv_create(device, ctypes.byref(h))
        if rc != ZV_OK:
            raise ZVError(rc, "no usable gfx950 device" if rc == ZV_ERR_NODEV else "zv_create")
        self.h = h
        self.basis = None
        if verbose:
            self._check(self.lib.zv_set_verbose(self.h, int(verbose)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.zv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != ZV_OK:
            buf = ctypes.create_string_buffer(512)
            self.lib.zv_last_error(self.h, buf, 512)
            raise ZVError(rc, buf.value.decode(errors="replace"))

    def set_tuning(self, key, value):
        """Tuning knob (TUNE_EST_SPLIT, TUNE_MC_CHUNK, TUNE_ERI_ITEMS,
        TUNE_ERI_STREAMS, TUNE_EST_TILES); 0 restores the library's
        automatic choice."""
        self._check(self.lib.zv_set_tuning(self.h, int(key), int(value)))

    # ------------------------------------------------------------ basis
    def set_basis(self, bas):
        self.basis = bas
        self._check(self.lib.zv_set_basis(
            self.h, bas.n, _p(np.ascontiguousarray(bas.center)), _p(np.ascontiguousarray(bas.npower)),
            _p(bas.n_sto), _p(bas.i_type), _p(bas.zeta), _p(bas.n_gauss), bas.ldg,
            _p(np.ascontiguousarray(bas.c_gauss)), _p(np.ascontiguousarray(bas.g_gauss)),
            float(bas.a_zv)))

    # ------------------------------------------------------------ analytic
    def gauss_eri(self, device_only=False):
        """(ijkl_gaus[nint], precond[n, n] with precond[k-1, i-1] = (ik|ik), n_screened);
        with device_only the ERIs stay on the device (get_result) and None is
        returned in their place."""
        n = self.basis.n
        out = None if device_only else np.zeros(self.basis.nint)
        pre = np.zeros((n, n))
        nscr = np.zeros(1, np.int64)
        self._check(self.lib.zv_gauss_eri(self.h, _p(out), _p(pre), _p(nscr)))
        return out, pre, int(nscr[0])

    def exact_monocenter(self, ijkl, err=None):
        """Overwrite the one-centre quartets of ijkl[nint] (kcp order) with
        their exact Slater integrals and err[kcp] with 0, in place
        (src/two_e_slater.F90:318-365); returns the number of one-centre
        quartets."""
        nint = self.basis.nint
        for name, a in (("ijkl", ijkl), ("err", err)):
            if a is not None and not (a.dtype == np.float64 and a.flags.c_contiguous
                                      and a.size == nint):
                raise ValueError(f"{name} must be a contiguous float64 array of {nint} entries")
        nm = np.zeros(1, np.int64)
        self._check(self.lib.zv_exact_monocenter(self.h, _p(ijkl), _p(err), _p(nm)))
        return int(nm[0])

    def one_e(self, charges=None, nuclei=None):
        """Overlap, nuclear attraction and kinetic integrals over the Gaussian
        fits (src/one_elect.F90): (S, V, K) as n x n arrays, [i-1, k-1] =
        (i|op|k).  Nuclei default to the basis geometry (charge = atomic
        number, src/read_geometry.F90:23)."""
        from . import basis as _B
        geo = self.basis.geometry
        if charges is None:
            charges = [float(_B.number_atom(lab)) for lab in geo.labels]
        if nuclei is None:
            nuclei = geo.coords
        q = np.ascontiguousarray(charges, np.float64)
        c = np.ascontiguousarray(nuclei, np.float64).reshape(-1, 3)
        n = self.basis.n
        S, V, K = (np.zeros((n, n), order="F") for _ in range(3))
        self._check(self.lib.zv_one_e(self.h, q.size, _p(q), _p(c), _p(S), _p(V), _p(K)))
        return S, V, K

    def get_result(self, which):
        """Device-resident result (RESULT_ERI, RESULT_MEAN, RESULT_STDERR) of
        the last gauss_eri / mc_blocks, copied to a new host array."""
        out = np.zeros(self.basis.nint)
        self._check(self.lib.zv_get_result(self.h, int(which), _p(out)))
        return out

    def result_max(self, which):
        """Largest entry of a device-resident result, reduced on the device."""
        v = np.zeros(1)
        self._check(self.lib.zv_result_max(self.h, int(which), _p(v)))
        return float(v[0])

    def gauss_ijkl(self, qi, qk, qj, ql):
        qi, qk, qj, ql = [np.ascontiguousarray(a, np.int32) for a in (qi, qk, qj, ql)]
        out = np.zeros(qi.size)
        self._check(self.lib.zv_gauss_ijkl_list(self.h, qi.size, _p(qi), _p(qk), _p(qj), _p(ql),
                                                _p(out)))
        return out

    def gauss_ijkl_noalloc(self, qi, qk, qj, ql):
        """gauss_ijkl through the allocation-free entry point (at most
        NOALLOC_MAXQ quartets; the primitive-pair tables must exist)."""
        qi, qk, qj, ql = [np.ascontiguousarray(a, np.int32) for a in (qi, qk, qj, ql)]
        out = np.zeros(qi.size)
        self._check(self.lib.zv_gauss_ijkl_list_noalloc(self.h, qi.size, _p(qi), _p(qk), _p(qj),
                                                        _p(ql), _p(out)))
        return out

    # ------------------------------------------------------------ MC
    def mc_blocks(self, nbl, npts, precision="fp64_parity", samples=None, seed=0, stream=0,
                  block_means=False, device_only=False):
        """samples = (r1[M,3], r2[M,3], pi0[M]) for the fixed-sample mode
        (M = nbl*npts), or None to draw on the device from (seed, stream).
        With device_only mean and err stay on the device (get_result,
        result_max) and (None, None) is returned in their place."""
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        nint = self.basis.nint
        mean, err = (None, None) if device_only else (np.zeros(nint), np.zeros(nint))
        blk = np.zeros((nbl, nint)) if block_means else None
        if samples is not None:
            r1, r2, pi0 = [np.ascontiguousarray(a, np.float64) for a in samples]
            m = nbl * npts
            for name, a, shape in (("r1", r1, (m, 3)), ("r2", r2, (m, 3)), ("pi0", pi0, (m,))):
                if a.shape != shape:
                    raise ValueError(f"{name} has shape {a.shape}, expected {shape} "
                                     f"(nbl * npts = {m} samples)")
            mode = SAMPLES_FIXED
        else:
            r1 = r2 = pi0 = None
            mode = SAMPLES_PHILOX
        self._check(self.lib.zv_mc_blocks(self.h, nbl, npts, mode, seed, stream, prec, _p(r1),
                                          _p(r2), _p(pi0), _p(mean), _p(err), _p(blk)))
        return (mean, err, blk) if block_means else (mean, err)

    def mc_refine(self, nbl, npts, precision="fp32", target=1e-6, seed=0, stream=1):
        """Refine the device-resident results: quartets with stderr > target
        get a fresh run of nbl x npts samples over the tiles holding them.
        Returns (quartets refined, largest standard error afterwards)."""
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        nref = np.zeros(1, np.int64)
        emax = np.zeros(1)
        self._check(self.lib.zv_mc_refine(self.h, int(nbl), int(npts), int(seed), int(stream),
                                          prec, float(target), _p(nref), _p(emax)))
        return int(nref[0]), float(emax[0])

    def draw_samples(self, m, seed=0, stream=0, m0=0):
        r1, r2, pi0 = np.zeros((m, 3)), np.zeros((m, 3)), np.zeros(m)
        self._check(self.lib.zv_draw_samples(self.h, seed, stream, m0, m, _p(r1), _p(r2),
                                             _p(pi0)))
        return r1, r2, pi0

    def pair_samples(self, i, k, r1, r2):
        ns = r1.shape[0]
        r1 = np.ascontiguousarray(r1, np.float64)
        r2 = np.ascontiguousarray(r2, np.float64)
        out = {name: np.zeros(ns) for name in ("rho1", "rho2", "rhog1", "rhog2", "jac1", "jac2")}
        ut1, ut2 = np.zeros((ns, 3)), np.zeros((ns, 3))
        self._check(self.lib.zv_pair_samples(self.h, int(i), int(k), ns, _p(r1), _p(r2), _p(ut1),
                                             _p(ut2), _p(out["rho1"]), _p(out["rho2"]),
                                             _p(out["rhog1"]), _p(out["rhog2"]), _p(out["jac1"]),
                                             _p(out["jac2"])))
        out["ut1"], out["ut2"] = ut1, ut2
        return out

    # ------------------------------------------------------------ cleanup
    def svd_clean(self, eri, rank, q=12, seed=0):
        e = np.array(eri, np.float64, copy=True)
        nrem = np.zeros(1, np.int32)
        self._check(self.lib.zv_svd_clean(self.h, rank, q, seed, _p(e), _p(nrem)))
        return e, int(nrem[0])

    def davidson_clean(self, eri):
        e = np.array(eri, np.float64, copy=True)
        nrem = np.zeros(1, np.int32)
        self._check(self.lib.zv_davidson_clean(self.h, _p(e), _p(nrem)))
        return e, int(nrem[0])

    def rsvd_clean(self, eri, rank, q=2, seed=0, sketch="fp32"):
        """Randomized Rayleigh-Ritz cleanup of the pair supermatrix: (cleaned
        eri, Ritz values ascending, number of negative pairs removed)."""
        e = np.array(eri, np.float64, copy=True)
        r = min(int(rank), self.basis.npair)
        ritz = np.zeros(r)
        nrem = np.zeros(1, np.int32)
        self._check(self.lib.zv_rsvd_clean(self.h, int(rank), int(q), int(seed),
                                           {"fp64": 0, "fp32": 1}[sketch], _p(e), _p(ritz),
                                           _p(nrem)))
        return e, ritz, int(nrem[0])

    def davidson_lowest(self, eri, nev=10, tol=1e-9, max_iter=1000, precond="none"):
        """Block Davidson on the pair supermatrix: (cleaned eri, the nev
        lowest Ritz values, their residual norms, number removed,
        iterations).  precond "shift-invert" uses the Cholesky factor of
        V - sigma I below the spectrum (one more N x N matrix)."""
        code = {"none": 0, "shift-invert": 1}[precond]
        e = np.array(eri, np.float64, copy=True)
        ev, res = np.zeros(int(nev)), np.zeros(int(nev))
        nrem = np.zeros(1, np.int32)
        its = np.zeros(1, np.int32)
        self._check(self.lib.zv_davidson_lowest(self.h, int(nev), float(tol), int(max_iter), code,
                                                _p(e), _p(ev), _p(res), _p(nrem), _p(its)))
        return e, ev, res, int(nrem[0]), int(its[0])

    def eri_stats(self):
        """Per ERI kernel of the last gauss_eri / gauss_ijkl: {key: dict(ms,
        launches, units, flops)} with key ("component", LB, LK) or ("shell",
        LU, LL, lane_batch, reduction)."""
        n_stat = 236
        ms, un, fl = np.zeros(n_stat), np.zeros(n_stat), np.zeros(n_stat)
        n = np.zeros(n_stat, np.int64)
        self._check(self.lib.zv_last_eri_stats(self.h, _p(ms), _p(n), _p(un), _p(fl)))
        out = {}
        for k in range(n_stat):
            if not (n[k] or un[k]):
                continue
            if k < 36:
                key = ("component", k // 6, k % 6)
            else:
                v, c = divmod(k - 36, 25)
                key = ("shell", c // 5, c % 5, v // 2 + 1, bool(v & 1))
            out[key] = dict(ms=float(ms[k]), launches=int(n[k]), units=float(un[k]),
                            flops=float(fl[k]))
        return out

    def timings(self):
        ms = np.zeros(8)
        n = np.zeros(8, np.int64)
        self._check(self.lib.zv_last_timings(self.h, _p(ms), _p(n)))
        return {name: (float(ms[t]), int(n[t])) for t, name in enumerate(PHASES)}
