"""qchem_amd -- MI355X-native Fock-build engine for qchem's molecular Gaussian path.

Python face of the two native libraries built in-tree by ``qchem_amd.build``:

* ``libqcjk.so``: the HIP engine behind the C ABI in ``include/qcjk.h``.  It computes
  McMurchie-Davidson ERIs digested into J and K, plus the one-electron matrices.
* ``libqchem_host.so``: the C++20 host side in ``include/qchem_host.h``, namely the basis
  conventions and the RHF driver that reaches J/K only through the HF faces.

There is no CPU fallback.  If the libraries are not built, every call raises.  Without a GPU,
engine creation fails with ``QCJK_ERR_NODEVICE``.
"""
import ctypes
import os

import numpy as np

from . import systems  # noqa: F401  (re-exported generators)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_HOST = None

QCJK_J, QCJK_K, QCJK_MIXED, QCJK_SCREEN = 1, 2, 4, 8
_ERRORS = {0: "ok", 1: "invalid argument", 2: "unsupported basis or option", 3: "HIP runtime error",
           4: "out of memory", 5: "no usable HIP device"}


class EngineError(RuntimeError):
    def __init__(self, where, code, detail=""):
        self.code = code
        super().__init__("%s failed: %s%s" % (where, _ERRORS.get(code, "error %d" % code),
                                              (" (%s)" % detail) if detail else ""))


class _Basis(ctypes.Structure):
    _fields_ = [("nshell", ctypes.c_int32), ("shell_l", ctypes.POINTER(ctypes.c_int32)),
                ("shell_nprim", ctypes.POINTER(ctypes.c_int32)), ("shell_center", ctypes.POINTER(ctypes.c_double)),
                ("prim_exp", ctypes.POINTER(ctypes.c_double)), ("prim_coef", ctypes.POINTER(ctypes.c_double)),
                ("bf_norm", ctypes.POINTER(ctypes.c_double)), ("order", ctypes.c_int32)]


class _Opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("screen_tol", ctypes.c_double), ("mixed_tol", ctypes.c_double),
                ("reserved", ctypes.c_int32 * 8)]


class Stats(ctypes.Structure):
    _fields_ = [("quartets_total", ctypes.c_int64), ("quartets_computed", ctypes.c_int64),
                ("quartets_fp32", ctypes.c_int64), ("prim_quartets", ctypes.c_int64),
                ("flops", ctypes.c_double), ("device_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ClassStat(ctypes.Structure):
    _fields_ = [("la", ctypes.c_int32), ("lb", ctypes.c_int32), ("lc", ctypes.c_int32), ("ld", ctypes.c_int32),
                ("wavefront_per_quartet", ctypes.c_int32), ("launches", ctypes.c_int64), ("quartets", ctypes.c_int64),
                ("prim_quartets", ctypes.c_int64), ("flops", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("alt_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _Shells(ctypes.Structure):
    _fields_ = [("nshell", ctypes.c_int32), ("l", ctypes.POINTER(ctypes.c_int32)),
                ("nprim", ctypes.POINTER(ctypes.c_int32)), ("exps", ctypes.POINTER(ctypes.c_double)),
                ("coefs", ctypes.POINTER(ctypes.c_double)), ("contracted", ctypes.POINTER(ctypes.c_int32)),
                ("centers", ctypes.POINTER(ctypes.c_double))]


class _ScfParams(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int32), ("e_tol", ctypes.c_double), ("fd_tol", ctypes.c_double),
                ("nproj", ctypes.c_int32), ("fd_max", ctypes.c_double), ("sv_tol", ctypes.c_double),
                ("jk_flags", ctypes.c_uint32), ("screen_tol", ctypes.c_double), ("mixed_tol", ctypes.c_double),
                ("device", ctypes.c_int32), ("verbose", ctypes.c_int32), ("guess", ctypes.c_int32),
                ("incremental", ctypes.c_int32)]

GUESS_CORE, GUESS_SAD = 0, 1


class _ScfResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("e_nuc", ctypes.c_double), ("e_one", ctypes.c_double),
                ("e_coul", ctypes.c_double), ("e_exch", ctypes.c_double), ("iterations", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("nbf", ctypes.c_int32), ("t_total", ctypes.c_double),
                ("t_jk", ctypes.c_double), ("t_diag", ctypes.c_double), ("t_setup", ctypes.c_double),
                ("quartets", ctypes.c_int64), ("quartets_fp32", ctypes.c_int64)]


QCJK_SYMBOLS = ["qcjk_create", "qcjk_create_batch", "qcjk_nbf", "qcjk_layout", "qcjk_build", "qcjk_build_device",
                "qcjk_graph_build", "qcjk_buffers", "qcjk_one_electron", "qcjk_one_electron_mol", "qcjk_schwarz",
                "qcjk_profile", "qcjk_class_report", "qcjk_set_variant", "qcjk_set_streams", "qcjk_streams", "qcjk_autotune",
                "qcjk_destroy",
                "qcjk_strerror"]
HOST_SYMBOLS = ["qchost_basis_prepare", "qcscf_rhf", "qcscf_rhf_batch", "qcscf_sad", "qchost_last_error"]


class _BatchResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("nconverged", ctypes.c_int32), ("t_total", ctypes.c_double),
                ("t_setup", ctypes.c_double), ("t_jk", ctypes.c_double), ("t_host", ctypes.c_double),
                ("jk_device_ms", ctypes.c_double), ("quartets", ctypes.c_int64), ("quartets_fp32", ctypes.c_int64)]


def lib_paths():
    return os.path.join(_HERE, "libqcjk.so"), os.path.join(_HERE, "libqchem_host.so")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def lib():
    """The engine library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        path = lib_paths()[0]
        if not os.path.exists(path):
            raise RuntimeError("qchem_amd: %s is not built (run __graft_entry__.build())" % path)
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.qcjk_create.argtypes = [ctypes.POINTER(_Basis), ctypes.POINTER(_Opts), ctypes.POINTER(vp)]
        L.qcjk_create_batch.argtypes = [ctypes.POINTER(_Basis), ctypes.c_int, ctypes.POINTER(_Opts), ctypes.POINTER(vp)]
        L.qcjk_nbf.argtypes = [vp]
        L.qcjk_layout.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                  ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.qcjk_graph_build.argtypes = [vp, vp, vp, vp, ctypes.c_uint, vp]
        L.qcjk_buffers.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L.qcjk_one_electron_mol.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.qcjk_build.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double), ctypes.c_uint, ctypes.POINTER(Stats)]
        L.qcjk_build_device.argtypes = [vp, vp, vp, vp, ctypes.c_uint, vp, ctypes.POINTER(Stats)]
        L.qcjk_one_electron.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.qcjk_schwarz.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.qcjk_profile.argtypes = [vp, ctypes.c_int]
        L.qcjk_class_report.argtypes = [vp, ctypes.POINTER(ClassStat), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.qcjk_set_variant.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.qcjk_set_streams.argtypes = [vp, ctypes.c_int]
        L.qcjk_streams.argtypes = [vp]
        L.qcjk_autotune.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(ClassStat), ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int)]
        L.qcjk_destroy.argtypes = [vp]
        L.qcjk_strerror.argtypes = [ctypes.c_int]
        L.qcjk_strerror.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def host_lib():
    global _HOST
    if _HOST is None:
        lib()
        path = lib_paths()[1]
        if not os.path.exists(path):
            raise RuntimeError("qchem_amd: %s is not built (run __graft_entry__.build())" % path)
        H = ctypes.CDLL(path)
        H.qchost_basis_prepare.argtypes = [ctypes.POINTER(_Shells), ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]
        H.qcscf_rhf.argtypes = [ctypes.POINTER(_Shells), ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double), ctypes.c_int32, ctypes.POINTER(_ScfParams),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_ScfResult),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]
        H.qcscf_rhf_batch.argtypes = [ctypes.POINTER(_Shells), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_ScfParams),
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_int32,
                                      ctypes.POINTER(_ScfResult), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_BatchResult)]
        H.qcscf_sad.argtypes = [ctypes.POINTER(_Shells), ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double), ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        H.qchost_last_error.restype = ctypes.c_char_p
        _HOST = H
    return _HOST


class Basis:
    """Molecular Cartesian basis: a list of shells ``dict(L, exps, coefs, center, contracted)``.

    ``contracted`` selects the reference's contracted-radial convention (coefficients times the
    primitive norm of x^L, GaussianRF.C:566-577); otherwise the single-exponent radial (coefficient 1).
    Folded coefficients and the per-component ns come from the C++ host library.
    """

    def __init__(self, shells):
        self.shells = [dict(s) for s in shells]
        for s in self.shells:
            s.setdefault("contracted", len(s["exps"]) > 1)
        self.l = np.array([s["L"] for s in self.shells], dtype=np.int32)
        self.nprim = np.array([len(s["exps"]) for s in self.shells], dtype=np.int32)
        self.exps = np.array([e for s in self.shells for e in s["exps"]], dtype=np.float64)
        self.coefs = np.array([c for s in self.shells for c in s["coefs"]], dtype=np.float64)
        self.contracted = np.array([1 if s["contracted"] else 0 for s in self.shells], dtype=np.int32)
        self.centers = np.ascontiguousarray(np.array([s["center"] for s in self.shells], dtype=np.float64).reshape(-1))
        self.nbf = int(sum((s["L"] + 1) * (s["L"] + 2) // 2 for s in self.shells))
        self.folded = np.zeros(len(self.exps))
        self.ns = np.zeros(self.nbf)
        nb = ctypes.c_int32(0)
        rc = host_lib().qchost_basis_prepare(ctypes.byref(self._shells()), _dp(self.folded), _dp(self.ns), ctypes.byref(nb))
        if rc != 0:
            raise EngineError("qchost_basis_prepare", rc, host_lib().qchost_last_error().decode())

    def _shells(self):
        return _Shells(len(self.shells), _ip(self.l), _ip(self.nprim), _dp(self.exps), _dp(self.coefs),
                       _ip(self.contracted), _dp(self.centers))

    def _cbasis(self):
        return _Basis(len(self.shells), _ip(self.l), _ip(self.nprim), _dp(self.centers), _dp(self.exps),
                      _dp(self.folded), _dp(self.ns), 0)


class JKEngine:
    """One engine handle (one HIP stream).  ``build`` overwrites J and K for a symmetric D.

    ``basis`` is one Basis, or a list of Bases for a batch handle (qcjk_create_batch): then D, J and
    K are lists of per-molecule matrices on the host, or packed device arrays of ``nelem`` doubles.
    """

    def __init__(self, basis, device=0, screen_tol=0.0, mixed_tol=0.0):
        self.batch = isinstance(basis, (list, tuple))
        self.bases = list(basis) if self.batch else [basis]
        self.basis = self.bases[0]
        opts = _Opts(device, float(screen_tol), float(mixed_tol))
        h = ctypes.c_void_p()
        if self.batch:
            arr = (_Basis * len(self.bases))(*[b._cbasis() for b in self.bases])
            rc = lib().qcjk_create_batch(arr, len(self.bases), ctypes.byref(opts), ctypes.byref(h))
        else:
            rc = lib().qcjk_create(ctypes.byref(basis._cbasis()), ctypes.byref(opts), ctypes.byref(h))
        if rc != 0:
            raise EngineError("qcjk_create_batch" if self.batch else "qcjk_create", rc)
        self.h = h
        nm = len(self.bases)
        nbf = np.zeros(nm, dtype=np.int32)
        off = np.zeros(nm, dtype=np.int64)
        nmol, nelem = ctypes.c_int32(0), ctypes.c_int64(0)
        rc = lib().qcjk_layout(h, ctypes.byref(nmol), _ip(nbf), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                               ctypes.byref(nelem))
        if rc != 0:
            raise EngineError("qcjk_layout", rc)
        self.nbfs, self.offsets, self.nelem = nbf.tolist(), off.tolist(), int(nelem.value)
        self.n = lib().qcjk_nbf(h)

    def close(self):
        if getattr(self, "h", None):
            lib().qcjk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pack(self, mats):
        flat = np.zeros(self.nelem)
        for M, o, n in zip(mats, self.offsets, self.nbfs):
            assert M.shape == (n, n)
            flat[o:o + n * n] = np.asarray(M, dtype=np.float64).reshape(-1)
        return flat

    def unpack(self, flat):
        return [np.asarray(flat[o:o + n * n]).reshape(n, n) for o, n in zip(self.offsets, self.nbfs)]

    def build(self, D, flags=QCJK_J | QCJK_K):
        flat = self.pack(D) if self.batch else np.ascontiguousarray(D, dtype=np.float64)
        if not self.batch:
            assert flat.shape == (self.n, self.n)
        J = np.zeros(self.nelem)
        K = np.zeros(self.nelem)
        st = Stats()
        rc = lib().qcjk_build(self.h, _dp(flat), _dp(J), _dp(K), flags, ctypes.byref(st))
        if rc != 0:
            raise EngineError("qcjk_build", rc)
        if self.batch:
            return self.unpack(J), self.unpack(K), st
        return J.reshape(self.n, self.n), K.reshape(self.n, self.n), st

    @staticmethod
    def _ptr(x):
        return None if x is None else (x if isinstance(x, int) else x.data_ptr())

    @staticmethod
    def _stream(stream):
        return None if stream is None else (stream if isinstance(stream, int) else stream.cuda_stream)

    def build_device(self, dD, dJ, dK, flags=QCJK_J | QCJK_K, stream=None, stats=False):
        """Device pointers (ints) or torch CUDA tensors; asynchronous on ``stream`` unless stats."""
        st = Stats() if stats else None
        rc = lib().qcjk_build_device(self.h, self._ptr(dD), self._ptr(dJ), self._ptr(dK), flags,
                                     self._stream(stream), ctypes.byref(st) if st is not None else None)
        if rc != 0:
            raise EngineError("qcjk_build_device", rc)
        return st

    def graph_build(self, dD, dJ, dK, flags=QCJK_J | QCJK_K, stream=None):
        """The same build replayed from a captured hipGraph (captured on first use)."""
        rc = lib().qcjk_graph_build(self.h, self._ptr(dD), self._ptr(dJ), self._ptr(dK), flags, self._stream(stream))
        if rc != 0:
            raise EngineError("qcjk_graph_build", rc)

    def buffers(self):
        """Device addresses of the handle's own D, J, K buffers (nelem doubles each)."""
        p = [ctypes.c_void_p() for _ in range(3)]
        rc = lib().qcjk_buffers(self.h, *[ctypes.byref(x) for x in p])
        if rc != 0:
            raise EngineError("qcjk_buffers", rc)
        return [x.value for x in p]

    def one_electron(self, Z, R, mol=None):
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
        n = self.nbfs[mol or 0]
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        if mol is None:
            rc = lib().qcjk_one_electron(self.h, len(Z), _dp(Z), _dp(R), _dp(S), _dp(T), _dp(V))
        else:
            rc = lib().qcjk_one_electron_mol(self.h, mol, len(Z), _dp(Z), _dp(R), _dp(S), _dp(T), _dp(V))
        if rc != 0:
            raise EngineError("qcjk_one_electron", rc)
        return S, T, V

    def profile(self, enable=True):
        rc = lib().qcjk_profile(self.h, 1 if enable else 0)
        if rc != 0:
            raise EngineError("qcjk_profile", rc)

    def class_report(self):
        cnt = ctypes.c_int(0)
        arr = (ClassStat * 64)()
        rc = lib().qcjk_class_report(self.h, arr, 64, ctypes.byref(cnt))
        if rc != 0:
            raise EngineError("qcjk_class_report", rc)
        return [arr[i].as_dict() for i in range(min(cnt.value, 64))]

    def set_variant(self, la, lb, lc, ld, variant):
        rc = lib().qcjk_set_variant(self.h, la, lb, lc, ld, variant)
        if rc != 0:
            raise EngineError("qcjk_set_variant", rc)

    def set_streams(self, nstreams):
        """Spread each build's class launches over nstreams concurrent streams (1-4)."""
        rc = lib().qcjk_set_streams(self.h, int(nstreams))
        if rc != 0:
            raise EngineError("qcjk_set_streams", rc)

    def streams(self):
        """Concurrent class streams per build."""
        return int(lib().qcjk_streams(self.h))

    def autotune(self, dD, reps=3):
        """Time every kernel variant per class on device density dD; keep the fastest (see tune.py)."""
        cnt = ctypes.c_int(0)
        arr = (ClassStat * 64)()
        ptr = dD if isinstance(dD, int) else dD.data_ptr()
        rc = lib().qcjk_autotune(self.h, ptr, reps, arr, 64, ctypes.byref(cnt))
        if rc != 0:
            raise EngineError("qcjk_autotune", rc)
        return [arr[i].as_dict() for i in range(min(cnt.value, 64))]

    def apply_tuning(self, table):
        """table: {"la lb lc ld": variant} (as written by qchem_amd.tune)."""
        for key, v in table.items():
            la, lb, lc, ld = (int(x) for x in key.split())
            self.set_variant(la, lb, lc, ld, int(v))

    def schwarz(self):
        ns = len(self.basis.shells)
        Q = np.zeros((ns, ns))
        rc = lib().qcjk_schwarz(self.h, _dp(Q))
        if rc != 0:
            raise EngineError("qcjk_schwarz", rc)
        return Q


def rhf(basis, Z, R, charge=0, max_iter=100, e_tol=1e-9, fd_tol=1e-6, nproj=4, fd_max=-1.0, sv_tol=5e-9,
        jk_flags=QCJK_J | QCJK_K, screen_tol=0.0, mixed_tol=0.0, device=0, D0=None, verbose=False, guess=GUESS_CORE,
        incremental=0):
    """Closed-shell RHF through the C++ driver; returns a dict with energy, density and history."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
    par = _ScfParams(max_iter, e_tol, fd_tol, nproj, fd_max, sv_tol, jk_flags, screen_tol, mixed_tol, device,
                     1 if verbose else 0, int(guess), int(incremental))
    res = _ScfResult()
    n = basis.nbf
    Dout = np.zeros((n, n))
    eh = np.zeros(max_iter)
    th = np.zeros(max_iter)
    d0 = None
    if D0 is not None:
        d0 = np.ascontiguousarray(D0, dtype=np.float64)
    rc = host_lib().qcscf_rhf(ctypes.byref(basis._shells()), len(Z), _dp(Z), _dp(R), int(charge), ctypes.byref(par),
                              _dp(d0) if d0 is not None else None, ctypes.byref(res), _dp(Dout), _dp(eh), _dp(th))
    if rc != 0:
        raise EngineError("qcscf_rhf", rc, host_lib().qchost_last_error().decode())
    it = res.iterations
    return {"energy": res.energy, "e_nuc": res.e_nuc, "e_one": res.e_one, "e_coul": res.e_coul,
            "e_exch": res.e_exch, "iterations": it, "converged": bool(res.converged), "D": Dout,
            "energy_history": eh[:it].copy(), "time_history": th[:it].copy(), "t_total": res.t_total,
            "t_jk": res.t_jk, "t_diag": res.t_diag, "t_setup": res.t_setup, "quartets": res.quartets,
            "fp32_fraction": res.quartets_fp32 / max(1, res.quartets)}


def sad_density(basis, Z, R, device=0):
    """SAD seed density of a molecule (qcscf_sad)."""
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
    D = np.zeros((basis.nbf, basis.nbf))
    rc = host_lib().qcscf_sad(ctypes.byref(basis._shells()), len(Z), _dp(Z), _dp(R), device, _dp(D))
    if rc != 0:
        raise EngineError("qcscf_sad", rc, host_lib().qchost_last_error().decode())
    return D


def rhf_batch(bases, mols, max_iter=100, e_tol=1e-9, fd_tol=1e-6, nproj=4, fd_max=-1.0, sv_tol=5e-9,
              jk_flags=QCJK_J | QCJK_K, screen_tol=0.0, mixed_tol=0.0, device=0, nthreads=16, guess=GUESS_CORE,
              D0=None, verbose=False):
    """Batched closed-shell RHF (qcscf_rhf_batch): one batched J/K build per iteration for every
    molecule.  ``mols``: dicts with Z, R (and optional charge).  Returns (per-molecule list, summary)."""
    nmol = len(bases)
    shells = (_Shells * nmol)(*[b._shells() for b in bases])
    natom = np.array([len(m["Z"]) for m in mols], dtype=np.int32)
    Z = np.ascontiguousarray(np.concatenate([np.asarray(m["Z"], dtype=np.float64) for m in mols]))
    R = np.ascontiguousarray(np.concatenate([np.asarray(m["R"], dtype=np.float64).reshape(-1) for m in mols]))
    charge = np.array([int(m.get("charge", 0)) for m in mols], dtype=np.int32)
    par = _ScfParams(max_iter, e_tol, fd_tol, nproj, fd_max, sv_tol, jk_flags, screen_tol, mixed_tol, device,
                     1 if verbose else 0, int(guess), 0)
    d0 = None
    if D0 is not None:  # per-molecule seeds, packed back to back
        d0 = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in D0]))
    res = (_ScfResult * nmol)()
    eh = np.zeros(nmol * max_iter)
    nbfs = [b.nbf for b in bases]
    dout = np.zeros(sum(n * n for n in nbfs))
    br = _BatchResult()
    rc = host_lib().qcscf_rhf_batch(shells, nmol, _ip(natom), _dp(Z), _dp(R), _ip(charge), ctypes.byref(par),
                                    _dp(d0) if d0 is not None else None, int(nthreads), res, _dp(eh), _dp(dout),
                                    ctypes.byref(br))
    if rc != 0:
        raise EngineError("qcscf_rhf_batch", rc, host_lib().qchost_last_error().decode())
    out = []
    off = 0
    for m in range(nmol):
        r = res[m]
        n = nbfs[m]
        out.append({"energy": r.energy, "iterations": r.iterations, "converged": bool(r.converged), "nbf": r.nbf,
                    "energy_history": eh[m * max_iter:m * max_iter + r.iterations].copy(),
                    "D": dout[off:off + n * n].reshape(n, n).copy()})
        off += n * n
    summary = {k: getattr(br, k) for k, _ in _BatchResult._fields_}
    return out, summary

