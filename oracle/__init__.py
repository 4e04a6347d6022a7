"""CPU oracle for the McMurchie-Davidson J/K path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import this
package, and only as the checker (or the timed CPU baseline).  The product package ``qchem_amd``
never imports it.

``mnd_oracle.c`` restates the reference algorithm (file:line citations in its header); this module
is a thin ctypes binding plus a numpy restatement of the reference RHF/DIIS loop (``scf.py``).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_boys_ref.argtypes = [ctypes.c_double, ctypes.c_int, dp]
        L.or_set_boys_mode.argtypes = [ctypes.c_int]
        L.or_boys_series.argtypes = [ctypes.c_int, ctypes.c_double]
        L.or_boys_series.restype = ctypes.c_double
        L.or_rnlm.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
        L.or_hermite2.argtypes = [ctypes.c_double, dp, dp, ctypes.c_int, ctypes.c_int, dp]
        L.or_cart_comp.argtypes = [ctypes.c_int, ctypes.c_int, ip]
        L.or_basis_create.argtypes = [ctypes.c_int, ip, ip, dp, dp, ip, dp]
        L.or_basis_create.restype = ctypes.c_void_p
        L.or_basis_destroy.argtypes = [ctypes.c_void_p]
        L.or_basis_nbf.argtypes = [ctypes.c_void_p]
        L.or_basis_ns.argtypes = [ctypes.c_void_p, dp]
        L.or_basis_coef.argtypes = [ctypes.c_void_p, dp]
        L.or_one_electron.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp, dp, dp]
        L.or_shell_quartet.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.or_jk_refalgo.argtypes = [ctypes.c_void_p, dp, dp, dp, dp]
        L.or_jk_refalgo.restype = ctypes.c_int
        L.or_jk_direct.argtypes = [ctypes.c_void_p, dp, dp, dp, ctypes.c_int, ctypes.c_longlong]
        L.or_jk_direct.restype = ctypes.c_longlong
        L.or_schwarz.argtypes = [ctypes.c_void_p, dp]
        L.or_eri4_create.argtypes = [ctypes.c_void_p]
        L.or_eri4_create.restype = ctypes.c_void_p
        L.or_eri4_build_seconds.argtypes = [ctypes.c_void_p]
        L.or_eri4_build_seconds.restype = ctypes.c_double
        L.or_eri4_matmul.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.or_eri4_destroy.argtypes = [ctypes.c_void_p]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def set_boys_mode(mode):
    """0: the reference's Boys algorithm (default); 1: series-exact F_j (test diagnostics)."""
    lib().or_set_boys_mode(int(mode))


def boys_ref(T, jmax):
    out = np.zeros(jmax + 1)
    lib().or_boys_ref(float(T), int(jmax), _dp(out))
    return out


def boys_series(j, T):
    return lib().or_boys_series(int(j), float(T))


def rnlm(L, alpha, dR):
    dR = np.ascontiguousarray(dR, dtype=np.float64)
    out = np.zeros((L + 1,) * 3)
    lib().or_rnlm(int(L), float(alpha), _dp(dR), _dp(out))
    return out


def hermite2(p, PA, PB, la, lb):
    PA = np.ascontiguousarray(PA, dtype=np.float64)
    PB = np.ascontiguousarray(PB, dtype=np.float64)
    out = np.zeros((3, la + lb + 1, la + 1, lb + 1))
    lib().or_hermite2(float(p), _dp(PA), _dp(PB), la, lb, _dp(out))
    return out


def cart_comp(L, k):
    out = np.zeros(3, dtype=np.int32)
    lib().or_cart_comp(L, k, _ip(out))
    return tuple(int(x) for x in out)


class Basis:
    """Oracle basis: same flat shell description the product C-ABI consumes (raw coefficients)."""

    def __init__(self, shells):
        # shells: list of dicts {L, exps, coefs, center, fold}
        self.shells = shells
        L = np.array([s["L"] for s in shells], dtype=np.int32)
        nprim = np.array([len(s["exps"]) for s in shells], dtype=np.int32)
        exps = np.array([e for s in shells for e in s["exps"]], dtype=np.float64)
        coefs = np.array([c for s in shells for c in s["coefs"]], dtype=np.float64)
        fold = np.array([1 if s.get("fold", len(s["exps"]) > 1) else 0 for s in shells], dtype=np.int32)
        cen = np.array([s["center"] for s in shells], dtype=np.float64).reshape(-1)
        self._keep = (L, nprim, exps, coefs, fold, cen)
        self.h = lib().or_basis_create(len(shells), _ip(L), _ip(nprim), _dp(exps), _dp(coefs), _ip(fold), _dp(cen))
        self.nbf = lib().or_basis_nbf(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_basis_destroy(self.h)
            self.h = None

    def ns(self):
        out = np.zeros(self.nbf)
        lib().or_basis_ns(self.h, _dp(out))
        return out

    def folded_coefs(self):
        out = np.zeros(len(self._keep[2]))
        lib().or_basis_coef(self.h, _dp(out))
        return out

    def one_electron(self, Z, R):
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
        n = self.nbf
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        lib().or_one_electron(self.h, len(Z), _dp(Z), _dp(R), _dp(S), _dp(T), _dp(V))
        return S, T, V

    def shell_quartet(self, a, b, c, d):
        dims = [(self.shells[s]["L"] + 1) * (self.shells[s]["L"] + 2) // 2 for s in (a, b, c, d)]
        out = np.zeros(dims)
        lib().or_shell_quartet(self.h, a, b, c, d, _dp(out))
        return out

    def jk_refalgo(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        n = self.nbf
        J, K, t = np.zeros((n, n)), np.zeros((n, n)), np.zeros(3)
        rc = lib().or_jk_refalgo(self.h, _dp(D), _dp(J), _dp(K), _dp(t))
        if rc != 0:
            raise RuntimeError("reference-algorithm J/K: primitive count too large for the R cache")
        return J, K, t

    def jk_direct(self, D, nthreads=0, pair_limit=-1):
        D = np.ascontiguousarray(D, dtype=np.float64)
        n = self.nbf
        J, K = np.zeros((n, n)), np.zeros((n, n))
        cnt = lib().or_jk_direct(self.h, _dp(D), _dp(J), _dp(K), int(nthreads), int(pair_limit))
        return J, K, cnt

    def schwarz(self):
        ns = len(self.shells)
        Q = np.zeros((ns, ns))
        lib().or_schwarz(self.h, _dp(Q))
        return Q


class ERI4:
    """The reference algorithm's cached J4/K4 tensors: built once, contracted per SCF iteration."""

    def __init__(self, basis):
        self.basis = basis
        self.h = lib().or_eri4_create(basis.h)
        if not self.h:
            raise RuntimeError("ERI4: primitive count too large for the reference algorithm's R cache")
        self.build_seconds = lib().or_eri4_build_seconds(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_eri4_destroy(self.h)
            self.h = None

    def jk(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        n = self.basis.nbf
        J, K = np.zeros((n, n)), np.zeros((n, n))
        lib().or_eri4_matmul(self.h, _dp(D), _dp(J), _dp(K))
        return J, K
