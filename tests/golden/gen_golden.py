#!/usr/bin/env python3
"""Generate the golden fixtures for the M&D path with an INDEPENDENT engine (mpmath, 30 digits).

Method: Obara & Saika, J. Chem. Phys. 84, 3963 (1986) vertical recursion for [a0|c0]^(m) plus the
Head-Gordon & Pople, J. Chem. Phys. 89, 5777 (1988) horizontal transfer; one-electron overlap /
kinetic / nuclear-attraction integrals by the Obara-Saika one-electron recursions; the Boys
function by its closed form F_j(T) = gamma(j+1/2, T) / (2 T^(j+1/2)).  None of this shares code or
algorithm with the McMurchie-Davidson oracle or the GPU kernels it checks.

Sample points follow the reference's own oracle specification (tools/oracle/generate_reference.py:
EXPS/CENTERS at :35-36, the ERI sweep at :67-90, the nuclei at :24), whose output file is absent
from the reference tree; values are normalised per Cartesian component to unit self-overlap, the
reference convention (generate_reference.py:4-7, PGData.C:27-31).  Components are listed in the
reference order (m outer, l inner, n = L-m-l; PG_Cart/Imp/IrrepBasisSet.C:46-54).

Run:  python tests/golden/gen_golden.py   (writes boys.json, eri_sweep.json, one_electron.json)
"""
import itertools
import json
import os
from functools import lru_cache

import mpmath as mp

mp.mp.dps = 30
HERE = os.path.dirname(os.path.abspath(__file__))


def comps(L):
    return [(L - m - l, l, m) for m in range(L + 1) for l in range(L - m + 1)]


def dfact(n):
    r = mp.mpf(1)
    while n > 1:
        r *= n
        n -= 2
    return r


def boys(m, T):
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * m + 1)
    return mp.gammainc(m + mp.mpf(1) / 2, 0, T) / (2 * T ** (m + mp.mpf(1) / 2))


def norm(alpha, c):
    """1/sqrt(<c|c>) for a primitive x^n y^l z^m exp(-alpha r^2)."""
    a = mp.mpf(alpha)
    L = sum(c)
    S = (mp.pi / (2 * a)) ** mp.mpf(1.5) * dfact(2 * c[0] - 1) * dfact(2 * c[1] - 1) * dfact(2 * c[2] - 1) / (4 * a) ** L
    return 1 / mp.sqrt(S)


def _sub(v, i, k=1):
    v = list(v)
    v[i] -= k
    return tuple(v)


def _add(v, i, k=1):
    v = list(v)
    v[i] += k
    return tuple(v)


def _first_nonzero(v):
    for i in range(3):
        if v[i] > 0:
            return i
    return -1


class ERI:
    """Primitive (ab|cd), Obara-Saika VRR + HGP HRR, one instance per (exponents, centres)."""

    def __init__(self, ea, A, eb, B, ec, C, ed, D):
        a, b, c, d = (mp.mpf(x) for x in (ea, eb, ec, ed))
        A, B, C, D = ([mp.mpf(x) for x in v] for v in (A, B, C, D))
        z, e = a + b, c + d
        P = [(a * A[i] + b * B[i]) / z for i in range(3)]
        Q = [(c * C[i] + d * D[i]) / e for i in range(3)]
        rho = z * e / (z + e)
        W = [(z * P[i] + e * Q[i]) / (z + e) for i in range(3)]
        AB2 = sum((A[i] - B[i]) ** 2 for i in range(3))
        CD2 = sum((C[i] - D[i]) ** 2 for i in range(3))
        PQ2 = sum((P[i] - Q[i]) ** 2 for i in range(3))
        self.pref = 2 * mp.pi ** mp.mpf(2.5) / (z * e * mp.sqrt(z + e)) * mp.exp(-a * b / z * AB2) * mp.exp(-c * d / e * CD2)
        self.T = rho * PQ2
        self.z, self.e, self.rho = z, e, rho
        self.PA = [P[i] - A[i] for i in range(3)]
        self.WP = [W[i] - P[i] for i in range(3)]
        self.QC = [Q[i] - C[i] for i in range(3)]
        self.WQ = [W[i] - Q[i] for i in range(3)]
        self.AB = [A[i] - B[i] for i in range(3)]
        self.CD = [C[i] - D[i] for i in range(3)]
        self.vrr = lru_cache(maxsize=None)(self._vrr)
        self.val = lru_cache(maxsize=None)(self._val)
        self.F = {}

    def _F(self, m):
        if m not in self.F:
            self.F[m] = boys(m, self.T)
        return self.F[m]

    def _vrr(self, a, c, m):
        if min(a) < 0 or min(c) < 0:
            return mp.mpf(0)
        if sum(a) == 0 and sum(c) == 0:
            return self.pref * self._F(m)
        i = _first_nonzero(a)
        if i >= 0:
            am = _sub(a, i)
            r = self.PA[i] * self.vrr(am, c, m) + self.WP[i] * self.vrr(am, c, m + 1)
            if am[i] > 0:
                amm = _sub(am, i)
                r += am[i] / (2 * self.z) * (self.vrr(amm, c, m) - self.rho / self.z * self.vrr(amm, c, m + 1))
            if c[i] > 0:
                r += c[i] / (2 * (self.z + self.e)) * self.vrr(am, _sub(c, i), m + 1)
            return r
        i = _first_nonzero(c)
        cm = _sub(c, i)
        r = self.QC[i] * self.vrr(a, cm, m) + self.WQ[i] * self.vrr(a, cm, m + 1)
        if cm[i] > 0:
            cmm = _sub(cm, i)
            r += cm[i] / (2 * self.e) * (self.vrr(a, cmm, m) - self.rho / self.e * self.vrr(a, cmm, m + 1))
        return r

    def _val(self, a, b, c, d):
        i = _first_nonzero(b)
        if i >= 0:  # (a,b|  = (a+1i,b-1i| + AB_i (a,b-1i|
            bm = _sub(b, i)
            return self.val(_add(a, i), bm, c, d) + self.AB[i] * self.val(a, bm, c, d)
        i = _first_nonzero(d)
        if i >= 0:
            dm = _sub(d, i)
            return self.val(a, b, _add(c, i), dm) + self.CD[i] * self.val(a, b, c, dm)
        return self.vrr(a, c, 0)


# ---- one-electron Obara-Saika --------------------------------------------------------------------
class OneE:
    def __init__(self, ea, A, eb, B, nuclei):
        a, b = mp.mpf(ea), mp.mpf(eb)
        A, B = [mp.mpf(x) for x in A], [mp.mpf(x) for x in B]
        p = a + b
        P = [(a * A[i] + b * B[i]) / p for i in range(3)]
        self.a, self.b, self.p = a, b, p
        self.PA = [P[i] - A[i] for i in range(3)]
        self.PB = [P[i] - B[i] for i in range(3)]
        self.P = P
        AB2 = sum((A[i] - B[i]) ** 2 for i in range(3))
        self.KAB = mp.exp(-a * b / p * AB2)
        self.nuclei = [(mp.mpf(n["Z"]), [mp.mpf(x) for x in n["center"]]) for n in nuclei]
        self.s1 = lru_cache(maxsize=None)(self._s1)
        self.vv = lru_cache(maxsize=None)(self._v)

    def _s1(self, i, na, nb):  # 1D overlap factor
        if na < 0 or nb < 0:
            return mp.mpf(0)
        if na == 0 and nb == 0:
            return mp.mpf(1)
        if na > 0:
            return self.PA[i] * self.s1(i, na - 1, nb) + ((na - 1) * self.s1(i, na - 2, nb) + nb * self.s1(i, na - 1, nb - 1)) / (2 * self.p)
        return self.PB[i] * self.s1(i, na, nb - 1) + (na * self.s1(i, na - 1, nb - 1) + (nb - 1) * self.s1(i, na, nb - 2)) / (2 * self.p)

    def overlap(self, ca, cb):
        r = (mp.pi / self.p) ** mp.mpf(1.5) * self.KAB
        for i in range(3):
            r *= self.s1(i, ca[i], cb[i])
        return r

    def kinetic(self, ca, cb):
        """<a| -1/2 nabla^2 |b> via the derivative identity on the b side."""
        b = self.b
        tot = mp.mpf(0)
        for i in range(3):
            t = 2 * b * b * self.overlap(ca, _add(cb, i, 2)) - b * (2 * cb[i] + 1) * self.overlap(ca, cb)
            if cb[i] >= 2:
                t += mp.mpf(cb[i] * (cb[i] - 1)) / 2 * self.overlap(ca, _sub(cb, i, 2))
            tot -= t
        return tot

    def _v(self, ca, cb, m, k):
        if min(ca) < 0 or min(cb) < 0:
            return mp.mpf(0)
        Z, C = self.nuclei[k]
        PC = [self.P[i] - C[i] for i in range(3)]
        if sum(ca) == 0 and sum(cb) == 0:
            T = self.p * sum(x * x for x in PC)
            return -Z * 2 * mp.pi / self.p * self.KAB * boys(m, T)
        i = _first_nonzero(ca)
        if i >= 0:
            am = _sub(ca, i)
            r = self.PA[i] * self.vv(am, cb, m, k) - PC[i] * self.vv(am, cb, m + 1, k)
            r += am[i] / (2 * self.p) * (self.vv(_sub(am, i), cb, m, k) - self.vv(_sub(am, i), cb, m + 1, k))
            r += cb[i] / (2 * self.p) * (self.vv(am, _sub(cb, i), m, k) - self.vv(am, _sub(cb, i), m + 1, k))
            return r
        i = _first_nonzero(cb)
        bm = _sub(cb, i)
        r = self.PB[i] * self.vv(ca, bm, m, k) - PC[i] * self.vv(ca, bm, m + 1, k)
        r += bm[i] / (2 * self.p) * (self.vv(ca, _sub(bm, i), m, k) - self.vv(ca, _sub(bm, i), m + 1, k))
        return r

    def nuclear(self, ca, cb):
        return sum(self.vv(tuple(ca), tuple(cb), 0, k) for k in range(len(self.nuclei)))


# ---- sample points (reference tools/oracle/generate_reference.py) -----------------------------------
EXPS = [0.30, 1.0, 3.7]
CENTERS = [[0.0, 0.0, 0.0], [0.0, 0.0, 1.3], [0.7, -1.1, 0.4]]
NUCLEI = [{"Z": 6.0, "center": [0.0, 0.0, 0.4]}, {"Z": 2.0, "center": [-0.6, 0.5, -0.3]}]
ERI_EXP = [1.1, 0.9, 1.3, 0.7]
ERI_CENTER = [[0, 0, 0], [0, 0, 1.1], [0.6, 0, 0.3], [0.2, -0.8, 0.5]]
LERI = 2
L1E = 2
BOYS_T = [0.0, 1e-9, 1e-3, 0.049, 0.05, 0.11, 0.5, 1.0, 2.345, 4.9, 7.77, 9.99, 11.94, 11.96, 12.0,
          13.3, 14.99, 15.0, 17.5, 18.0, 23.9, 24.0, 29.99, 30.0, 33.3, 45.0, 80.0, 250.0]


def gen_boys():
    out = {"T": BOYS_T, "jmax": 16, "F": []}
    for T in BOYS_T:
        out["F"].append([float(boys(j, T)) for j in range(17)])
    return out


def gen_eri():
    out = []
    for La, Lb, Lc, Ld in itertools.product(range(LERI + 1), repeat=4):
        e = ERI(ERI_EXP[0], ERI_CENTER[0], ERI_EXP[1], ERI_CENTER[1], ERI_EXP[2], ERI_CENTER[2], ERI_EXP[3], ERI_CENTER[3])
        vals = []
        for ca in comps(La):
            for cb in comps(Lb):
                for cc in comps(Lc):
                    for cd in comps(Ld):
                        v = e.val(ca, cb, cc, cd)
                        v *= norm(ERI_EXP[0], ca) * norm(ERI_EXP[1], cb) * norm(ERI_EXP[2], cc) * norm(ERI_EXP[3], cd)
                        vals.append(float(v))
        out.append({"L": [La, Lb, Lc, Ld], "values": vals})
    return {"exps": ERI_EXP, "centers": ERI_CENTER, "quartets": out}


def gen_one():
    recs = []
    for La, Lb in itertools.product(range(L1E + 1), repeat=2):
        for ea, eb in itertools.product(EXPS, repeat=2):
            for ia, ib in itertools.product(range(len(CENTERS)), repeat=2):
                o = OneE(ea, CENTERS[ia], eb, CENTERS[ib], NUCLEI)
                S, T, V = [], [], []
                for ca in comps(La):
                    for cb in comps(Lb):
                        nn = norm(ea, ca) * norm(eb, cb)
                        S.append(float(o.overlap(ca, cb) * nn))
                        T.append(float(o.kinetic(ca, cb) * nn))
                        V.append(float(o.nuclear(ca, cb) * nn))
                recs.append({"L": [La, Lb], "exps": [ea, eb], "centers": [ia, ib], "S": S, "T": T, "V": V})
    return {"centers": CENTERS, "nuclei": NUCLEI, "pairs": recs}


def main():
    with open(os.path.join(HERE, "boys.json"), "w") as f:
        json.dump(gen_boys(), f)
    print("boys.json")
    with open(os.path.join(HERE, "eri_sweep.json"), "w") as f:
        json.dump(gen_eri(), f)
    print("eri_sweep.json")
    with open(os.path.join(HERE, "one_electron.json"), "w") as f:
        json.dump(gen_one(), f)
    print("one_electron.json")


if __name__ == "__main__":
    main()
