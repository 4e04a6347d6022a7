"""CPU: pin the oracle before trusting it.

Checks run against the independent mpmath/Obara-Saika fixtures in tests/golden/ (the reference's own
fixture file is absent from its tree), against closed forms (Boys, Hermite expansion identity,
R_tuv as derivatives of F_0), and against invariants (8-fold symmetry, Schwarz, the two J/K routes
of the oracle, brute force on a tiny basis).  Each check is chosen so that a dropped term, a wrong
sign or a transposed index fails it.
"""
import json
import os

import mpmath as mp
import numpy as np
import numpy.polynomial.hermite as H
import pytest

import oracle
from qchem_amd import systems as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_boys_reference_algorithm_vs_closed_form():
    g = load("boys.json")
    for T, row in zip(g["T"], g["F"]):
        f = oracle.boys_ref(T, 16)
        for j in range(17):
            assert abs(f[j] - row[j]) <= 2e-12 * abs(row[j]), (T, j)


def test_boys_series_table_regeneration():
    g = load("boys.json")
    for T, row in zip(g["T"], g["F"]):
        if T > 40:
            continue
        for j in (0, 5, 16):
            assert abs(oracle.boys_series(j, T) - row[j]) <= 1e-15 * abs(row[j])


def test_hermite_expansion_identity():
    """(x-A)^i (x-B)^j = sum_t E^{ij}_t p^{t/2} H_t(sqrt(p)(x-P)) for every i, j, t."""
    rng = np.random.default_rng(0)
    a, b, A, B = 0.8, 1.7, -0.3, 0.9
    p = a + b
    P = (a * A + b * B) / p
    E = oracle.hermite2(p, [P - A] * 3, [P - B] * 3, 2, 2)[0]  # x axis: [t, i, j]
    for x in rng.uniform(-2, 2, 7):
        for i in range(3):
            for j in range(3):
                lhs = (x - A) ** i * (x - B) ** j
                rhs = sum(E[t, i, j] * p ** (t / 2) * H.hermval(np.sqrt(p) * (x - P), [0] * t + [1]) for t in range(i + j + 1))
                assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_rtuv_are_derivatives_of_f0():
    """R^{(0)}_{tuv} = d^t/dX^t d^u/dY^u d^v/dZ^v F_0(alpha (X^2+Y^2+Z^2))."""
    mp.mp.dps = 40
    alpha, X, Y, Z = 0.7, 0.4, -0.9, 1.3
    R = oracle.rnlm(4, alpha, [X, Y, Z])

    def F0(x, y, z):
        T = alpha * (x * x + y * y + z * z)
        return mp.sqrt(mp.pi / T) * mp.erf(mp.sqrt(T)) / 2

    for (t, u, v) in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (0, 1, 2), (2, 1, 1), (4, 0, 0), (1, 2, 1)]:
        d = mp.diff(F0, (X, Y, Z), (t, u, v))
        assert abs(R[t, u, v] - float(d)) <= 1e-10 * max(1.0, abs(float(d))), (t, u, v)


def test_eri_sweep_vs_obara_saika():
    g = load("eri_sweep.json")
    worst = 0.0
    for q in g["quartets"]:
        L = q["L"]
        sh = [dict(L=L[s], exps=[g["exps"][s]], coefs=[1.0], center=g["centers"][s]) for s in range(4)]
        v = oracle.Basis(sh).shell_quartet(0, 1, 2, 3).reshape(-1)
        ref = np.array(q["values"])
        worst = max(worst, np.abs(v - ref).max())
    assert worst < 1e-14


def test_one_electron_vs_obara_saika():
    g = load("one_electron.json")
    Z = [n["Z"] for n in g["nuclei"]]
    R = [n["center"] for n in g["nuclei"]]
    for pr in g["pairs"][::7]:
        La, Lb = pr["L"]
        ia, ib = pr["centers"]
        sh = [dict(L=La, exps=[pr["exps"][0]], coefs=[1.0], center=g["centers"][ia]),
              dict(L=Lb, exps=[pr["exps"][1]], coefs=[1.0], center=g["centers"][ib])]
        Sm, Tm, Vm = oracle.Basis(sh).one_electron(Z, R)
        na = (La + 1) * (La + 2) // 2
        for M, key, tol in ((Sm, "S", 1e-14), (Tm, "T", 1e-13), (Vm, "V", 1e-12)):
            assert np.abs(M[:na, na:].reshape(-1) - np.array(pr[key])).max() < tol, key


def _tiny_basis(seed=3):
    rng = S.SplitMix64(seed)
    Zs, xyz = S.random_molecule(2, rng)
    return [s for Z, c in zip(Zs, xyz) for s in S.random_contracted_shells(Z, c, rng, nprim=2)]


def _full_eri(ob, shells):
    nb = [(s["L"] + 1) * (s["L"] + 2) // 2 for s in shells]
    off = np.concatenate([[0], np.cumsum(nb)])
    n = off[-1]
    G = np.zeros((n, n, n, n))
    ns = len(shells)
    for a in range(ns):
        for b in range(ns):
            for c in range(ns):
                for d in range(ns):
                    G[off[a]:off[a + 1], off[b]:off[b + 1], off[c]:off[c + 1], off[d]:off[d + 1]] = ob.shell_quartet(a, b, c, d)
    return G


def test_eri_eightfold_symmetry_and_schwarz():
    shells = _tiny_basis()
    ob = oracle.Basis(shells)
    G = _full_eri(ob, shells)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(G - G.transpose(perm)).max() < 1e-13
    n = G.shape[0]
    diag = np.einsum("abab->ab", G)
    assert (diag > -1e-14).all()
    bound = np.sqrt(np.abs(diag)).reshape(n, n)
    assert (np.abs(G) <= np.einsum("ab,cd->abcd", bound, bound) + 1e-12).all()


def test_jk_routes_vs_brute_force():
    shells = _tiny_basis(5)
    ob = oracle.Basis(shells)
    G = _full_eri(ob, shells)
    n = G.shape[0]
    rng = np.random.default_rng(1)
    C = rng.normal(size=(n, 3))
    D = C @ C.T
    Jb = np.einsum("abcd,cd->ab", G, D)
    Kb = np.einsum("abcd,bd->ac", G, D)
    J1, K1, _ = ob.jk_refalgo(D)
    J2, K2, _ = ob.jk_direct(D)
    J3, K3 = oracle.ERI4(ob).jk(D)  # the reference's materialised ERI4 + MatMul route
    for J, K in ((J1, K1), (J2, K2), (J3, K3)):
        assert np.abs(J - Jb).max() < 1e-12 * np.abs(Jb).max()
        assert np.abs(K - Kb).max() < 1e-12 * np.abs(Kb).max()
    # a plausible packing mistake (K built with the J index pairing) must fail
    assert np.abs(J1 - Kb).max() > 1e-3


def test_jk_translation_invariance():
    shells = _tiny_basis(7)
    moved = [dict(s, center=[x + 0.37 for x in s["center"]]) for s in shells]
    ob, om = oracle.Basis(shells), oracle.Basis(moved)
    n = ob.nbf
    rng = np.random.default_rng(2)
    D = rng.normal(size=(n, n))
    D = D + D.T
    J0, K0, _ = ob.jk_direct(D)
    J1, K1, _ = om.jk_direct(D)
    assert np.abs(J0 - J1).max() < 1e-12 * np.abs(J0).max()
    assert np.abs(K0 - K1).max() < 1e-12 * np.abs(K0).max()


def test_normalisation_convention():
    """ns makes every Cartesian component unit-normalised (PGData.C:27-31)."""
    shells = _tiny_basis(9)
    ob = oracle.Basis(shells)
    S_, _, _ = ob.one_electron([1.0], [[0, 0, 0]])
    assert np.allclose(np.diag(S_), 1.0, atol=1e-13)


@pytest.mark.slow
def test_oracle_scf_n2():
    import oracle.scf as oscf
    m = S.n2_even_tempered()
    ob = oracle.Basis(m["shells"])
    r = oscf.rhf(ob, m["Z"], m["R"], lambda D: ob.jk_direct(D)[:2])
    assert r["converged"]
    assert abs(r["energy"] - (-106.9323789865)) < 1e-8


DZVP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dzvp_HNO.bsd")
# The reference's HF/DZVP anchors (IntegrationTests/M_HF_U.C:21-51): geometry in bohr, energy in Eh,
# and the relative tolerance that test uses.
ANCHORS = {
    "water": ([8, 1, 1], [[0, 0, 0], [0, 1.431, 1.107], [0, -1.431, 1.107]], -76.022903, 1e-5),
    "N2": ([7, 7], [[-1.03, 0, 0], [1.04, 0, 0]], -108.94524, 1e-4),
}


@pytest.mark.parametrize("name", sorted(ANCHORS))
def test_oracle_scf_dzvp_anchor(name):
    """Basis reader + integrals + normalisation + SCF against the reference's own converged energies.
    The anchors are printed to 6 and 5 decimals; the oracle lands within their rounding."""
    import oracle.scf as oscf
    from qchem_amd import basisfile as bf
    Z, R, e_ref, rel = ANCHORS[name]
    ob = oracle.Basis(bf.molecule_shells(DZVP, Z, R))
    r = oscf.rhf(ob, Z, R, lambda D: ob.jk_direct(D)[:2], e_tol=1e-10, fd_tol=1e-7)
    assert r["converged"]
    assert abs(r["energy"] - e_ref) < rel * abs(e_ref)
    assert abs(r["energy"] - e_ref) < 5e-6  # the anchor's last printed digit
