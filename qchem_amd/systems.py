"""Seeded synthetic molecules and basis sets for the benchmark and test workloads.

Every generator is deterministic: it uses a splitmix64 stream (Steele, Lea & Flood, OOPSLA 2014),
seeded as ``20261202 + workload index`` unless a seed is given.

Workloads:

* ``n2_even_tempered``: N2 at z = +-1.037 bohr.  Each atom carries uncontracted even-tempered
  Cartesian shells: 8 s (0.1 * 2.5^k), 5 p (0.15 * 2.5^k) and 2 d (0.4, 1.2), giving 70
  functions.  The atom positions follow the reference CLI's N2 (CLIapps/scfrun.C:134).
* ``random_cluster``: 16 atoms, Z uniform in {1, 6, 7, 8}, positions uniform in a 12-bohr cube
  with rejection below 2.0 bohr.  An odd electron count is made even by removing one H, or by
  adding one if there is none.  Heavy atoms get 3s2p1d and H gets 2s1p.  Each shell is a
  3-primitive contraction with exponents log-uniform in [0.1, 50] and N(0,1) coefficients,
  renormalised by the contracted-radial convention.
* ``alkane_chain``: a C16H32-like zig-zag chain (C-C 2.91, C-H 2.06 bohr, tetrahedral) with
  N(0, 0.05^2) bohr jitter.  C uses the N2 rules.  H uses 4 s (0.1 * 2.5^k) and 1 p (0.15),
  giving 784 functions.
* ``batch``: ``random_cluster`` rules with 2 to 12 atoms per molecule.
* ``acetone_like``: a C3H6O-like 10-atom structure displaced by N(0, 0.15^2) bohr, with the N2
  rules on C/O and the alkane-chain H rule.  That gives 182 functions.
"""
import math

SEED0 = 20261202


class SplitMix64:
    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = seed & self.MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def normal(self):
        u1 = self.uniform()
        while u1 <= 0.0:
            u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def choice(self, seq):
        return seq[int(self.uniform() * len(seq)) % len(seq)]

    def randint(self, lo, hi):  # inclusive
        return lo + int(self.uniform() * (hi - lo + 1)) % (hi - lo + 1)


def even_tempered_shells(Z, center):
    """N2 rules for heavy atoms (uncontracted 8s5p2d) and the 4s1p H rule."""
    out = []
    if Z == 1:
        for k in range(4):
            out.append(dict(L=0, exps=[0.1 * 2.5 ** k], coefs=[1.0], center=list(center), contracted=False))
        out.append(dict(L=1, exps=[0.15], coefs=[1.0], center=list(center), contracted=False))
        return out
    for k in range(8):
        out.append(dict(L=0, exps=[0.1 * 2.5 ** k], coefs=[1.0], center=list(center), contracted=False))
    for k in range(5):
        out.append(dict(L=1, exps=[0.15 * 2.5 ** k], coefs=[1.0], center=list(center), contracted=False))
    for a in (0.4, 1.2):
        out.append(dict(L=2, exps=[a], coefs=[1.0], center=list(center), contracted=False))
    return out


def random_contracted_shells(Z, center, rng, nprim=3):
    """3s2p1d (heavy) or 2s1p (H), each a 3-primitive contraction, log-uniform exponents in [0.1, 50]."""
    spec = [0, 0, 1] if Z == 1 else [0, 0, 0, 1, 1, 2]
    out = []
    for L in spec:
        exps = [math.exp(rng.uniform(math.log(0.1), math.log(50.0))) for _ in range(nprim)]
        coefs = [rng.normal() for _ in range(nprim)]
        out.append(dict(L=L, exps=exps, coefs=coefs, center=list(center), contracted=True))
    return out


def _dist2(a, b):
    return sum((x - y) ** 2 for x, y in zip(a, b))


def _place(rng, atoms_xyz, edge, rmin=2.0, tries=100000):
    for _ in range(tries):
        p = [rng.uniform(0.0, edge) for _ in range(3)]
        if all(_dist2(p, q) >= rmin * rmin for q in atoms_xyz):
            return p
    raise RuntimeError("could not place atom")


def random_molecule(natom, rng, edge=None):
    """Random cluster rules: Z in {1,6,7,8}, 2.0-bohr rejection, even electron count."""
    if edge is None:
        edge = 12.0 * (natom / 16.0) ** (1.0 / 3.0)
    Zs, xyz = [], []
    for _ in range(natom):
        Zs.append(rng.choice([1, 6, 7, 8]))
        xyz.append(_place(rng, xyz, edge))
    if sum(Zs) % 2:
        if 1 in Zs:
            i = len(Zs) - 1 - Zs[::-1].index(1)
            del Zs[i]
            del xyz[i]
        else:
            Zs.append(1)
            xyz.append(_place(rng, xyz, edge))
    return Zs, xyz


def n2_even_tempered():
    Zs = [7, 7]
    xyz = [[0.0, 0.0, -1.037], [0.0, 0.0, 1.037]]
    shells = [s for Z, c in zip(Zs, xyz) for s in even_tempered_shells(Z, c)]
    return dict(name="N2 even-tempered 8s5p2d (70 functions)", Z=Zs, R=xyz, shells=shells, charge=0)


def random_cluster(natom=16, seed=None):
    rng = SplitMix64(SEED0 + 1 if seed is None else seed)
    Zs, xyz = random_molecule(natom, rng, edge=12.0)
    shells = [s for Z, c in zip(Zs, xyz) for s in random_contracted_shells(Z, c, rng)]
    return dict(name="random %d-atom cluster, 3s2p1d/2s1p K=3" % len(Zs), Z=Zs, R=xyz, shells=shells, charge=0)


def alkane_chain(ncarbon=16, seed=None, jitter=0.05):
    rng = SplitMix64(SEED0 + 2 if seed is None else seed)
    dcc, dch = 2.91, 2.06
    half = math.radians(109.4712206) / 2.0
    cs = []
    for i in range(ncarbon):
        cs.append([i * dcc * math.sin(half), (i % 2) * dcc * math.cos(half), 0.0])
    Zs, xyz = [], []
    for i, c in enumerate(cs):
        Zs.append(6)
        xyz.append(list(c))
        nb = [cs[j] for j in (i - 1, i + 1) if 0 <= j < ncarbon]
        mid = [sum(p[k] for p in nb) / len(nb) for k in range(3)]
        u = [c[k] - mid[k] for k in range(3)]
        nu = math.sqrt(sum(x * x for x in u)) or 1.0
        u = [x / nu for x in u]
        for sgn in (1.0, -1.0):
            d = [u[0] * math.cos(half), u[1] * math.cos(half), sgn * math.sin(half)]
            Zs.append(1)
            xyz.append([c[k] + dch * d[k] for k in range(3)])
    xyz = [[x + jitter * rng.normal() for x in p] for p in xyz]
    shells = [s for Z, c in zip(Zs, xyz) for s in even_tempered_shells(Z, c)]
    return dict(name="C%dH%d zig-zag chain (%d functions)" % (ncarbon, 2 * ncarbon, nbf(shells)), Z=Zs, R=xyz,
                shells=shells, charge=0)


def batch(nmol=512, seed=None, lo=2, hi=12):
    rng = SplitMix64(SEED0 + 3 if seed is None else seed)
    mols = []
    for _ in range(nmol):
        na = rng.randint(lo, hi)
        Zs, xyz = random_molecule(na, rng)
        shells = [s for Z, c in zip(Zs, xyz) for s in random_contracted_shells(Z, c, rng)]
        mols.append(dict(Z=Zs, R=xyz, shells=shells, charge=0))
    return mols


# C3H6O-like reference structure (acetone-shaped), bohr
_ACETONE = [
    (8, [0.0, 0.0, 2.28]),
    (6, [0.0, 0.0, 0.0]),
    (6, [2.43, 0.0, -1.53]),
    (6, [-2.43, 0.0, -1.53]),
    (1, [4.07, 0.0, -0.29]),
    (1, [2.45, 1.68, -2.72]),
    (1, [2.45, -1.68, -2.72]),
    (1, [-4.07, 0.0, -0.29]),
    (1, [-2.45, 1.68, -2.72]),
    (1, [-2.45, -1.68, -2.72]),
]


def acetone_like(index=0, seed=None, sigma=0.15):
    rng = SplitMix64((SEED0 + 4 if seed is None else seed) * 1000003 + index)
    Zs = [z for z, _ in _ACETONE]
    xyz = [[x + sigma * rng.normal() for x in r] for _, r in _ACETONE]
    shells = [s for Z, c in zip(Zs, xyz) for s in even_tempered_shells(Z, c)]
    return dict(name="C3H6O-like geometry %d" % index, Z=Zs, R=xyz, shells=shells, charge=0)


def nbf(shells):
    return sum((s["L"] + 1) * (s["L"] + 2) // 2 for s in shells)
