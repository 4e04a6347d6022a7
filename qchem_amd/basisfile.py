"""Gaussian94 basis-set files (.bsd) -> shell lists for ``qchem_amd.Basis``.

Follows the reference reader (Readers/Imp/Gaussian94.C:28-225): the atom block starts with the
upper-case element symbol and a charge field and ends at ``****``; each shell line is the angular
momentum letter(s), the primitive count and a scale factor (read and ignored); a one-primitive shell
is a single-exponent radial (coefficient 1), a longer one a contracted radial whose coefficients
are used as given (the primitive norms are folded in by ``Basis``).  Shells sharing letters (SP)
must share coefficients and yield the highest letter, as there.
"""
import re

_SYMBOLS = {1: "H", 2: "HE", 3: "LI", 4: "BE", 5: "B", 6: "C", 7: "N", 8: "O", 9: "F", 10: "NE"}
_L = {"S": 0, "P": 1, "D": 2, "F": 3}


def _num(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def read_atom(path, Z):
    """Shell tuples (L, exps, coefs, contracted) of element Z in a Gaussian94 file."""
    sym = _SYMBOLS[Z]
    lines = open(path).read().split("\n")
    start = next(i for i, l in enumerate(lines) if l.strip() == "!")  # body starts after the "!" line
    i = start + 2
    while i < len(lines):
        head = lines[i].split()
        if len(head) >= 2 and head[0].upper() == sym and re.fullmatch(r"-?\d+", head[1]):
            break
        while i < len(lines) and lines[i].strip() != "****":
            i += 1
        i += 1
    else:
        raise KeyError("element %s not in %s" % (sym, path))
    shells = []
    i += 1
    while lines[i].strip() != "****":
        tok = lines[i].split()
        letters, nprim = tok[0].upper(), int(tok[1])
        Ls = sorted(_L[c] for c in letters)
        exps, coefs = [], []
        for k in range(nprim):
            row = lines[i + 1 + k].split()
            exps.append(_num(row[0]))
            cs = [_num(x) for x in row[1:1 + len(Ls)]]
            if len(set(cs)) > 1 and nprim > 1:
                raise ValueError("contraction coefficients vary with L (not handled, as in the reference)")
            coefs.append(cs[0])
        shells.append((Ls[-1], exps, coefs if nprim > 1 else [1.0], nprim > 1))
        i += 1 + nprim
    return shells


def molecule_shells(path, Z, R):
    """Shell dicts (qchem_amd.Basis format) for atoms Z at positions R (bohr), in atom order."""
    out = []
    for z, r in zip(Z, R):
        for L, exps, coefs, contracted in read_atom(path, int(z)):
            out.append(dict(L=L, exps=list(exps), coefs=list(coefs), center=list(r), contracted=contracted))
    return out
