"""Basis-set ingestion: NWChem text -> normalised Cartesian CGFs.

Mirrors the reference's ingestion semantics:

* primitives are (coefficient, exponent) pairs (ParsecTools/ParserBasis.hs:107);
* each primitive coefficient is divided by the primitive norm
  sqrt(prod_k (2k-1)!! / (4e)^L * (pi/2e)^1.5) (BasisSet/NormalizeBasis.hs:52-59);
* no contracted renormalisation happens: normGlobal's factor evaluates to 1
  because it builds the probe CGF with (exponent, coefficient) swapped and
  sijContracted short-circuits self overlap to 1.0
  (BasisSet/NormalizeBasis.hs:37-45, Integrals/IntegralsEvaluation.hs:183-184);
* CGF order per shell follows FetchBasis.hs:129-137: SP -> S,Px,Py,Pz;
  P -> Px,Py,Pz; D -> Dxx,Dxy,Dxz,Dyy,Dyz,Dzz.

Unlike ParserBasis.hs:116-118 (`triple` packs each SP primitive as
(c_s, c_p, 1), so every SP exponent becomes 1.0), SP blocks keep their
exponent; sp_exponent_bug=True reproduces the reference's behaviour.
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# Funtype -> Cartesian exponents (GlobalTypes.hs:302-323)
LXYZ = {
    "S": (0, 0, 0),
    "Px": (1, 0, 0), "Py": (0, 1, 0), "Pz": (0, 0, 1),
    "Dxx": (2, 0, 0), "Dxy": (1, 1, 0), "Dxz": (1, 0, 1),
    "Dyy": (0, 2, 0), "Dyz": (0, 1, 1), "Dzz": (0, 0, 2),
}
SHELL_FUNCS = {
    "S": ("S",),
    "P": ("Px", "Py", "Pz"),
    "D": ("Dxx", "Dxy", "Dxz", "Dyy", "Dyz", "Dzz"),   # FetchBasis.hs:137
}
# SampleProjects.hs:98 orders a d shell xx,yy,zz,xy,xz,yz
D_ORDER_SAMPLE = ("Dxx", "Dyy", "Dzz", "Dxy", "Dxz", "Dyz")

Z_OF = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8,
        "F": 9, "Ne": 10, "Na": 11, "Mg": 12}          # GlobalTypes.hs:337-341

BOHR_PER_ANGSTROM = 1.0 / 0.529177                       # Units.hs:13


def parse_nwchem(text):
    """Parse an EMSL/NWChem 'ao basis' block.

    Returns {element: [(shell_type, [(c, e)] or [(cs, cp, e)]), ...]} with
    shell_type one of 'S', 'SP', 'P', 'D' (ParserBasis.hs:190-220, with the SP
    exponent kept).
    """
    lines = text.splitlines()
    start = None
    for i, ln in enumerate(lines):
        if ln.strip().startswith('BASIS "ao basis"'):
            start = i + 1
            break
    if start is None:
        raise ValueError("no 'BASIS \"ao basis\"' block")
    out = {}
    cur = None
    for ln in lines[start:]:
        s = ln.strip()
        if not s or s.startswith("#"):
            continue
        if s.upper() == "END":
            break
        tok = s.split()
        if tok[0].isalpha() and len(tok) == 2 and tok[1].isalpha():
            el, kind = tok[0], tok[1].upper()
            cur = (kind, [])
            out.setdefault(el, []).append(cur)
            continue
        vals = [float(t.replace("D", "E")) for t in tok]
        kind = cur[0]
        if kind == "SP":
            e, cs, cp = vals
            cur[1].append((cs, cp, e))
        else:
            e, c = vals
            cur[1].append((c, e))
    return out


def _fac_odd(i):
    # facOdd (Integrals/IntegralsEvaluation.hs:530-535) for the small odd
    # arguments normalisation uses: (2k-1)!! with k <= 2.
    if i < 2:
        return 1.0
    k = (1 + i) // 2
    return math.factorial(2 * k) / (2.0 ** k * math.factorial(k))


def norm_coeff(prims, funtype):
    """normCoeff: NormalizeBasis.hs:52-59."""
    idx = LXYZ[funtype]
    prod = 1.0
    for k in idx:
        prod *= _fac_odd(2 * k - 1)
    L = sum(idx)
    out = []
    for c, e in prims:
        ang = prod / (4.0 * e) ** L
        out.append((c / math.sqrt(ang * (math.pi / (2.0 * e)) ** 1.5), e))
    return out


def load_basis_sets():
    with open(os.path.join(HERE, "data", "basis_sets.json")) as f:
        return json.load(f)


def element_cgfs(shells, d_order=SHELL_FUNCS["D"], sp_exponent_bug=False):
    """createCGF (FetchBasis.hs:128-137): shells -> list of (funtype, prims).
    sp_exponent_bug: every SP primitive gets exponent 1.0, as the reference's
    parser stores it (ParserBasis.hs:116-118)."""
    cgfs = []
    for kind, prims in shells:
        if kind == "SP":
            s = [(cs, 1.0 if sp_exponent_bug else e) for cs, cp, e in prims]
            p = [(cp, 1.0 if sp_exponent_bug else e) for cs, cp, e in prims]
            cgfs.append(("S", norm_coeff(s, "S")))
            for ft in SHELL_FUNCS["P"]:
                cgfs.append((ft, norm_coeff(p, ft)))
        elif kind == "S":
            cgfs.append(("S", norm_coeff(prims, "S")))
        elif kind == "P":
            for ft in SHELL_FUNCS["P"]:
                cgfs.append((ft, norm_coeff(prims, ft)))
        elif kind == "D":
            for ft in d_order:
                cgfs.append((ft, norm_coeff(prims, ft)))
        else:
            raise ValueError("unsupported shell type %r" % kind)
    return cgfs


class Molecule:
    """[AtomData] in array form (GlobalTypes.hs:92-105).

    atoms: list of (symbol, (x, y, z) in bohr, [(funtype, [(c, e)])]).
    Arrays follow the C-ABI basis contract: one row per reference CGF in
    [AtomData] order and getBasis order.
    """

    def __init__(self, atoms, charge=0.0):
        self.atoms = atoms
        self.charge = float(charge)
        center, lxyz, off, coef, expo, atom_of = [], [], [0], [], [], []
        for ia, (_sym, xyz, cgfs) in enumerate(atoms):
            for ft, prims in cgfs:
                center.append(xyz)
                lxyz.append(LXYZ[ft])
                for c, e in prims:
                    coef.append(c)
                    expo.append(e)
                off.append(len(coef))
                atom_of.append(ia)
        self.center = np.ascontiguousarray(center, dtype=np.float64).reshape(-1, 3)
        self.lxyz = np.ascontiguousarray(lxyz, dtype=np.int32).reshape(-1, 3)
        self.prim_off = np.ascontiguousarray(off, dtype=np.int32)
        self.coef = np.ascontiguousarray(coef, dtype=np.float64)
        self.expo = np.ascontiguousarray(expo, dtype=np.float64)
        self.atom_of_bf = np.ascontiguousarray(atom_of, dtype=np.int32)
        self.Z = np.ascontiguousarray([Z_OF[a[0]] for a in atoms], dtype=np.float64)
        self.R = np.ascontiguousarray([a[1] for a in atoms], dtype=np.float64).reshape(-1, 3)

    @property
    def nbf(self):
        return self.center.shape[0]

    @property
    def natom(self):
        return len(self.atoms)

    @property
    def occupied(self):
        # scfHF: floor ((sum Z - charge) / 2)   (HartreeFock.hs:75-76)
        return int(math.floor((self.Z.sum() - self.charge) / 2.0))

    def nuclear_repulsion(self):
        """nuclearRep (HartreeFock.hs:300-305): each atom pair once."""
        e = 0.0
        for a in range(self.natom):
            for b in range(a + 1, self.natom):
                rab = math.sqrt(float(((self.R[a] - self.R[b]) ** 2).sum()))
                e += self.Z[a] * self.Z[b] / rab
        return e

    def with_coords(self, R):
        """Same basis, new nuclear positions (bohr): for gradients / FD."""
        R = np.asarray(R, dtype=np.float64).reshape(-1, 3)
        atoms = [(sym, tuple(R[i]), cgfs) for i, (sym, _xyz, cgfs) in enumerate(self.atoms)]
        return Molecule(atoms, self.charge)


def build_molecule(symbols, coords_bohr, basis_name, charge=0.0, d_order=SHELL_FUNCS["D"], sp_exponent_bug=False):
    sets = load_basis_sets()[basis_name.lower()]
    atoms = []
    for sym, xyz in zip(symbols, coords_bohr):
        cgfs = element_cgfs([(k, [tuple(p) for p in prims]) for k, prims in sets[sym]], d_order, sp_exponent_bug)
        atoms.append((sym, tuple(float(v) for v in xyz), cgfs))
    return Molecule(atoms, charge)


def parse_xyz(text):
    """parserXYZ (ParsecTools/ParseXYZ.hs:33-43): Angstrom -> bohr (Units.hs:13).

    As in the reference: leading whitespace, the atom count, the rest of
    that line and then exactly one more line (the comment, which may be
    empty) are consumed; then n atoms follow, each a symbol of ASCII letters
    and three numbers, with any whitespace (blank lines included) between
    them (ParseUtils.hs:41-57).  Fewer than n atoms raise ValueError."""
    s = text.lstrip()
    head, _, rest = s.partition("\n")
    tok = head.split()
    if not tok or not tok[0].isdigit():
        raise ValueError("parse_xyz: no atom count on the first line")
    n = int(tok[0])
    _comment, _, body = rest.partition("\n")
    t = body.split()
    if len(t) < 4 * n:
        raise ValueError("parse_xyz: %d atoms declared, %d found" % (n, len(t) // 4))
    syms, xyz = [], []
    for k in range(n):
        sym = t[4 * k]
        if not (sym.isascii() and sym.isalpha()):
            raise ValueError("parse_xyz: atom %d has symbol %r" % (k, sym))
        syms.append(sym)
        xyz.append([float(v) / 0.529177 for v in t[4 * k + 1:4 * k + 4]])
    return syms, xyz
