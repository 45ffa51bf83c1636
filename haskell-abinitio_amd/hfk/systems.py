"""Workload geometries: the water monomer, seeded water clusters, the C12H26
alkane and a few textbook/sample molecules.  All coordinates in bohr.

Cluster generator for n waters (seed s, default s = n):
  rng = numpy.random.default_rng(s); lattice edge L = smallest integer with
  L^3 >= 2n; sites = sorted(rng.choice(L^3, n, replace=False)); for each site
  in order: offset = rng.uniform(-0.5, 0.5, 3), q = rng.standard_normal(4)
  normalised (a unit quaternion) rotating the monomer about the origin;
  position = 5.5 * (ix, iy, iz) + offset with site = (ix*L + iy)*L + iz.
  Atom order per water: O, H, H.
"""
import math

import numpy as np

from .basis import build_molecule

# Water monomer: O (0,0,0.2217), H (0,+-1.4309,-0.8867) bohr (OH 1.810, HOH 104.5)
WATER = (("O", (0.0, 0.0, 0.2217)), ("H", (0.0, 1.4309, -0.8867)), ("H", (0.0, -1.4309, -0.8867)))
SPACING = 5.5


def quat_matrix(q):
    q = np.asarray(q, dtype=np.float64)
    q = q / np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def lattice_edge(n):
    L = 1
    while L ** 3 < 2 * n:
        L += 1
    return L


def water_cluster_geometry(nwat, seed=None):
    seed = nwat if seed is None else seed
    rng = np.random.default_rng(seed)
    L = lattice_edge(nwat)
    sites = sorted(int(s) for s in rng.choice(L ** 3, nwat, replace=False))
    syms, xyz = [], []
    tmpl = np.array([p for _, p in WATER])
    for s in sites:
        ix, iy, iz = s // (L * L), (s // L) % L, s % L
        off = rng.uniform(-0.5, 0.5, 3)
        R = quat_matrix(rng.standard_normal(4))
        pos = SPACING * np.array([ix, iy, iz], dtype=np.float64) + off
        for (sym, _), r in zip(WATER, tmpl):
            syms.append(sym)
            xyz.append(R @ r + pos)
    return syms, np.array(xyz)


def water(basis="sto-3g"):
    syms = [s for s, _ in WATER]
    return build_molecule(syms, [p for _, p in WATER], basis)


def water_cluster(nwat, basis="6-31g", seed=None):
    syms, xyz = water_cluster_geometry(nwat, seed)
    return build_molecule(syms, xyz, basis)


def alkane_geometry(nc=12, seed=12, noise=0.05):
    """All-trans zig-zag CnH(2n+2): C-C 2.9102, C-H 2.0598 bohr, tetrahedral
    angles, carbons first then hydrogens; seeded uniform noise per coordinate."""
    cc, ch = 2.9102, 2.0598
    half = math.radians(109.4712206) / 2.0
    dx, dy = cc * math.sin(half), cc * math.cos(half)
    C = [np.array([i * dx, (i % 2) * dy, 0.0]) for i in range(nc)]
    H = []
    tet = math.radians(109.4712206)
    for i in range(nc):
        nbrs = [C[j] for j in (i - 1, i + 1) if 0 <= j < nc]
        u = sum((n - C[i]) / np.linalg.norm(n - C[i]) for n in nbrs)
        if len(nbrs) == 2:
            b = -u / np.linalg.norm(u)                  # bisector away from the chain
            z = np.array([0.0, 0.0, 1.0])
            for sgn in (1, -1):
                d = math.cos(tet / 2) * b + sgn * math.sin(tet / 2) * z
                H.append(C[i] + ch * d / np.linalg.norm(d))
        else:
            a = nbrs[0] - C[i]
            a = a / np.linalg.norm(a)
            # one H anti to the chain bond in the carbon plane, two out of plane
            perp = np.array([-a[1], a[0], 0.0])
            perp = perp / np.linalg.norm(perp)
            if np.dot(perp, C[i] - np.mean(C, axis=0)) < 0:
                perp = -perp
            ct = math.cos(math.pi - tet)
            d0 = ct * a + math.sqrt(1 - ct * ct) * perp
            H.append(C[i] + ch * d0)
            for sgn in (1, -1):
                d = ct * a - 0.5 * math.sqrt(1 - ct * ct) * perp + sgn * math.sqrt(0.75) * math.sqrt(1 - ct * ct) * np.array([0, 0, 1.0])
                H.append(C[i] + ch * d / np.linalg.norm(d))
    xyz = np.array(C + H)
    rng = np.random.default_rng(seed)
    xyz = xyz + rng.uniform(-noise, noise, xyz.shape)
    return ["C"] * nc + ["H"] * len(H), xyz


def alkane(nc=12, basis="6-31g*", seed=12):
    syms, xyz = alkane_geometry(nc, seed)
    return build_molecule(syms, xyz, basis)


def sto3g_s(zeta):
    """STO-3G 1s contraction for Slater exponent zeta (Szabo & Ostlund 3.5)."""
    base = ((0.444635, 0.109818), (0.535328, 0.405771), (0.154329, 2.22766))
    return [(c, a * zeta * zeta) for c, a in base]


def custom_molecule(atoms):
    """atoms: [(symbol, xyz_bohr, [(funtype, [(c, e)])])], primitives normalised here."""
    from .basis import Molecule, norm_coeff
    out = []
    for sym, xyz, cgfs in atoms:
        out.append((sym, tuple(xyz), [(ft, norm_coeff(prims, ft)) for ft, prims in cgfs]))
    return Molecule(out)


def h2_szabo(R=1.4, zeta=1.24):
    return custom_molecule([("H", (0, 0, 0), [("S", sto3g_s(zeta))]), ("H", (0, 0, R), [("S", sto3g_s(zeta))])])


def heh_szabo(R=1.4632, zeta_he=2.0925, zeta_h=1.24):
    m = custom_molecule([("He", (0, 0, 0), [("S", sto3g_s(zeta_he))]), ("H", (0, 0, R), [("S", sto3g_s(zeta_h))])])
    m.charge = 1.0
    return m
