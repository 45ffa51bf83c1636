"""Full finite-difference check of the analytic gradient of a water cluster.

    python tests/fd_gradient_cluster.py [nwat] [h]

For (H2O)_nwat/6-31G (default 16 waters: 48 atoms, 144 coordinates) it
converges the SCF (accurate Boys, no screen, max|e| < 1e-10), takes
hfk_gradient, then central differences of the variational SCF energy at +-h
(default 1e-4 bohr) along every coordinate: 2 * 3 * natom displaced SCF runs, each
started from the undisplaced converged density.  Prints one line per atom
and the maximum deviation; exits non-zero above 1e-6 Eh/bohr.  Needs a GPU.
"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "haskell-abinitio_amd"))
import hfk  # noqa: E402
from hfk import systems  # noqa: E402

SCF = dict(converge=1e-10, diis_start=float("inf"), diis_size=8, max_steps=100)


def variational_energy(mol, D, H, F):
    """1/2 sum_ij D_ij (H_ij + F(D)_ij) + V_nn from flat upper triangles.
    scfHF's own energy uses the DIIS-extrapolated Fock matrix, which differs
    from F(D) to first order in the final residual (about 1e-9 Eh here); with
    F(D) the error is second order, so differences at h = 1e-4 bohr resolve
    the gradient to about 1e-8 Eh/bohr."""
    n = mol.nbf
    w = np.full(len(D), 2.0)
    w[np.cumsum(np.r_[0, np.arange(n, 1, -1)])] = 1.0   # diagonal entries count once
    return 0.5 * float(np.sum(w * D * (H + F))) + mol.nuclear_repulsion()


def main():
    nwat = int(sys.argv[1]) if len(sys.argv) > 1 else 16
    h = float(sys.argv[2]) if len(sys.argv) > 2 else 1e-4
    opts = hfk.options(screen_mode=hfk.HFK_SCREEN_NONE, boys_mode=hfk.HFK_BOYS_ACCURATE)
    mol = systems.water_cluster(nwat, "6-31g")
    print("(H2O)_%d 6-31G: %d atoms, %d functions, h = %g bohr" % (nwat, mol.natom, mol.nbf, h), flush=True)
    t0 = time.perf_counter()
    with hfk.Context(mol, opts) as ctx:
        r = ctx.scfHF(guess_density=ctx.fragment_guess(), **SCF)
        if r.max_error > 1e-9:
            sys.exit("SCF not converged: max|e| %.2e" % r.max_error)
        g = ctx.gradient(r.getDensity)
    D0 = r.getDensity
    print("E %.10f (%d steps), analytic gradient done in %.1f s" % (r.getEnergy, r.steps, time.perf_counter() - t0),
          flush=True)

    def energy(R):
        m = mol.with_coords(R)
        with hfk.Context(m, opts) as c:
            rr = c.scfHF(guess_density=D0, **SCF)
            if rr.max_error > 1e-9:
                sys.exit("displaced SCF not converged: max|e| %.2e" % rr.max_error)
            H = c.hcore()
            F = c.fock(H, rr.getDensity)
        return variational_energy(m, rr.getDensity, H, F)

    fd = np.zeros_like(g)
    for a in range(mol.natom):
        for k in range(3):
            R = mol.R.copy()
            R[a, k] += h
            ep = energy(R)
            R[a, k] -= 2 * h
            em = energy(R)
            fd[a, k] = (ep - em) / (2 * h)
        print("atom %3d %-2s analytic % .9f % .9f % .9f  max|analytic - fd| %.2e" % (
            a, mol.atoms[a][0], g[a, 0], g[a, 1], g[a, 2], np.max(np.abs(g[a] - fd[a]))), flush=True)
    err = np.max(np.abs(g - fd))
    print("max |analytic - fd| over %d coordinates: %.3e Eh/bohr (max |g| %.3e); sum_A g_A %.2e; %.1f s" % (
        g.size, err, np.max(np.abs(g)), np.max(np.abs(g.sum(axis=0))), time.perf_counter() - t0), flush=True)
    sys.exit(0 if err < 1e-6 else 1)


if __name__ == "__main__":
    main()
