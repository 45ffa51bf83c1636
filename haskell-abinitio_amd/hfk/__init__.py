"""hfk: MI355X restricted Hartree-Fock integral engine (Python host side).

Mirrors the Repa-facing functions of Haskell-abinitio's RHF path on top of
the C ABI of libhfk.so (include/hfk.h):

=====================  ===================================================
hfk                    reference (paths relative to the reference tree)
=====================  ===================================================
Context.calcIntegrals  calcIntegrals   HartreeFock/HartreeFock.hs:166-181
Context.calcGmatrix    calcGmatrix     HartreeFock/HartreeFock.hs:225-241
Context.fock           fock            HartreeFock/HartreeFock.hs:213-222
Context.calcDensity    calcDensity     HartreeFock/HartreeFock.hs:195-208
Context.symmOrtho      symmOrtho       HartreeFock/BasisOrthogonalization.hs:26-35
Context.unitaryTransf  unitaryTransf   NumericalTools/LinearAlgebra.hs:155-159
Context.mtxOverlap     mtxOverlap      Integrals/IntegralsEvaluation.hs:169-177
Context.hcore          hcore           Integrals/IntegralsEvaluation.hs:227-240
Context.scfHF          scfHF           HartreeFock/HartreeFock.hs:70-123
=====================  ===================================================

Flat matrices are upper-triangular row-major (LinearAlgebra.hs:240-253);
square matrices are row-major numpy arrays.  Errors raise HfkError (the
reference throws HSFOCKException / calls error).
"""
import ctypes as C
from collections import namedtuple

import numpy as np

from . import _lib
from ._lib import (HFK_BOYS_ACCURATE, HFK_BOYS_REFERENCE, HFK_GUESS_CORE, HFK_GUESS_DENSITY, HFK_GUESS_INTENDED,
                   HFK_GUESS_REFERENCE,
                   HFK_SCREEN_NONE, HFK_SCREEN_REFERENCE, HfkError, check)
from .basis import Molecule, build_molecule
from . import guess as _guess

__all__ = [
    "Context", "HFData", "HfkError", "Molecule", "build_molecule", "options",
    "HFK_BOYS_ACCURATE", "HFK_BOYS_REFERENCE", "HFK_SCREEN_NONE", "HFK_SCREEN_REFERENCE",
    "HFK_GUESS_CORE", "HFK_GUESS_DENSITY", "HFK_GUESS_INTENDED", "HFK_GUESS_REFERENCE", "flat_len", "eri_flat_len",
]

# HFData (GlobalTypes.hs:122-128) plus run statistics
HFData = namedtuple("HFData", "getFock getCoeff getDensity getOrbE getEnergy steps max_error seconds fock_seconds "
                               "delta_energy")

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def flat_len(n):
    return n * (n + 1) // 2


def eri_flat_len(n):
    m = flat_len(n)
    return m * (m + 1) // 2


def options(screen_mode=HFK_SCREEN_REFERENCE, boys_mode=HFK_BOYS_REFERENCE, screen_tau=1e-8,
            schwarz_cut=0.0, fp32_threshold=0.0):
    o = _lib.HfkOptions()
    _lib.lib().hfk_default_options(C.byref(o))
    o.screen_mode = screen_mode
    o.boys_mode = boys_mode
    o.screen_tau = screen_tau
    o.schwarz_cut = schwarz_cut
    o.fp32_threshold = fp32_threshold
    return o


class Context:
    """An hfk_ctx bound to one molecule/basis (one HIP device, one stream)."""

    def __init__(self, mol, opts=None):
        L = _lib.lib()
        self.mol = mol
        self.n = mol.nbf
        self._keep = [mol.center, mol.lxyz, mol.prim_off, mol.coef, mol.expo, mol.atom_of_bf, mol.Z, mol.R]
        b = _lib.HfkBasis()
        b.nbf = mol.nbf
        b.nprim_total = int(mol.prim_off[-1])
        b.natom = mol.natom
        b.center = _ptr(mol.center)
        b.lxyz = mol.lxyz.ctypes.data_as(C.POINTER(C.c_int32))
        b.prim_off = mol.prim_off.ctypes.data_as(C.POINTER(C.c_int32))
        b.coef = _ptr(mol.coef)
        b.expo = _ptr(mol.expo)
        b.atom_of_bf = mol.atom_of_bf.ctypes.data_as(C.POINTER(C.c_int32))
        b.Z = _ptr(mol.Z)
        b.R = _ptr(mol.R)
        self._opts = opts if opts is not None else options()
        h = C.c_void_p()
        check(L.hfk_create(C.byref(b), C.byref(self._opts), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hfk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_options(self, opts):
        check(_lib.lib().hfk_set_options(self._h, C.byref(opts)))
        self._opts = opts

    # --- the reference's functions ---------------------------------------
    def calcIntegrals(self):
        out = np.zeros(eri_flat_len(self.n), dtype=np.float64)
        check(_lib.lib().hfk_eri_flat(self._h, _ptr(out), out.size))
        return out

    def calcGmatrix(self, D_flat):
        D = np.ascontiguousarray(D_flat, dtype=np.float64)
        G = np.zeros(flat_len(self.n), dtype=np.float64)
        check(_lib.lib().hfk_fock_G(self._h, _ptr(D), _ptr(G)))
        return G

    def fock(self, H_flat, D_flat):
        H = np.ascontiguousarray(H_flat, dtype=np.float64)
        D = np.ascontiguousarray(D_flat, dtype=np.float64)
        F = np.zeros(flat_len(self.n), dtype=np.float64)
        check(_lib.lib().hfk_fock(self._h, _ptr(H), _ptr(D), _ptr(F)))
        return F

    def fock_dev(self, H_ptr, D_ptr, F_ptr):
        """Device-resident Fock build on raw device pointers (ints)."""
        check(_lib.lib().hfk_fock_dev(self._h, H_ptr, D_ptr, F_ptr))

    def calcDensity(self, Cmo, occ):
        Cm = np.ascontiguousarray(Cmo, dtype=np.float64)
        D = np.zeros(flat_len(self.n), dtype=np.float64)
        check(_lib.lib().hfk_density(self._h, _ptr(Cm), int(occ), _ptr(D)))
        return D

    def symmOrtho(self, S):
        S = np.ascontiguousarray(S, dtype=np.float64)
        X = np.zeros((self.n, self.n), dtype=np.float64)
        check(_lib.lib().hfk_symm_ortho(self._h, _ptr(S), _ptr(X)))
        return X

    def unitaryTransf(self, X, F_flat):
        X = np.ascontiguousarray(X, dtype=np.float64)
        F = np.ascontiguousarray(F_flat, dtype=np.float64)
        Fp = np.zeros((self.n, self.n), dtype=np.float64)
        check(_lib.lib().hfk_unitary(self._h, _ptr(X), _ptr(F), _ptr(Fp)))
        return Fp

    def mtxOverlap(self):
        S = np.zeros(flat_len(self.n), dtype=np.float64)
        check(_lib.lib().hfk_overlap(self._h, _ptr(S)))
        return S

    def hcore(self):
        H = np.zeros(flat_len(self.n), dtype=np.float64)
        check(_lib.lib().hfk_hcore(self._h, _ptr(H)))
        return H

    def scfHF(self, charge=None, guess=HFK_GUESS_CORE, max_steps=30, diis_size=5, converge=1e-7, diis_start=0.1,
              verbose=False, guess_density=None, incremental_tau=0.0, rebuild_every=16, energy_converge=0.0):
        """scfHF.  Defaults are the reference's (30 steps, 5 DIIS vectors,
        DIIS switched on once max|e| < 0.1, converged at max|e| < 1e-7);
        diis_start=float('inf') switches DIIS on from the second step.
        guess_density: flat starting density (overrides guess).
        incremental_tau > 0: Fock builds from the density change, skipping
        shell quartets with Q_ab Q_cd max|dD| < tau, with a full build every
        rebuild_every steps (the reference rebuilds every step).
        energy_converge > 0: stop only when the last step's energy change is
        also below it (the reference tests max|e| alone)."""
        o = _lib.HfkScfOptions()
        _lib.lib().hfk_default_scf_options(C.byref(o))
        o.guess = guess
        if guess_density is not None:
            Dg = np.ascontiguousarray(guess_density, dtype=np.float64)
            if Dg.shape != (flat_len(self.n),):
                raise ValueError("guess_density has shape %s, expected (%d,)" % (Dg.shape, flat_len(self.n)))
            o.guess = HFK_GUESS_DENSITY
            o.guess_density = _ptr(Dg)
        o.max_steps = max_steps
        o.diis_size = diis_size
        o.converge = converge
        o.diis_start = diis_start
        o.verbose = 1 if verbose else 0
        o.incremental_tau = float(incremental_tau)
        o.rebuild_every = int(rebuild_every)
        o.energy_converge = float(energy_converge)
        o.charge = self.mol.charge if charge is None else float(charge)
        r = _lib.HfkScfResult()
        n = self.n
        Cm = np.zeros((n, n))
        eps = np.zeros(n)
        D = np.zeros(flat_len(n))
        F = np.zeros(flat_len(n))
        check(_lib.lib().hfk_scf(self._h, C.byref(o), C.byref(r), _ptr(Cm), _ptr(eps), _ptr(D), _ptr(F)))
        return HFData(F, Cm, D, eps, r.energy, r.steps, r.max_error, r.seconds, r.fock_seconds, r.delta_energy)

    def fragment_guess(self, **scf_kw):
        """Block-diagonal starting density from SCF runs on each bonded
        fragment (same options as this context), or None for a single
        fragment; pass it to scfHF(guess_density=...)."""
        return _guess.fragment_guess(self.mol, lambda sub: Context(sub, self._opts), **scf_kw)

    def gradient(self, D_flat, W_flat=None):
        """Nuclear gradient dE/dR (natom x 3, Eh/bohr) at density D_flat.
        W_flat: energy-weighted density; None computes 1/2 D F D, which is
        W at SCF convergence."""
        D = np.ascontiguousarray(D_flat, dtype=np.float64)
        g = np.zeros((self.mol.natom, 3))
        if W_flat is None:
            check(_lib.lib().hfk_gradient(self._h, _ptr(D), None, _ptr(g)))
        else:
            W = np.ascontiguousarray(W_flat, dtype=np.float64)
            check(_lib.lib().hfk_gradient(self._h, _ptr(D), _ptr(W), _ptr(g)))
        return g

    def class_stats(self):
        """Per-launch records of the last Fock build: class, precision,
        surviving shell/primitive/function quartets, kernel ms."""
        n = C.c_int32(0)
        _lib.lib().hfk_get_class_stats(self._h, None, 0, C.byref(n))
        arr = (_lib.HfkClassStat * max(1, n.value))()
        check(_lib.lib().hfk_get_class_stats(self._h, arr, n.value, C.byref(n)))
        return [{k: getattr(arr[i], k) for k, _ in arr[i]._fields_} for i in range(n.value)]

    def stats(self):
        s = _lib.HfkStats()
        check(_lib.lib().hfk_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}
