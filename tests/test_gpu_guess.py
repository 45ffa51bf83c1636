"""Starting densities: HFK_GUESS_DENSITY (caller-supplied D) and the
fragment guess built on it converge to the same SCF state as the core
guess, in fewer steps for a cluster."""
import sys

import numpy as np
import pytest

import hfk
from hfk import systems

pytestmark = pytest.mark.gpu

SCF = dict(converge=1e-9, diis_start=float("inf"), diis_size=8, max_steps=100)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_guess_density_converged_start(require_gpu):
    """Starting from the converged density the SCF stops after one step at
    the same energy."""
    _say("case: water 6-31g, SCF restarted from its own converged density")
    mol = systems.water("6-31g")
    with hfk.Context(mol) as ctx:
        r0 = ctx.scfHF(**SCF)
        r1 = ctx.scfHF(guess_density=r0.getDensity, **SCF)
    assert r1.steps <= 2, r1.steps
    assert abs(r1.getEnergy - r0.getEnergy) < 1e-10
    assert np.max(np.abs(r1.getDensity - r0.getDensity)) < 1e-7


def test_fragment_guess_water_cluster(require_gpu):
    _say("case: (H2O)8 6-31g, fragment guess vs core guess")
    mol = systems.water_cluster(8, "6-31g")
    with hfk.Context(mol) as ctx:
        r_core = ctx.scfHF(**SCF)
        Dg = ctx.fragment_guess()
        assert Dg is not None and Dg.shape == (hfk.flat_len(mol.nbf),)
        r_frag = ctx.scfHF(guess_density=Dg, **SCF)
    _say("  core: %d steps E %.10f; fragment: %d steps E %.10f" % (
        r_core.steps, r_core.getEnergy, r_frag.steps, r_frag.getEnergy))
    assert r_frag.max_error < 1e-9 and r_core.max_error < 1e-9
    assert abs(r_frag.getEnergy - r_core.getEnergy) < 1e-8
    assert r_frag.steps < r_core.steps


def test_fragment_guess_single_fragment(require_gpu):
    mol = systems.water("sto-3g")
    with hfk.Context(mol) as ctx:
        assert ctx.fragment_guess() is None


def test_guess_density_shape_checked(require_gpu):
    mol = systems.water("sto-3g")
    with hfk.Context(mol) as ctx:
        with pytest.raises(ValueError):
            ctx.scfHF(guess_density=np.zeros(3))
