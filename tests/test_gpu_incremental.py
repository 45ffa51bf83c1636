"""Incremental Fock builds (hfk_scf_options.incremental_tau): the SCF built
from density changes, with Schwarz x max|dD| screening and periodic full
rebuilds, converges to the energy of the full build every step."""
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


@pytest.mark.parametrize("tau,rebuild", [(1e-12, 8), (1e-10, 4)])
def test_incremental_matches_full_build(require_gpu, tau, rebuild):
    _say("case: (H2O)8 6-31g incremental tau=%g rebuild=%d vs full builds" % (tau, rebuild))
    mol = systems.water_cluster(8, "6-31g")
    with hfk.Context(mol) as ctx:
        Dg = ctx.fragment_guess()
        full = ctx.scfHF(guess_density=Dg, **SCF)
        inc = ctx.scfHF(guess_density=Dg, incremental_tau=tau, rebuild_every=rebuild, **SCF)
    _say("  full %d steps E %.12f; incremental %d steps E %.12f" % (full.steps, full.getEnergy, inc.steps,
                                                                   inc.getEnergy))
    assert full.max_error < 1e-9 and inc.max_error < 1e-9
    assert abs(inc.getEnergy - full.getEnergy) < 1e-8
    assert np.max(np.abs(inc.getDensity - full.getDensity)) < 1e-6


def test_incremental_rejects_negative_tau(require_gpu):
    mol = systems.water("sto-3g")
    with hfk.Context(mol) as ctx:
        with pytest.raises(hfk.HfkError):
            ctx.scfHF(incremental_tau=-1.0)
