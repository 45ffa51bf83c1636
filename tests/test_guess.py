"""Fragment detection for the cluster starting density (host logic, CPU)."""
import numpy as np

from hfk import systems
from hfk.basis import Molecule
from hfk.guess import fragments


def test_water_cluster_fragments():
    for nwat in (2, 8, 64):
        mol = systems.water_cluster(nwat, "sto-3g")
        frags = fragments(mol)
        assert len(frags) == nwat
        for f in frags:
            assert sorted(mol.atoms[a][0] for a in f) == ["H", "H", "O"]
        assert sorted(a for f in frags for a in f) == list(range(mol.natom))


def test_alkane_single_fragment():
    mol = systems.alkane(12, "sto-3g")
    assert len(fragments(mol)) == 1


def test_separated_molecules_split():
    w = systems.water("sto-3g")
    atoms = list(w.atoms) + [(s, tuple(np.asarray(x) + [20.0, 0.0, 0.0]), c) for s, x, c in w.atoms]
    assert fragments(Molecule(atoms)) == [[0, 1, 2], [3, 4, 5]]


def _never(_sub):
    raise AssertionError("fragment_guess must not build a context here")


def test_fragment_guess_declines_without_gpu_work():
    """Cases where the fragment guess returns None before any SCF: a single
    fragment, a charged molecule, an open-shell fragment."""
    from hfk.guess import fragment_guess
    w = systems.water("sto-3g")
    assert fragment_guess(w, _never) is None                      # one fragment
    two = Molecule(list(w.atoms) + [(s, tuple(np.asarray(x) + [20.0, 0.0, 0.0]), c) for s, x, c in w.atoms],
                   charge=1.0)
    assert fragment_guess(two, _never) is None                    # charged
    h = [a for a in w.atoms if a[0] == "H"][0]
    odd = Molecule(list(w.atoms) + [(h[0], tuple(np.asarray(h[1]) + [20.0, 0.0, 0.0]), h[2])])
    assert len(fragments(odd)) == 2
    assert fragment_guess(odd, _never) is None                    # open-shell H atom
