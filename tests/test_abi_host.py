"""CPU-only checks of the boundary and the host logic (no GPU calls)."""
import os
import re

import numpy as np
import pytest

import hfk
from hfk import _lib, basis, flops, systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "hfk.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(hfk_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_default_options_are_reference():
    o = hfk.options()
    assert o.screen_mode == hfk.HFK_SCREEN_REFERENCE and o.boys_mode == hfk.HFK_BOYS_REFERENCE
    assert o.screen_tau == 1e-8 and o.fp32_threshold == 0.0


def test_flops_model_anchors():
    # Head-Gordon-Pople flop counts: VRR per primitive quartet (f_prim), HRR per
    # contracted quartet (f_hrr); (ss|ss) 43 and (ps|ss) 72 price the headline kernels
    assert flops.f_prim(0, 0, 0, 0) == 43
    assert flops.f_prim(1, 0, 0, 0) == 72
    assert (flops.f_prim(1, 1, 0, 0), flops.f_hrr(1, 1, 0, 0)) == (153, 18)
    assert flops.f_prim(1, 0, 1, 0) == 225
    assert (flops.f_prim(1, 1, 1, 0), flops.f_hrr(1, 1, 1, 0)) == (606, 54)
    assert (flops.f_prim(1, 1, 1, 1), flops.f_hrr(1, 1, 1, 1)) == (1695, 324)
    assert (flops.f_prim(2, 2, 2, 2), flops.f_hrr(2, 2, 2, 2)) == (30355, 11256)
    # gradient model: three centres at the energy class's cost plus 9 x 4
    # contraction flops per component quartet
    assert flops.f_grad_prim(0, 0, 0, 0) == 3 * 43 + 36
    assert flops.f_grad_prim(1, 0, 1, 0) == 3 * 225 + 36 * 9


def test_basis_parser_keeps_sp_exponents():
    txt = 'BASIS "ao basis" PRINT\n#c\nC    SP\n      7.8682724  -0.1193324   0.0689991\n      1.8812885  -0.1608542   0.3164240\nEND\n'
    out = basis.parse_nwchem(txt)
    (kind, prims), = out["C"]
    assert kind == "SP" and prims[0] == (-0.1193324, 0.0689991, 7.8682724)


def test_sp_exponent_bug_flag():
    """sp_exponent_bug=True reproduces ParserBasis.hs:116-118 (every SP
    primitive gets exponent 1.0); S and P shells are untouched."""
    shells = [("S", [(0.5, 10.0)]), ("SP", [(-0.12, 0.07, 7.87), (-0.16, 0.32, 1.88)])]
    good = basis.element_cgfs(shells)
    bug = basis.element_cgfs(shells, sp_exponent_bug=True)
    assert [ft for ft, _ in good] == [ft for ft, _ in bug] == ["S", "S", "Px", "Py", "Pz"]
    assert good[0] == bug[0]
    for (_, pg), (_, pb) in zip(good[1:], bug[1:]):
        assert [e for _, e in pg] == [7.87, 1.88] and [e for _, e in pb] == [1.0, 1.0]
    m = basis.build_molecule(["O"], [(0.0, 0.0, 0.0)], "6-31g", sp_exponent_bug=True)
    # two SP shells of 3 and 1 primitives, stored once per S, Px, Py, Pz function
    assert m.nbf == 9 and np.count_nonzero(m.expo == 1.0) == (3 + 1) * 4


def test_normalisation_matches_reference_formula():
    # normCoeff (NormalizeBasis.hs:52-59): c / sqrt(prod (2k-1)!! / (4e)^L * (pi/2e)^1.5)
    for ft, prod in (("S", 1.0), ("Px", 1.0), ("Dxx", 3.0), ("Dxy", 1.0)):
        L = sum(basis.LXYZ[ft])
        (c, e), = basis.norm_coeff([(0.7, 1.3)], ft)
        ref = 0.7 / np.sqrt(prod / (4 * 1.3) ** L * (np.pi / (2 * 1.3)) ** 1.5)
        assert abs(c - ref) < 1e-15 * ref and e == 1.3


def test_water_cluster_generator_is_seeded_and_sized():
    for n, nbf, L in ((8, 104, 3), (16, 208, 4), (32, 416, 4), (64, 832, 6)):
        assert systems.lattice_edge(n) == L
        m = systems.water_cluster(n, "6-31g")
        assert m.nbf == nbf and m.occupied == 5 * n
    a = systems.water_cluster_geometry(8)[1]
    b = systems.water_cluster_geometry(8)[1]
    assert np.array_equal(a, b)
    # rigid waters: O-H 1.810 bohr
    for w in range(8):
        o, h1, h2 = a[3 * w], a[3 * w + 1], a[3 * w + 2]
        assert abs(np.linalg.norm(h1 - o) - 1.8101) < 1e-3 and abs(np.linalg.norm(h2 - o) - 1.8101) < 1e-3


def test_alkane_size():
    m = systems.alkane()
    assert m.natom == 38 and m.nbf == 232


def test_function_order_follows_fetchbasis():
    m = systems.water("6-31g")
    # O: 1s, 2s, 2px, 2py, 2pz, 3s, 3px, 3py, 3pz; H: 1s, 2s (x2)
    assert [tuple(v) for v in m.lxyz[:9]] == [(0, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                                              (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    assert m.nbf == 13


def test_no_gpu_fallback_without_library(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError):
        _lib.lib()


def test_parse_xyz_skips_exactly_two_lines():
    """parserXYZ consumes the count line and exactly one comment line, which
    may be empty (ParseXYZ.hs:33-43); coordinates are Angstrom -> bohr."""
    atoms = "O 0.0 0.0 0.1173\nH 0.0 0.7572 -0.4692\nH 0.0 -0.7572 -0.4692\n"
    for comment in ("", "water, comment line", "   "):
        syms, xyz = basis.parse_xyz("3\n%s\n%s" % (comment, atoms))
        assert syms == ["O", "H", "H"]
        assert abs(xyz[1][1] - 0.7572 / 0.529177) < 1e-14 and abs(xyz[0][2] - 0.1173 / 0.529177) < 1e-14
    # leading blank lines before the count and blank lines between atoms are whitespace
    syms, _ = basis.parse_xyz("\n\n  3\ncomment\nO 0 0 0\n\nH 0 0 1\nH 0 1 0")
    assert syms == ["O", "H", "H"]
    with pytest.raises(ValueError):
        basis.parse_xyz("3\n\nO 0 0 0\nH 0 0 1\n")
    with pytest.raises(ValueError):
        basis.parse_xyz("comment only\n")
