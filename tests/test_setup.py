"""Host set-up tables against what the reference printed for the same decks
(basis-function listing, momentum symmetry table and inverses, reference
determinant and H00), plus closed-form checks."""
import math

import numpy as np
import pytest

from hande_amd import api, _native as N
from helpers import goldens

CASES = ["hubbard_k_1d_n4", "hubbard_k_nontilted_2d_fciqmc", "hubbard_k_1d_renorm", "hubbard_k_1d_ifciqmc"]


@pytest.mark.parametrize("name", CASES)
def test_hubbard_basis_matches_reference_listing(name):
    e = goldens()[name]
    system, _ = api.load_deck(e["input"], lib=N.lib())
    ms, label, eps = system.basis()
    assert len(e["basis"]) == system.nbasis
    for i, line in enumerate(e["basis"]):
        parts = line.replace("(", " ").replace(")", " ").replace(",", " ").split()
        # index, k components..., ms, <i|f|i>
        assert int(parts[0]) == i + 1
        assert int(parts[-2]) == ms[i]
        assert abs(float(parts[-1]) - eps[i]) <= 1e-8 * max(1.0, abs(eps[i]))


@pytest.mark.parametrize("name", CASES)
def test_hubbard_symmetry_and_reference(name):
    e = goldens()[name]
    system, _ = api.load_deck(e["input"], lib=N.lib())
    v = system.view
    nsym = v.nsym
    inv = [v.inv_sym[i] for i in range(nsym)]
    assert inv == e["inv_sym"]
    printed = e["sym_table_printed"]  # row i, column j = sym_table(j, i)
    for i in range(nsym):
        for j in range(nsym):
            assert v.sym_table[j * nsym + i] == printed[i][j]
    assert sorted(system.reference_occ()) == sorted(e["json"]["det"])
    assert abs(system.H00 - e["json"]["H00"]) < 1e-7


def test_hubbard_closed_forms():
    # 4x4 lattice, 10 electrons: U/N = 0.25, closed shells (-4 x1, -2 x4) per spin.
    s = api.hubbard_k(electrons=10, lattice=[[4, 0], [0, 4]], ms=0, sym="tot_sym", U=4, t=1)
    assert s.view.coulomb_k == 0.25
    assert s.view.nbasis == 32
    _, _, eps = s.basis()
    assert sorted(set(np.round(eps, 12)))[:2] == [-4.0, -2.0]
    # H00 = nalpha*nbeta*U/N + sum of occupied kinetic energies = 25*0.25 + 2*(-4 - 8)
    assert abs(s.H00 - (25 * 0.25 + 2 * (-4.0 - 4 * 2.0))) < 1e-12
    # diagonal element of the reference from the host Slater-Condon rule
    assert s.diagonal(s.f0()) == s.H00


def test_read_in_h2o_cas_matches_reference(tmp_path):
    """FCIDUMP reader with a CAS(8,13): frozen-core energy folded into E_core
    and the one-body integrals; H00 and pattempt_single as printed by the
    reference (H00 -76.02403856, pattempt_single 0.04938272)."""
    from helpers import deck
    e = goldens()["H2O-RHF-cc-pVTZ"]
    system, _ = deck("H2O-RHF-cc-pVTZ", str(tmp_path), lib=N.lib())
    v = system.view
    assert (v.nbasis, v.nel) == (26, 8)
    assert abs(v.H00 - e["json"]["H00"]) < 1e-8
    assert abs(v.pattempt_single - e["json"]["pattempt_single"]) < 1e-8
    assert abs(v.pattempt_single + v.pattempt_double - 1.0) < 1e-15
    # the host diagonal rule agrees with H00 and the eight-fold ERI symmetry holds
    assert system.diagonal(system.f0()) == v.H00
    ms, sym, eps = system.basis()
    assert ms[:4] == [1, -1, 1, -1]
    assert all(eps[2 * i] == eps[2 * i + 1] for i in range(13))
    assert all(eps[2 * i] <= eps[2 * i + 2] + 1e-12 for i in range(12))
