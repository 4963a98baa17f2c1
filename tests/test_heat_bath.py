"""Heat-bath excitation generator of the oracle (src/excit_gen_heat_bath_mol.F90).

The reference holds no heat-bath trajectory golden that can run here, so the
oracle's generator is pinned by what the mathematics fixes:
  * the alias tables (lib/local/alias.f90) reproduce weights/total exactly;
  * the tables equal |<ij||ab>| sums built independently from the integrals;
  * every reported pgen equals the probability, summed over all random paths
    of the generator (i, j, a, single/double decision, b), of producing that
    excitation, and every reported matrix element equals the Slater--Condon
    value computed here from the integrals.
"""
import ctypes as C
import os

import numpy as np
import pytest

import helpers
from hande_amd import _native as N
from hande_amd import api, synthetic

NORB, NEL, SEED = 6, 4, 611


@pytest.fixture(scope="module")
def orc():
    return helpers.oracle()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_alias_tables_reproduce_distribution(orc, seed):
    rng = np.random.default_rng(seed)
    n = 37
    w = rng.random(n) ** 3
    w[rng.choice(n, 9, replace=False)] = 0.0
    tot = float(w.sum())
    u = np.zeros(n)
    k = np.zeros(n, dtype=np.int32)
    orc.orc_alias_tables(n, w.ctypes.data_as(C.POINTER(C.c_double)), tot, u.ctypes.data_as(C.POINTER(C.c_double)),
                         k.ctypes.data_as(C.POINTER(C.c_int)))
    # outcome of column j: j itself with probability min(1, U_j), alias K_j otherwise
    keep = np.clip(u, 0.0, 1.0)
    prob = np.zeros(n)
    for j in range(n):
        prob[j] += keep[j] / n
        prob[k[j] - 1] += (1.0 - keep[j]) / n
    np.testing.assert_allclose(prob, w / tot, rtol=0, atol=1e-14)
    assert np.all(prob[w == 0.0] == 0.0)


def _system(tmp_path):
    path = synthetic.write_fcidump(os.path.join(tmp_path, "FCIDUMP"), NORB, NEL, SEED)
    return api.read_in(int_file=path, nel=NEL, ms=0)


def _params(system):
    p = N.FqParams()
    p.nspaces = 1
    p.initiator = 0
    p.real_amplitudes = 0
    p.pop_bit_shift = 0
    p.spawn_cutoff = 0
    p.excit_gen = N.FQ_EXCIT_HEAT_BATH
    p.rng_mode = N.FQ_RNG_REPLAY
    p.main_capacity = p.spawn_capacity = 1 << 12
    return p


def _integrals():
    _eps, h, eri = synthetic.synthetic_integrals(NORB, SEED)
    return h, eri


def _spatial(o):
    return (o - 1) // 2


def _ms(o):
    return 1 if o % 2 == 1 else -1


def _phys(eri, p, q, r, s):
    """<pq|rs> = (pr|qs) over spin orbitals with the spatial integrals."""
    return eri[_spatial(p), _spatial(r), _spatial(q), _spatial(s)]


def _h2(eri, i, j, a, b):
    """<ij||ab> with HANDE's spin rules (src/hamiltonian_molecular.f90:294-346)."""
    h = 0.0
    if _ms(i) == _ms(a):
        h = _phys(eri, i, j, a, b)
    if _ms(i) == _ms(b):
        h = h - _phys(eri, i, j, b, a)
    return h


def _h1(h1, eri, occ, i, a):
    v = h1[_spatial(i), _spatial(a)]
    for o in occ:
        if o != i:
            v += _phys(eri, i, o, a, o)
            if _ms(o) == _ms(i):
                v -= _phys(eri, i, o, o, a)
    return v


def _tables(orc, h):
    nb = 2 * NORB
    out = {}
    for name, shape in [("i_w", (nb,)), ("ij_w", (nb, nb)), ("ija_w", (nb, nb, nb)), ("ija_tot", (nb, nb)),
                        ("ijab_w", (nb, nb, nb, nb)), ("ijab_tot", (nb, nb, nb))]:
        n = orc.orc_hb_table(h, name.encode(), None, 0)
        a = np.zeros(n)
        orc.orc_hb_table(h, name.encode(), a.ctypes.data_as(C.POINTER(C.c_double)), n)
        out[name] = a.reshape(shape)
    return out


def test_heat_bath_tables_match_integrals(orc, tmp_path):
    system = _system(tmp_path)
    p = _params(system)
    h = orc.orc_create(system.view, C.byref(p), 7)
    try:
        assert orc.orc_hb_biased(h) == 0
        t = _tables(orc, h)
    finally:
        orc.orc_destroy(h)
    _h1m, eri = _integrals()
    nb = 2 * NORB
    ref = np.zeros((nb, nb, nb, nb))
    for i in range(1, nb + 1):
        for j in range(1, nb + 1):
            if i == j:
                continue
            it, jt = min(i, j), max(i, j)
            for a in range(1, nb + 1):
                for b in range(1, nb + 1):
                    if len({i, j, a, b}) < 4:
                        continue
                    if not ((_ms(it) == _ms(a) and _ms(jt) == _ms(b)) or (_ms(it) == _ms(b) and _ms(jt) == _ms(a))):
                        continue
                    ref[i - 1, j - 1, a - 1, b - 1] = abs(_h2(eri, it, jt, min(a, b), max(a, b)))
    np.testing.assert_allclose(t["ijab_w"], ref, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(t["ijab_tot"], ref.sum(axis=3), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(t["ija_w"], ref.sum(axis=3), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(t["ija_tot"], ref.sum(axis=(2, 3)), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(t["ij_w"], ref.sum(axis=(2, 3)), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(t["i_w"], ref.sum(axis=(1, 2, 3)), rtol=1e-12, atol=1e-15)


def _enumerate(t, h1, eri, occ):
    """Exact probability of every outcome of the generator from determinant occ,
    by summing over its random paths (weights/totals stand for the alias draws)."""
    nb = 2 * NORB
    occset = set(occ)
    outcomes = {}
    null = 0.0
    iw = np.array([t["i_w"][o - 1] for o in occ])
    for ii, i in enumerate(occ):
        pi = iw[ii] / iw.sum()
        ijw = np.array([t["ij_w"][i - 1, o - 1] for o in occ])
        if not ijw.sum() > 0:
            null += pi
            continue
        for jj, j in enumerate(occ):
            pj = ijw[jj] / ijw.sum()
            if pj == 0.0:
                continue
            tot_a = t["ija_tot"][i - 1, j - 1]
            if not tot_a > 0:
                null += pi * pj
                continue
            for a in range(1, nb + 1):
                pa = t["ija_w"][i - 1, j - 1, a - 1] / tot_a
                if pa == 0.0:
                    continue
                p = pi * pj * pa
                if a in occset:
                    null += p
                    continue
                if _ms(a) == _ms(i):  # C1 symmetry: every orbital spans the same irrep
                    hm = abs(_h1(h1, eri, occ, i, a))
                    tt = t["ijab_tot"][i - 1, j - 1, a - 1]
                    ps = hm / (tt + hm) if hm < tt else 0.5
                else:
                    ps = 0.0
                key = (1, (i,), (a,))
                outcomes[key] = outcomes.get(key, 0.0) + p * ps
                for b in range(1, nb + 1):
                    pb = t["ijab_w"][i - 1, j - 1, a - 1, b - 1] / t["ijab_tot"][i - 1, j - 1, a - 1]
                    if pb == 0.0:
                        continue
                    if b in occset:
                        null += p * (1 - ps) * pb
                        continue
                    key = (2, (min(i, j), max(i, j)), (min(a, b), max(a, b)))
                    outcomes[key] = outcomes.get(key, 0.0) + p * (1 - ps) * pb
    return outcomes, null


def _parity_sign(occ, frm, to):
    """Sign of the excited determinant relative to occ (creation order by orbital index)."""
    cur = list(occ)
    sign = 1
    for f, t in zip(frm, to):
        k = cur.index(f)
        sign *= (-1) ** k
        cur.pop(k)
        pos = sum(1 for o in cur if o < t)
        sign *= (-1) ** pos
        cur.insert(pos, t)
    return sign


def test_heat_bath_pgen_and_hmatel_match_path_enumeration(orc, tmp_path):
    system = _system(tmp_path)
    p = _params(system)
    h = orc.orc_create(system.view, C.byref(p), 7)
    h1, eri = _integrals()
    rng = np.random.default_rng(5)
    nb = 2 * NORB
    try:
        t = _tables(orc, h)
        f0 = int(system.f0()[0])
        dets = [[o for o in range(1, nb + 1) if (f0 >> (o - 1)) & 1]]
        while len(dets) < 4:
            al = sorted(rng.choice(np.arange(1, nb + 1, 2), NEL // 2, replace=False))
            be = sorted(rng.choice(np.arange(2, nb + 1, 2), NEL // 2, replace=False))
            dets.append(sorted(int(x) for x in al + be))
        ires = (C.c_int * 6)()
        dres = (C.c_double * 2)()
        for occ in dets:
            outcomes, null = _enumerate(t, h1, eri, occ)
            assert abs(sum(outcomes.values()) + null - 1.0) < 1e-12
            f = np.zeros(1, dtype=np.uint64)
            for o in occ:
                f[0] |= np.uint64(1) << np.uint64(o - 1)
            seen = set()
            nallowed = 0
            ntry = 1500
            for _ in range(ntry):
                draws = rng.random(8)
                orc.orc_hb_generate(h, f.ctypes.data_as(C.POINTER(C.c_uint64)),
                                    draws.ctypes.data_as(C.POINTER(C.c_double)), 8, ires, dres)
                if not ires[0]:
                    continue
                nallowed += 1
                nx = ires[1]
                key = (nx, tuple(ires[2:2 + nx]), tuple(ires[4:4 + nx]))
                assert key in outcomes, key
                np.testing.assert_allclose(dres[0], outcomes[key], rtol=1e-10)
                if nx == 1:
                    ref_h = _h1(h1, eri, occ, key[1][0], key[2][0])
                else:
                    ref_h = _h2(eri, *key[1], *key[2])
                ref_h *= _parity_sign(occ, key[1], key[2])
                np.testing.assert_allclose(dres[1], ref_h, rtol=1e-12, atol=1e-14)
                seen.add(key)
            # the acceptance rate is binomial with mean 1 - P(null)
            sd = np.sqrt(ntry * null * (1 - null))
            assert abs(nallowed - ntry * (1 - null)) < 5 * sd + 1
            assert len(seen) > 10
    finally:
        orc.orc_destroy(h)


def test_exact_small_reference_energy_matches_host_tables(tmp_path):
    """The brute-force Hamiltonian of tests/exact_small.py and the host's
    integral reader agree on <D0|H|D0> (the energy zero of E_proj)."""
    import exact_small
    path = synthetic.write_fcidump(os.path.join(tmp_path, "FCIDUMP"), 8, 4, 811)
    system = api.read_in(int_file=path, nel=4, ms=0)
    e0, hf = exact_small.ground_state(8, 4, 811)
    assert abs(hf - system.H00) < 1e-12
    assert e0 < hf
