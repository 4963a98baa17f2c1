"""The oracle's production-mode draws.

Philox4x32-10 is pinned by the known-answer vectors published with Random123
(Salmon, Moraes, Dror, Shaw, SC'11; kat_vectors, philox4x32 with 10 rounds);
the double built from two words is checked against its definition on the
2^-52 grid, and a production-mode oracle run is checked to be a function of
the seed alone."""
import ctypes as C

import numpy as np
import pytest

import helpers
from hande_amd import _native as N
from hande_amd import api

U32 = C.c_uint32 * 4
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.fixture(scope="module")
def orc():
    return helpers.oracle()


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(orc, ctr, key, want):
    out = U32()
    orc.orc_philox4x32_10(U32(*ctr), (C.c_uint32 * 2)(*key), out)
    assert tuple(out) == want


def _hub_params(seed, real=False, initiator=False):
    p = N.FqParams()
    p.nspaces = 1
    p.initiator = int(initiator)
    p.initiator_pop = 3.0
    p.real_amplitudes = int(real)
    p.pop_bit_shift = 31 if real else 0
    p.spawn_cutoff = int(np.ceil(0.01 * (1 << 31))) if real else 0
    p.excit_gen = N.FQ_EXCIT_RENORM
    p.rng_mode = N.FQ_RNG_PHILOX
    p.philox_seed = seed
    p.main_capacity = p.spawn_capacity = 1 << 16
    return p


def _run(orc, system, p, ncycles, seed):
    h = orc.orc_create(system.view, C.byref(p), 1)
    orc.orc_set_production(h, seed)
    try:
        rf = 1 << p.pop_bit_shift
        f0 = np.array([system.view.f0[0]], dtype=np.uint64)
        pop0 = np.array([20 * rf], dtype=np.int64)
        orc.orc_load(h, f0.ctypes.data_as(C.POINTER(C.c_uint64)), pop0.ctypes.data_as(C.POINTER(C.c_int64)),
                     np.zeros(1).ctypes.data_as(C.POINTER(C.c_double)), 1)
        sh = (C.c_double * 2)(0.5, 0.5)
        eo = (C.c_double * 2)(0.0, 0.0)
        for _ in range(ncycles):
            o = N.FqCycleOut()
            orc.orc_cycle(h, sh, eo, 0.01, C.byref(o))
            assert o.rng_consumed == 0  # no dSFMT draws in production mode
        return helpers.odownload(orc, h, 1, 1)
    finally:
        orc.orc_destroy(h)


@pytest.mark.parametrize("real", [False, True])
def test_production_oracle_depends_on_seed_only(orc, real):
    system = api.hubbard_k(electrons=6, lattice=[[10]], ms=0, sym=1, U=4.0, t=1.0)
    a = _run(orc, system, _hub_params(7, real, real), 40, 7)
    b = _run(orc, system, _hub_params(7, real, real), 40, 7)
    c = _run(orc, system, _hub_params(8, real, real), 40, 8)
    assert len(a[2]) > 20
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not (len(a[2]) == len(c[2]) and all(np.array_equal(x, y) for x, y in zip(a, c)))
