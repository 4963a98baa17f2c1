"""Semi-stochastic projection on the GPU (fq_semi_stoch_init and the cycle).

* replay: the reference's two He2 semi-stochastic runs (combined
  annihilation; separate annihilation with a replica pair) through the host
  driver, every report row against the golden table (1e-10);
* production: engine and oracle side by side on Philox streams, a core made
  mid-run on both (same determinants chosen, same number of stored
  Hamiltonian elements), then the walker lists bit-identical after every
  cycle, for both annihilation modes, on a molecular system (single and double
  Slater--Condon elements, a core wider than one 256-column chunk of the
  device Hamiltonian build) and on the momentum-space Hubbard model."""
import ctypes as C
import os

import numpy as np
import pytest

from hande_amd import _native as N
from hande_amd import api, synthetic
from helpers import compare_tables, deck, goldens, production_lockstep

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["He2-aug-cc-pVDZ_real_32_SS", "He2-aug-cc-pVDZ_real_32_SS_replica"])
def test_replay_semi_stochastic_golden(name, tmp_path):
    e = goldens()[name]
    print(f"[semi-stoch replay] {name}", flush=True)
    system, opts = deck(name, str(tmp_path))
    got = api.fciqmc(system, opts["qmc"], opts.get("fciqmc"), rng_mode="replay", semi_stoch=opts["semi_stoch"])
    compare_tables(got, e["report"])


def _params(seed, spaces=1, cap=1 << 18):
    p = N.FqParams()
    p.nspaces = spaces
    p.initiator = 1
    p.initiator_pop = 3.0
    p.quadrature_initiator = 1
    p.real_amplitudes = 1
    p.pop_bit_shift = 31
    p.spawn_cutoff = int(np.ceil(0.01 * (1 << 31)))
    p.excit_gen = N.FQ_EXCIT_RENORM
    p.rng_mode = N.FQ_RNG_PHILOX
    p.philox_seed = seed
    p.main_capacity = cap
    p.spawn_capacity = cap
    p.device = 0
    return p


@pytest.mark.parametrize("separate", [0, 1])
def test_production_semi_stochastic_molecular(tmp_path, separate):
    path = synthetic.write_fcidump(os.path.join(tmp_path, "FCIDUMP"), 12, 6, 1206, chol_scale=0.1)
    system = api.read_in(int_file=path, nel=6, ms=0)
    print(f"[semi-stoch production] 12o6e separate={separate}", flush=True)
    outs = production_lockstep(system, _params(301 + separate), 70, 3000, shift=0.0, tau=0.02,
                               label=f"ss{separate}", core=(40, 600, separate))
    assert outs[-1].nstates > 1000


@pytest.mark.parametrize("separate", [0, 1])
def test_production_semi_stochastic_hubbard_replica(separate):
    system = api.hubbard_k(electrons=6, lattice=[[10]], ms=0, sym=1, U=3.0, t=1.0)
    print(f"[semi-stoch production] hubbard_k replica separate={separate}", flush=True)
    outs = production_lockstep(system, _params(311 + separate, spaces=2), 60, 30, tau=0.01,
                               label=f"hub{separate}", core=(30, 150, separate))
    assert outs[-1].nstates > 150
