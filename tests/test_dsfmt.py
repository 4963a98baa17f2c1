"""The replay stream generator against vectors written by the reference's own
dSFMT C source (lib/dSFMT-src-2.2.3/dSFMT.c, 50000-double refills as in
lib/dSFMT_F03_interface/dSFMT_interface.F90:387-409)."""
import ctypes as C
import hashlib
import json
import os
import struct

import numpy as np
import pytest

from hande_amd import _native as N

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dsfmt_vectors.json")))


def hexs(arr):
    return [struct.pack("<d", float(x))[::-1].hex() for x in arr]


@pytest.mark.parametrize("stream", GOLD["streams"], ids=lambda s: str(s["seed"]))
def test_dsfmt_stream_matches_reference(stream):
    lib = N.lib()
    n = stream["count"]
    out = np.zeros(n)
    assert lib.fqh_dsfmt_fill(stream["seed"], out.ctypes.data_as(C.POINTER(C.c_double)), n) == 0
    h = hexs(out)
    assert h[:64] == stream["first"]
    for k, v in stream["around_refill"].items():
        assert h[int(k)] == v, k
    assert h[-8:] == stream["last"]
    assert hashlib.sha256("\n".join(h).encode()).hexdigest() == stream["sha256"]
    assert np.all((out >= 0.0) & (out < 1.0))
