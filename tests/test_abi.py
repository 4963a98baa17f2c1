"""The C-ABI library loads without a GPU and exports every declared symbol."""
import os
import re

from hande_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in ("fqmc.h", "fqmc_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(fqh?_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    lib = N.lib()
    names = declared_symbols()
    assert "fq_cycle" in names and "fq_create" in names and "fqh_fciqmc" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_gpu_fails_cleanly():
    import ctypes as C
    from hande_amd import api
    lib = N.lib()
    s = api.hubbard_k(electrons=2, lattice=[[2, 0], [0, 2]], ms=0, sym=1)
    p = N.FqParams()
    p.nspaces = 1
    p.main_capacity = 16
    p.spawn_capacity = 16
    e = C.c_void_p()
    rc = lib.fq_create(s.view, C.byref(p), C.byref(e))
    assert rc in (N.FQ_OK, -6)  # -6 = FQ_ERR_NO_DEVICE on a GPU-less host
    if rc == N.FQ_OK:
        lib.fq_destroy(e)
    # argument errors are reported, not crashed on
    p.nspaces = 3
    assert lib.fq_create(s.view, C.byref(p), C.byref(e)) == -1
