"""CPU-side checks of the C ABI: the library loads and exports every function include/fvx.h
declares; host-only utilities (RCM ordering) work without a GPU."""
import ctypes
import os
import re

import numpy as np

from fvens_amd import fvx, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fvx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(fvx_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(fvx.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_rcm_is_a_permutation_and_reduces_bandwidth():
    raw = meshgen.airfoil_omesh(64, 16, 8)
    perm = fvx.rcm_permutation(raw)
    assert sorted(perm.tolist()) == list(range(raw.nelem))
    import pyoracle as po

    def bandwidth(r):
        m = po.Mesh(r)
        inter = m.intfac[m.nbface:]
        return int(np.max(inter[:, 1] - inter[:, 0]))

    assert bandwidth(raw.permuted(perm)) < bandwidth(raw)


def test_invalid_mesh_is_rejected_with_status():
    raw = meshgen.gaussian_bump(4, 2, 0)
    bad = meshgen.RawMesh(raw.coords, raw.inpoel.copy(), raw.nnode.copy(), raw.bface[:-1],
                          raw.btag[:-1])
    try:
        fvx.rcm_permutation(bad)          # topology without the last boundary face still pairs
    except fvx.FvxError:
        pass
    bad2 = meshgen.RawMesh(raw.coords, raw.inpoel.copy(), raw.nnode.copy() * 0 + 5, raw.bface,
                           raw.btag)
    try:
        fvx.rcm_permutation(bad2)
        raise AssertionError("expected an error")
    except fvx.FvxError as e:
        assert e.code == 1
