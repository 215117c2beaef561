"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares; host-only entry points work on the CPU."""
import ctypes as C
import glob
import os
import re

import numpy as np

from iceicle_amd import abi, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(ice_[a-z_0-9]+)\s*\(", text, flags=re.M))
    return names


def test_exports_every_declared_symbol():
    lib = C.CDLL(engine.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes front end binds exactly the declared set
    assert set(engine._SIGS) == names


def test_version_and_defaults():
    L = engine.lib()
    assert L.ice_version() == 1
    p = abi.DiscParams()
    engine.check(L.ice_disc_params_default(abi.FLUX_NAVIER_STOKES, C.byref(p)))
    ref = abi.default_params(abi.FLUX_NAVIER_STOKES)
    assert p.gamma == ref.gamma and p.Rgas == ref.Rgas and p.Pr == ref.Pr and list(p.ref) == [1.0] * 8


def test_mesh_export_host_only_and_errors():
    m = abi.mesh_desc(2, [3, 2], [0, 0], [1, 1], ["DIRICHLET"] * 4)
    coord, conn, faces, cnt = engine.mesh_export(m)
    assert cnt.nelem == 6 and cnt.nnode == 12 and cnt.nface == 7 + 10
    bad = abi.mesh_desc(2, [3, 0], [0, 0], [1, 1], ["DIRICHLET"] * 4)
    cnt2 = abi.MeshCounts()
    rc = engine.lib().ice_mesh_export(C.byref(bad), C.byref(cnt2), None, None, None)
    assert rc == abi.ICE_ERR_ARG and b"nelem" in engine.lib().ice_last_error()
    # one-sided periodic boundaries are rejected
    bad2 = abi.mesh_desc(2, [3, 2], [0, 0], [1, 1], ["PERIODIC", "DIRICHLET", "DIRICHLET", "DIRICHLET"])
    assert engine.lib().ice_mesh_export(C.byref(bad2), C.byref(cnt2), None, None, None) == abi.ICE_ERR_ARG


def test_field_eval_host():
    f = abi.Field.make("SINUSOIDS", [2, 0.5, 1, 0, 0, -1.0, 1, 2, 0])
    x = np.array([[0.1, 0.2], [0.7, 0.3]])
    v = engine.field_eval(f, 2, 1, x)[:, 0]
    expect = 0.5 * np.sin(2 * np.pi * x[:, 0]) - np.sin(2 * np.pi * (x[:, 0] + 2 * x[:, 1]))
    np.testing.assert_allclose(v, expect, atol=1e-15)
