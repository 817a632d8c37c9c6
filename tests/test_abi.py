"""C ABI checks that need no GPU: the library loads, exports every entry point include/raspa3_amd/rg.h declares, and
its host-side set-up (SimulationBox, grid dimensions, force-field mixing and shift) agrees with the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import raspa3_amd as ra
from raspa3_amd import _lib
from raspa3_amd import workloads as w

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include", "raspa3_amd")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", "raspa3_amd", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(rg_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, missing
    # the ctypes table covers the header exactly
    assert names == set(_lib.SIGNATURES), sorted(names ^ set(_lib.SIGNATURES))
    ra.lib()  # loads, resolves every signature


def test_box_matches_oracle():
    for args in [(20.0, 19.9, 13.4, 90, 90, 90), (40.0, 40.0, 30.0, 90, 90, 100), (13.0, 14.0, 12.5, 80, 95, 105)]:
        a = ra.box(*args)
        b = oracle.box(*args)
        assert a["rect"] == b["rect"]
        assert np.array_equal(a["cell"], b["cell"])
        assert np.allclose(a["inv"], b["inv"], rtol=1e-15, atol=0)
        assert np.allclose(a["widths"], b["widths"], rtol=1e-15)
    c = ra.box(11.8671, 11.8671, 11.8671, lengths_only=True)
    assert np.array_equal(c["cell"], np.diag([11.8671] * 3))


def test_grid_dims_truncation():
    """19.9/0.2 + 0.5 = 99.99999999999997 truncates to 99 (src/raspakit/system_interpolation_grids.cpp:222-226)."""
    fw0 = w.framework_config0()
    assert ra.grid_dims(fw0.cell, 0.2) == [100, 99, 67]
    fw1 = w.framework_config1()
    assert ra.grid_dims(fw1.cell, 0.1) == [394, 394, 300]
    n = np.zeros(3, np.int32)
    oracle.lib().orc_grid_dims(oracle._d(np.ascontiguousarray(fw1.cell.T.reshape(9))), 0.1, oracle._i(n))
    assert list(n) == [394, 394, 300]


@pytest.mark.parametrize("make", [w.forcefield_config0, w.forcefield_config1])
def test_forcefield_pairs_match_oracle(make):
    spec = make()
    ff = ra.ForceField(None, spec)
    eps, sig, sh, lj = ff.pairs()
    o = oracle.pairs_from_spec(spec)
    n = len(spec.eps_K)
    assert np.array_equal(eps.reshape(-1), o["eps"])
    assert np.array_equal(sig.reshape(-1), o["sigma"])
    assert np.array_equal(sh.reshape(-1), o["shift"])
    assert np.array_equal(lj.reshape(-1), o["lj"])
    assert lj.shape == (n, n)


def test_invalid_arguments_are_rejected_without_device():
    import ctypes as C
    h = C.c_void_p()
    # zero pseudo-atoms
    st = ra.lib().rg_forcefield_create(None, 0, None, None, None, 1, 12.0, 12.0, 12.0, 0.25, 1, 1e6, C.byref(h))
    assert st == _lib.RG_ERR_INVALID
    st = ra.lib().rg_grid_dims(None, 0.1, None)
    assert st == _lib.RG_ERR_INVALID
