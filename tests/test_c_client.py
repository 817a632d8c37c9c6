"""The C ABI from a plain C99 program (tests/c_client/rg_client.c), compiled with gcc against include/raspa3_amd/rg.h
and linked to libraspa3_amd.so, as a non-Python caller would use it.

CPU: the program compiles and links with -std=c99 -pedantic -Werror.  GPU: it computes the methane golden energy in
ITQ-29 (cbmc_interactions.cpp:28-51) by the direct path and the same energies from a grid it builds, checked against
the golden value and the oracle's grid interpolation.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from raspa3_amd import workloads as w

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_client", "rg_client.c")
LIBDIR = os.path.join(ROOT, "raspa3_amd")


def compile_client(tmp_path):
    exe = str(tmp_path / "rg_client")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", os.path.join(ROOT, "include"),
                    SRC, "-L", LIBDIR, "-lraspa3_amd", f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True)
    return exe


def test_c_client_compiles_against_header(tmp_path):
    exe = compile_client(tmp_path)
    assert os.path.exists(exe)


def write_input(path, cell, rect, spec, pos, typ, q, probe, spacing, sites, site_t, site_q):
    with open(path, "w") as f:
        f.write(" ".join(repr(float(x)) for x in cell.T.reshape(-1)) + f" {int(rect)}\n")
        f.write(f"{len(spec.eps_K)}\n")
        for e, s, fw in zip(spec.eps_K, spec.sigma, spec.is_framework):
            f.write(f"{e!r} {s!r} {int(fw)}\n")
        f.write(f"{int(spec.shifted)} {spec.rc_framework!r} {spec.rc_molecule!r} {spec.rc_coulomb!r} {spec.alpha!r} "
                f"{int(spec.use_charge)} {spec.overlap!r}\n")
        f.write(f"{len(typ)}\n")
        for p, t, c in zip(pos, typ, q):
            f.write(f"{p[0]!r} {p[1]!r} {p[2]!r} {int(t)} {float(c)!r}\n")
        f.write(f"{probe} {spacing!r} {len(sites)}\n")
        for p, t, c in zip(sites, site_t, site_q):
            f.write(f"{p[0]!r} {p[1]!r} {p[2]!r} {int(t)} {float(c)!r}\n")


@pytest.mark.gpu
def test_c_client_golden_and_grid(tmp_path, refvals, itq29):
    exe = compile_client(tmp_path)
    rows = refvals["zeolite_forcefield"]["pseudo_atoms"]
    spec = w._ff([(r[0], r[2], r[3], r[4], r[1]) for r in rows], alpha=0.25, use_charge=False)
    a = itq29["a"]
    cell = np.diag([a, a, a])
    typ = np.array([spec.index(e) for e in itq29["element"]], np.int32)
    ch4 = spec.index("CH4")
    r1 = refvals["cbmc_framework_molecule_1"]
    rng = np.random.default_rng(11)
    sites = np.vstack([np.array(r1["methane_fractional"]) * a, rng.uniform(0, a, (63, 3))])
    inp = tmp_path / "input.txt"
    write_input(inp, cell, True, spec, itq29["pos"], typ, np.zeros(len(typ)), ch4, 0.3, sites, [ch4] * 64,
                [0.0] * 64)
    out = subprocess.run([exe, str(inp)], check=True, capture_output=True, text=True, timeout=120).stdout.split("\n")
    dims = [int(x) for x in out[0].split()[1:]]
    rec = np.array([[float(x) for x in line.split()[2:]] for line in out[1:] if line.startswith("site")])
    assert rec.shape == (64, 4)
    e2k = oracle.units()["energy_to_kelvin"]
    # the golden energy by the direct path
    assert rec[0, 0] == 0 and abs(rec[0, 1] * e2k - r1["framework_vdw_K"]) < 1e-6
    # the grid the program built, interpolated at the sites, against the oracle's grid and interpolation
    pairs = oracle.pairs_from_spec(spec)
    assert dims == list(w.grid_dims(cell, 0.3)[0])
    g = oracle.grid_build(cell, True, dims, 0, itq29["pos"], typ, np.zeros(len(typ)), pairs, ch4, 12.0, 12.0, 0.25,
                          nthreads=8)
    eo = oracle.interpolate(cell, True, dims, g.reshape(-1), sites)
    tol = 1e-3 * oracle.units()["kelvin_to_energy"]
    assert np.all(np.abs(rec[:, 3] - eo) <= tol + 1e-9 * np.abs(eo))
