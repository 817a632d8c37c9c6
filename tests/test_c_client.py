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
    def f(x):
        return repr(float(x))

    with open(path, "w") as out:
        out.write(" ".join(f(x) for x in cell.T.reshape(-1)) + f" {int(rect)}\n")
        out.write(f"{len(spec.eps_K)}\n")
        for e, s, fw in zip(spec.eps_K, spec.sigma, spec.is_framework):
            out.write(f"{f(e)} {f(s)} {int(fw)}\n")
        out.write(f"{int(spec.shifted)} {f(spec.rc_framework)} {f(spec.rc_molecule)} {f(spec.rc_coulomb)} "
                  f"{f(spec.alpha)} {int(spec.use_charge)} {f(spec.overlap)}\n")
        out.write(f"{len(typ)}\n")
        for p, t, c in zip(pos, typ, q):
            out.write(f"{f(p[0])} {f(p[1])} {f(p[2])} {int(t)} {f(c)}\n")
        out.write(f"{int(probe)} {f(spacing)} {len(sites)}\n")
        for p, t, c in zip(sites, site_t, site_q):
            out.write(f"{f(p[0])} {f(p[1])} {f(p[2])} {int(t)} {f(c)}\n")


def methane_itq29_input(path, refvals, itq29):
    rows = refvals["zeolite_forcefield"]["pseudo_atoms"]
    spec = w._ff([(r[0], r[2], r[3], r[4], r[1]) for r in rows], alpha=0.25, use_charge=False)
    a = itq29["a"]
    cell = np.diag([a, a, a])
    typ = np.array([spec.index(e) for e in itq29["element"]], np.int32)
    ch4 = spec.index("CH4")
    r1 = refvals["cbmc_framework_molecule_1"]
    rng = np.random.default_rng(11)
    sites = np.vstack([np.array(r1["methane_fractional"]) * a, rng.uniform(0, a, (63, 3))])
    write_input(path, cell, True, spec, itq29["pos"], typ, np.zeros(len(typ)), ch4, 0.3, sites, [ch4] * 64,
                [0.0] * 64)
    return spec, cell, typ, ch4, r1, sites


def test_c_client_reads_its_input(tmp_path, refvals, itq29):
    """Without a GPU the program gets as far as rg_ctx_create (exit status 1 and the entry point named), so the input
    was read in full (a malformed input exits with 2 before any library call)."""
    import raspa3_amd  # noqa: F401  (the GPU test below needs the library built; here it only has to load)
    exe = compile_client(tmp_path)
    inp = tmp_path / "input.txt"
    methane_itq29_input(inp, refvals, itq29)
    r = subprocess.run([exe, str(inp)], capture_output=True, text=True, timeout=60)
    if r.returncode == 0:
        pytest.skip("a HIP device is visible: the GPU test covers the full run")
    assert r.returncode == 1 and "rg_ctx_create" in r.stderr, (r.returncode, r.stderr)


@pytest.mark.gpu
def test_c_client_golden_and_grid(tmp_path, refvals, itq29):
    exe = compile_client(tmp_path)
    inp = tmp_path / "input.txt"
    spec, cell, typ, ch4, r1, sites = methane_itq29_input(inp, refvals, itq29)
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
