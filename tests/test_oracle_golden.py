"""Pins the FP64 oracle against the reference's own unit-test values (tests/golden/reference_values.json).

These exercise units, Lorentz-Berthelot mixing, the cutoff shift, the minimum-image convention, the LJ and Ewald
real-space pair energies, automatic Ewald parameters, the Ewald Fourier sum with its self, exclusion and net-charge
terms, and the single-ion Fourier energy.
"""
import numpy as np
import pytest

import oracle

TOL_K = 1e-6  # the reference tests' EXPECT_NEAR tolerance


def zeolite(refvals, shifted=True, rc=12.0):
    rows = refvals["zeolite_forcefield"]["pseudo_atoms"]
    names = [r[0] for r in rows]
    isf = [int(r[1]) for r in rows]
    eps = [r[2] for r in rows]
    sig = [r[3] for r in rows]
    pairs = oracle.ff_pairs(eps, sig, isf, shifted, rc, rc)
    return names, pairs


def itq29_framework(itq29, names):
    t = np.array([names.index(e) for e in itq29["element"]], np.int32)
    return itq29["pos"], t, itq29["charge"]


def co2(names, positions, qC=0.6512, qO=-0.3256):
    p = np.array(positions, float)
    n = len(p) // 3
    t = np.array([names.index("O_co2"), names.index("C_co2"), names.index("O_co2")] * n, np.int32)
    q = np.array([qO, qC, qO] * n)
    mol = np.repeat(np.arange(n), 3).astype(np.int32)
    return dict(pos=p, type=t, charge=q, molecule=mol)


def test_units():
    u = oracle.units()
    assert abs(u["energy_to_kelvin"] - 1.2027242847) < 1e-9  # src/raspakit/units.cpp:98 (documented value)
    assert abs(u["coulomb"] - 138935.4835) < 1e-3  # src/raspakit/units.cpp:75-78
    assert abs(u["energy_to_kelvin"] * u["kelvin_to_energy"] - 1.0) < 1e-15


def test_cbmc_framework_methane(refvals, itq29):
    r = refvals["cbmc_framework_molecule_1"]
    names, pairs = zeolite(refvals)
    fpos, ftyp, fq = itq29_framework(itq29, names)
    b = oracle.box(itq29["a"], itq29["a"], itq29["a"], lengths_only=True)
    site = dict(pos=np.array(r["methane_fractional"]) * itq29["a"], type=[names.index("CH4")], charge=[0.0])
    e = oracle.framework_energy(b["cell"], True, fpos, ftyp, fq, pairs, 12.0, 12.0, 0.25, False, site)
    u = oracle.units()
    assert abs(e[0] * u["energy_to_kelvin"] - r["framework_vdw_K"]) < TOL_K


def test_cbmc_framework_two_methanes(refvals, itq29):
    r = refvals["cbmc_framework_molecule_2"]
    names, pairs = zeolite(refvals)
    fpos, ftyp, fq = itq29_framework(itq29, names)
    b = oracle.box(itq29["a"], itq29["a"], itq29["a"], lengths_only=True)
    pos = np.array(r["methane_fractional"]) * itq29["a"]
    sites = dict(pos=pos, type=[names.index("CH4")] * 2, charge=[0.0, 0.0], molecule=[0, 1])
    u = oracle.units()
    e = oracle.framework_energy(b["cell"], True, fpos, ftyp, fq, pairs, 12.0, 12.0, 0.25, False, sites)
    assert abs(e[0] * u["energy_to_kelvin"] - r["framework_vdw_K"]) < TOL_K
    # CBMC::computeInterMolecularEnergy of either molecule against the other (cbmc_interactions.cpp:76-89)
    for m in (0, 1):
        one = dict(pos=pos[m:m + 1], type=[names.index("CH4")], charge=[0.0], molecule=[m])
        ei = oracle.intermolecular_energy(b["cell"], True, sites, pairs, 12.0, 12.0, 0.25, False, one)
        assert abs(ei[0] * u["energy_to_kelvin"] - r["molecule_vdw_K"]) < TOL_K


def test_static_energy_co2_itq29(refvals, itq29):
    r = refvals["static_energy_co2_itq29"]
    names, pairs = zeolite(refvals)
    fpos, ftyp, fq = itq29_framework(itq29, names)
    b = oracle.box(itq29["a"], itq29["a"], itq29["a"], lengths_only=True)
    alpha, kmax = oracle.ewald_auto(1e-6, 12.0, b["widths"])
    mols = co2(names, r["co2_positions"])
    u = oracle.units()
    e = oracle.framework_energy(b["cell"], True, fpos, ftyp, fq, pairs, 12.0, 12.0, alpha, True, mols)
    assert abs(e[0] * u["energy_to_kelvin"] - r["framework_vdw_K"]) < TOL_K
    assert abs(e[1] * u["energy_to_kelvin"] - r["framework_charge_K"]) < TOL_K
    first = {k: v[:3] for k, v in mols.items()}
    ei = oracle.intermolecular_energy(b["cell"], True, mols, pairs, 12.0, 12.0, alpha, True, first)
    assert abs(ei[0] * u["energy_to_kelvin"] - r["molecule_vdw_K"]) < TOL_K
    assert abs(ei[1] * u["energy_to_kelvin"] - r["molecule_charge_K"]) < TOL_K


def test_ewald_two_co2_in_box(refvals):
    r = refvals["ewald_2co2_box10"]
    names, _ = zeolite(refvals)
    b = oracle.box(*r["box"], lengths_only=True)
    m = co2(names, r["co2_positions"])
    out, _ = oracle.ewald_total(b["cell"], True, r["alpha"], r["kmax"], np.zeros((0, 3)), np.zeros(0), m["pos"],
                                m["charge"], np.ones(6), m["molecule"])
    u = oracle.units()
    assert abs(out.sum() * u["energy_to_kelvin"] - r["ewald_fourier_K"]) < TOL_K


def test_ewald_two_co2_in_itq29(refvals, itq29):
    r = refvals["ewald_2co2_itq29"]
    names, _ = zeolite(refvals)
    fpos, ftyp, fq = itq29_framework(itq29, names)
    b = oracle.box(itq29["a"], itq29["a"], itq29["a"], lengths_only=True)
    u = oracle.units()
    single = oracle.ewald_single_ion(b["cell"], True, r["alpha"], r["kmax"])
    assert abs(single * u["energy_to_kelvin"] - r["single_ion_K"]) < TOL_K
    m = co2(names, refvals["static_energy_co2_itq29"]["co2_positions"])
    out, _ = oracle.ewald_total(b["cell"], True, r["alpha"], r["kmax"], fpos, fq, m["pos"], m["charge"], np.ones(6),
                                m["molecule"])
    assert abs(out.sum() * u["energy_to_kelvin"] - r["ewald_fourier_K"]) < TOL_K


def test_ewald_delta_equals_difference_of_totals(refvals, itq29):
    """energyDifferenceEwaldFourier = E(after) - E(before) of the golden-pinned total, for an insertion, a
    deletion and a displacement; the trial structure factor equals the total's stored one after the move."""
    names, _ = zeolite(refvals)
    fpos, ftyp, fq = itq29_framework(itq29, names)
    a = itq29["a"]
    b = oracle.box(a, a, a, lengths_only=True)
    alpha, kmax = 0.25, [8, 8, 8]
    m = co2(names, refvals["static_energy_co2_itq29"]["co2_positions"])
    before, sk_before = oracle.ewald_total(b["cell"], True, alpha, kmax, fpos, fq, m["pos"][:3], m["charge"][:3],
                                           np.ones(3), m["molecule"][:3])
    after, sk_after = oracle.ewald_total(b["cell"], True, alpha, kmax, fpos, fq, m["pos"], m["charge"], np.ones(6),
                                         m["molecule"])
    new = dict(pos=m["pos"][3:], q=m["charge"][3:], molecule=[1, 1, 1])
    empty = dict(pos=np.zeros((0, 3)), q=np.zeros(0), molecule=np.zeros(0))
    net = 0.0 + m["charge"][:3].sum()
    d, trial = oracle.ewald_delta(b["cell"], True, alpha, kmax, sk_before, new, empty, net)
    assert abs(d.sum() - (after.sum() - before.sum())) < 1e-9 * max(1.0, abs(after.sum()))
    assert np.allclose(trial, sk_after, rtol=0, atol=1e-10)
    # deletion is the reverse
    d2, trial2 = oracle.ewald_delta(b["cell"], True, alpha, kmax, sk_after, empty, new, net + m["charge"][3:].sum())
    assert abs(d2.sum() + d.sum()) < 1e-9 * max(1.0, abs(after.sum()))
    assert np.allclose(trial2, sk_before, atol=1e-10)
    # displacement of molecule 1
    moved = dict(pos=m["pos"][3:] + np.array([0.3, -0.2, 0.5]), q=m["charge"][3:], molecule=[1, 1, 1])
    d3, _ = oracle.ewald_delta(b["cell"], True, alpha, kmax, sk_after, moved, new, net + m["charge"][3:].sum())
    pos2 = m["pos"].copy()
    pos2[3:] += np.array([0.3, -0.2, 0.5])
    after2, _ = oracle.ewald_total(b["cell"], True, alpha, kmax, fpos, fq, pos2, m["charge"], np.ones(6),
                                   m["molecule"])
    assert abs(d3.sum() - (after2.sum() - after.sum())) < 1e-9 * max(1.0, abs(after.sum()))


def test_select_trial_exact():
    """selectTrialPosition (src/raspakit/cbmc/util.cpp:15-48): the cumulative search picks the first index whose
    cumulative shifted weight reaches u * sum."""
    lbf = np.array([-3.0, -1.0, -2.0, -1.5])
    w = np.exp(lbf - lbf.max())
    c = np.cumsum(w)
    for u in np.linspace(0.0, 0.999999, 37):
        expect = int(np.searchsorted(c, u * c[-1], side="left"))
        assert oracle.select_trial(lbf, u) == expect


def mfi_supercell(mfi, names, rep):
    shifts = np.array([[i, j, k] for i in range(rep[0]) for j in range(rep[1]) for k in range(rep[2])], float)
    pos = ((mfi["frac"][None, :, :] + shifts[:, None, :]) * mfi["abc"]).reshape(-1, 3)
    t = np.tile(np.array([names.index(e) for e in mfi["element"]], np.int32), len(shifts))
    return pos, t, np.tile(mfi["charge"], len(shifts)), mfi["abc"] * np.array(rep)


def test_static_energy_co2_mfi(refvals, mfi):
    """Two CO2 in MFI 2x2x2 (static_energy.cpp:48-73): pins the Pnma expansion of the reference's asymmetric unit
    (tests/golden/make_fixtures.py) together with the pair terms."""
    r = refvals["static_energy_co2_mfi"]
    names, pairs = zeolite(refvals)
    fpos, ftyp, fq, lengths = mfi_supercell(mfi, names, r["replicate"])
    b = oracle.box(*lengths, lengths_only=True)
    alpha, kmax = oracle.ewald_auto(1e-6, 12.0, b["widths"])
    mols = co2(names, r["co2_positions"])
    u = oracle.units()
    e = oracle.framework_energy(b["cell"], True, fpos, ftyp, fq, pairs, 12.0, 12.0, alpha, True, mols)
    assert abs(e[0] * u["energy_to_kelvin"] - r["framework_vdw_K"]) < TOL_K
    assert abs(e[1] * u["energy_to_kelvin"] - r["framework_charge_K"]) < TOL_K
    first = {k: v[:3] for k, v in mols.items()}
    ei = oracle.intermolecular_energy(b["cell"], True, mols, pairs, 12.0, 12.0, alpha, True, first)
    assert abs(ei[0] * u["energy_to_kelvin"] - r["molecule_vdw_K"]) < TOL_K
    assert abs(ei[1] * u["energy_to_kelvin"] - r["molecule_charge_K"]) < TOL_K


def test_ewald_two_co2_in_mfi(refvals, mfi):
    """Single-ion and total Ewald Fourier energy of two CO2 in MFI 2x2x2 (ewald.cpp:248-280)."""
    r = refvals["ewald_2co2_mfi"]
    names, _ = zeolite(refvals)
    fpos, ftyp, fq, lengths = mfi_supercell(mfi, names, r["replicate"])
    b = oracle.box(*lengths, lengths_only=True)
    u = oracle.units()
    single = oracle.ewald_single_ion(b["cell"], True, r["alpha"], r["kmax"])
    assert abs(single * u["energy_to_kelvin"] - r["single_ion_K"]) < TOL_K
    m = co2(names, refvals["static_energy_co2_mfi"]["co2_positions"])
    out, _ = oracle.ewald_total(b["cell"], True, r["alpha"], r["kmax"], fpos, fq, m["pos"], m["charge"], np.ones(6),
                                m["molecule"])
    assert abs(out.sum() * u["energy_to_kelvin"] - r["ewald_fourier_K"]) < TOL_K
