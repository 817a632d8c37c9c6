"""GPU parity of the trial-energy, Rosenbluth and Ewald operators (through the C ABI).

The direct framework/inter-molecular energies and the Ewald Fourier energies are checked against the reference's
own golden values (the same ones that pin the oracle), and then against the oracle on the synthetic
configuration-1 framework at sizes the oracle evaluates in seconds.
"""
import numpy as np
import pytest

import oracle
import raspa3_amd as ra
from raspa3_amd import workloads as w

pytestmark = pytest.mark.gpu

U = oracle.units()
E2K = U["energy_to_kelvin"]


def zeolite_spec(refvals, use_charge, alpha=0.25):
    rows = refvals["zeolite_forcefield"]["pseudo_atoms"]
    return w._ff([(r[0], r[2], r[3], r[4], r[1]) for r in rows], alpha=alpha, use_charge=use_charge)


def itq29_system(ctx, refvals, itq29, use_charge, alpha, kmax=(8, 8, 8)):
    spec = zeolite_spec(refvals, use_charge, alpha)
    ff = ra.ForceField(ctx, spec)
    a = itq29["a"]
    cell = np.diag([a, a, a])
    typ = np.array([spec.index(e) for e in itq29["element"]], np.int32)
    fw = w.Framework("itq29", cell, True, itq29["pos"], typ, itq29["charge"])
    f = ra.Framework(ctx, fw)
    sysm = ra.System(ctx, ff, cell, True, framework=f, kmax=kmax, capacity=64)
    return spec, sysm


def co2_atoms(spec, positions, qC=0.6512, qO=-0.3256):
    p = np.array(positions, float)
    n = len(p) // 3
    t = np.array([spec.index("O_co2"), spec.index("C_co2"), spec.index("O_co2")] * n, np.int32)
    q = np.array([qO, qC, qO] * n)
    m = np.repeat(np.arange(n), 3).astype(np.int32)
    return p, t, q, m


def test_golden_methane_in_itq29(gpu_ctx, refvals, itq29):
    """cbmc_interactions.cpp:28-51 and :53-90 through rg_trial_energies."""
    spec, s = itq29_system(gpu_ctx, refvals, itq29, False, 0.25)
    r1 = refvals["cbmc_framework_molecule_1"]
    ch4 = spec.index("CH4")
    e, ov = s.trial_energies(np.array([r1["methane_fractional"]]) * itq29["a"], [ch4], [0.0], [0])
    assert not ov.any()
    assert abs(e[0, 0] * E2K - r1["framework_vdw_K"]) < 1e-6
    r2 = refvals["cbmc_framework_molecule_2"]
    pos = np.array(r2["methane_fractional"]) * itq29["a"]
    s.set_adsorbates(pos, [ch4, ch4], [0.0, 0.0], [0, 1])
    e, ov = s.trial_energies(pos, [ch4, ch4], [0.0, 0.0], [0, 1])
    assert abs(e[:, 0].sum() * E2K - r2["framework_vdw_K"]) < 1e-6
    assert abs(e[0, 2] * E2K - r2["molecule_vdw_K"]) < 1e-6
    assert abs(e[1, 2] * E2K - r2["molecule_vdw_K"]) < 1e-6


def test_golden_co2_static_energies(gpu_ctx, refvals, itq29):
    """static_energy.cpp:21-47: framework/molecule VDW and Coulomb with automatic Ewald parameters."""
    a = itq29["a"]
    alpha, _ = ra.ewald_parameters(1e-6, 12.0, [a, a, a])
    spec, s = itq29_system(gpu_ctx, refvals, itq29, True, alpha)
    r = refvals["static_energy_co2_itq29"]
    p, t, q, m = co2_atoms(spec, r["co2_positions"])
    s.set_adsorbates(p, t, q, m)
    e, ov = s.trial_energies(p, t, q, m)
    assert not ov.any()
    assert abs(e[:, 0].sum() * E2K - r["framework_vdw_K"]) < 1e-6
    assert abs(e[:, 1].sum() * E2K - r["framework_charge_K"]) < 1e-6
    # each inter-molecular pair is seen from both molecules
    assert abs(0.5 * e[:, 2].sum() * E2K - r["molecule_vdw_K"]) < 1e-6
    assert abs(0.5 * e[:, 3].sum() * E2K - r["molecule_charge_K"]) < 1e-6


def test_golden_ewald_fourier(gpu_ctx, refvals, itq29):
    """ewald.cpp:25-55 (2 CO2 in a 10 A box) and ewald.cpp:146-178 (2 CO2 in ITQ-29, single-ion energy)."""
    r = refvals["ewald_2co2_box10"]
    spec = zeolite_spec(refvals, True, r["alpha"])
    ff = ra.ForceField(gpu_ctx, spec)
    cell = np.diag(r["box"])
    s = ra.System(gpu_ctx, ff, cell, True, kmax=r["kmax"], capacity=16)
    assert s.info()["nk"] == 2600
    p, t, q, m = co2_atoms(spec, r["co2_positions"])
    s.set_adsorbates(p, t, q, m)
    assert abs(s.ewald_energy().sum() * E2K - r["ewald_fourier_K"]) < 1e-6
    r2 = refvals["ewald_2co2_itq29"]
    spec2, s2 = itq29_system(gpu_ctx, refvals, itq29, True, r2["alpha"], r2["kmax"])
    info = s2.info()
    single = -(info["single_ion_sum"] - U["coulomb"] * r2["alpha"] / np.sqrt(np.pi))
    assert abs(single * E2K - r2["single_ion_K"]) < 1e-6
    p, t, q, m = co2_atoms(spec2, refvals["static_energy_co2_itq29"]["co2_positions"])
    s2.set_adsorbates(p, t, q, m)
    assert abs(s2.ewald_energy().sum() * E2K - r2["ewald_fourier_K"]) < 1e-6


def config1_co2_system(ctx, nmol, seed):
    fw, spec = w.framework_config1(), w.forcefield_config1()
    ff = ra.ForceField(ctx, spec)
    f = ra.Framework(ctx, fw)
    s = ra.System(ctx, ff, fw.cell, fw.rect, framework=f, kmax=(8, 8, 8), capacity=3 * 256)
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0, 1, (nmol, 3)) @ fw.cell.T
    axes = rng.normal(size=(nmol, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    pos = np.concatenate([np.stack([c + 1.16 * a, c, c - 1.16 * a]) for c, a in zip(centers, axes)])
    t = np.array([spec.index("O_co2"), spec.index("C_co2"), spec.index("O_co2")] * nmol, np.int32)
    q = np.array([-0.35, 0.70, -0.35] * nmol)
    m = np.repeat(np.arange(nmol), 3).astype(np.int32)
    return fw, spec, s, pos, t, q, m


def test_trial_energies_match_oracle_config1(gpu_ctx):
    """Random trial sites (including overlapping ones) in the 4095-site triclinic framework with 60 CO2."""
    fw, spec, s, pos, t, q, m = config1_co2_system(gpu_ctx, 60, 3)
    s.set_adsorbates(pos, t, q, m)
    pairs = oracle.pairs_from_spec(spec)
    rng = np.random.default_rng(8)
    ns = 200
    sp = rng.uniform(0, 1, (ns, 3)) @ fw.cell.T
    sp[:5] = fw.pos[1365:1370] + 0.3  # sites 0.52 A from framework O atoms overlap
    st = rng.choice([spec.index("C_co2"), spec.index("O_co2")], ns).astype(np.int32)
    sq = np.where(st == spec.index("C_co2"), 0.70, -0.35)
    smol = np.full(ns, 1000, np.int32)
    e, ov = s.trial_energies(sp, st, sq, smol, skip_molecule=7)
    keep = m != 7
    mol = dict(pos=pos[keep], type=t[keep], charge=q[keep], molecule=m[keep])
    for i in range(ns):
        site = dict(pos=sp[i:i + 1], type=st[i:i + 1], charge=sq[i:i + 1], molecule=smol[i:i + 1])
        fo = oracle.framework_energy(fw.cell, fw.rect, fw.pos, fw.type, fw.charge, pairs, 12.0, 12.0, spec.alpha,
                                     True, site)
        io = oracle.intermolecular_energy(fw.cell, fw.rect, mol, pairs, 12.0, 12.0, spec.alpha, True, site)
        if fo is None or io is None:
            assert ov[i], i
            continue
        assert not ov[i], i
        ref = np.array([fo[0], fo[1], io[0], io[1]])
        assert np.all(np.abs(e[i] - ref) <= 1e-9 * np.maximum(1.0, np.abs(ref))), (i, e[i], ref)
    assert ov[:5].all()


def test_trial_energies_with_cell_list_match_oracle(gpu_ctx):
    """Direct framework sums through the framework cell list: charged ITQ-29 5 x 5 x 5 (9000 sites, 59.4 A) has
    four cells of at least one cutoff per axis, so each site visits the 27 cells around it; random sites and sites
    0.4 A from framework O atoms (overlaps) against the oracle's scan over every atom (minimum image).  Configuration 1
    (30 A along c: two cells) keeps the full scan, covered by the test above."""
    print("case: trial energies, framework cell list (ITQ-29 5x5x5)", flush=True)
    spec = w.forcefield_zeolite()
    fw, _ = w.framework_itq29(spec, 5)
    ff = ra.ForceField(gpu_ctx, spec)
    s = ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=ra.Framework(gpu_ctx, fw), kmax=(0, 0, 0), capacity=64)
    pairs = oracle.pairs_from_spec(spec)
    rng = np.random.default_rng(12)
    ns = 400
    sp = rng.uniform(-0.2, 1.2, (ns, 3)) @ fw.cell.T  # also outside the cell: the site's cell is taken modulo 1
    oxygens = np.flatnonzero(fw.type == spec.index("O"))
    sp[:8] = fw.pos[rng.choice(oxygens, 8, replace=False)] + 0.4 / np.sqrt(3.0)
    st = rng.choice([spec.index("C_co2"), spec.index("O_co2")], ns).astype(np.int32)
    sq = np.where(st == spec.index("C_co2"), 0.6512, -0.3256)
    e, ov = s.trial_energies(sp, st, sq, np.full(ns, 1000, np.int32))
    checked = 0
    for i in range(ns):
        site = dict(pos=sp[i:i + 1], type=st[i:i + 1], charge=sq[i:i + 1], molecule=np.array([1000], np.int32))
        fo = oracle.framework_energy(fw.cell, fw.rect, fw.pos, fw.type, fw.charge, pairs, 12.0, 12.0, spec.alpha,
                                     True, site)
        if fo is None:
            assert ov[i], i
            continue
        assert not ov[i], i
        ref = np.array([fo[0], fo[1]])
        assert np.all(np.abs(e[i, :2] - ref) <= 1e-9 * np.maximum(1.0, np.abs(ref))), (i, e[i], ref)
        checked += 1
    assert ov[:8].all() and checked > 150  # about 40 % of random points in the zeolite overlap


def test_ewald_delta_accept_matches_oracle(gpu_ctx):
    """Insertion, displacement and deletion of CO2 in the charged framework at kmax = 8 (2600 k-vectors):
    rg_ewald_delta equals the oracle's energyDifferenceEwaldFourier, and after rg_ewald_accept the stored
    structure factor gives the oracle's total Fourier energy."""
    fw, spec, s, pos, t, q, m = config1_co2_system(gpu_ctx, 20, 5)
    alpha, kmax = spec.alpha, [8, 8, 8]
    s.set_adsorbates(pos[:-3], t[:-3], q[:-3], m[:-3])
    tot, stored = oracle.ewald_total(fw.cell, fw.rect, alpha, kmax, fw.pos, fw.charge, pos[:-3], q[:-3],
                                     np.ones(len(q) - 3), m[:-3])
    gpu_tot = s.ewald_energy()
    assert np.allclose(gpu_tot, tot, rtol=1e-10, atol=1e-6)
    net = q[:-3].sum()
    empty = dict(pos=np.zeros((0, 3)), q=np.zeros(0), molecule=np.zeros(0, np.int32))
    new = dict(pos=pos[-3:], q=q[-3:], molecule=m[-3:])
    d = s.ewald_delta(pos[-3:], q[-3:], m[-3:], np.zeros((0, 3)), [], [])
    do, trial = oracle.ewald_delta(fw.cell, fw.rect, alpha, kmax, stored, new, empty, net)
    assert np.allclose(d, do, rtol=1e-10, atol=1e-6), (d, do)
    s.ewald_accept()
    tot2, stored2 = oracle.ewald_total(fw.cell, fw.rect, alpha, kmax, fw.pos, fw.charge, pos, q, np.ones(len(q)), m)
    # the accepted structure factor carries the new molecule; the adsorbate list (self and exclusion terms) is the
    # caller's and still holds the old atoms until it is set again
    assert np.allclose(s.ewald_energy()[0], tot2[0], rtol=1e-10, atol=1e-6)
    s.set_adsorbates(pos, t, q, m)
    assert np.allclose(s.ewald_energy(), tot2, rtol=1e-10, atol=1e-6)
    # displacement of molecule 4
    sel = m == 4
    moved = pos[sel] + np.array([0.7, -0.4, 0.2])
    d = s.ewald_delta(moved, q[sel], m[sel], pos[sel], q[sel], m[sel])
    do, _ = oracle.ewald_delta(fw.cell, fw.rect, alpha, kmax, stored2, dict(pos=moved, q=q[sel], molecule=m[sel]),
                               dict(pos=pos[sel], q=q[sel], molecule=m[sel]), q.sum())
    assert np.allclose(d, do, rtol=1e-10, atol=1e-6)
    # deletion of molecule 2 (not accepted: the stored state stays)
    sel = m == 2
    d = s.ewald_delta(np.zeros((0, 3)), [], [], pos[sel], q[sel], m[sel])
    do, _ = oracle.ewald_delta(fw.cell, fw.rect, alpha, kmax, stored2, empty,
                               dict(pos=pos[sel], q=q[sel], molecule=m[sel]), q.sum())
    assert np.allclose(d, do, rtol=1e-10, atol=1e-6)


def test_rosenbluth_selection_exact(gpu_ctx):
    """Selection is identical to the reference's cumulative search for the same uniforms; log W matches the
    log-sum-exp of the non-overlapping Boltzmann factors; an all-overlap set selects nothing."""
    rng = np.random.default_rng(2)
    ns, k, beta = 5000, 10, 1.0 / (U["kB"] * 300.0)
    E = rng.normal(-1500.0, 900.0, (ns, k)) / E2K
    ov = rng.random((ns, k)) < 0.2
    ov[0] = True
    u = rng.random(ns)
    sel, lw = ra.rosenbluth_select(gpu_ctx, E, ov, beta, u)
    assert sel[0] == -1 and np.isneginf(lw[0])
    for s in range(1, ns):
        idx = np.flatnonzero(~ov[s])
        if len(idx) == 0:
            assert sel[s] == -1
            continue
        lbf = -beta * E[s, idx]
        assert idx[oracle.select_trial(lbf, u[s])] == sel[s]
        ref = np.log(np.exp(lbf - lbf.max()).sum()) + lbf.max() - np.log(k)
        assert abs(lw[s] - ref) <= 1e-12 * max(1.0, abs(ref))
