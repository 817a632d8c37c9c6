"""GPU Monte Carlo engine against the oracle's replay of the same moves (through the C ABI).

Both draw every trial from the same Philox4x32-10 stream positions, so a Widom insertion or a GCMC trajectory is
replayed move by move: the selected trials and every accept/reject decision must agree, and energies must agree to
FP64 summation-order differences.  Tolerances: 1e-9 relative on energies and log Rosenbluth weights, 1e-8 A on
adsorbate positions after a run; move counters exactly.
"""
import json
import os

import numpy as np
import pytest

import oracle
import raspa3_amd as ra
from raspa3_amd import workloads as w
from raspa3_amd._lib import RG_GRID_EWALD_REAL, RG_GRID_LENNARD_JONES

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def close(a, b, rel=1e-9, abs_=1e-9):
    a, b = np.asarray(a, float), np.asarray(b, float)
    with np.errstate(invalid="ignore"):  # -inf - -inf (two all-overlap trials) is NaN; the second term accepts it
        return np.all((np.abs(a - b) <= abs_ + rel * np.abs(b)) | (np.isneginf(a) & np.isneginf(b)))


def test_widom_methane_itq29_grid(gpu_ctx):
    """Config-0 style Widom (single-site CH4, grid path) in the reference's ITQ-29 fixture, 3x3x3 box."""
    print("case: widom CH4 ITQ-29", flush=True)
    d = json.load(open(os.path.join(GOLDEN, "itq29.json")))
    rows = json.load(open(os.path.join(GOLDEN, "reference_values.json")))["zeolite_forcefield"]["pseudo_atoms"]
    spec = w._ff([(r[0], r[2], r[3], 0.0, r[1]) for r in rows], use_charge=False)
    a = d["a"]
    frac = np.array(d["fractional"])
    typ = np.array([spec.index(e) for e in d["element"]], np.int32)
    unit = np.diag([a, a, a])
    ff = ra.ForceField(gpu_ctx, spec)
    fu = ra.Framework(gpu_ctx, w.Framework("itq29", unit, True, frac * a, typ, np.zeros(len(typ))))
    dims = ra.grid_dims(unit, 0.3)
    g = ra.Grid(gpu_ctx, unit, True, dims)
    ra.build_grids(gpu_ctx, ff, fu, [(g, RG_GRID_LENNARD_JONES, spec.index("CH4"), 12.0)])
    big = np.diag([3 * a] * 3)
    pos = np.vstack([frac * a + a * np.array([i, j, k]) for i in range(3) for j in range(3) for k in range(3)])
    fwb = w.Framework("itq29x3", big, True, pos, np.tile(typ, 27), np.zeros(len(pos)))
    fb = ra.Framework(gpu_ctx, fwb)
    s = ra.System(gpu_ctx, ff, big, True, framework=fb, grids={spec.index("CH4"): g}, kmax=(0, 0, 0), capacity=64)
    comp = w.methane_component(spec)
    params = w.mc_params(seed=1234)
    lw, u, ew = ra.widom(gpu_ctx, s, comp, params, 0, 4096)
    grid = g.download().reshape(dims[2], dims[1], dims[0], 8)
    om = oracle.MC(spec, fwb, [comp], params, kmax=(0, 0, 0), grids=({spec.index("CH4"): grid}, None, dims, unit))
    olw, ou, oew = om.widom(0, 0, 4096)
    assert close(lw, olw) and close(u, ou)
    W = np.exp(lw).mean()
    assert W > 1.0  # methane adsorbs in the LTA cages


def test_widom_config0_replays_oracle_with_chemical_potential(gpu_ctx):
    """The CPU-runnable workload of the reference's Widom path: CH4 in the 288-site orthorhombic silicate
    (20.0 x 19.9 x 13.4 A), CH4 grid at 0.2 A (100 x 99 x 67), 1e4 Widom insertions of 10 first-bead trials, seed 42,
    300 K.  Every insertion's Rosenbluth weight and energy replay the oracle; the excess chemical potential and its
    95 % block interval (5 blocks, rg_widom_statistics) equal the oracle's statistics of the oracle's weights.  The
    synthetic framework is non-porous for CH4 (no point of the cell lies more than 2.6 A from a framework atom), so
    <W> (~1e-17) is a rare-event average that 1e5 trials do not converge: the move-by-move replay, not a comparison
    with a quadrature of exp(-beta U), is what pins it here (the porous ITQ-29 Widom tests compare with Henry's law
    in the oracle's tests)."""
    print("case: widom CH4 config 0 (1e4 insertions, seed 42)", flush=True)
    from oracle import statistics as so
    fw, spec = w.framework_config0(), w.forcefield_config0()
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    dims = ra.grid_dims(fw.cell, 0.2)
    ch4 = spec.index("CH4")
    g = ra.Grid(gpu_ctx, fw.cell, fw.rect, dims)
    ra.build_grids(gpu_ctx, ff, f, [(g, RG_GRID_LENNARD_JONES, ch4, 12.0)])
    s = ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids={ch4: g}, kmax=(0, 0, 0), capacity=64)
    comp = w.methane_component(spec)
    params = w.mc_params(seed=42)
    n = 10000
    lw, u, _ = ra.widom(gpu_ctx, s, comp, params, 0, n)
    host = g.download().reshape(dims[2], dims[1], dims[0], 8)
    om = oracle.MC(spec, fw, [comp], params, kmax=(0, 0, 0), grids=({ch4: host}, None, dims))
    olw, ou, _ = om.widom(0, 0, n)
    assert close(lw, olw) and close(u, ou)
    wts = np.exp(lw).reshape(5, -1)
    ow = np.exp(olw).reshape(5, -1)
    blocks = np.full(5, float(wts.shape[1]))
    vol = abs(np.linalg.det(fw.cell))
    st = ra.widom_statistics(blocks, wts.sum(1), np.zeros(5) + 1.0 / vol * blocks, 300.0)
    ost = so.widom_statistics(blocks, ow.sum(1), np.zeros(5) + 1.0 / vol * blocks, 300.0)
    e2k = oracle.units()["energy_to_kelvin"]
    print(f"config 0 Widom: <W> = {st['rosenbluth_weight']:.6e} +- {st['rosenbluth_weight_ci']:.2e}, mu_ex = "
          f"{st['mu_excess'] * e2k:.3f} +- {st['mu_excess_ci'] * e2k:.3f} K", flush=True)
    assert st["mu_excess"] == pytest.approx(ost["mu_excess"], rel=1e-9)
    assert st["mu_excess_ci"] == pytest.approx(ost["mu_excess_ci"], rel=1e-6)


def test_widom_methane_into_loaded_system(gpu_ctx):
    """Widom insertions of CH4 into ITQ-29 holding CH4 molecules from a replayed GCMC run: the single-atom Widom
    kernel (several insertions per wavefront) with inter-molecular LJ, against the oracle holding the same
    molecules."""
    print("case: widom CH4 into a loaded system", flush=True)
    spec = w.forcefield_zeolite()
    t = spec.index("CH4")
    fw, unit, ff, f, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, [t])
    comp = w.methane_component(spec)
    params = w.mc_params(pressure_pa=1e6, seed=77, max_molecules=256, production_cycles=3)
    s = ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids=grids, kmax=(0, 0, 0), capacity=1024)
    mc = ra.Mc(gpu_ctx, [s], [comp], params)
    mc.run(6, sample=False, graph=False)
    om = oracle.MC(spec, fw, [comp], params, kmax=(0, 0, 0), system=0, grids=(host, None, dims, unit))
    om.run(6)
    gp, gt, gc = mc.adsorbates(0)
    op, ot, oc = om.adsorbates()
    assert len(gt) == len(ot) > 20 and np.allclose(gp, op, rtol=0, atol=1e-8)
    loaded = ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids=grids, kmax=(0, 0, 0), capacity=1024)
    loaded.set_adsorbates(gp, gt, np.zeros(len(gt)), np.arange(len(gt)))
    lw, u, ew = ra.widom(gpu_ctx, loaded, comp, params, 1000, 3000)
    olw, ou, oew = om.widom(0, 1000, 3000)
    assert close(lw, olw) and close(u, ou) and np.all(ew == 0.0)
    assert np.isfinite(lw).sum() > 1000
    mc.close()


def test_widom_device_outputs_equal_host_outputs(gpu_ctx):
    """rg_widom_device writes the same insertions as rg_widom, for the single-atom kernel and the general one.  The
    device output arrays are the data of a grid object (device memory the library owns and can copy back)."""
    print("case: widom device outputs", flush=True)
    spec = w.forcefield_zeolite()
    types = [spec.index("CH4"), spec.index("C_co2"), spec.index("O_co2")]
    fw, unit, ff, f, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, types)
    s = ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids=grids, coulomb_grid=coul, kmax=(6, 6, 6),
                  capacity=64)
    params = w.mc_params(seed=3)
    n = 700
    buf = ra.Grid(gpu_ctx, unit, True, (7, 7, 7))  # 2744 doubles >= 3 n
    base = ra.lib().rg_grid_device_data(buf.h)
    for comp in (w.methane_component(spec), w.co2_reference(spec)):
        ref = ra.widom(gpu_ctx, s, comp, params, 50, n)
        buf.upload(np.full(buf.count, 7.0))
        ra.widom_device(gpu_ctx, s, comp, params, 50, n, base, base + 8 * n, base + 16 * n)
        got = buf.download(0, 3 * n).reshape(3, n)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)


def test_widom_co2_config1_direct_ewald(gpu_ctx):
    """CO2 Widom in the 4095-site charged triclinic framework: direct framework sums, orientation trials and the
    Ewald Fourier difference at kmax = 8."""
    print("case: widom CO2 config1", flush=True)
    fw, spec = w.framework_config1(), w.forcefield_config1()
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    s = ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, kmax=(8, 8, 8), capacity=64)
    comp = w.co2_component(spec)
    params = w.mc_params(seed=99)
    lw, u, ew = ra.widom(gpu_ctx, s, comp, params, 100, 32)
    om = oracle.MC(spec, fw, [comp], params, kmax=(8, 8, 8))
    olw, ou, oew = om.widom(0, 100, 32)
    assert close(lw, olw) and close(u, ou) and close(ew, oew, abs_=1e-7)
    assert np.isfinite(lw).sum() > 0


def small_charged_setup(ctx, spec, types, spacing=0.3):
    fw = w.framework_small_charged()
    ff = ra.ForceField(ctx, spec)
    f = ra.Framework(ctx, fw)
    dims = ra.grid_dims(fw.cell, spacing)
    grids = {t: ra.Grid(ctx, fw.cell, fw.rect, dims) for t in types}
    coul = ra.Grid(ctx, fw.cell, fw.rect, dims)
    ra.build_grids(ctx, ff, f, [(g, RG_GRID_LENNARD_JONES, t, 12.0) for t, g in grids.items()] +
                   [(coul, RG_GRID_EWALD_REAL, 0, 12.0)])
    host = {t: g.download().reshape(dims[2], dims[1], dims[0], 8) for t, g in grids.items()}
    hcoul = coul.download().reshape(dims[2], dims[1], dims[0], 8)
    return fw, ff, f, dims, grids, coul, host, hcoul


def itq29_setup(ctx, spec, types, spacing=0.3, scheme=8):
    """Charged ITQ-29, 3 x 3 x 3 cells (35.6 A), grids built on the unit cell (tricubic or triquintic)."""
    fw, unit = w.framework_itq29(spec, 3)
    ff = ra.ForceField(ctx, spec)
    fu = ra.Framework(ctx, w.Framework("itq29", unit, True, fw.pos[:72], fw.type[:72], fw.charge[:72]))
    dims = ra.grid_dims(unit, spacing)
    grids = {t: ra.Grid(ctx, unit, True, dims, scheme=scheme) for t in types}
    coul = ra.Grid(ctx, unit, True, dims, scheme=scheme)
    ra.build_grids(ctx, ff, fu, [(g, RG_GRID_LENNARD_JONES, t, 12.0) for t, g in grids.items()] +
                   [(coul, RG_GRID_EWALD_REAL, 0, 12.0)])
    host = {t: g.download().reshape(dims[2], dims[1], dims[0], scheme) for t, g in grids.items()}
    hcoul = coul.download().reshape(dims[2], dims[1], dims[0], scheme)
    return fw, unit, ff, ra.Framework(ctx, fw), dims, grids, coul, host, hcoul


def compare_with_oracle(mc, sysidx, spec, fw, comps, params, kmax, host, hcoul, dims, gcell, cycles, proposal=None):
    om = oracle.MC(spec, fw, comps, params, kmax=kmax, system=sysidx, grids=(host, hcoul, dims, gcell))
    if proposal is not None:
        # (n3, prob) for every component, or a list of (component, n3, prob)
        for item in (proposal if isinstance(proposal, list) else [(-1,) + tuple(proposal)]):
            om.set_proposal(item[1], item[2], component=item[0])
    om.run(cycles[0])
    om.run(cycles[1], sample=True)
    o = om.state()
    g = mc.state(sysidx)
    assert g["n_molecules"] == o["n_molecules"]
    assert list(g["trials"]) == list(o["trials"]) and list(g["accepted"]) == list(o["accepted"])
    assert close(g["energy"], o["energy"], rel=1e-8, abs_=1e-6)
    gp, gt, gc = mc.adsorbates(sysidx)
    op, ot, oc = om.adsorbates()
    assert np.array_equal(gt, ot) and np.array_equal(gc, oc)
    assert np.allclose(gp, op, rtol=0, atol=1e-8)
    return g


@pytest.mark.parametrize("scheme", [8, 27])
def test_gcmc_co2_replays_oracle_trajectory(gpu_ctx, scheme):
    """Two independent CO2 systems in charged ITQ-29 (grid path, tricubic or triquintic grids, Ewald kmax 6) for
    5 + 15 cycles at 1e5 Pa."""
    print(f"case: gcmc CO2 replay, {scheme} values per grid point", flush=True)
    spec = w.forcefield_zeolite()
    types = [spec.index("C_co2"), spec.index("O_co2")]
    fw, unit, ff, f, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, types, scheme=scheme)
    comp = w.co2_reference(spec)
    params = w.mc_params(pressure_pa=1e5, seed=5, max_molecules=128, production_cycles=15)
    systems = [ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids=grids, coulomb_grid=coul, kmax=(6, 6, 6),
                         capacity=8 * 128) for _ in range(2)]
    mc = ra.Mc(gpu_ctx, systems, [comp], params)
    mc.run(5, sample=False, graph=False)
    mc.run(15, sample=True, graph=True)
    for sidx in range(2):
        g = compare_with_oracle(mc, sidx, spec, fw, [comp], params, (6, 6, 6), host, hcoul, dims, unit, (5, 15))
        assert g["accepted"][0] > 0 and g["accepted"][2] > 0 and g["accepted"][3] > 0
    # drift: running energy equals the energy recomputed from scratch
    tot = mc.total_energy()
    for sidx in range(2):
        e = mc.state(sidx)["energy"]
        assert abs(tot[sidx] - e) <= 1e-7 * max(1.0, abs(e)), (tot[sidx], e)
    mc.close()


def test_gcmc_many_trials_replays_oracle_trajectory(gpu_ctx):
    """Trial sets with more grid terms than lanes: 40 first-bead trials (80 grid terms: the site's grid and the
    Coulomb grid per trial) and 20 orientations of CO2's two other atoms (80 terms), so each set's gathers take two
    rounds of 64 lanes.  Two systems in charged ITQ-29, Ewald kmax 6, 3 + 6 cycles against the oracle's replay."""
    print("case: gcmc CO2 replay, 40 first-bead trials and 20 orientations", flush=True)
    spec = w.forcefield_zeolite()
    types = [spec.index("C_co2"), spec.index("O_co2")]
    fw, unit, ff, f, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, types)
    comp = w.co2_reference(spec)
    params = w.mc_params(pressure_pa=1e5, seed=11, k_first=40, k_orient=20, max_molecules=128, production_cycles=6)
    systems = [ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids=grids, coulomb_grid=coul, kmax=(6, 6, 6),
                         capacity=8 * 128) for _ in range(2)]
    mc = ra.Mc(gpu_ctx, systems, [comp], params)
    mc.run(3, sample=False, graph=False)
    mc.run(6, sample=True, graph=True)
    for sidx in range(2):
        g = compare_with_oracle(mc, sidx, spec, fw, [comp], params, (6, 6, 6), host, hcoul, dims, unit, (3, 6))
        assert g["accepted"][0] > 0
    mc.close()


def test_gcmc_reinsertion_replays_oracle_trajectory(gpu_ctx):
    """CBMC reinsertion (MC_Moves::reinsertionMove, src/raspakit/mc_moves/component/reinsertion.cpp:34-214) on the
    device against the oracle's replay: CO2/N2 80:20 in charged ITQ-29 with Ewald kmax 6, 30 % reinsertions,
    a tabulated first-bead proposal of its own for each component of one of the two systems (the stored weight
    carries the proposal's bias)."""
    print("case: gcmc reinsertion replay", flush=True)
    spec = w.forcefield_zeolite([("N_n2", 36.0, 3.31, -0.482), ("N_com", 0.0, 1.0, 0.964)])
    types = [spec.index(n) for n in ("C_co2", "O_co2", "N_n2", "N_com")]
    fw, unit, ff, f, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, types)
    comps = [w.co2_reference(spec, 0.8), w.n2_component(spec, 0.2)]
    params = w.mc_params(pressure_pa=5e5, seed=91, max_molecules=128, production_cycles=10, p_translation=0.2,
                         p_rotation=0.2, p_reinsertion=0.3)
    systems = [ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids=grids, coulomb_grid=coul, kmax=(6, 6, 6),
                         capacity=8 * 128) for _ in range(2)]
    mc = ra.Mc(gpu_ctx, systems, comps, params)
    n3 = [8, 8, 8]
    rng = np.random.default_rng(5)
    props = [(0, n3, rng.uniform(0.2, 1.0, 512)), (1, [6, 7, 8], rng.uniform(0.2, 1.0, 336))]
    for comp_index, dims3, prob in props:
        mc.set_proposal(1, dims3, prob, component=comp_index)
    mc.run(4, sample=False, graph=True)
    mc.run(10, sample=True, graph=True)
    for sidx in range(2):
        g = compare_with_oracle(mc, sidx, spec, fw, comps, params, (6, 6, 6), host, hcoul, dims, unit, (4, 10),
                                proposal=props if sidx == 1 else None)
        assert g["trials"][4] > 100 and g["accepted"][4] > 10
    tot = mc.total_energy()
    for sidx in range(2):
        e = mc.state(sidx)["energy"]
        assert abs(tot[sidx] - e) <= 1e-7 * max(1.0, abs(e)), (tot[sidx], e)
    mc.close()


def test_gcmc_hexane_reinsertion_replays_oracle_trajectory(gpu_ctx):
    """CBMC reinsertion of flexible chains (MC_Moves::reinsertionMove, src/raspakit/mc_moves/component/reinsertion.cpp
    :34-214): two n-hexane systems in ITQ-29 with 40 % reinsertion moves, each chain regrown bead by bead with its own
    atoms left out and its old configuration retraced with Esselink's stored first-bead weight, replay the oracle
    move by move for 4 + 12 cycles; the running energy equals the recomputed total."""
    print("case: gcmc hexane reinsertion replay", flush=True)
    spec, fw, unit, ff, f, dims, grids, host = hexane_itq29(gpu_ctx)
    comp = w.hexane_component(spec, 0.3)
    params = w.mc_params(pressure_pa=10.0, seed=17, max_molecules=64, production_cycles=12, p_translation=0.15,
                         p_rotation=0.15, p_reinsertion=0.4)
    systems = [ra.System(gpu_ctx, ff, fw.cell, True, framework=f, grids=grids, kmax=(0, 0, 0), capacity=8 * 64)
               for _ in range(2)]
    mc = ra.Mc(gpu_ctx, systems, [comp], params)
    mc.run(4, sample=False, graph=False)
    mc.run(12, sample=True, graph=True)
    for sidx in range(2):
        g = compare_with_oracle(mc, sidx, spec, fw, [comp], params, (0, 0, 0), host, None, dims, unit, (4, 12))
        assert g["n_molecules"] > 0 and g["trials"][4] > 0
    assert sum(mc.state(s)["accepted"][4] for s in range(2)) > 0
    tot = mc.total_energy()
    for sidx in range(2):
        e = mc.state(sidx)["energy"]
        assert abs(tot[sidx] - e) <= 1e-7 * max(1.0, abs(e)), (tot[sidx], e)
    mc.close()


def test_gcmc_mixture_with_proposal_replays_oracle(gpu_ctx):
    """CO2/N2 80:20 in charged ITQ-29 with a tabulated first-bead proposal (random positive voxel probabilities)."""
    print("case: gcmc mixture + proposal replay", flush=True)
    spec = w.forcefield_zeolite([("N_n2", 36.0, 3.31, -0.482), ("N_com", 0.0, 1.0, 0.964)])
    types = [spec.index(n) for n in ("C_co2", "O_co2", "N_n2", "N_com")]
    fw, unit, ff, f, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, types)
    comps = [w.co2_reference(spec, 0.8), w.n2_component(spec, 0.2)]
    params = w.mc_params(pressure_pa=5e5, seed=77, max_molecules=128, production_cycles=10)
    s = ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids=grids, coulomb_grid=coul, kmax=(6, 6, 6),
                  capacity=8 * 128)
    mc = ra.Mc(gpu_ctx, [s], comps, params)
    rng = np.random.default_rng(3)
    n3 = [8, 8, 8]
    prob = rng.uniform(0.2, 1.0, 512)
    mc.set_proposal(0, n3, prob)
    mc.run(4, sample=False, graph=True)
    mc.run(10, sample=True, graph=True)
    compare_with_oracle(mc, 0, spec, fw, comps, params, (6, 6, 6), host, hcoul, dims, unit, (4, 10),
                        proposal=(n3, prob))
    st = mc.state(0)
    assert st["n_per_component"][0] > 0 and st["n_molecules"] == sum(st["n_per_component"])
    mc.close()


def test_graph_replay_equals_eager(gpu_ctx):
    """Replaying one captured MC cycle gives the same bits as launching it eagerly, with host synchronisation and
    state queries between the replays."""
    print("case: graph vs eager", flush=True)
    spec = w.forcefield_zeolite()
    types = [spec.index("C_co2"), spec.index("O_co2")]
    fw, unit, ff, f, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, types, spacing=0.4)
    comp = w.co2_reference(spec)
    params = w.mc_params(pressure_pa=1e5, seed=8, max_molecules=64, production_cycles=6)
    runs = []
    for graph in (False, True):
        systems = [ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids=grids, coulomb_grid=coul,
                             kmax=(6, 6, 6), capacity=8 * 64) for _ in range(3)]
        mc = ra.Mc(gpu_ctx, systems, [comp], params)
        trace = []
        for _ in range(6):
            mc.run(1, sample=True, graph=graph)
            trace.append([mc.state(i)["energy"] for i in range(3)])
        runs.append((trace, [mc.adsorbates(i)[0] for i in range(3)]))
        mc.close()
    assert runs[0][0] == runs[1][0]
    assert all(len(p) > 0 for p in runs[0][1])
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)


def test_molecule_weights_match_oracle(gpu_ctx):
    """Training data of the learned proposal: orientation-averaged Boltzmann factors of CO2 from the grids."""
    print("case: molecule weights", flush=True)
    spec = w.forcefield_config1()
    types = [spec.index("C_co2"), spec.index("O_co2")]
    fw, ff, f, dims, grids, coul, host, hcoul = small_charged_setup(gpu_ctx, spec, types, spacing=0.4)
    comp = w.co2_component(spec)
    s = ra.System(gpu_ctx, ff, fw.cell, fw.rect, framework=f, grids=grids, coulomb_grid=coul, kmax=(6, 6, 6),
                  capacity=64)
    frac = np.random.default_rng(2).random((777, 3))
    for m in (1, 16, 64):
        lw = ra.molecule_weights(gpu_ctx, s, comp, 300.0, frac, n_orient=m, seed=17)
        om = oracle.MC(spec, fw, [comp], w.mc_params(), kmax=(6, 6, 6), grids=(host, hcoul, dims))
        olw = om.molecule_weights(0, frac, m, 17)
        assert close(lw, olw, rel=1e-9, abs_=1e-9), m
    assert np.isfinite(lw).any() and np.isneginf(lw).any()


def hexane_itq29(ctx, spacing=0.3):
    spec = w.forcefield_hexane_zeolite()
    spec.use_charge = False
    fw, unit = w.framework_itq29(spec, 3)
    ff = ra.ForceField(ctx, spec)
    fu = ra.Framework(ctx, w.Framework("itq29", unit, True, fw.pos[:72], fw.type[:72], fw.charge[:72]))
    dims = ra.grid_dims(unit, spacing)
    types = [spec.index("CH3"), spec.index("CH2")]
    grids = {t: ra.Grid(ctx, unit, True, dims) for t in types}
    ra.build_grids(ctx, ff, fu, [(g, RG_GRID_LENNARD_JONES, t, 12.0) for t, g in grids.items()])
    host = {t: g.download().reshape(dims[2], dims[1], dims[0], 8) for t, g in grids.items()}
    return spec, fw, unit, ff, ra.Framework(ctx, fw), dims, grids, host


def test_widom_hexane_replays_oracle(gpu_ctx):
    """Flexible-chain growth (n-hexane, 6 beads, bend and torsion sampled by rejection) in ITQ-29: the GPU's Widom
    insertions equal the oracle's, trial for trial."""
    print("case: widom hexane", flush=True)
    spec, fw, unit, ff, f, dims, grids, host = hexane_itq29(gpu_ctx)
    comp = w.hexane_component(spec, 0.3)
    s = ra.System(gpu_ctx, ff, fw.cell, True, framework=f, grids=grids, kmax=(0, 0, 0), capacity=64)
    params = w.mc_params(seed=31)
    lw, u, _ = ra.widom(gpu_ctx, s, comp, params, 5, 256)
    om = oracle.MC(spec, fw, [comp], params, kmax=(0, 0, 0), grids=(host, None, dims, unit))
    olw, ou, _ = om.widom(0, 5, 256)
    assert close(lw, olw, rel=1e-9, abs_=1e-9) and close(u, ou, rel=1e-9, abs_=1e-6)
    assert np.isfinite(lw).sum() > 100


def test_gcmc_hexane_replays_oracle_trajectory(gpu_ctx):
    """Two n-hexane GCMC systems in ITQ-29 (chain insertion, deletion with retrace, translation, rotation) for
    4 + 12 cycles replay the oracle move by move; the running energy (with the chains' bend, torsion and
    intra-molecular LJ) equals the recomputed total."""
    print("case: gcmc hexane replay", flush=True)
    spec, fw, unit, ff, f, dims, grids, host = hexane_itq29(gpu_ctx)
    comp = w.hexane_component(spec, 0.3)
    params = w.mc_params(pressure_pa=10.0, seed=13, max_molecules=64, production_cycles=12)
    systems = [ra.System(gpu_ctx, ff, fw.cell, True, framework=f, grids=grids, kmax=(0, 0, 0), capacity=8 * 64)
               for _ in range(2)]
    mc = ra.Mc(gpu_ctx, systems, [comp], params)
    mc.run(4, sample=False, graph=False)
    mc.run(12, sample=True, graph=True)
    for sidx in range(2):
        g = compare_with_oracle(mc, sidx, spec, fw, [comp], params, (0, 0, 0), host, None, dims, unit, (4, 12))
        assert g["accepted"][0] > 0 and g["n_molecules"] > 0
    tot = mc.total_energy()
    for sidx in range(2):
        e = mc.state(sidx)["energy"]
        assert abs(tot[sidx] - e) <= 1e-7 * max(1.0, abs(e)), (tot[sidx], e)
    mc.close()


def test_gcmc_uneven_batch_replays_oracle(gpu_ctx):
    """Systems of different sizes in one launch: n-hexane (flexible) + CO2 (Ewald kmax 6) at 1e5 Pa and mole
    fractions 0.1 / 0.9 in ITQ-29 3x3x3 (1944 sites) and 3x3x4 (2592 sites), interpolation grids shared from the
    unit cell.  Each system replays its own oracle trajectory for 3 + 8 cycles, and its running energy equals the
    recomputed total."""
    print("case: gcmc uneven batch (hexane + CO2, 3x3x3 and 3x3x4)", flush=True)
    spec = w.forcefield_hexane_zeolite()
    types = [spec.index(n) for n in ("CH3", "CH2", "C_co2", "O_co2")]
    fw3, unit, ff, f3, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, types)
    fw4, _ = w.framework_itq29(spec, (3, 3, 4))
    f4 = ra.Framework(gpu_ctx, fw4)
    comps = [w.hexane_component(spec, 0.3, 0.1), w.co2_reference(spec, 0.9)]
    params = w.mc_params(pressure_pa=1e5, seed=19, max_molecules=96, production_cycles=8)
    fws = [fw3, fw4]
    systems = [ra.System(gpu_ctx, ff, fw.cell, True, framework=fo, grids=grids, coulomb_grid=coul, kmax=(6, 6, 6),
                         capacity=6 * 96) for fw, fo in zip(fws, [f3, f4])]
    mc = ra.Mc(gpu_ctx, systems, comps, params)
    mc.run(3, sample=False, graph=True)
    mc.run(8, sample=True, graph=True)
    for sidx, fw in enumerate(fws):
        g = compare_with_oracle(mc, sidx, spec, fw, comps, params, (6, 6, 6), host, hcoul, dims, unit, (3, 8))
        assert g["accepted"][0] > 0 and g["n_per_component"][0] > 0 and g["n_per_component"][1] > 0
    tot = mc.total_energy()
    for sidx in range(2):
        e = mc.state(sidx)["energy"]
        assert abs(tot[sidx] - e) <= 1e-7 * max(1.0, abs(e)), (tot[sidx], e)
    mc.close()


def replay_to_steps(om, target, sample):
    """Runs the oracle's cycles one at a time until its step counter reaches target (a balanced launch runs whole
    cycles, so the counter must land on it exactly)."""
    while om.state()["steps"] < target:
        om.run(1, sample=sample)
    assert om.state()["steps"] == target


def test_balanced_launches_replay_oracle(gpu_ctx):
    """Balanced launches (rg_mc_set_balance) over uneven systems: CO2 (Ewald kmax 6) at 1e5 Pa in charged ITQ-29
    3x3x3, 3x3x4 and 4x4x4 in one launch per cycle.  Each system replays the oracle's trajectory: the oracle runs the
    same cycles one after another, matched launch by launch through the step counter; the samples of sampled launch i
    land in block i with exactly the oracle's sums over that launch's cycles; and the less loaded systems run more
    cycles per launch than the most loaded one."""
    print("case: balanced launches, CO2 in ITQ-29 3x3x3 / 3x3x4 / 4x4x4", flush=True)
    spec = w.forcefield_zeolite()
    types = [spec.index("C_co2"), spec.index("O_co2")]
    fw3, unit, ff, f3, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, types)
    fws = [fw3] + [w.framework_itq29(spec, c)[0] for c in ((3, 3, 4), (4, 4, 4))]
    fos = [f3] + [ra.Framework(gpu_ctx, fw) for fw in fws[1:]]
    comp = w.co2_reference(spec)
    params = w.mc_params(pressure_pa=1e5, seed=29, max_molecules=160, production_cycles=4)
    systems = [ra.System(gpu_ctx, ff, fw.cell, True, framework=fo, grids=grids, coulomb_grid=coul, kmax=(6, 6, 6),
                         capacity=8 * 160) for fw, fo in zip(fws, fos)]
    mc = ra.Mc(gpu_ctx, systems, [comp], params)
    # cost = steps x molecules: the 4x4x4 system (2.4 times the 3x3x3 volume) sets the budget once loaded (while
    # filling, systems of any size gain about the same number of molecules per cycle, so loadings part only later)
    mc.set_balance(True, 0.0)
    mc.run(60, sample=False, graph=True)
    eq_steps = [mc.state(s)["steps"] for s in range(3)]
    launch_steps = []
    for _ in range(4):
        mc.run(1, sample=True, graph=True)
        launch_steps.append([mc.state(s)["steps"] for s in range(3)])
    cycles = [mc.state(s)["cycles"] for s in range(3)]
    print("sampled cycles per system over 4 launches:", cycles, "molecules:",
          [mc.state(s)["n_molecules"] for s in range(3)], flush=True)
    for sidx, fw in enumerate(fws):
        om = oracle.MC(spec, fw, [comp], params, kmax=(6, 6, 6), system=sidx, grids=(host, hcoul, dims, unit))
        replay_to_steps(om, eq_steps[sidx], False)
        g = mc.state(sidx)
        for i in range(4):
            before = om.state()["blocks"].sum(0)
            replay_to_steps(om, launch_steps[i][sidx], True)
            d = om.state()["blocks"].sum(0) - before
            assert g["block_count"][i] == d[0] and g["block_n"][i][0] == d[2] and g["block_nn"][i][0][0] == d[6]
            assert close(g["block_energy"][i], d[1], rel=1e-8, abs_=1e-6)
            assert close(g["block_energy_n"][i][0], d[4], rel=1e-8, abs_=1e-6)
        assert g["block_count"][4] == 0
        o = om.state()
        assert g["n_molecules"] == o["n_molecules"] and g["cycles"] == o["cycles"]
        assert list(g["trials"]) == list(o["trials"]) and list(g["accepted"]) == list(o["accepted"])
        assert close(g["energy"], o["energy"], rel=1e-8, abs_=1e-6)
        gp, gt, gc = mc.adsorbates(sidx)
        op, ot, oc = om.adsorbates()
        assert np.array_equal(gt, ot) and np.array_equal(gc, oc)
        assert np.allclose(gp, op, rtol=0, atol=1e-8)
    assert min(cycles) >= 4 and max(cycles) > 4
    tot = mc.total_energy()
    for sidx in range(3):
        e = mc.state(sidx)["energy"]
        assert abs(tot[sidx] - e) <= 1e-7 * max(1.0, abs(e)), (tot[sidx], e)
    mc.close()


def test_balanced_launches_are_refused_for_chains(gpu_ctx):
    """The balanced kernel is built for rigid molecules only."""
    spec, fw, unit, ff, f, dims, grids, host = hexane_itq29(gpu_ctx)
    s = ra.System(gpu_ctx, ff, fw.cell, True, framework=f, grids=grids, kmax=(0, 0, 0), capacity=8 * 64)
    mc = ra.Mc(gpu_ctx, [s], [w.hexane_component(spec, 0.3)], w.mc_params(max_molecules=64))
    with pytest.raises(ra.RGError, match="rigid"):
        mc.set_balance(True)
    mc.close()


@pytest.mark.parametrize("case", ["co2", "hexane"])
def test_launch_shapes_give_identical_trajectories(gpu_ctx, case):
    """Every launch shape (one wavefront per system with 1, 2 or 4 systems per workgroup, or a team of four
    wavefronts per system splitting the adsorbate chunks and the Ewald k-vector passes) computes the same
    trajectory bit for bit: the heavy loops are cut into fixed partitions combined in a fixed order."""
    print(f"case: launch shapes {case}", flush=True)
    if case == "co2":
        spec = w.forcefield_zeolite()
        types = [spec.index("C_co2"), spec.index("O_co2")]
        fw, unit, ff, f, dims, grids, coul, host, hcoul = itq29_setup(gpu_ctx, spec, types)
        comps, kmax = [w.co2_reference(spec)], (6, 6, 6)
        params = w.mc_params(pressure_pa=3e5, seed=61, max_molecules=128, production_cycles=6)
    else:
        spec, fw, unit, ff, f, dims, grids, host = hexane_itq29(gpu_ctx)
        coul, comps, kmax = None, [w.hexane_component(spec, 0.3)], (0, 0, 0)
        params = w.mc_params(pressure_pa=20.0, seed=62, max_molecules=64, production_cycles=6)
    states = []
    for shape in range(4):
        systems = [ra.System(gpu_ctx, ff, fw.cell, True, framework=f, grids=grids, coulomb_grid=coul, kmax=kmax,
                             capacity=8 * 128) for _ in range(5)]
        mc = ra.Mc(gpu_ctx, systems, comps, params)
        mc.set_launch_shape(shape)
        mc.run(6, sample=False, graph=shape % 2 == 0)
        mc.run(6, sample=True, graph=shape % 2 == 1)
        states.append([(mc.state(s), mc.adsorbates(s)) for s in range(5)])
        mc.close()
    for other in states[1:]:
        for (a, pa), (b, pb) in zip(states[0], other):
            assert a["energy"] == b["energy"] and list(a["accepted"]) == list(b["accepted"])
            assert np.array_equal(a["block_energy"], b["block_energy"])
            assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])
    assert sum(s[0]["n_molecules"] for s in states[0]) > 0
