"""What the trained network gains over plain CBMC, on the device (through the C ABI).

CO2 in MFI (2 x 2 x 2 cells, the reference's silicalite test structure, charged zeolite force field, Ewald) at 300 K
and 1e5 Pa, 32 systems in one launch: plain CBMC (uniform first beads) against the learned first-bead proposal (an
MLP trained on the device on the orientation-averaged Boltzmann factors the grids give, tabulated on 32^3 voxels).
The proposal must raise the acceptance of insertions, and the loading and isosteric heat must agree with plain CBMC,
since the proposal's density enters the acceptance rule exactly.  The loading is compared through the 32 systems'
own production means, which are independent (sigma = their standard deviation / sqrt(32), within 3 sigma: the five
pooled blocks of the reference's estimator understate the scatter between runs at this loading); the isosteric heat
through the reference's estimators (rg_enthalpy_of_adsorption on the production blocks pooled over the systems,
sigma = 95 % interval / t_4), within 3 sigma.  Each sigma is a five-block estimate (4 degrees of freedom), so the
difference of the two heats over its combined sigma follows roughly Student's t with 8 degrees of freedom: a 2-sigma
bar fails about 8 % of correct runs, a 3-sigma bar about 2 %.  The learned run is not the same draw on every run (the
network's weight gradients are summed with FP32 atomics, so training is not bit-reproducible), so the bar has to hold
for fresh draws.
"""
import numpy as np
import pytest

import raspa3_amd as ra
from raspa3_amd import workloads as w
from raspa3_amd._lib import RG_GRID_EWALD_REAL, RG_GRID_LENNARD_JONES

pytestmark = pytest.mark.gpu

T95_4 = 2.776
KJ = 8.314462618e-3


def pooled(mc, nsys, cells, temperature):
    st = [mc.state(k) for k in range(nsys)]
    wgt = sum(s["block_count"] for s in st)
    u = sum(s["block_energy"] for s in st)
    n = sum(s["block_n"][:, :1] for s in st)
    un = sum(s["block_energy_n"][:, :1] for s in st)
    nn = sum(s["block_nn"][:, :1, :1] for s in st)
    load, load_ci = ra.block_average(n[:, 0] / cells, wgt)
    dh, dh_ci = ra.enthalpy_of_adsorption(wgt, u, n, un, nn, temperature)
    acc = sum(int(s["accepted"][0]) for s in st)
    tri = sum(int(s["trials"][0]) for s in st)
    per = np.array([s["block_n"][:, 0].sum() / s["block_count"].sum() / cells for s in st])
    return dict(loading=per.mean(), loading_sigma=per.std(ddof=1) / np.sqrt(len(per)), pooled_loading=load,
                pooled_loading_sigma=load_ci / T95_4, qst=-dh[0] * ra.ENERGY_TO_KELVIN * KJ,
                qst_sigma=dh_ci[0] * ra.ENERGY_TO_KELVIN * KJ / T95_4, acceptance=acc / tri)


def test_learned_proposal_raises_acceptance_at_matched_loading(gpu_ctx):
    print("case: learned vs plain, CO2 in MFI", flush=True)
    T, P, nsys = 300.0, 1e5, 32
    spec = w.forcefield_zeolite()
    fw, unit = w.framework_mfi(spec, (2, 2, 2))
    ff = ra.ForceField(gpu_ctx, spec)
    fu = ra.Framework(gpu_ctx, w.Framework("mfi", unit, True, fw.pos[:288], fw.type[:288], fw.charge[:288]))
    dims = ra.grid_dims(unit, 0.2)
    types = [spec.index("C_co2"), spec.index("O_co2")]
    grids = {t: ra.Grid(gpu_ctx, unit, True, dims) for t in types}
    coul = ra.Grid(gpu_ctx, unit, True, dims)
    ra.build_grids(gpu_ctx, ff, fu, [(g, RG_GRID_LENNARD_JONES, t, 12.0) for t, g in grids.items()] +
                   [(coul, RG_GRID_EWALD_REAL, 0, 12.0)])
    fb = ra.Framework(gpu_ctx, fw)
    _, kmax = ra.ewald_parameters(1e-6, 12.0, np.diag(fw.cell))
    comp = w.co2_reference(spec)
    # the proposal: Boltzmann factors of CO2 averaged over 16 orientations at 2^19 points, a network trained on them
    s0 = ra.System(gpu_ctx, ff, fw.cell, True, framework=fb, grids=grids, coulomb_grid=coul, kmax=tuple(kmax),
                   capacity=64)
    frac = np.random.default_rng(5).random((1 << 19, 3))
    lw = ra.molecule_weights(gpu_ctx, s0, comp, T, frac, 16, 11)
    beta = ra.beta(T)
    net = ra.Mlp(gpu_ctx, seed=3)
    losses = net.train(frac, np.where(np.isfinite(lw), -lw / beta, 1e7), beta, batch=4096, epochs=3, lr=1e-3, seed=9)
    assert losses[-1] < losses[0]
    n3 = [32, 32, 32]
    prob = net.proposal(n3, 0.1)
    net.close()
    s0.close()

    def run(seed, proposal):
        systems = [ra.System(gpu_ctx, ff, fw.cell, True, framework=fb, grids=grids, coulomb_grid=coul,
                             kmax=tuple(kmax), capacity=8 * 256) for _ in range(nsys)]
        mc = ra.Mc(gpu_ctx, systems, [comp], w.mc_params(temperature=T, pressure_pa=P, seed=seed, max_molecules=256,
                                                         production_cycles=400))
        if proposal is not None:
            for k in range(nsys):
                mc.set_proposal(k, *proposal)
        mc.run(600, sample=False, graph=True)
        base = [mc.state(k) for k in range(nsys)]  # acceptance over the production cycles only
        mc.run(400, sample=True, graph=True)
        r = pooled(mc, nsys, 8.0, T)
        acc = sum(int(mc.state(k)["accepted"][0] - base[k]["accepted"][0]) for k in range(nsys))
        tri = sum(int(mc.state(k)["trials"][0] - base[k]["trials"][0]) for k in range(nsys))
        r["acceptance"] = acc / tri
        mc.close()
        for sy in systems:
            sy.close()
        return r

    plain = run(401, None)
    learned = run(402, (n3, prob))
    print(f"plain: acceptance {plain['acceptance']:.4f}, N/cell {plain['loading']:.4f} +- {plain['loading_sigma']:.4f}, "
          f"Qst {plain['qst']:.3f} +- {plain['qst_sigma']:.3f} kJ/mol", flush=True)
    print(f"learned: acceptance {learned['acceptance']:.4f}, N/cell {learned['loading']:.4f} +- "
          f"{learned['loading_sigma']:.4f}, Qst {learned['qst']:.3f} +- {learned['qst_sigma']:.3f} kJ/mol", flush=True)
    assert learned["acceptance"] > 1.3 * plain["acceptance"]
    assert abs(learned["loading"] - plain["loading"]) <= 3.0 * np.hypot(learned["loading_sigma"], plain["loading_sigma"])
    assert abs(learned["qst"] - plain["qst"]) <= 3.0 * np.hypot(learned["qst_sigma"], plain["qst_sigma"])
