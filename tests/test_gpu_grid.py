"""GPU parity of the grid builder and tricubic interpolation against the FP64 oracle (through the C ABI).

Tolerance (per stored value, internal energy units): |gpu - oracle| <= 1e-3 K + 1e-9 |oracle|.  The absolute
term is the bar the GPU path is held to on energies; the relative term covers the derivative channels, whose
values near framework atoms reach 1e8 and more, where FP64 summation order alone moves the last digits.
"""
import numpy as np
import pytest

import oracle
import raspa3_amd as ra
from raspa3_amd import workloads as w
from raspa3_amd._lib import RG_GRID_EWALD_REAL, RG_GRID_LENNARD_JONES, RGError

pytestmark = pytest.mark.gpu

K_TO_E = oracle.units()["kelvin_to_energy"]
ABS = 1e-3 * K_TO_E
REL = 1e-9
NTHREADS = 16


def assert_close(gpu, ref, what):
    err = np.abs(gpu - ref)
    bound = ABS + REL * np.abs(ref)
    bad = np.flatnonzero(err > bound)
    assert bad.size == 0, f"{what}: {bad.size} values out of tolerance; worst {err.reshape(-1)[bad[:5]]} " \
                          f"ref {np.asarray(ref).reshape(-1)[bad[:5]]}"


def test_config0_full_grid_matches_oracle(gpu_ctx):
    """CH4 LJ grid, orthorhombic 20.0 x 19.9 x 13.4, h = 0.2 -> 100 x 99 x 67, 2x2x2 supercell: every point."""
    print("case: config0 full LJ grid", flush=True)
    fw, spec = w.framework_config0(), w.forcefield_config0()
    dims = ra.grid_dims(fw.cell, 0.2)
    assert dims == [100, 99, 67]
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    g = ra.Grid(gpu_ctx, fw.cell, fw.rect, dims)
    ra.build_grids(gpu_ctx, ff, f, [(g, RG_GRID_LENNARD_JONES, spec.index("CH4"), spec.rc_framework)])
    got = g.download().reshape(dims[2], dims[1], dims[0], 8)
    ref = oracle.grid_build(fw.cell, fw.rect, dims, 0, fw.pos, fw.type, fw.charge, oracle.pairs_from_spec(spec),
                            spec.index("CH4"), 12.0, 12.0, spec.alpha, nthreads=NTHREADS)
    m, r = channel_errors(got, ref)
    print(f"config0 CH4 grid vs oracle, every point: max |d| K per channel {np.array2string(m, precision=2)}, "
          f"rms {np.array2string(r, precision=2)}", flush=True)
    assert_close(got, ref, "config0")
    # overlap clamps are present and reproduced exactly
    clamped = ref[..., 0] == 2.0 * spec.overlap
    assert clamped.any()
    assert np.array_equal(got[..., 0][clamped], ref[..., 0][clamped])


def config1_grids(ctx):
    fw, spec = w.framework_config1(), w.forcefield_config1()
    dims = ra.grid_dims(fw.cell, 0.1)
    ff = ra.ForceField(ctx, spec)
    f = ra.Framework(ctx, fw)
    grids = [ra.Grid(ctx, fw.cell, fw.rect, dims) for _ in range(3)]
    specs = [(grids[0], RG_GRID_LENNARD_JONES, spec.index("C_co2"), 12.0),
             (grids[1], RG_GRID_LENNARD_JONES, spec.index("O_co2"), 12.0),
             (grids[2], RG_GRID_EWALD_REAL, 0, 12.0)]
    return fw, spec, dims, ff, f, grids, specs


def sample_points(dims, n, seed):
    rng = np.random.default_rng(seed)
    ijk = np.stack([rng.integers(0, d, n) for d in dims], 1)
    corners = np.array([[i, j, k] for i in (0, dims[0] - 1) for j in (0, dims[1] - 1) for k in (0, dims[2] - 1)])
    return np.vstack([ijk, corners]).astype(np.int32)


def gather(grid, dims, ijk):
    out = np.empty((len(ijk), 8))
    for q, (i, j, k) in enumerate(ijk):
        off = 8 * (int(i) + dims[0] * (int(j) + dims[1] * int(k)))
        out[q] = grid.download(off, 8)
    return out


def test_config1_fused_grids_match_oracle(gpu_ctx):
    """Triclinic 40 x 40 x 30, gamma 100, 4095 charged sites, h = 0.1 -> 394 x 394 x 300 (46.6M points per grid):
    C_co2 and O_co2 LJ grids and the unit-charge Ewald real-space grid, built in one fused pass at full size;
    2,000 sampled points plus the 8 corners are compared with the oracle."""
    print("case: config1 fused grids", flush=True)
    fw, spec, dims, ff, f, grids, specs = config1_grids(gpu_ctx)
    ra.build_grids(gpu_ctx, ff, f, specs)
    ijk = sample_points(dims, 2000, 5)
    pairs = oracle.pairs_from_spec(spec)
    for (g, gt, probe, _), name in zip(specs, ["C_co2", "O_co2", "ewald"]):
        ref = oracle.grid_points(fw.cell, fw.rect, dims, gt, fw.pos, fw.type, fw.charge, pairs, probe, 12.0, 12.0,
                                 spec.alpha, ijk, nthreads=NTHREADS)
        got = gather(g, dims, ijk)
        m, r = channel_errors(got, ref)
        print(f"config1 {name} vs oracle, 2008 points: max |d| K per channel {np.array2string(m, precision=2)}, "
              f"rms {np.array2string(r, precision=2)}", flush=True)
        assert_close(got, ref, f"config1 {name}")


def channel_errors(gpu, ref):
    """Max and RMS |gpu - ref| per stored channel, in K."""
    d = np.abs(np.asarray(gpu) - np.asarray(ref)).reshape(-1, 8) / K_TO_E
    return d.max(0), np.sqrt((d * d).mean(0))


SPLIT_RADIUS = 7.0


def test_config1_split_precision_ewald_grid(gpu_ctx):
    """The Ewald real-space grid of configuration 1 with FP32 pair terms beyond 7 A (rg_grid_build_many_precision):
    every one of the 46.6M points within 1e-3 K of the FP64 build on every channel, and 2,000 sampled points plus the
    corners within the same bar of the oracle."""
    print("case: config1 split-precision Ewald grid", flush=True)
    fw, spec, dims, ff, f, grids, specs = config1_grids(gpu_ctx)
    g64, g32 = grids[2], grids[0]
    ra.build_grids(gpu_ctx, ff, f, [(g64, RG_GRID_EWALD_REAL, 0, 12.0)])
    ra.build_grids(gpu_ctx, ff, f, [(g32, RG_GRID_EWALD_REAL, 0, 12.0)], split_radius=SPLIT_RADIUS)
    n = 8 * dims[0] * dims[1] * dims[2]
    chunk = 8 * dims[0] * dims[1] * 25
    mx, s2 = np.zeros(8), np.zeros(8)
    for off in range(0, n, chunk):
        c = min(chunk, n - off)
        m, r = channel_errors(g32.download(off, c), g64.download(off, c))
        mx = np.maximum(mx, m)
        s2 += r * r * (c // 8)
    rms = np.sqrt(s2 / (n // 8))
    print(f"split {SPLIT_RADIUS} A vs FP64, all points: max |d| K per channel {np.array2string(mx, precision=2)}, "
          f"rms {np.array2string(rms, precision=2)}", flush=True)
    assert np.all(mx <= 1e-3), mx
    ijk = sample_points(dims, 2000, 6)
    ref = oracle.grid_points(fw.cell, fw.rect, dims, 3, fw.pos, fw.type, fw.charge, oracle.pairs_from_spec(spec), 0,
                             12.0, 12.0, spec.alpha, ijk, nthreads=NTHREADS)
    got = gather(g32, dims, ijk)
    m, r = channel_errors(got, ref)
    print(f"split vs oracle, 2008 points: max |d| K {np.array2string(m, precision=2)}", flush=True)
    assert np.all(np.abs(got - ref) <= ABS + REL * np.abs(ref))


def test_split_precision_rejects_an_unfit_radius(gpu_ctx):
    """A split radius whose alpha r range is too wide for the FP32 erfcx polynomial is refused, not approximated."""
    fw, spec, dims, ff, f, grids, specs = config1_grids(gpu_ctx)
    with pytest.raises(RGError, match="split radius"):
        ra.build_grids(gpu_ctx, ff, f, [(grids[2], RG_GRID_EWALD_REAL, 0, 12.0)], split_radius=0.5)


def test_fused_equals_single_and_is_deterministic(gpu_ctx):
    """Size-independent properties at full config-1 size: a grid built alone equals the same grid from the fused
    pass bit for bit, and rebuilding gives identical bits."""
    print("case: fused vs single", flush=True)
    fw, spec, dims, ff, f, grids, specs = config1_grids(gpu_ctx)
    ra.build_grids(gpu_ctx, ff, f, specs)
    fused = [g.download(0, 8 * dims[0] * dims[1] * 20) for g in grids]
    single = ra.Grid(gpu_ctx, fw.cell, fw.rect, dims)
    for q, (g, gt, probe, cut) in enumerate(specs):
        ra.build_grids(gpu_ctx, ff, f, [(single, gt, probe, cut)])
        assert np.array_equal(single.download(0, fused[q].size), fused[q])
    ra.build_grids(gpu_ctx, ff, f, specs)
    assert np.array_equal(grids[0].download(0, fused[0].size), fused[0])


def test_small_triclinic_with_supercell(gpu_ctx):
    """A small triclinic cell (supercell 2x2x2 for rc = 12) with a 2-point axis, all grid types, every point."""
    print("case: small triclinic", flush=True)
    cell, rect = w.cell_from_lengths_angles(9.0, 10.0, 11.0, 80.0, 95.0, 105.0)
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, 1, (10, 3)) @ cell.T
    typ = np.array([0, 1] * 5, np.int32)
    q = np.array([1.3, -1.3] * 5)
    fw = w.Framework("small", cell, rect, pos, typ, q)
    spec = w.forcefield_config1()
    dims = [5, 2, 7]
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    gs = [ra.Grid(gpu_ctx, cell, rect, dims) for _ in range(2)]
    ra.build_grids(gpu_ctx, ff, f, [(gs[0], RG_GRID_LENNARD_JONES, spec.index("O_co2"), 12.0),
                                    (gs[1], RG_GRID_EWALD_REAL, 0, 12.0)])
    pairs = oracle.pairs_from_spec(spec)
    for g, gt, pr in ((gs[0], 0, spec.index("O_co2")), (gs[1], 3, 0)):
        ref = oracle.grid_build(cell, rect, dims, gt, pos, typ, q, pairs, pr, 12.0, 12.0, spec.alpha)
        assert_close(g.download().reshape(ref.shape), ref, f"small type {gt}")


def test_supercell_narrower_than_cutoff_is_rejected(gpu_ctx):
    fw, spec = w.framework_config0(), w.forcefield_config0()
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    g = ra.Grid(gpu_ctx, fw.cell, fw.rect, [10, 10, 10])
    with pytest.raises(RGError, match="supercell"):
        ra.build_grids(gpu_ctx, ff, f, [(g, RG_GRID_LENNARD_JONES, spec.index("CH4"), 5.0)])


def test_interpolation_matches_oracle(gpu_ctx):
    """Tricubic interpolation of the config-0 grid at 20,000 random positions (inside and outside the cell)."""
    print("case: interpolation", flush=True)
    fw, spec = w.framework_config0(), w.forcefield_config0()
    dims = ra.grid_dims(fw.cell, 0.2)
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    g = ra.Grid(gpu_ctx, fw.cell, fw.rect, dims)
    ra.build_grids(gpu_ctx, ff, f, [(g, RG_GRID_LENNARD_JONES, spec.index("CH4"), 12.0)])
    data = g.download()
    rng = np.random.default_rng(9)
    pos = rng.uniform(-0.5, 1.5, (20000, 3)) @ fw.cell.T
    e, grad = g.interpolate(pos, gradient=True)
    eo, go = oracle.interpolate(fw.cell, fw.rect, dims, data, pos, gradient=True)
    assert_close(e, eo, "interpolate")
    assert np.allclose(grad, go, rtol=1e-9, atol=1e-6)


def test_interpolation_hessian_matches_oracle(gpu_ctx):
    """Tricubic energies, gradients and Hessians of the interpolant at 5,000 random positions."""
    print("case: tricubic hessian", flush=True)
    cell, rect = w.cell_from_lengths_angles(12.0, 13.0, 12.5, 85.0, 95.0, 100.0)
    rng = np.random.default_rng(21)
    dims = (9, 10, 8)
    data = rng.normal(size=(dims[2], dims[1], dims[0], 8))
    g = ra.Grid(gpu_ctx, cell, rect, dims)
    g.upload(data)
    pos = rng.uniform(-0.5, 1.5, (5000, 3)) @ cell.T
    e, gr, h = g.interpolate_hessian(pos)
    eo, go, ho = oracle.interpolate_hessian(cell, rect, dims, data, pos)
    np.testing.assert_allclose(e, eo, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(gr, go, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(h, ho, rtol=1e-9, atol=1e-9)


def _quintic_case(triclinic):
    if triclinic:
        cell, rect = w.cell_from_lengths_angles(9.0, 10.0, 11.0, 80.0, 95.0, 105.0)
    else:
        cell, rect = np.diag([9.5, 10.5, 11.0]), True
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, 1, (10, 3)) @ cell.T
    typ = np.array([0, 1] * 5, np.int32)
    q = np.array([1.3, -1.3] * 5)
    return cell, rect, pos, typ, q


@pytest.mark.parametrize("triclinic,dims", [(False, (6, 5, 7)), (True, (6, 5, 7)), (True, (5, 2, 7))])
def test_triquintic_grids_match_oracle(gpu_ctx, triclinic, dims):
    """Triquintic LJ and Ewald real-space grids (27 values per point) at every point of a small framework with a
    supercell, against the oracle's values; per channel the absolute bound also allows 1e-12 of the channel's
    largest magnitude (the moments of the device builder group the same pair terms differently)."""
    print(f"case: triquintic grids triclinic={triclinic} dims={dims}", flush=True)
    cell, rect, pos, typ, q = _quintic_case(triclinic)
    fw = w.Framework("small", cell, rect, pos, typ, q)
    spec = w.forcefield_config1()
    dims = list(dims)
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    gs = [ra.Grid(gpu_ctx, cell, rect, dims, scheme=27) for _ in range(2)]
    ra.build_grids(gpu_ctx, ff, f, [(gs[0], RG_GRID_LENNARD_JONES, spec.index("O_co2"), 12.0),
                                    (gs[1], RG_GRID_EWALD_REAL, 0, 12.0)])
    pairs = oracle.pairs_from_spec(spec)
    ijk = np.array([(i, j, k) for k in range(dims[2]) for j in range(dims[1]) for i in range(dims[0])])
    for g, gt, pr in ((gs[0], 0, spec.index("O_co2")), (gs[1], 3, 0)):
        ref = oracle.grid_points_q(cell, rect, dims, gt, pos, typ, q, pairs, pr, 12.0, 12.0, spec.alpha, ijk,
                                   nthreads=NTHREADS)
        got = g.download().reshape(ref.shape)
        bound = ABS + REL * np.abs(ref) + 1e-12 * np.abs(ref).max(axis=0)
        bad = np.argwhere(np.abs(got - ref) > bound)
        assert bad.size == 0, f"type {gt}: {len(bad)} values out of tolerance, first {bad[:5]}"
        # interpolation: energy, gradient and Hessian at random positions against the oracle on the same data
        rng = np.random.default_rng(3 + gt)
        pts = rng.uniform(-0.2, 1.2, (3000, 3)) @ cell.T
        e, gr, h = g.interpolate_hessian(pts)
        eo, go, ho = oracle.interpolate_q(cell, rect, dims, got.reshape(dims[2], dims[1], dims[0], 27), pts)
        for a, b, what in ((e, eo, "energy"), (gr, go, "gradient"), (h, ho, "hessian")):
            np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(b).max(), err_msg=what)
        e2 = g.interpolate(pts)
        np.testing.assert_allclose(e2, e, rtol=1e-13, atol=1e-13 * np.abs(e).max())


def test_trial_energies_read_triquintic_grids(gpu_ctx):
    """Trial-site framework energies (rg_trial_energies) from triquintic LJ and Coulomb grids equal the oracle's
    triquintic interpolation of the same grid data (Coulomb scaled by the site charge)."""
    print("case: trial energies on triquintic grids", flush=True)
    cell, rect, pos, typ, q = _quintic_case(True)
    fw = w.Framework("small", cell, rect, pos, typ, q)
    spec = w.forcefield_config1()
    dims = [9, 10, 11]
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    probe = spec.index("C_co2")
    lj, coul = ra.Grid(gpu_ctx, cell, rect, dims, 27), ra.Grid(gpu_ctx, cell, rect, dims, 27)
    ra.build_grids(gpu_ctx, ff, f, [(lj, RG_GRID_LENNARD_JONES, probe, 12.0), (coul, RG_GRID_EWALD_REAL, 0, 12.0)])
    s = ra.System(gpu_ctx, ff, cell, rect, framework=f, grids={probe: lj}, coulomb_grid=coul, kmax=(0, 0, 0),
                  capacity=16)
    rng = np.random.default_rng(12)
    n = 500
    sp = rng.uniform(0, 1, (n, 3)) @ cell.T
    st = np.full(n, probe, np.int32)
    sq = np.full(n, 0.7)
    e, ov = s.trial_energies(sp, st, sq, np.full(n, 1000, np.int32))
    shape = (dims[2], dims[1], dims[0], 27)
    elj = oracle.interpolate_q(cell, rect, dims, lj.download().reshape(shape), sp)[0]
    eq = 0.7 * oracle.interpolate_q(cell, rect, dims, coul.download().reshape(shape), sp)[0]
    ok = ~np.asarray(ov, bool)
    assert ok.sum() > n // 2
    assert np.allclose(e[ok, 0], elj[ok], rtol=1e-9, atol=1e-9)
    assert np.allclose(e[ok, 1], eq[ok], rtol=1e-9, atol=1e-9)
    assert np.all(elj[~ok] > spec.overlap)


def test_mixed_schemes_are_rejected(gpu_ctx):
    cell, rect, pos, typ, q = _quintic_case(True)
    fw = w.Framework("small", cell, rect, pos, typ, q)
    spec = w.forcefield_config1()
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    g3 = ra.Grid(gpu_ctx, cell, rect, [4, 4, 4])
    g5 = ra.Grid(gpu_ctx, cell, rect, [4, 4, 4], scheme=27)
    with pytest.raises(RGError, match="scheme"):
        ra.build_grids(gpu_ctx, ff, f, [(g3, RG_GRID_LENNARD_JONES, spec.index("O_co2"), 12.0),
                                        (g5, RG_GRID_LENNARD_JONES, spec.index("O_co2"), 12.0)])
    with pytest.raises(RGError, match="scheme"):
        ra.grid_check(gpu_ctx, ff, f, g5, spec.index("O_co2"), coulomb_grid=g3, n_points=10)
    with pytest.raises(RGError):
        ra.Grid(gpu_ctx, cell, rect, [4, 4, 4], scheme=1)


def _check_close(g, o):
    assert g["n_points"] == o["n_points"] and g["n_coulomb_points"] == o["n_coulomb_points"]
    for part in ("lennard_jones", "coulomb"):
        a, b = g[part], o[part]
        for key in ("weight", "energy_table", "energy_full"):
            assert abs(a[key] - b[key]) <= 1e-9 * abs(b[key]) + 1e-9, (part, key, a[key], b[key])
        for key in ("gradient_table", "gradient_full"):
            assert np.allclose(a[key], b[key], rtol=1e-8, atol=1e-8), (part, key, a[key], b[key])
        assert abs(a["energy_relative_error"] - b["energy_relative_error"]) <= 1e-6 * b["energy_relative_error"]
        assert np.allclose(a["gradient_relative_error"], b["gradient_relative_error"], rtol=1e-6, atol=0)
        for key in ("hessian_table", "hessian_full"):
            assert np.allclose(a[key], b[key], rtol=1e-8, atol=1e-8 * max(1.0, np.abs(b[key]).max())), \
                (part, key, a[key], b[key])
        assert np.allclose(a["hessian_relative_error"], b["hessian_relative_error"], rtol=1e-6, atol=0)


@pytest.mark.parametrize("scheme", [8, 27])
@pytest.mark.parametrize("case", ["itq29", "triclinic"])
def test_grid_check_matches_oracle(gpu_ctx, case, scheme):
    """rg_grid_check (the reference's test of a built grid, system_interpolation_grids.cpp:240-840) against the
    oracle's restatement on the same grids and 2000 given points: Boltzmann weights, averages of the interpolated
    and analytic energies, gradients and (triquintic) Hessians, relative errors, for the LJ grid and the
    charge-scaled Ewald grid."""
    print(f"case: grid check {case}, {scheme} values per point", flush=True)
    if case == "itq29":
        spec = w.forcefield_zeolite()
        fw3, cell = w.framework_itq29(spec, 3)
        pos, typ, q, rect = fw3.pos[:72], fw3.type[:72], fw3.charge[:72], True
        dims = ra.grid_dims(cell, 0.3)
    else:
        spec = w.forcefield_config1()
        cell, rect = w.cell_from_lengths_angles(9.0, 10.0, 11.0, 80.0, 95.0, 105.0)
        rng = np.random.default_rng(4)
        pos = rng.uniform(0, 1, (10, 3)) @ cell.T
        typ = np.array([spec.index("Si"), spec.index("O")] * 5, np.int32)
        q = np.array([1.3, -1.3] * 5)
        dims = [12, 13, 14]
    probe = spec.index("C_co2")
    fw = w.Framework(case, cell, rect, pos, typ, q)
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    lj, coul = ra.Grid(gpu_ctx, cell, rect, dims, scheme), ra.Grid(gpu_ctx, cell, rect, dims, scheme)
    ra.build_grids(gpu_ctx, ff, f, [(lj, RG_GRID_LENNARD_JONES, probe, 12.0), (coul, RG_GRID_EWALD_REAL, 0, 12.0)])
    frac = np.random.default_rng(21).uniform(0, 1, (2000, 3))
    qp = float(spec.charge[probe])
    g = ra.grid_check(gpu_ctx, ff, f, lj, probe, coul, qp, 300.0, frac=frac)
    o = oracle.grid_check(cell, rect, dims, pos, typ, q, oracle.pairs_from_spec(spec), probe, spec.rc_framework,
                          12.0, spec.alpha, lj.download(), frac, 300.0, coulomb_grid=coul.download(),
                          probe_charge=qp, overlap=spec.overlap)
    _check_close(g, o)
    assert 0 < g["n_coulomb_points"] < g["n_points"] == 2000
    # the default test: 1e5 Philox points, as the reference runs it after every build
    d = ra.grid_check(gpu_ctx, ff, f, lj, probe, coul, qp, 300.0, n_points=100000, seed=5)
    assert d["n_points"] == 100000 and np.isfinite(d["lennard_jones"]["energy_relative_error"])


def test_grid_cache_round_trip_and_cube_export(gpu_ctx, tmp_path):
    """rg_grid_save / rg_grid_load give back the built grid bit for bit and refuse a file built from other inputs;
    rg_grid_write_cube writes the reference's cube header and every stored value in memory order, in K
    (InterpolationEnergyGrid::writeOutput, interpolation_energy_grid.cpp:4760-4790)."""
    cell, rect = w.cell_from_lengths_angles(12.0, 13.0, 12.5, 85.0, 95.0, 100.0)
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, 1, (24, 3)) @ cell.T
    typ = np.array([0, 1, 1] * 8, np.int32)
    q = np.array([2.05, -1.025, -1.025] * 8)
    fw = w.Framework("cache", cell, rect, pos, typ, q)
    spec = w.forcefield_config1()
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    dims = [11, 12, 9]
    g = ra.Grid(gpu_ctx, cell, rect, dims)
    ra.build_grids(gpu_ctx, ff, f, [(g, RG_GRID_EWALD_REAL, 0, 12.0)])
    key = ra.grid_key(ff, f, RG_GRID_EWALD_REAL, 0, 12.0)
    assert key != ra.grid_key(ff, f, RG_GRID_EWALD_REAL, 0, 12.0, split_radius=7.0)
    assert key != ra.grid_key(ff, f, RG_GRID_LENNARD_JONES, spec.index("O_co2"), 12.0)
    path = tmp_path / "ewald.grid"
    ra.grid_save(gpu_ctx, g, key, path)
    h = ra.Grid(gpu_ctx, cell, rect, dims)
    ra.grid_load(gpu_ctx, h, key, path)
    assert np.array_equal(h.download(), g.download())
    with pytest.raises(RGError, match="different grid"):
        ra.grid_load(gpu_ctx, h, key + 1, path)
    with pytest.raises(RGError, match="different grid"):
        ra.grid_load(gpu_ctx, ra.Grid(gpu_ctx, cell, rect, [11, 12, 10]), key, path)
    cube = tmp_path / "grid.cube"
    ra.grid_write_cube(gpu_ctx, g, cube)
    lines = cube.read_text().splitlines()
    assert lines[0] == "Cube interpolation file" and lines[1] == "Written by RASPA-3"
    assert [float(x) for x in lines[2].split()] == [1.0, 0.0, 0.0, 0.0]
    for a in range(3):
        row = lines[3 + a].split()
        assert int(row[0]) == -dims[a]
        assert np.allclose([float(x) for x in row[1:]], cell[:, a] / dims[a], rtol=0, atol=0)
    assert lines[6] == "1 0.0 0.0 0.0 0.0"
    vals = np.array([float(x) for x in lines[7:]])
    assert vals.size == 8 * np.prod(dims)
    assert np.allclose(vals, g.download() * oracle.units()["energy_to_kelvin"], rtol=1e-15, atol=0)
