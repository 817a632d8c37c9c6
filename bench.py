"""Benchmark of the framework-energy path on one MI355X.

Headline (value): interpolation-grid build rate.  Workload: CO2 grids with charges — triclinic cell a = b = 40,
c = 30 A, gamma = 100 deg, 4095 framework sites (1365 Si +2.05 e, 2730 O -1.025 e), LJ on O (44 K, 3.3 A) with
Lorentz-Berthelot mixing to C_co2 (27 K, 2.80 A) and O_co2 (79 K, 3.05 A), Ewald alpha = 0.265 A^-1, 12 A cutoff,
h = 0.1 A -> 394 x 394 x 300 points per grid.  One step builds the three grids (C_co2 LJ, O_co2 LJ, unit-charge
Ewald real space: 139.7M grid points, 8.94 GB FP64) with inputs resident in HBM: one launch for the two LJ grids and
one for the Ewald grid.  The Ewald grid is built in split precision: pair terms of images beyond 7 A (where the
pair energy and its derivatives are small) have their radial factors in FP32, every sum is FP64, and the grid is
compared point by point with the all-FP64 build of the same kernel, which is timed beside it (--split-radius 0
makes the all-FP64 build the headline).
value = grid points per second (framework-energy evaluations with their tricubic derivatives) over the timed steps.
roofline = vector FLOP/s of the dominant kernel (the Ewald grid) against the MI355X FP64 vector peak (78.6 TFLOP/s;
the fraction of the FP32 vector peak, 157.3 TFLOP/s, is given beside it), with the algorithmic FLOP count of the
frozen convention: 79 FLOP per in-cutoff LJ pair and 86 FLOP (+3 transcendentals, not counted) per in-cutoff Ewald
real-space pair, triclinic cell; the LJ kernel and the all-FP64 Ewald kernel are reported beside it.
cpu_baseline = the FP64 oracle (the reference's loop order and arithmetic restated in C++) on the host cores, on a
bounded sample of z-planes of the same three grids.

Second half of the metric, in the same JSON line:
* trial_energies: Widom insertions through the device Monte Carlo kernel — united-atom CH4 in the configuration-0
  framework (grid at 0.2 A) and rigid CO2 in the configuration-1 framework (grids at 0.1 A, Ewald kmax 8); trial
  site energies per second, with the grid-gather roofline (512 B per interpolated site).
* gcmc: CO2 GCMC at 300 K and 1e5 Pa in charged ITQ-29 (3 x 3 x 3 cells, the reference's test structure and zeolite
  force field), many independent systems in one launch per cycle (hipGraph replay), plain CBMC against the learned
  first-bead proposal: accepted insertions per second over the production cycles, loading and isosteric heat with
  their standard errors (over systems), and whether they agree within 2 sigma.
* mlp: training throughput of the proposal network (BF16 MFMA) on 2^21 points of the configuration-1 CO2 grids at
  300 K, samples/s and MFMA TFLOP/s at 1.08 MFLOP per training sample.
* grid_check: the test the reference runs after building grids (1e5 random points per LJ grid, Boltzmann-weighted
  relative errors of the interpolated energy and gradient, and of the Ewald grid scaled by the probe charge).
* gcmc_mfi: the same comparison in MFI (2 x 2 x 2 cells, the reference's silicalite test structure), proposal trained
  at the simulation temperature; near saturation at 1e5 Pa.
* gcmc_uneven (bench_gcmc_uneven, switched off): batched CBMC over 64 systems of uneven size and loading (ITQ-29 supercells of
  1944-7200 sites), flexible n-hexane chains and rigid CO2 with the Ewald difference at kmax 8, one launch per
  cycle.  No option runs it: its launch-shape tuning pass took a GPU memory access fault whose cause has not been
  found.
* gcmc_mixture: CO2/N2 80:20 GCMC at 300 K and 1e5 Pa in 256 systems of uneven size (charged ITQ-29 supercells of
  2592-15552 sites, Ewald kmax 8) in balanced launches (every system at least one cycle per launch, the smaller ones
  more while the most loaded runs one), plain CBMC against the learned first-bead proposal (one network per
  adsorbate, trained on the unit cell, tiled over each supercell): accepted insertions/s, loading per cell of both
  components and the isosteric heat, and whether they agree within 2 sigma.
* grids_triquintic: the three grids of the headline workload in the triquintic scheme (27 values per point, one
  launch per grid), per-kernel HIP-event times, FP64 FLOP rates, and the grid test with Hessians.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
FLOP_LJ_TRIC = 79
FLOP_EWALD_TRIC = 86
# FP64 FLOP per in-cutoff pair of the triclinic triquintic kernels, counted as for the tricubic ones (FMA = 2,
# reciprocal 1, square root, exp and erfc not counted): distance 8, fractional offsets 3, radial factors f0..f6 44
# (LJ) or 64 (Ewald), monomials 26, 61 moment FMAs 122
FLOP_LJ_QUINTIC = 203
FLOP_EWALD_QUINTIC = 223
GRID_BYTES_PER_SITE = 512      # 8 corner chunks of 64 B per tricubic interpolation
MLP_FLOP_PER_SAMPLE = 1.08e6   # forward 2 * 180.5k + backward 4 * 180.5k (GEMM work of a training sample)
# HBM bytes per launch of the split-precision Ewald grid kernel on the headline workload, from two rocprofv3 --pmc
# runs of "bench.py --grids-only --steps 1 --warmup 1 --no-cpu-baseline" (FETCH_SIZE 25,752 KiB, WRITE_SIZE
# 2,910,675 KiB for the first dispatch), each scaled by what the same counter reads for a 1 GiB device-to-device copy ("bench.py
# --copy-calibration 1024": FETCH_SIZE 524,300 KiB -> x2.0, WRITE_SIZE 1,048,576 KiB -> x1.0).  The algorithmic
# bytes are the 64 B written per point (2.98e9) and the 1.2 MB image list.
EWALD_GRID_HBM_BYTES = 25752 * 1024 * 2.0 + 2910675 * 1024 * 1.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 64))


def incutoff_counts(fw, lj_types, rc, nsample=20000, seed=0):
    """Mean number of in-cutoff framework images per grid point (all sites, and LJ-active sites), estimated on
    random points of the cell with an explicit image list."""
    from scipy.spatial import cKDTree

    H = fw.cell
    Hinv = np.linalg.inv(H)
    w = 1.0 / np.linalg.norm(Hinv, axis=1)
    m = rc / w
    f = fw.pos @ Hinv.T
    imgs, lj = [], []
    rng = [range(int(np.floor(-m[a] - 1)), int(np.ceil(1 + m[a] + 1)) + 1) for a in range(3)]
    for n0 in rng[0]:
        for n1 in rng[1]:
            for n2 in rng[2]:
                g = f + np.array([n0, n1, n2])
                keep = np.all((g >= -m - 1e-9) & (g <= 1 + m + 1e-9), axis=1)
                imgs.append(g[keep] @ H.T)
                lj.append(np.isin(fw.type[keep], lj_types))
    imgs = np.vstack(imgs)
    lj = np.concatenate(lj)
    pts = np.random.default_rng(seed).uniform(0, 1, (nsample, 3)) @ H.T
    t_all = cKDTree(imgs)
    t_lj = cKDTree(imgs[lj])
    n_all = np.array([len(x) for x in t_all.query_ball_point(pts, rc - 1e-12)])
    n_lj = np.array([len(x) for x in t_lj.query_ball_point(pts, rc - 1e-12)])
    return n_all.mean(), n_lj.mean()


def cpu_baseline(fw, spec, dims, target_s=12.0, gpu_grids=None):
    """Times the FP64 oracle (the reference's loop order and arithmetic) on every host core over whole z-planes of the
    three grids: median of 3 runs of the baseline-ISA build (the reference forbids -march) and of a -march=native
    build.  With gpu_grids ({name: Grid}) the oracle's planes of the native run are also compared with the same planes
    of the GPU grids: max and RMS |gpu - oracle| per stored channel in K."""
    import oracle

    pairs = oracle.pairs_from_spec(spec)
    cores = host_cores()
    grids = [("C_co2", 0, spec.index("C_co2")), ("O_co2", 0, spec.index("O_co2")), ("ewald", 3, 0)]

    def run(planes, native):
        t = time.perf_counter()
        out = {}
        for name, gt, probe in grids:
            out[name] = oracle.grid_build(fw.cell, fw.rect, dims, gt, fw.pos, fw.type, fw.charge, pairs, probe,
                                          spec.rc_framework, spec.rc_framework, spec.alpha, k_range=(0, planes),
                                          nthreads=cores, native=native)
        return time.perf_counter() - t, out

    t1, _ = run(1, False)
    planes = int(max(1, min(dims[2], round(target_s / 6.0 / max(t1, 1e-3)))))
    pts = 3 * planes * dims[0] * dims[1]
    rates, ref = {}, None
    for native in (False, True):
        ts = []
        for _ in range(3):
            t, out = run(planes, native)
            ts.append(t)
            if native and ref is None:
                ref = out
        rates["native" if native else "baseline_isa"] = dict(median_s=float(np.median(ts)), runs_s=ts,
                                                              points_per_s=pts / float(np.median(ts)))
    value = rates["baseline_isa"]["points_per_s"]
    base = dict(value=value, unit="grid points/s", cores=cores, kind="port", builds=rates,
                sample=f"FP64 oracle (reference loop order, all framework atoms per point), {planes} z-plane(s) x "
                       f"{dims[0]}x{dims[1]} points x 3 grids = {pts} points per run, median of 3 runs on {cores} "
                       f"threads (rows of points split over the threads); value = the baseline-ISA build, "
                       f"-march=native beside it")
    if gpu_grids:
        k2e = oracle.units()["kelvin_to_energy"]
        err = {}
        for name, g in gpu_grids.items():
            r = ref["ewald" if name.startswith("ewald") else name].reshape(-1, 8)
            got = g.download(0, r.size).reshape(-1, 8)
            d = np.abs(got - r) / k2e
            err[name] = dict(max_abs_K=d.max(0).tolist(), rms_abs_K=np.sqrt((d * d).mean(0)).tolist())
        base["grid_error_vs_oracle"] = dict(points_per_grid=planes * dims[0] * dims[1], planes=[0, planes], **err)
    return base


def grid_channel_errors(a, b, chunk_points, n_points, k_to_e):
    """Max and RMS |a - b| per stored channel (K) of two tricubic grids, streamed through host memory in chunks."""
    mx, s2 = np.zeros(8), np.zeros(8)
    for p0 in range(0, n_points, chunk_points):
        c = min(chunk_points, n_points - p0)
        d = np.abs(a.download(8 * p0, 8 * c) - b.download(8 * p0, 8 * c)).reshape(-1, 8) / k_to_e
        mx = np.maximum(mx, d.max(0))
        s2 += (d * d).sum(0)
    return mx, np.sqrt(s2 / n_points)


def bench_grids(ctx, args):
    import torch

    import raspa3_amd as ra
    from raspa3_amd import workloads as w
    from raspa3_amd._lib import RG_GRID_EWALD_REAL, RG_GRID_LENNARD_JONES

    fw, spec = w.framework_config1(), w.forcefield_config1()
    dims = ra.grid_dims(fw.cell, 0.1)
    npts = dims[0] * dims[1] * dims[2]
    ff = ra.ForceField(ctx, spec)
    f = ra.Framework(ctx, fw)
    grids = [ra.Grid(ctx, fw.cell, fw.rect, dims) for _ in range(3)]
    g64 = ra.Grid(ctx, fw.cell, fw.rect, dims)  # the same Ewald grid with every pair term in FP64
    specs = [(grids[0], RG_GRID_LENNARD_JONES, spec.index("C_co2"), spec.rc_framework),
             (grids[1], RG_GRID_LENNARD_JONES, spec.index("O_co2"), spec.rc_framework),
             (grids[2], RG_GRID_EWALD_REAL, 0, spec.rc_coulomb)]
    split = args.split_radius
    log(f"grids: 3 x {dims} ({3 * npts} points), {fw.n} framework sites, Ewald pair terms in FP32 beyond {split} A")
    # one step = the two LJ grids (one launch, FP64) and the Ewald real-space grid (one launch, FP32 pair terms
    # beyond the split radius, FP64 within it and for every accumulation), each call bracketed by HIP events on the
    # library's stream, so the per-kernel averages come from the timed region itself
    lj_specs, q_specs = specs[:2], specs[2:]
    q64_specs = [(g64, RG_GRID_EWALD_REAL, 0, spec.rc_coulomb)]
    for _ in range(max(1, args.warmup)):
        ra.build_grids(ctx, ff, f, lj_specs)
        ra.build_grids(ctx, ff, f, q_specs, split_radius=split)
        ra.build_grids(ctx, ff, f, q64_specs)
    stream = torch.cuda.ExternalStream(ctx.stream, device=torch.device("cuda", 0))
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        ra.build_grids(ctx, ff, f, lj_specs, sync=False)
        ev[k][1].record(stream)
        ra.build_grids(ctx, ff, f, q_specs, sync=False, split_radius=split)
        ev[k][2].record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    lj_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / args.steps
    q_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / args.steps
    value = 3 * npts / (ms_per_step / 1000.0)
    log(f"grids: ms/step {ms_per_step:.3f} (LJ kernel {lj_ms:.3f}, Ewald kernel {q_ms:.3f}); {value:.4e} grid points/s")
    # the FP64 Ewald kernel, timed the same way outside the headline step
    ev64 = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    ctx.sync()
    for k in range(args.steps):
        ev64[k].record(stream)
        ra.build_grids(ctx, ff, f, q64_specs, sync=False)
    ev64[args.steps].record(stream)
    ctx.sync()
    q64_ms = sum(ev64[k].elapsed_time(ev64[k + 1]) for k in range(args.steps)) / args.steps
    log(f"grids: FP64 Ewald kernel {q64_ms:.3f} ms")

    n_all, n_lj = incutoff_counts(fw, [spec.index("O")], spec.rc_framework)
    flop_q = npts * FLOP_EWALD_TRIC * n_all
    flop_lj = npts * 2 * FLOP_LJ_TRIC * n_lj
    achieved = flop_q / (q_ms / 1000.0) / 1e12
    lj_achieved = flop_lj / (lj_ms / 1000.0) / 1e12
    a64 = flop_q / (q64_ms / 1000.0) / 1e12
    roof = dict(bound="valu", achieved=round(achieved, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                frac=round(achieved / FP64_PEAK_TFLOPS, 4), traffic=EWALD_GRID_HBM_BYTES,
                traffic_algorithmic=float(npts * 64), traffic_note="HBM bytes per Ewald launch (rocprofv3 "
                "FETCH_SIZE and WRITE_SIZE scaled by a 1 GiB copy); the kernel writes 64 B per point and is bound "
                "by VALU issue, not HBM",
                peak_note="FP64 vector peak: every sum and every pair term within the split radius is FP64, the "
                          "radial factors of the pairs beyond it are FP32 (unpacked FP32 FMAs measured at the FP64 "
                          "issue rate, 4.3 against 4.25 cycles per wave64 instruction); frac_vs_fp32_peak prices the "
                          "same FLOP count against the 157.3 TFLOP/s FP32 figure, which needs packed FP32",
                frac_vs_fp32_peak=round(achieved / 157.3, 4),
                kernel=f"k_grid_build<0,true,true,true> (Ewald real-space grid, triclinic, FP32 pair terms beyond "
                       f"{split} A), the dominant kernel of the step",
                flop_per_launch=float(flop_q), kernel_ms=round(q_ms, 4),
                incutoff_pairs_per_point=dict(lj_active=round(n_lj, 2), charged=round(n_all, 2)),
                other_kernels=dict(lj=dict(kernel="k_grid_build<2,false,true> (C_co2 + O_co2 LJ grids, FP64)",
                                           kernel_ms=round(lj_ms, 4), achieved=round(lj_achieved, 3),
                                           frac=round(lj_achieved / FP64_PEAK_TFLOPS, 4),
                                           flop_per_launch=float(flop_lj)),
                                   ewald_fp64=dict(kernel="k_grid_build<0,true,true> (Ewald grid, every pair term "
                                                          "FP64)", kernel_ms=round(q64_ms, 4), achieved=round(a64, 3),
                                                   frac=round(a64 / FP64_PEAK_TFLOPS, 4))),
                step_achieved=round((flop_q + flop_lj) / ((lj_ms + q_ms) / 1000.0) / 1e12, 3))
    # per-point error of the split Ewald grid against the FP64 one, every point of the grid
    import oracle
    k2e = oracle.units()["kelvin_to_energy"]
    t = time.perf_counter()
    mx, rms = grid_channel_errors(grids[2], g64, dims[0] * dims[1] * 25, npts, k2e)
    split_err = dict(split_radius_A=split, points=npts, max_abs_K=mx.tolist(), rms_abs_K=rms.tolist(),
                     fp64_kernel_ms=round(q64_ms, 4), split_kernel_ms=round(q_ms, 4),
                     speedup=round(q64_ms / q_ms, 4), seconds=round(time.perf_counter() - t, 2),
                     channels="E, dE/ds0, dE/ds1, dE/ds2, d2E/ds0ds1, d2E/ds0ds2, d2E/ds1ds2, d3E/ds0ds1ds2 (stored, "
                              "scaled by the cell counts)")
    log(f"split vs FP64 Ewald grid, all points: max |d| K {np.array2string(mx, precision=2)}")
    state = dict(fw=fw, spec=spec, ff=ff, f=f, grids=grids, g64=g64, dims=dims, pairs=(n_all, n_lj),
                 split_error=split_err)
    # the test the reference runs after building (1e5 random points per LJ grid, with the Ewald grid scaled by the
    # probe's charge); outside the timed region
    check = {}
    t = time.perf_counter()
    for g, name in ((grids[0], "C_co2"), (grids[1], "O_co2")):
        pr = spec.index(name)
        c = ra.grid_check(ctx, ff, f, g, pr, grids[2], float(spec.charge[pr]), 300.0, n_points=100000, seed=7)
        check[name] = dict(lj_energy_relative_error=c["lennard_jones"]["energy_relative_error"],
                           lj_gradient_relative_error=c["lennard_jones"]["gradient_relative_error"].tolist(),
                           ewald_energy_relative_error=c["coulomb"]["energy_relative_error"],
                           ewald_points=c["n_coulomb_points"])
    check["seconds"] = round(time.perf_counter() - t, 4)
    state["grid_check"] = check
    log(f"grid check (1e5 points per LJ grid): {json.dumps(check)}")
    return value, ms_per_step, npts, roof, state


def bench_grids_quintic(ctx, g1, args):
    """Triquintic grids (27 values per point) of the headline workload: C_co2 LJ, O_co2 LJ and the unit-charge Ewald
    real-space grid, one launch each (3 x 46.6M points, 30.2 GB FP64).  Per-kernel HIP-event times over the timed
    steps; FP64 FLOP/s with the counted FLOP per in-cutoff pair of the triclinic kernels (FLOP_*_QUINTIC)."""
    import torch

    import raspa3_amd as ra
    from raspa3_amd._lib import RG_GRID_EWALD_REAL, RG_GRID_LENNARD_JONES, RG_SCHEME_TRIQUINTIC

    fw, spec, ff, f, dims = g1["fw"], g1["spec"], g1["ff"], g1["f"], g1["dims"]
    npts = dims[0] * dims[1] * dims[2]
    grids = [ra.Grid(ctx, fw.cell, fw.rect, dims, scheme=RG_SCHEME_TRIQUINTIC) for _ in range(3)]
    specs = [(grids[0], RG_GRID_LENNARD_JONES, spec.index("C_co2"), spec.rc_framework),
             (grids[1], RG_GRID_LENNARD_JONES, spec.index("O_co2"), spec.rc_framework),
             (grids[2], RG_GRID_EWALD_REAL, 0, spec.rc_coulomb)]
    for _ in range(max(1, args.warmup)):
        ra.build_grids(ctx, ff, f, specs)
    stream = torch.cuda.ExternalStream(ctx.stream, device=torch.device("cuda", 0))
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        for g in range(3):
            ev[k][g].record(stream)
            ra.build_grids(ctx, ff, f, specs[g:g + 1], sync=False)
        ev[k][3].record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1000.0 / args.steps
    kms = [sum(e[g].elapsed_time(e[g + 1]) for e in ev) / args.steps for g in range(3)]
    n_all, n_lj = g1["pairs"]
    flop_lj = npts * FLOP_LJ_QUINTIC * n_lj
    flop_q = npts * FLOP_EWALD_QUINTIC * n_all
    out = dict(points_per_s=3 * npts / (ms_per_step / 1000.0), ms_per_step=round(ms_per_step, 3),
               points_per_step=3 * npts, bytes_written_per_step=3 * npts * 27 * 8,
               kernels={"lj_C_co2": round(kms[0], 3), "lj_O_co2": round(kms[1], 3), "ewald": round(kms[2], 3)},
               flop_per_pair=dict(lj=FLOP_LJ_QUINTIC, ewald=FLOP_EWALD_QUINTIC))
    out["ewald_achieved_tflops"] = round(flop_q / (kms[2] / 1000.0) / 1e12, 3)
    out["ewald_frac"] = round(out["ewald_achieved_tflops"] / FP64_PEAK_TFLOPS, 4)
    out["lj_achieved_tflops"] = round(flop_lj / (kms[0] / 1000.0) / 1e12, 3)
    out["lj_frac"] = round(out["lj_achieved_tflops"] / FP64_PEAK_TFLOPS, 4)
    log(f"triquintic grids: ms/step {ms_per_step:.3f} (kernels {kms}); {out['points_per_s']:.4e} grid points/s")
    # the reference's grid test on the triquintic grids (1e5 points, energies, gradients and Hessians), outside the
    # timed region
    pr = spec.index("C_co2")
    t = time.perf_counter()
    c = ra.grid_check(ctx, ff, f, grids[0], pr, grids[2], float(spec.charge[pr]), 300.0, n_points=100000, seed=7)
    out["grid_check_C_co2"] = dict(
        lj_energy_relative_error=c["lennard_jones"]["energy_relative_error"],
        lj_gradient_relative_error=c["lennard_jones"]["gradient_relative_error"].tolist(),
        lj_hessian_relative_error=c["lennard_jones"]["hessian_relative_error"].tolist(),
        ewald_energy_relative_error=c["coulomb"]["energy_relative_error"],
        seconds=round(time.perf_counter() - t, 4))
    log(f"triquintic grid check: {json.dumps(out['grid_check_C_co2'])}")
    del grids
    return out


def widom_mu(lw, ew, temperature):
    """The reference's Widom statistics (rg_widom_statistics: <W> and the excess chemical potential with their 95 %
    intervals over 5 blocks) of per-insertion log Rosenbluth weights lw and Ewald Fourier differences ew (rigid
    molecules, W_IG = 1); energies in K."""
    import raspa3_amd as ra

    wts = np.exp(np.asarray(lw) - ra.beta(temperature) * np.asarray(ew))
    nb = len(wts) // 5
    blocks = wts[:5 * nb].reshape(5, nb)
    st = ra.widom_statistics(np.full(5, float(nb)), blocks.sum(1), np.zeros(5), temperature)
    return dict(rosenbluth_weight=st["rosenbluth_weight"], rosenbluth_weight_ci=st["rosenbluth_weight_ci"],
                mu_excess_K=st["mu_excess"] * ra.ENERGY_TO_KELVIN, mu_excess_ci_K=st["mu_excess_ci"] * ra.ENERGY_TO_KELVIN)


def bench_trial_energies(ctx, g1, args):
    """Widom insertions through the device kernel: trial site energies per second, and the reference's Widom
    statistics of the insertions (Rosenbluth weight, excess chemical potential)."""
    import torch

    import raspa3_amd as ra
    from raspa3_amd import workloads as w
    from raspa3_amd._lib import RG_GRID_LENNARD_JONES

    out = {}
    # configuration 0: CH4 in the 288-site framework, grid at 0.2 A, 10 first-bead trials per insertion
    fw0, spec0 = w.framework_config0(), w.forcefield_config0()
    ff0 = ra.ForceField(ctx, spec0)
    f0 = ra.Framework(ctx, fw0)
    d0 = ra.grid_dims(fw0.cell, 0.2)
    g0 = ra.Grid(ctx, fw0.cell, fw0.rect, d0)
    ra.build_grids(ctx, ff0, f0, [(g0, RG_GRID_LENNARD_JONES, spec0.index("CH4"), 12.0)])
    s0 = ra.System(ctx, ff0, fw0.cell, fw0.rect, framework=f0, grids={spec0.index("CH4"): g0}, kmax=(0, 0, 0),
                   capacity=64)
    comp0 = w.methane_component(spec0)
    p0 = w.mc_params(seed=7)
    n = args.widom_insertions
    # outputs resident on the device, timed with HIP events on the library's stream; the host-buffer call (results
    # copied back, as rg_widom returns them) is timed beside it
    stream = torch.cuda.ExternalStream(ctx.stream, device=torch.device("cuda", 0))
    outs = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    ptrs = [o.data_ptr() for o in outs]
    ra.widom_device(ctx, s0, comp0, p0, 0, 4096, *ptrs)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ctx.sync()
    ev0.record(stream)
    ra.widom_device(ctx, s0, comp0, p0, 4096, n, *ptrs)
    ev1.record(stream)
    ctx.sync()
    dt = ev0.elapsed_time(ev1) / 1000.0
    lw = outs[0].cpu().numpy()
    stats0 = widom_mu(lw, outs[2].cpu().numpy(), 300.0)
    t = time.perf_counter()
    ra.widom(ctx, s0, comp0, p0, 4096, n)
    dt_host = time.perf_counter() - t
    sites = n * p0["k_first"]
    out["ch4_config0"] = dict(insertions=n, seconds=round(dt, 5), sites_per_s=sites / dt,
                              sites_per_s_host_buffers=sites / dt_host, widom=stats0,
                              roofline=dict(bound="hbm", achieved=sites / dt * GRID_BYTES_PER_SITE / 1e9,
                                            peak=HBM_PEAK_GBS, unit="GB/s",
                                            frac=sites / dt * GRID_BYTES_PER_SITE / 1e9 / HBM_PEAK_GBS))
    log(f"trial energies CH4 config 0: {sites / dt:.4e} sites/s ({sites / dt_host:.4e} with the results copied "
        "to the host)")
    # configuration 1: rigid CO2 on the three configuration-1 grids, Ewald Fourier difference at kmax 8
    fw1, spec1 = g1["fw"], g1["spec"]
    s1 = ra.System(ctx, g1["ff"], fw1.cell, fw1.rect, framework=g1["f"],
                   grids={spec1.index("C_co2"): g1["grids"][0], spec1.index("O_co2"): g1["grids"][1]},
                   coulomb_grid=g1["grids"][2], kmax=(8, 8, 8), capacity=64)
    comp1 = w.co2_component(spec1)
    p1 = w.mc_params(seed=9)
    n1 = max(1024, n // 8)
    ra.widom_device(ctx, s1, comp1, p1, 0, 1024, *ptrs)
    ctx.sync()
    ev0.record(stream)
    ra.widom_device(ctx, s1, comp1, p1, 1024, n1, *ptrs)
    ev1.record(stream)
    ctx.sync()
    dt = ev0.elapsed_time(ev1) / 1000.0
    sites = n1 * (p1["k_first"] + 2 * p1["k_orient"])
    stats1 = widom_mu(outs[0][:n1].cpu().numpy(), outs[2][:n1].cpu().numpy(), 300.0)
    out["co2_config1"] = dict(insertions=n1, seconds=round(dt, 5), sites_per_s=sites / dt,
                              ewald_deltas_per_s=n1 / dt, n_kvectors=s1.info()["nk"], widom=stats1)
    log(f"trial energies CO2 config 1: {sites / dt:.4e} sites/s, {n1 / dt:.4e} Ewald deltas/s")
    out["unit"] = "trial site energies/s"
    return out


T95_4 = 2.776  # Student-t deviate at 95 % for the 4 degrees of freedom of 5 blocks (the reference's table)
KJ_PER_MOL_PER_K = 8.314462618e-3


def pooled_stats(states, ncomp, temperature, groups, cells):
    """The reference's estimators (rg_block_average, rg_enthalpy_of_adsorption: block averages, the per-component
    enthalpy of adsorption from the inverted fluctuation matrix, 95 % block intervals) on the production blocks of
    many independent systems.  Systems of one group sample the same state, so each of the 5 blocks pools their
    block sums; loading per unit cell and the isosteric heat Q_st = -Delta H are computed per group, then averaged
    over the groups.  sigma = the block interval / t_4 (a standard error); the group sigmas combine in quadrature."""
    import raspa3_amd as ra

    out = dict(loading_per_cell=[], loading_per_cell_sigma=[], qst_kJmol=[], qst_sigma=[], groups=len(groups))
    per_group = []
    for g in groups:
        w = sum(states[k]["block_count"] for k in g)
        u = sum(states[k]["block_energy"] for k in g)
        n = sum(states[k]["block_n"][:, :ncomp] for k in g)
        un = sum(states[k]["block_energy_n"][:, :ncomp] for k in g)
        nn = sum(states[k]["block_nn"][:, :ncomp, :ncomp] for k in g)
        c = cells[g[0]]
        lo = [ra.block_average(n[:, i] / c, w) for i in range(ncomp)]
        dh, dh_ci = ra.enthalpy_of_adsorption(w, u, n, un, nn, temperature)
        per_group.append(dict(loading=[m for m, _ in lo], loading_ci=[ci for _, ci in lo],
                              qst=(-dh * ra.ENERGY_TO_KELVIN * KJ_PER_MOL_PER_K).tolist(),
                              qst_ci=(dh_ci * ra.ENERGY_TO_KELVIN * KJ_PER_MOL_PER_K).tolist()))
    ng = len(per_group)
    for i in range(ncomp):
        out["loading_per_cell"].append(float(np.mean([p["loading"][i] for p in per_group])))
        out["loading_per_cell_sigma"].append(
            float(np.sqrt(sum((p["loading_ci"][i] / T95_4) ** 2 for p in per_group)) / ng))
        out["qst_kJmol"].append(float(np.mean([p["qst"][i] for p in per_group])))
        out["qst_sigma"].append(float(np.sqrt(sum((p["qst_ci"][i] / T95_4) ** 2 for p in per_group)) / ng))
    out["by_group"] = per_group
    return out


def within_2sigma(a, b, key, sigma_key):
    """Per component: |a - b| <= 2 sqrt(sigma_a^2 + sigma_b^2)."""
    return [bool(abs(x - y) <= 2.0 * np.hypot(sx, sy))
            for x, y, sx, sy in zip(a[key], b[key], a[sigma_key], b[sigma_key])]


def gcmc_stats(mc, nsys, base, temperature, ncomp=1, groups=None, cells=None):
    """Accepted insertions of the production cycles and the reference's loading and isosteric-heat estimators."""
    states = [mc.state(s) for s in range(nsys)]
    acc = sum(int(states[s]["accepted"][0] - base[s][0]) for s in range(nsys))
    tri = sum(int(states[s]["trials"][0] - base[s][1]) for s in range(nsys))
    r = pooled_stats(states, ncomp, temperature, groups or [list(range(nsys))], cells or [1.0] * nsys)
    r.update(accepted_insertions=acc, insertion_trials=tri)
    return r


def bench_gcmc(ctx, args, framework="itq29", tau=None, pressure=1e5):
    """Plain CBMC against the learned first-bead proposal, CO2 at 300 K and 1e5 Pa in charged ITQ-29 (3 x 3 x 3
    cells), MFI (2 x 2 x 2 cells) or a seeded silicate with two channels (about 5 % of its volume accessible).
    tau: training temperature factor of the proposal (default args.tau)."""
    import torch

    import raspa3_amd as ra
    from raspa3_amd import workloads as w
    from raspa3_amd._lib import RG_GRID_EWALD_REAL, RG_GRID_LENNARD_JONES

    T, P = 300.0, pressure
    n2 = args.gcmc_component == "n2"
    spec = w.forcefield_zeolite([("N_n2", 36.0, 3.31, -0.482), ("N_com", 0.0, 1.0, 0.964)] if n2 else ())
    tau = args.tau if tau is None else tau
    if framework == "mfi":
        fw, unit = w.framework_mfi(spec, (2, 2, 2))
        nuc, label = 288, "MFI 2x2x2 (2304 sites)"
    elif framework == "channels":
        fw = w.framework_channels(spec)
        unit, nuc = fw.cell, fw.n
        label = f"a seeded silicate with two 5 A channels (40 x 40 x 30 A, {fw.n} sites)"
    else:
        fw, unit = w.framework_itq29(spec, 3)
        nuc, label = 72, "charged ITQ-29 3x3x3 (1944 sites)"
    ff = ra.ForceField(ctx, spec)
    fu = ra.Framework(ctx, w.Framework(framework, unit, True, fw.pos[:nuc], fw.type[:nuc], fw.charge[:nuc]))
    dims = ra.grid_dims(unit, 0.2)
    types = [spec.index(n) for n in (("N_n2", "N_com") if n2 else ("C_co2", "O_co2"))]
    grids = {t: ra.Grid(ctx, unit, True, dims) for t in types}
    coul = ra.Grid(ctx, unit, True, dims)
    ra.build_grids(ctx, ff, fu, [(g, RG_GRID_LENNARD_JONES, t, 12.0) for t, g in grids.items()] +
                   [(coul, RG_GRID_EWALD_REAL, 0, 12.0)])
    fb = ra.Framework(ctx, fw)
    _, kmax = ra.ewald_parameters(1e-6, 12.0, np.diag(fw.cell))
    comp = w.n2_component(spec) if n2 else w.co2_reference(spec)
    nsys = args.gcmc_systems
    stream = torch.cuda.ExternalStream(ctx.stream, device=torch.device("cuda", 0))

    cells = {"mfi": 8.0, "channels": 1.0}.get(framework, 27.0)

    def run(label, seed, proposal=None, k_first=10):
        systems = [ra.System(ctx, ff, fw.cell, True, framework=fb, grids=grids, coulomb_grid=coul,
                             kmax=tuple(kmax), capacity=8 * 256) for _ in range(nsys)]
        params = w.mc_params(temperature=T, pressure_pa=P, seed=seed, max_molecules=256, k_first=k_first,
                             production_cycles=args.gcmc_production)
        mc = ra.Mc(ctx, systems, [comp], params)
        if proposal is not None:
            for s in range(nsys):
                mc.set_proposal(s, proposal[0], proposal[1])
        mc.run(args.gcmc_equilibration, sample=False, graph=True)
        base = [(mc.state(s)["accepted"][0], mc.state(s)["trials"][0]) for s in range(nsys)]
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ctx.sync()
        ev0.record(stream)
        t = time.perf_counter()
        mc.run(args.gcmc_production, sample=True, graph=True)
        dt = time.perf_counter() - t
        ev1.record(stream)
        ctx.sync()
        r = gcmc_stats(mc, nsys, base, T, 1, None, [cells] * nsys)
        tot = mc.total_energy()
        r["max_energy_drift"] = float(max(abs(tot[s] - mc.state(s)["energy"]) for s in range(nsys)))
        r.update(k_first=k_first, seconds=round(dt, 4), accepted_insertions_per_s=r["accepted_insertions"] / dt,
                 insertion_acceptance=r["accepted_insertions"] / max(1, r["insertion_trials"]),
                 kernel_ms_per_cycle=ev0.elapsed_time(ev1) / args.gcmc_production)
        del r["by_group"]
        log(f"gcmc {framework} {P:g} Pa {label} k_first {k_first}: {r['accepted_insertions_per_s']:.4e} accepted "
            f"insertions/s, acc {r['insertion_acceptance']:.4f}, N/cell = {r['loading_per_cell'][0]:.4f} +- "
            f"{r['loading_per_cell_sigma'][0]:.4f}, Qst = {r['qst_kJmol'][0]:.3f} +- {r['qst_sigma'][0]:.3f} kJ/mol")
        mc.close()
        for s in systems:
            s.close()
        return r

    plain = run("plain", 1001)
    sweep = sorted({int(x) for x in args.kfirst_sweep.split(",")}, reverse=True)
    # plain CBMC with fewer first-bead trials too, so that the learned proposal's gain is measured against the best
    # plain setting and not only against the reference's default of 10
    plain_k = {}
    for k in [k for k in sweep if k != 10]:
        r = run("plain", 1101 + k, k_first=k)
        r["loading_within_2sigma"] = within_2sigma(r, plain, "loading_per_cell", "loading_per_cell_sigma")
        r["qst_within_2sigma"] = within_2sigma(r, plain, "qst_kJmol", "qst_sigma")
        r["over_plain"] = r["accepted_insertions_per_s"] / plain["accepted_insertions_per_s"]
        plain_k[f"k_first_{k}"] = r
    # learned proposal: orientation-averaged Boltzmann factors of CO2 from the grids at 2^21 random points, tempered
    # (trained at tau * beta), tabulated on voxel centres with a uniform floor
    s0 = ra.System(ctx, ff, fw.cell, True, framework=fb, grids=grids, coulomb_grid=coul, kmax=tuple(kmax), capacity=64)
    rng = np.random.default_rng(5)
    frac = rng.random((1 << 21, 3))
    t = time.perf_counter()
    lw = ra.molecule_weights(ctx, s0, comp, T, frac, 16, 11)
    t_data = time.perf_counter() - t
    beta = ra.beta(T)
    u_eff = np.where(np.isfinite(lw), -lw / beta, 1e7)
    net = ra.Mlp(ctx, seed=3)
    t = time.perf_counter()
    losses = net.train(frac, u_eff, tau * beta, batch=4096, epochs=args.mlp_epochs, lr=1e-3, seed=9)
    t_train = time.perf_counter() - t
    n3 = [args.voxels] * 3
    prob = net.proposal(n3, args.floor)
    net.close()
    # the learned proposal with fewer first-bead trials per insertion (the acceptance rule is exact for any number)
    learned = {}
    for k in sweep:
        r = run("learned", 1002 + k, (n3, prob), k_first=k)
        r["loading_within_2sigma"] = within_2sigma(r, plain, "loading_per_cell", "loading_per_cell_sigma")
        r["qst_within_2sigma"] = within_2sigma(r, plain, "qst_kJmol", "qst_sigma")
        r["over_plain"] = r["accepted_insertions_per_s"] / plain["accepted_insertions_per_s"]
        learned[f"k_first_{k}"] = r
    def is_matched(r):
        return all(r["loading_within_2sigma"]) and all(r["qst_within_2sigma"])

    matched = [r for r in learned.values() if is_matched(r)]
    best = max(matched or learned.values(), key=lambda r: r["accepted_insertions_per_s"])
    best_plain = max([plain] + [r for r in plain_k.values() if is_matched(r)],
                     key=lambda r: r["accepted_insertions_per_s"])
    return dict(workload=f"{'N2 (TraPPE, grown from an N site)' if n2 else 'CO2 (reference rigid model)'} in {label}, "
                         f"{T:g} K, {P:g} Pa, "
                         f"{nsys} systems, {args.gcmc_equilibration} + {args.gcmc_production} cycles, grids at 0.2 A "
                         f"on the unit cell, Ewald kmax {list(kmax)}, 10 orientation trials; plain CBMC with 10 "
                         f"first-bead trials (and with each other count of {args.kfirst_sweep}), the learned "
                         f"proposal with {args.kfirst_sweep}",
                plain=plain, plain_k_first=plain_k, learned=learned, best_learned_k_first=best["k_first"],
                learned_over_plain=best["over_plain"],
                best_plain_k_first=best_plain["k_first"],
                learned_over_best_plain=best["accepted_insertions_per_s"] / best_plain["accepted_insertions_per_s"],
                learned_over_plain_same_k_first=learned.get("k_first_10", best)["over_plain"],
                loading_within_2sigma=best["loading_within_2sigma"], qst_within_2sigma=best["qst_within_2sigma"],
                statistics="the reference's block averages and per-component enthalpy of adsorption (rg_block_average, "
                           "rg_enthalpy_of_adsorption) on the 5 production blocks pooled over the systems; sigma = "
                           "95 % block interval / t_4",
                proposal=dict(training_points=len(frac), orientations=16, tau=tau, floor=args.floor,
                              voxels=n3, epochs=args.mlp_epochs, final_loss=float(losses[-1]),
                              data_seconds=round(t_data, 4), train_seconds=round(t_train, 4)),
                unit="accepted insertions/s")


def bench_gcmc_uneven(ctx, args):
    """Batched CBMC over uneven systems: n-hexane (flexible chains) + CO2 (Ewald kmax 8) at 300 K and 1000 Pa
    (mole fractions 0.2 / 0.8) in ITQ-29 supercells of six sizes (1944-7200 sites), one launch per cycle."""
    import raspa3_amd as ra
    from raspa3_amd import workloads as w
    from raspa3_amd._lib import RG_GRID_EWALD_REAL, RG_GRID_LENNARD_JONES

    spec = w.forcefield_hexane_zeolite()
    ff = ra.ForceField(ctx, spec)
    fw3, unit = w.framework_itq29(spec, 3)
    fu = ra.Framework(ctx, w.Framework("itq29", unit, True, fw3.pos[:72], fw3.type[:72], fw3.charge[:72]))
    dims = ra.grid_dims(unit, 0.2)
    types = [spec.index(n) for n in ("CH3", "CH2", "C_co2", "O_co2")]
    grids = {t: ra.Grid(ctx, unit, True, dims) for t in types}
    coul = ra.Grid(ctx, unit, True, dims)
    ra.build_grids(ctx, ff, fu, [(g, RG_GRID_LENNARD_JONES, t, 12.0) for t, g in grids.items()] +
                   [(coul, RG_GRID_EWALD_REAL, 0, 12.0)])
    # ideal-gas Rosenbluth weight of the hexane growth: Widom insertions into an empty box
    empty = ra.System(ctx, ff, np.diag([40.0] * 3), True, kmax=(0, 0, 0), capacity=64)
    lw0, _, _ = ra.widom(ctx, empty, w.hexane_component(spec), w.mc_params(seed=3), 0, 1 << 16)
    log_wig = float(np.log(np.exp(lw0).mean()))
    comps = [w.hexane_component(spec, log_wig, 0.2), w.co2_reference(spec, 0.8)]
    shapes = [(3, 3, 3), (3, 3, 4), (3, 4, 4), (4, 4, 4), (4, 4, 5), (4, 5, 5)]
    frameworks = {}
    for sh in shapes:
        fwk, _ = w.framework_itq29(spec, sh)
        frameworks[sh] = (fwk, ra.Framework(ctx, fwk))
    nsys = args.uneven_systems
    systems = []
    for k in range(nsys):
        fwk, fo = frameworks[shapes[k % len(shapes)]]
        systems.append(ra.System(ctx, ff, fwk.cell, True, framework=fo, grids=grids, coulomb_grid=coul,
                                 kmax=(8, 8, 8), capacity=8 * 256))
    params = w.mc_params(pressure_pa=1000.0, seed=77, max_molecules=256,
                         production_cycles=args.uneven_production)
    mc = ra.Mc(ctx, systems, comps, params)
    mc.run(args.uneven_equilibration, sample=False, graph=True)
    s0 = [mc.state(k) for k in range(nsys)]
    t = time.perf_counter()
    mc.run(args.uneven_production, sample=True, graph=True)
    dt = time.perf_counter() - t
    s1 = [mc.state(k) for k in range(nsys)]
    acc = sum(s1[k]["accepted"][0] - s0[k]["accepted"][0] for k in range(nsys))
    steps = sum(s1[k]["steps"] - s0[k]["steps"] for k in range(nsys))
    n = np.array([s["n_per_component"] for s in s1])
    tot = mc.total_energy()
    drift = max(abs(tot[k] - s1[k]["energy"]) for k in range(nsys))
    mc.close()
    log(f"gcmc uneven: {acc / dt:.4e} accepted insertions/s, {steps / dt:.4e} MC steps/s, hexane loadings "
        f"{n[:, 0].min()}-{n[:, 0].max()}")
    return dict(workload=f"{nsys} systems, n-hexane (flexible, 0.2) + CO2 (rigid, Ewald kmax 8, 0.8) at 300 K and "
                         f"1000 Pa in ITQ-29 supercells {shapes} (1944-7200 sites), {args.uneven_equilibration} + "
                         f"{args.uneven_production} cycles",
                accepted_insertions_per_s=acc / dt, mc_steps_per_s=steps / dt, seconds=round(dt, 4),
                loadings_hexane=[int(n[:, 0].min()), int(n[:, 0].max())],
                loadings_co2=[int(n[:, 1].min()), int(n[:, 1].max())], log_ideal_rosenbluth_hexane=log_wig,
                max_energy_drift=float(drift), unit="accepted insertions/s")


def bench_gcmc_mixture(ctx, args):
    """Combined workload: CO2/N2 80:20 GCMC in many systems of uneven size (charged ITQ-29 supercells of 2592-15552
    sites) in balanced launches (smaller systems run extra cycles while the largest runs one), plain CBMC against the
    learned first-bead proposal.  Each adsorbate has a proposal
    network of its own, trained on that molecule's orientation-averaged Boltzmann factors in the 11.87 A unit cell,
    whose table is tiled over each supercell (any positive density keeps the acceptance rule exact)."""
    import raspa3_amd as ra
    from raspa3_amd import workloads as w
    from raspa3_amd._lib import RG_GRID_EWALD_REAL, RG_GRID_LENNARD_JONES

    T, P, v = 300.0, args.mixture_pressure, args.mixture_voxels
    spec = w.forcefield_zeolite([("N_n2", 36.0, 3.31, -0.482), ("N_com", 0.0, 1.0, 0.964)])
    ff = ra.ForceField(ctx, spec)
    fw3, unit = w.framework_itq29(spec, 3)
    fu = ra.Framework(ctx, w.Framework("itq29", unit, True, fw3.pos[:72], fw3.type[:72], fw3.charge[:72]))
    dims = ra.grid_dims(unit, 0.2)
    grids = {spec.index(n): ra.Grid(ctx, unit, True, dims) for n in ("C_co2", "O_co2", "N_n2", "N_com")}
    coul = ra.Grid(ctx, unit, True, dims)
    ra.build_grids(ctx, ff, fu, [(g, RG_GRID_LENNARD_JONES, t, 12.0) for t, g in grids.items()] +
                   [(coul, RG_GRID_EWALD_REAL, 0, 12.0)])
    comps = [w.co2_reference(spec, 0.8), w.n2_component(spec, 0.2)]
    shapes = [(3, 3, 4), (3, 4, 4), (4, 4, 4), (4, 4, 5), (4, 5, 5), (5, 5, 5), (5, 5, 6), (5, 6, 6), (6, 6, 6)]
    frameworks = {}
    for sh in shapes:
        fwk, _ = w.framework_itq29(spec, sh)
        frameworks[sh] = (fwk, ra.Framework(ctx, fwk))
    nsys, mmax = args.mixture_systems, args.mixture_max_molecules
    shape_of = [shapes[k % len(shapes)] for k in range(nsys)]
    cells = np.array([np.prod(sh) for sh in shape_of], float)

    # one proposal per adsorbate: orientation-averaged Boltzmann factors of CO2 and of N2 at 2^20 points of the unit
    # cell, tempered, each learned by its own network and tabulated on v^3 voxels of the unit cell (x fastest),
    # repeated over each supercell
    s0 = ra.System(ctx, ff, unit, True, framework=fu, grids=grids, coulomb_grid=coul, kmax=(0, 0, 0), capacity=64)
    beta = ra.beta(T)
    tables, losses, t_train = [], [], 0.0
    for ci, comp in enumerate(comps):
        frac = np.random.default_rng(6 + ci).random((1 << 20, 3))
        lw = ra.molecule_weights(ctx, s0, comp, T, frac, 16, 13 + ci)
        net = ra.Mlp(ctx, seed=4 + ci)
        t = time.perf_counter()
        losses.append(net.train(frac, np.where(np.isfinite(lw), -lw / beta, 1e7), args.tau * beta, batch=4096,
                                epochs=args.mlp_epochs, lr=1e-3, seed=10 + ci))
        t_train += time.perf_counter() - t
        cell_table = net.proposal([v, v, v], args.floor).reshape(v, v, v)
        net.close()
        tables.append({sh: ([v * sh[0], v * sh[1], v * sh[2]], np.tile(cell_table, (sh[2], sh[1], sh[0])).reshape(-1))
                       for sh in shapes})
    s0.close()
    groups = [[k for k in range(nsys) if shape_of[k] == sh] for sh in shapes]
    groups = [g for g in groups if g]

    def run(label, seed, learned, k_first=10):
        systems = [ra.System(ctx, ff, frameworks[sh][0].cell, True, framework=frameworks[sh][1], grids=grids,
                             coulomb_grid=coul, kmax=(8, 8, 8), capacity=3 * mmax) for sh in shape_of]
        params = w.mc_params(temperature=T, pressure_pa=P, seed=seed, max_molecules=mmax, k_first=k_first,
                             production_cycles=args.mixture_production)
        mc = ra.Mc(ctx, systems, comps, params)
        if not args.no_mixture_balance:
            # lightly loaded (smaller) supercells run extra whole cycles while the most loaded one runs its first
            mc.set_balance(True, args.mixture_balance_overhead)
        if learned:
            for k, sh in enumerate(shape_of):
                for ci in range(len(comps)):
                    mc.set_proposal(k, *tables[ci][sh], component=ci)
        # launches in chunks of at most 1000, with a progress line after each
        chunk = 1000
        for done in range(0, args.mixture_equilibration, chunk):
            mc.run(min(chunk, args.mixture_equilibration - done), sample=False, graph=True)
            log(f"gcmc mixture {label}: {min(done + chunk, args.mixture_equilibration)} equilibration launches done")
        st0 = [mc.state(k) for k in range(nsys)]
        base = [(st0[k]["accepted"][0], st0[k]["trials"][0]) for k in range(nsys)]
        ctx.sync()
        t = time.perf_counter()
        for done in range(0, args.mixture_production, chunk):
            mc.run(min(chunk, args.mixture_production - done), sample=True, graph=True)
            if args.mixture_production > chunk:
                log(f"gcmc mixture {label}: {min(done + chunk, args.mixture_production)} production launches done")
        ctx.sync()
        dt = time.perf_counter() - t
        r = gcmc_stats(mc, nsys, base, T, 2, groups, list(cells))
        st1 = [mc.state(k) for k in range(nsys)]
        # CO2 per cell in each block of the production cycles (mean over systems): flat once equilibrated
        by_block = np.mean([st["block_n"][:, 0] / np.maximum(st["block_count"], 1) / cells[k]
                            for k, st in enumerate(st1)], axis=0)
        tot = mc.total_energy()
        drift = max(abs(tot[k] - st1[k]["energy"]) for k in range(nsys))
        n_final = np.array([st["n_per_component"] for st in st1])
        ncyc = np.array([st["cycles"] for st in st1])
        mc.close()
        for s in systems:
            s.close()
        r["qst_by_shape"] = [[round(float(x), 3) for x in g["qst"]] for g in r.pop("by_group")]
        r.update(k_first=k_first, seconds=round(dt, 4), accepted_insertions_per_s=r["accepted_insertions"] / dt,
                 insertion_acceptance=r["accepted_insertions"] / max(1, r["insertion_trials"]),
                 co2_per_cell_by_block=[round(float(x), 4) for x in by_block],
                 molecules_final=[int(n_final.sum(1).min()), int(n_final.sum(1).max())],
                 production_cycles_per_system=[int(ncyc.min()), round(float(ncyc.mean()), 1), int(ncyc.max())],
                 max_energy_drift=float(drift))
        log(f"gcmc mixture {label}: {r['accepted_insertions_per_s']:.4e} accepted insertions/s, acc "
            f"{r['insertion_acceptance']:.4f}, CO2/N2 per cell {r['loading_per_cell'][0]:.4f} +- "
            f"{r['loading_per_cell_sigma'][0]:.4f} / {r['loading_per_cell'][1]:.4f} +- "
            f"{r['loading_per_cell_sigma'][1]:.4f}, Qst {r['qst_kJmol'][0]:.3f} +- {r['qst_sigma'][0]:.3f} / "
            f"{r['qst_kJmol'][1]:.3f} +- {r['qst_sigma'][1]:.3f} kJ/mol")
        return r

    if args.mixture_method != "both":
        # one method alone (long runs split over separate invocations); no comparison in this line
        r = run(args.mixture_method, 2001 if args.mixture_method == "plain" else 2002,
                args.mixture_method == "learned", 10 if args.mixture_method == "plain" else args.mixture_kfirst)
        return {args.mixture_method: r, "workload": f"CO2/N2 80:20 at {T:g} K and {P:g} Pa, {nsys} systems in charged "
                f"ITQ-29 supercells {shapes}, {args.mixture_equilibration} + {args.mixture_production} launches, "
                f"balanced={not args.no_mixture_balance}", "unit": "accepted insertions/s"}
    plain = run("plain", 2001, False)
    learned = run("learned", 2002, True, args.mixture_kfirst)
    extra = {}
    if args.mixture_plain_same_k and args.mixture_kfirst != 10:
        # plain CBMC with the learned run's first-bead trial count (off by default: it adds a third full run)
        r = run("plain", 2003, False, args.mixture_kfirst)
        extra["plain_same_k_first"] = dict(
            r, loading_within_2sigma=within_2sigma(r, plain, "loading_per_cell", "loading_per_cell_sigma"),
            qst_within_2sigma=within_2sigma(r, plain, "qst_kJmol", "qst_sigma"))
        extra["learned_over_plain_same_k_first"] = (learned["accepted_insertions_per_s"] /
                                                   r["accepted_insertions_per_s"])
    balance = ("one cycle per system and launch" if args.no_mixture_balance else
               f"balanced launches (each system at least one cycle per launch, more while its cost, steps x "
               f"(molecules + {args.mixture_balance_overhead:g}), stays within the largest single-cycle cost)")
    return dict(**extra, workload=f"CO2/N2 80:20 at {T:g} K and {P:g} Pa, {nsys} systems in charged ITQ-29 "
                                  f"supercells {shapes} (2592-15552 sites), Ewald kmax 8, "
                                  f"{args.mixture_equilibration} + {args.mixture_production} launches, {balance}, "
                                  f"grids at 0.2 A on the unit cell; plain CBMC with 10 first-bead trials, the "
                                  f"learned proposal with {args.mixture_kfirst}",
                plain=plain, learned=learned,
                learned_over_plain=learned["accepted_insertions_per_s"] / plain["accepted_insertions_per_s"],
                loading_within_2sigma=within_2sigma(learned, plain, "loading_per_cell", "loading_per_cell_sigma"),
                qst_within_2sigma=within_2sigma(learned, plain, "qst_kJmol", "qst_sigma"),
                statistics="the reference's block averages and per-component enthalpy of adsorption on the 5 "
                           "production blocks pooled over the systems of each supercell shape, averaged over the "
                           "shapes; sigma = 95 % block interval / t_4, combined over the shapes",
                proposal=dict(networks="one per adsorbate (CO2, N2)", training_points=1 << 20,
                              voxels_per_unit_cell=v, tau=args.tau, floor=args.floor, epochs=args.mlp_epochs,
                              final_loss=[float(x[-1]) for x in losses], train_seconds=round(t_train, 4)),
                unit="accepted insertions/s")


def bench_mlp(ctx, g1, args):
    """Proposal-network training throughput on configuration-1 CO2 grid data (2^21 points, 300 K)."""
    import raspa3_amd as ra
    from raspa3_amd import workloads as w

    fw, spec = g1["fw"], g1["spec"]
    s = ra.System(ctx, g1["ff"], fw.cell, fw.rect, framework=g1["f"],
                  grids={spec.index("C_co2"): g1["grids"][0], spec.index("O_co2"): g1["grids"][1]},
                  coulomb_grid=g1["grids"][2], kmax=(0, 0, 0), capacity=64)
    comp = w.co2_component(spec)
    frac = np.random.default_rng(1).random((1 << 21, 3))
    beta = ra.beta(300.0)
    lw = ra.molecule_weights(ctx, s, comp, 300.0, frac, 16, 3)
    u = np.where(np.isfinite(lw), -lw / beta, 1e7)
    net = ra.Mlp(ctx, seed=1)
    net.train(frac[:1 << 16], u[:1 << 16], beta, batch=args.mlp_batch, epochs=1, lr=1e-3, seed=1)
    t = time.perf_counter()
    losses = net.train(frac, u, beta, batch=args.mlp_batch, epochs=args.mlp_epochs, lr=1e-3, seed=2)
    dt = time.perf_counter() - t
    samples = (len(frac) // args.mlp_batch) * args.mlp_batch * args.mlp_epochs
    tflops = samples / dt * MLP_FLOP_PER_SAMPLE / 1e12
    net.close()
    log(f"mlp: {samples / dt:.4e} samples/s, {tflops:.2f} TFLOP/s BF16 MFMA")
    return dict(samples_per_s=samples / dt, batch=args.mlp_batch, epochs=args.mlp_epochs, seconds=round(dt, 4),
                finite_fraction=float(np.isfinite(lw).mean()), losses=[float(x) for x in losses],
                roofline=dict(bound="mfma", achieved=tflops, peak=BF16_PEAK_TFLOPS, unit="TFLOP/s",
                              frac=tflops / BF16_PEAK_TFLOPS))


def parallel_timed(nthreads, make, step, target_s):
    """Runs make(i) then step(obj) repeatedly on nthreads host threads (the oracle's ctypes calls release the GIL),
    from a common start until each thread has run target_s seconds; returns (objects, per-thread counts, wall s)."""
    import threading

    objs = [None] * nthreads
    counts = [0] * nthreads
    errors = []
    ready = threading.Barrier(nthreads + 1)
    done = threading.Barrier(nthreads + 1)

    def work(i):
        try:
            objs[i] = make(i)
        except Exception as e:  # reported after the barrier
            errors.append(e)
        ready.wait()
        t0 = time.perf_counter()
        while not errors and time.perf_counter() - t0 < target_s:
            counts[i] += step(objs[i])
        done.wait()

    threads = [threading.Thread(target=work, args=(i,)) for i in range(nthreads)]
    for t in threads:
        t.start()
    ready.wait()
    t0 = time.perf_counter()
    done.wait()
    wall = time.perf_counter() - t0
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return objs, counts, wall


def cpu_baseline_gcmc(target_s=10.0):
    """The oracle's plain GCMC of the same workload as the GPU run, one independent system per host thread (each
    with its own seed), 100 untimed cycles first; accepted insertions of all threads over the wall time."""
    import oracle
    from raspa3_amd import workloads as w

    spec = w.forcefield_zeolite()
    fw, unit = w.framework_itq29(spec, 3)
    dims = w.grid_dims(unit, 0.2)[0]
    pairs = oracle.pairs_from_spec(spec)
    cores = host_cores()
    g = {t: oracle.grid_build(unit, True, dims, 0, fw.pos[:72], fw.type[:72], fw.charge[:72], pairs, t, 12.0, 12.0,
                              spec.alpha, nthreads=cores)
         for t in (spec.index("C_co2"), spec.index("O_co2"))}
    coul = oracle.grid_build(unit, True, dims, 3, fw.pos[:72], fw.type[:72], fw.charge[:72], pairs, 0, 12.0, 12.0,
                             spec.alpha, nthreads=cores)
    _, kmax = oracle.ewald_auto(1e-6, 12.0, np.diag(fw.cell))

    def make(i):
        mc = oracle.MC(spec, fw, [w.co2_reference(spec)], w.mc_params(seed=3 + i, max_molecules=256),
                       kmax=tuple(int(k) for k in kmax), grids=(g, coul, dims, unit))
        mc.run(100)  # loads the system part of the way
        mc.base = mc.state()["accepted"][0]
        return mc

    def step(mc):
        mc.run(5)
        return 5

    mcs, cycles, wall = parallel_timed(cores, make, step, target_s)
    acc = sum(int(mc.state()["accepted"][0] - mc.base) for mc in mcs)
    nmol = [mc.state()["n_molecules"] for mc in mcs]
    return dict(value=acc / wall, unit="accepted insertions/s", cores=cores, kind="port",
                sample=f"oracle plain GCMC, {cores} independent systems on {cores} threads, {sum(cycles)} cycles after "
                       f"100 untimed ones ({acc} accepted insertions in {wall:.1f} s, N = {min(nmol)}-{max(nmol)})")


def cpu_baseline_widom(target_s=10.0):
    """The oracle's Widom insertions of CH4 in configuration 0, the trial-energy workload of the GPU run (grid at
    0.2 A, 10 first-bead trials per insertion), on every host thread (disjoint insertion ranges of one seed)."""
    import oracle
    from raspa3_amd import workloads as w

    fw, spec = w.framework_config0(), w.forcefield_config0()
    pairs = oracle.pairs_from_spec(spec)
    t = spec.index("CH4")
    dims = w.grid_dims(fw.cell, 0.2)[0]
    cores = host_cores()
    g = oracle.grid_build(fw.cell, fw.rect, dims, 0, fw.pos, fw.type, fw.charge, pairs, t, 12.0, 12.0, spec.alpha,
                          nthreads=cores)
    params = w.mc_params(seed=7)
    chunk = 20000

    def make(i):
        mc = oracle.MC(spec, fw, [w.methane_component(spec)], params, kmax=(0, 0, 0), grids=({t: g}, None, dims))
        mc.next_first = i * (1 << 30)
        return mc

    def step(mc):
        mc.widom(0, mc.next_first, chunk)
        mc.next_first += chunk
        return chunk

    _, counts, wall = parallel_timed(cores, make, step, target_s)
    n = sum(counts)
    sites = n * params["k_first"]
    return dict(value=sites / wall, unit="trial site energies/s", cores=cores, kind="port",
                sample=f"oracle Widom, CH4 in configuration 0, {n} insertions x {params['k_first']} trials on "
                       f"{cores} threads in {wall:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--split-radius", type=float, default=7.0,
                    help="Ewald real-space pair terms beyond this distance (A) in FP32 in the headline grid step")
    ap.add_argument("--grids-only", action="store_true", help="measure the headline grid build only")
    ap.add_argument("--gcmc-only", default="", metavar="NAMES",
                    help="measure the headline grid build and only the GCMC comparisons named (itq29, mfi, channels; "
                         "name@pressure_pa sets the pressure)")
    ap.add_argument("--mlp-only", action="store_true",
                    help="measure the headline grid build and the proposal-network training only")
    ap.add_argument("--quintic-only", action="store_true",
                    help="measure the headline grid build and the triquintic grids of the same workload only")
    ap.add_argument("--widom-insertions", type=int, default=1 << 20)
    ap.add_argument("--gcmc-systems", type=int, default=256)
    ap.add_argument("--gcmc-component", choices=("co2", "n2"), default="co2",
                    help="adsorbate of the --gcmc-only comparisons (the default legs use CO2)")
    ap.add_argument("--gcmc-equilibration", type=int, default=1000)
    ap.add_argument("--gcmc-production", type=int, default=400)
    ap.add_argument("--tau", type=float, default=0.5)
    ap.add_argument("--kfirst-sweep", default="10,4,2,1",
                    help="first-bead trial counts of the learned-proposal GCMC runs (plain CBMC uses 10)")
    ap.add_argument("--floor", type=float, default=0.1)
    ap.add_argument("--voxels", type=int, default=64)
    ap.add_argument("--mlp-epochs", type=int, default=4)
    ap.add_argument("--mlp-batch", type=int, default=4096)
    ap.add_argument("--uneven-systems", type=int, default=64)
    # a cycle has max(20, N) steps per component (src/raspakit/drivers/monte_carlo.cpp:301), so the two-component legs
    # run twice the steps of a one-component cycle
    ap.add_argument("--uneven-equilibration", type=int, default=150)
    ap.add_argument("--uneven-production", type=int, default=50)
    ap.add_argument("--mixture-only", action="store_true",
                    help="measure the headline grid build and the CO2/N2 mixture GCMC over uneven systems only")
    ap.add_argument("--mixture-systems", type=int, default=256)
    ap.add_argument("--mixture-equilibration", type=int, default=1000)
    ap.add_argument("--mixture-production", type=int, default=400)
    ap.add_argument("--mixture-pressure", type=float, default=1e5)
    ap.add_argument("--mixture-voxels", type=int, default=16, help="proposal voxels per unit-cell axis")
    ap.add_argument("--mixture-kfirst", type=int, default=2, help="first-bead trials of the learned mixture run")
    ap.add_argument("--mixture-method", choices=("both", "plain", "learned"), default="both",
                    help="run both methods of the mixture and compare them (default), or one of them alone")
    ap.add_argument("--no-mixture-balance", action="store_true",
                    help="one cycle per system and launch in the mixture (no balanced launches)")
    ap.add_argument("--mixture-balance-overhead", type=float, default=256.0,
                    help="balanced launches: a cycle costs its steps x (molecules + this)")
    ap.add_argument("--mixture-plain-same-k", action="store_true",
                    help="also run plain CBMC in the mixture with the learned run's first-bead trial count")
    ap.add_argument("--mixture-max-molecules", type=int, default=512)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 3 times (the reference kernel that scales the HBM "
                         "byte counters of the traffic figure) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (a multi-GPU build would shard the "
              "grids into z-slabs, one per GPU, with no communication; only the single-GPU case is built and run)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration > 0:
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_bytes=n, reads_per_launch=n, writes_per_launch=n)), flush=True)
        return

    import raspa3_amd as ra

    torch.cuda.init()  # the HIP events of the timed regions are torch events on the library's stream
    torch.cuda.set_device(0)
    ctx = ra.Context(0)
    value, ms_per_step, npts, roof, g1 = bench_grids(ctx, args)
    extra = {"grid_check": g1["grid_check"], "grid_split_error": g1["split_error"]}
    if args.mlp_only:
        extra["mlp"] = bench_mlp(ctx, g1, args)
    if args.quintic_only:
        extra["grids_triquintic"] = bench_grids_quintic(ctx, g1, args)
    if args.gcmc_only:
        for item in args.gcmc_only.split(","):
            name, _, pa = item.partition("@")
            extra["gcmc_" + item] = bench_gcmc(ctx, args, name, tau=1.0 if name != "itq29" else None,
                                               pressure=float(pa) if pa else 1e5)
    elif args.mixture_only:
        extra["gcmc_mixture"] = bench_gcmc_mixture(ctx, args)
    elif not args.grids_only and not args.quintic_only and not args.mlp_only:
        extra["grids_triquintic"] = bench_grids_quintic(ctx, g1, args)
        extra["trial_energies"] = bench_trial_energies(ctx, g1, args)
        extra["mlp"] = bench_mlp(ctx, g1, args)
        extra["gcmc"] = bench_gcmc(ctx, args)
        extra["gcmc_mfi"] = bench_gcmc(ctx, args, "mfi", tau=1.0)
        # the uneven-batch leg (hexane + CO2 over 64 ITQ-29 supercells, bench_gcmc_uneven) is not run: its last run
        # took a GPU memory access fault in the launch-shape tuning pass of rg_mc_run and the cause is not found
        extra["gcmc_mixture"] = bench_gcmc_mixture(ctx, args)
    base = None
    if not args.no_cpu_baseline:
        base = cpu_baseline(g1["fw"], g1["spec"], g1["dims"], args.cpu_seconds,
                            gpu_grids={"C_co2": g1["grids"][0], "O_co2": g1["grids"][1],
                                       f"ewald_split_{args.split_radius:g}A": g1["grids"][2],
                                       "ewald_fp64": g1["g64"]})
        log(f"cpu baseline {base['value']:.4e} points/s ({base['sample']})")
        if "gcmc" in extra:
            extra["gcmc"]["cpu_baseline"] = cpu_baseline_gcmc()
            log(f"cpu baseline gcmc {extra['gcmc']['cpu_baseline']['value']:.4e} accepted insertions/s")
        if "trial_energies" in extra:
            extra["trial_energies"]["cpu_baseline"] = cpu_baseline_widom(5.0)
            log(f"cpu baseline widom {extra['trial_energies']['cpu_baseline']['value']:.4e} trial site energies/s")
    fw, spec = g1["fw"], g1["spec"]
    out = dict(metric="framework-energy evals/s; GCMC accepted insertions/s at matched loading",
               value=value, unit="grid points/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=ms_per_step, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64 (Ewald pair terms beyond the split radius in f32, accumulated in f64)",
               data="synthetic",
               config=dict(workload="CO2 grids with charges: triclinic 40x40x30 A gamma=100, 4095 sites, "
                                    "C_co2 LJ + O_co2 LJ + Ewald-real grids at h=0.1 A (394x394x300 each), "
                                    "one LJ launch for both LJ grids (FP64) + one Ewald launch per step (FP32 pair "
                                    f"terms beyond {args.split_radius:g} A, FP64 within it and for every sum)",
                           grid_points_per_step=3 * npts, framework_sites=fw.n, cutoff=spec.rc_framework,
                           alpha=spec.alpha, split_radius_A=args.split_radius),
               roofline=roof, cpu_baseline=base, **extra)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
