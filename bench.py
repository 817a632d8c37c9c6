"""Benchmark: interpolation-grid build rate of the framework-energy path on one MI355X.

Workload (the configuration the metric is quoted on): CO2 grids with charges — triclinic cell a = b = 40, c = 30 A,
gamma = 100 deg, 4095 framework sites (1365 Si +2.05 e, 2730 O -1.025 e), LJ on O (44 K, 3.3 A) with
Lorentz-Berthelot mixing to C_co2 (27 K, 2.80 A) and O_co2 (79 K, 3.05 A), Ewald alpha = 0.265 A^-1, 12 A cutoff,
h = 0.1 A -> 394 x 394 x 300 points per grid.  One step builds the three grids (C_co2 LJ, O_co2 LJ, unit-charge
Ewald real space: 139.7M grid points, 8.94 GB FP64) in one fused pass with inputs resident in HBM.

value = grid points per second (framework-energy evaluations with their tricubic derivatives) over the timed steps.
roofline = FP64 vector FLOP/s of the build kernel against the MI355X FP64 vector peak (78.6 TFLOP/s), with the
algorithmic FLOP count of the frozen convention: 79 FLOP per in-cutoff LJ pair and 86 FLOP (+3 transcendentals,
not counted) per in-cutoff Ewald real-space pair, triclinic cell.
cpu_baseline = the FP64 oracle (the reference's loop order and arithmetic restated in C++) on the host cores, on a
bounded sample of z-planes of the same three grids.
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
FLOP_LJ_TRIC = 79
FLOP_EWALD_TRIC = 86


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def cpu_baseline(fw, spec, dims, target_s=12.0):
    """Times the FP64 oracle on the host cores over whole z-planes of the three grids."""
    import oracle

    pairs = oracle.pairs_from_spec(spec)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 64))
    grids = [(0, spec.index("C_co2")), (0, spec.index("O_co2")), (3, 0)]

    def run(planes):
        t = time.perf_counter()
        for gt, probe in grids:
            oracle.grid_build(fw.cell, fw.rect, dims, gt, fw.pos, fw.type, fw.charge, pairs, probe,
                              spec.rc_framework, spec.rc_framework, spec.alpha, k_range=(0, planes), nthreads=cores)
        return time.perf_counter() - t

    t1 = run(1)
    planes = int(max(1, min(dims[2], round(target_s / max(t1, 1e-3)))))
    t = run(planes) if planes > 1 else t1
    pts = 3 * planes * dims[0] * dims[1]
    return dict(value=pts / t, unit="grid points/s", cores=cores, kind="port",
                sample=f"FP64 oracle (reference loop order, all framework atoms per point), {planes} z-plane(s) x "
                       f"{dims[0]}x{dims[1]} points x 3 grids = {pts} points in {t:.2f} s on {cores} threads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the build path shards over "
              "z-slabs across GPUs by design, see DESIGN.md, but only the single-GPU case is run)", file=sys.stderr)
        sys.exit(2)

    import torch

    import raspa3_amd as ra
    from raspa3_amd import workloads as w
    from raspa3_amd._lib import RG_GRID_EWALD_REAL, RG_GRID_LENNARD_JONES

    fw, spec = w.framework_config1(), w.forcefield_config1()
    dims = ra.grid_dims(fw.cell, 0.1)
    npts = dims[0] * dims[1] * dims[2]
    ctx = ra.Context(0)
    ff = ra.ForceField(ctx, spec)
    f = ra.Framework(ctx, fw)
    grids = [ra.Grid(ctx, fw.cell, fw.rect, dims) for _ in range(3)]
    specs = [(grids[0], RG_GRID_LENNARD_JONES, spec.index("C_co2"), spec.rc_framework),
             (grids[1], RG_GRID_LENNARD_JONES, spec.index("O_co2"), spec.rc_framework),
             (grids[2], RG_GRID_EWALD_REAL, 0, spec.rc_coulomb)]
    log(f"workload: 3 grids {dims} ({3 * npts} points), {fw.n} framework sites")
    for _ in range(max(1, args.warmup)):
        ra.build_grids(ctx, ff, f, specs)
    stream = torch.cuda.ExternalStream(ctx.stream)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ctx.sync()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        ra.build_grids(ctx, ff, f, specs, sync=False)
    ev1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    value = 3 * npts / (ms_per_step / 1000.0)
    log(f"ms/step {ms_per_step:.3f} (events {kernel_ms:.3f}); {value:.4e} grid points/s")

    n_all, n_lj = incutoff_counts(fw, [spec.index("O")], spec.rc_framework)
    flop = npts * (2 * FLOP_LJ_TRIC * n_lj + FLOP_EWALD_TRIC * n_all)
    achieved = flop / (kernel_ms / 1000.0) / 1e12
    roof = dict(bound="valu", achieved=round(achieved, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                frac=round(achieved / FP64_PEAK_TFLOPS, 4), traffic=None,
                kernel="k_grid_build<2,true,true> (fused C_co2 LJ + O_co2 LJ + Ewald real, triclinic)",
                flop_per_launch=float(flop), incutoff_pairs_per_point=dict(lj_active=round(n_lj, 2),
                                                                           charged=round(n_all, 2)),
                kernel_ms=round(kernel_ms, 4))
    base = None
    if not args.no_cpu_baseline:
        base = cpu_baseline(fw, spec, dims, args.cpu_seconds)
        log(f"cpu baseline {base['value']:.4e} points/s ({base['sample']})")
    out = dict(metric="framework-energy evals/s; GCMC accepted insertions/s at matched loading",
               value=value, unit="grid points/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=ms_per_step, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
               data="synthetic",
               config=dict(workload="CO2 grids with charges: triclinic 40x40x30 A gamma=100, 4095 sites, "
                                    "C_co2 LJ + O_co2 LJ + Ewald-real grids at h=0.1 A (394x394x300 each), "
                                    "fused build", grid_points_per_step=3 * npts, framework_sites=fw.n,
                           cutoff=spec.rc_framework, alpha=spec.alpha),
               roofline=roof, cpu_baseline=base)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
