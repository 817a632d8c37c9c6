"""Pins the oracle's triquintic grid values and triquintic interpolation (no reference test pins them by value; the
reference's triquintic tests, tests/raspakit-tests/triquintic_derivative_inter_lennard_jones.cpp, check its
derivatives by finite differences).

* The 27 fractional derivatives at a point equal mpmath's mixed partial derivatives (30 digits) of the explicit sum
  over periodic images within the cutoff, for LJ and the unit-charge Ewald real-space grid, in a triclinic and a
  non-cubic rectangular cell.  A wrong radial factor, sign, metric term or value order fails here.
* Stored grid values are those derivatives divided by the cell counts per derivative axis, clamped on overlap.
* The triquintic interpolant reproduces every polynomial of degree <= 5 per axis exactly, with its gradient and
  Hessian, and at the nodes of a real grid it returns mpmath's Cartesian energy, gradient and Hessian.
* On a real LJ grid its error at random points falls as h^6.
"""
import itertools

import mpmath as mp
import numpy as np
import pytest

import oracle

U = oracle.units()
ORDERS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1),
          (0, 0, 2), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2), (2, 2, 0),
          (2, 0, 2), (0, 2, 2), (2, 1, 1), (1, 2, 1), (1, 1, 2), (2, 2, 1), (2, 1, 2), (1, 2, 2), (2, 2, 2)]


def system(triclinic):
    if triclinic:
        cell = oracle.box(7.3, 8.1, 6.7, 80.0, 95.0, 105.0)["cell"]
        rect = False
    else:
        cell, rect = np.diag([7.3, 8.1, 6.7]), True
    rng = np.random.default_rng(3)
    fpos = rng.random((3, 3)) @ cell.T
    ftype = np.array([0, 1, 0], np.int32)
    fq = np.array([0.7, -1.1, 0.4])
    pairs = oracle.ff_pairs([50.0, 80.0, 120.0], [3.0, 3.4, 3.7], [1, 1, 0], True, 8.0, 8.0)
    return cell, rect, fpos, ftype, fq, pairs


def image_terms(p, cell, fpos, rc):
    """(atom, dr) for every periodic image within rc of p (brute force over lattice translations)."""
    out = []
    for m in range(len(fpos)):
        for n in itertools.product(range(-4, 5), repeat=3):
            dr = p - (fpos[m] + cell @ np.array(n))
            if dr @ dr < rc * rc:
                out.append((m, dr))
    return out


def mp_energy(terms, cell, ftype, fq, pairs, probe, alpha, grid_type, shift_frac=None, shift_cart=None):
    """Energy of the fixed image set displaced by H ds (fractional) or dr (Cartesian), in mpmath."""
    H = [[mp.mpf(cell[r, c]) for c in range(3)] for r in range(3)]

    def E(*d):
        tot = mp.mpf(0)
        for m, dr in terms:
            if shift_frac:
                v = [mp.mpf(dr[r]) + H[r][0] * d[0] + H[r][1] * d[1] + H[r][2] * d[2] for r in range(3)]
            else:
                v = [mp.mpf(dr[r]) + d[r] for r in range(3)]
            rr = v[0] ** 2 + v[1] ** 2 + v[2] ** 2
            if grid_type == 0:
                k = probe * pairs["n"] + ftype[m]
                t3 = (mp.mpf(pairs["sigma"][k]) ** 2 / rr) ** 3
                tot += 4 * mp.mpf(pairs["eps"][k]) * (t3 * t3 - t3) - mp.mpf(pairs["shift"][k])
            else:
                r = mp.sqrt(rr)
                tot += mp.mpf(U["coulomb"]) * mp.mpf(fq[m]) * mp.erfc(mp.mpf(alpha) * r) / r
        return tot

    return E


@pytest.mark.parametrize("triclinic", [False, True])
@pytest.mark.parametrize("grid_type", [0, 3])
def test_fractional_derivatives_match_mpmath(triclinic, grid_type):
    cell, rect, fpos, ftype, fq, pairs = system(triclinic)
    probe, rc, alpha = 2, 8.0, 0.3
    p = np.array([0.31, 0.47, 0.62]) @ cell.T
    got = oracle.quintic_at_positions(cell, rect, grid_type, fpos, ftype, fq, pairs, probe, rc, rc, alpha, p[None])[0]
    mp.mp.dps = 30
    E = mp_energy(image_terms(p, cell, fpos, rc), cell, ftype, fq, pairs, probe, alpha, grid_type, shift_frac=True)
    for o, d in enumerate(ORDERS):
        ref = float(mp.diff(E, (0, 0, 0), d))
        assert abs(got[o] - ref) <= 1e-11 * max(1.0, abs(ref)), (o, d, got[o], ref)


def test_stored_values_are_scaled_and_clamped():
    cell, rect, fpos, ftype, fq, pairs = system(True)
    dims = (9, 10, 8)
    nc = np.array(dims) - 1.0
    ijk = np.array([[2, 3, 4], [0, 0, 0], [8, 9, 7]])
    for gt in (0, 3):
        stored = oracle.grid_points_q(cell, rect, dims, gt, fpos, ftype, fq, pairs, 2, 8.0, 8.0, 0.3, ijk)
        pos = (ijk / nc) @ cell.T
        raw = oracle.quintic_at_positions(cell, rect, gt, fpos, ftype, fq, pairs, 2, 8.0, 8.0, 0.3, pos)
        scale = np.array([np.prod(nc ** np.array(d)) for d in ORDERS])
        np.testing.assert_allclose(stored, raw / scale, rtol=1e-12, atol=0)
    # a node on top of a framework atom: energy 2 * overlap, first derivatives clamped, the rest zero
    cell, rect = np.diag([8.0, 8.0, 8.0]), True
    fpos = np.array([[2.0, 2.0, 2.0]])
    overlap = 1e4
    stored = oracle.grid_points_q(cell, rect, (9, 9, 9), 0, fpos, np.array([1], np.int32), np.zeros(1), pairs, 2, 8.0,
                                  8.0, 0.3, np.array([[2, 2, 2], [2, 3, 2]]), overlap=overlap)
    assert stored[0, 0] == 2 * overlap and np.all(np.abs(stored[0, 1:4]) <= overlap) and np.all(stored[0, 4:] == 0)
    assert stored[1, 0] > overlap and np.all(stored[1, 4:] == 0) and np.abs(stored[1, 1:4]).max() == overlap


def random_polynomial_grid(cell, dims, seed):
    """Grid of a random polynomial of degree 5 per fractional axis (values scaled as stored) and the polynomial."""
    coef = np.random.default_rng(seed).normal(size=(6, 6, 6))
    nc = np.array(dims) - 1.0

    def P(s, d):
        tot = 0.0
        for i, j, k in itertools.product(range(6), repeat=3):
            if i < d[0] or j < d[1] or k < d[2]:
                continue
            f = np.prod([np.prod(range(e - dd + 1, e + 1)) for e, dd in zip((i, j, k), d)])
            tot += coef[i, j, k] * f * s[0] ** (i - d[0]) * s[1] ** (j - d[1]) * s[2] ** (k - d[2])
        return tot

    grid = np.zeros((dims[2], dims[1], dims[0], 27))
    for k, j, i in itertools.product(range(dims[2]), range(dims[1]), range(dims[0])):
        s = np.array([i, j, k]) / nc
        for o, d in enumerate(ORDERS):
            grid[k, j, i, o] = P(s, d) / np.prod(nc ** np.array(d))
    return grid, P


def test_interpolation_reproduces_quintic_polynomials():
    cell = oracle.box(7.3, 8.1, 6.7, 80.0, 95.0, 105.0)["cell"]
    dims = (5, 4, 6)
    grid, P = random_polynomial_grid(cell, dims, 1)
    S = np.random.default_rng(2).uniform(0.01, 0.99, (25, 3))
    e, g, h = oracle.interpolate_q(cell, False, dims, grid, S @ cell.T)
    inv = np.linalg.inv(cell)
    for q, s in enumerate(S):
        gs = np.array([P(s, (1, 0, 0)), P(s, (0, 1, 0)), P(s, (0, 0, 1))])
        hs = np.array([[P(s, tuple(int(x == a) + int(x == b) for x in range(3))) for b in range(3)] for a in range(3)])
        assert abs(e[q] - P(s, (0, 0, 0))) <= 1e-12 * max(1, abs(P(s, (0, 0, 0))))
        np.testing.assert_allclose(g[q], inv.T @ gs, rtol=1e-11, atol=1e-11 * np.abs(gs).max())
        np.testing.assert_allclose(h[q], inv.T @ hs @ inv, rtol=1e-10, atol=1e-10 * np.abs(hs).max())


@pytest.mark.parametrize("grid_type", [0, 3])
def test_interpolant_at_nodes_is_the_analytic_energy_gradient_and_hessian(grid_type):
    cell, rect, fpos, ftype, fq, pairs = system(True)
    dims = (9, 10, 8)
    nc = np.array(dims) - 1.0
    node = np.array([3, 4, 5])
    ijk = np.array([node + np.array(c) for c in itertools.product((0, 1), repeat=3)])
    stored = oracle.grid_points_q(cell, rect, dims, grid_type, fpos, ftype, fq, pairs, 2, 8.0, 8.0, 0.3, ijk)
    grid = np.zeros((dims[2], dims[1], dims[0], 27))
    for (i, j, k), v in zip(ijk, stored):
        grid[k, j, i] = v
    p = ((node + 0.25) / nc) @ cell.T  # inside the cell whose corners are filled
    pn = (node / nc) @ cell.T
    mp.mp.dps = 30
    for pt in (pn + 1e-12, p):
        e, g, h = oracle.interpolate_q(cell, rect, dims, grid, pt[None])
        if pt is p:
            break
        E = mp_energy(image_terms(pt, cell, fpos, 8.0), cell, ftype, fq, pairs, 2, 0.3, grid_type, shift_cart=True)
        ref_g = np.array([float(mp.diff(E, (0, 0, 0), tuple(int(a == b) for b in range(3)))) for a in range(3)])
        ref_h = np.array([[float(mp.diff(E, (0, 0, 0), tuple(int(a == c) + int(b == c) for c in range(3))))
                           for b in range(3)] for a in range(3)])
        scale = max(1.0, abs(float(E(0, 0, 0))))
        assert abs(e[0] - float(E(0, 0, 0))) <= 1e-9 * scale
        np.testing.assert_allclose(g[0], ref_g, rtol=1e-8, atol=1e-9 * np.abs(ref_g).max())
        np.testing.assert_allclose(h[0], ref_h, rtol=1e-7, atol=1e-8 * np.abs(ref_h).max())
    # the Hessian of the interpolant is symmetric
    np.testing.assert_allclose(h[0], h[0].T, rtol=1e-13, atol=1e-13 * np.abs(h[0]).max())


def test_interpolation_error_falls_as_h6():
    """Relative RMS error of the interpolated LJ energy at random points more than 2.5 A from every framework atom,
    at two grid spacings: halving h divides it by about 64 (at least 32 is required here).  The cutoff of 12 A
    keeps the kink of the truncated potential (its derivative jumps at the cutoff) below the interpolation error;
    with a 4.4 A cutoff the kink dominates and the error falls only about 5 times."""
    cell, rect = np.diag([9.0, 9.0, 9.0]), True
    fpos = np.array([[1.0, 2.0, 3.0], [5.5, 6.0, 4.0]])
    ftype = np.array([1, 1], np.int32)
    rc = 12.0
    pairs = oracle.ff_pairs([50.0, 80.0, 120.0], [3.0, 3.4, 3.7], [1, 1, 0], True, rc, rc)
    pts = np.random.default_rng(5).uniform(0, 1, (400, 3)) @ cell.T
    dist = np.array([min(np.linalg.norm(p - x - cell @ np.array(n)) for x in fpos
                         for n in itertools.product((-1, 0, 1), repeat=3)) for p in pts])
    sel = dist > 2.5
    ref = np.array([oracle.energy_at_position(cell, rect, 0, fpos, ftype, np.zeros(2), pairs, 2, rc, rc, 0.3, p)
                    for p in pts[sel]])
    errs = []
    for n in (25, 49):
        dims = (n, n, n)
        ijk = np.array([(i, j, k) for k in range(n) for j in range(n) for i in range(n)])
        vals = oracle.grid_points_q(cell, rect, dims, 0, fpos, ftype, np.zeros(2), pairs, 2, rc, rc, 0.3, ijk,
                                    nthreads=8)
        e, _, _ = oracle.interpolate_q(cell, rect, dims, vals.reshape(n, n, n, 27), pts[sel])
        errs.append(np.sqrt(np.mean((e - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    assert errs[1] < errs[0] / 32.0, errs


def _quintic_grids(dims, triclinic, overlap=1e30):
    cell, rect, fpos, ftype, fq, pairs = system(triclinic)
    ijk = np.array([(i, j, k) for k in range(dims[2]) for j in range(dims[1]) for i in range(dims[0])])
    shape = (dims[2], dims[1], dims[0], 27)
    lj = oracle.grid_points_q(cell, rect, dims, 0, fpos, ftype, fq, pairs, 2, 8.0, 8.0, 0.3, ijk, overlap=overlap,
                              nthreads=8).reshape(shape)
    q = oracle.grid_points_q(cell, rect, dims, 3, fpos, ftype, fq, pairs, 0, 8.0, 8.0, 0.3, ijk, overlap=overlap,
                             nthreads=8).reshape(shape)
    return cell, rect, fpos, ftype, fq, pairs, lj, q


@pytest.mark.parametrize("triclinic", [False, True])
def test_grid_check_quintic_is_exact_at_nodes(triclinic):
    """The grid test with triquintic grids (system_interpolation_grids.cpp:328-441) at grid nodes: the interpolant
    reproduces the analytic energy, gradient and Hessian there, so every relative error vanishes; the weights are
    exp(-beta E) of the independent per-point energy."""
    dims = (7, 8, 6)
    cell, rect, fpos, ftype, fq, pairs, lj, q = _quintic_grids(dims, triclinic)
    nc = np.array(dims) - 1.0
    frac = np.array([[i, j, k] for i in (1, 3, 5) for j in (2, 4, 6) for k in (1, 2, 4)]) / nc
    r = oracle.grid_check(cell, rect, dims, fpos, ftype, fq, pairs, 2, 8.0, 8.0, 0.3, lj, frac, 300.0,
                          coulomb_grid=q, probe_charge=0.8, overlap=1e30)
    for part in ("lennard_jones", "coulomb"):
        st = r[part]
        assert st["energy_relative_error"] < 1e-10
        assert np.all(st["gradient_relative_error"] < 1e-9) and np.all(st["hessian_relative_error"] < 1e-9), st
        assert np.allclose(st["hessian_table"], st["hessian_full"], rtol=1e-9, atol=1e-9 * np.abs(st["hessian_full"]).max())
    ek = np.array([oracle.energy_at_position(cell, rect, 0, fpos, ftype, fq, pairs, 2, 8.0, 8.0, 0.3, f @ cell.T)
                   for f in frac]) * U["energy_to_kelvin"]
    w = np.exp(-ek / (U["kB"] * 300.0))
    assert abs(r["lennard_jones"]["weight"] - w.sum()) <= 1e-12 * w.sum()


def test_grid_check_quintic_beats_tricubic_at_random_points():
    """At random points of the same grid spacing the triquintic energy, gradient and Hessian errors of the grid test
    are below the tricubic energy and gradient errors (the tricubic scheme reports no Hessian)."""
    dims = (13, 14, 12)
    cell, rect, fpos, ftype, fq, pairs, lj, q = _quintic_grids(dims, True, overlap=1e6)
    lj3 = oracle.grid_build(cell, rect, dims, 0, fpos, ftype, fq, pairs, 2, 8.0, 8.0, 0.3)
    frac = np.random.default_rng(4).uniform(0, 1, (300, 3))
    r5 = oracle.grid_check(cell, rect, dims, fpos, ftype, fq, pairs, 2, 8.0, 8.0, 0.3, lj, frac, 300.0)
    r3 = oracle.grid_check(cell, rect, dims, fpos, ftype, fq, pairs, 2, 8.0, 8.0, 0.3, lj3, frac, 300.0)
    a, b = r5["lennard_jones"], r3["lennard_jones"]
    assert a["energy_relative_error"] < b["energy_relative_error"] / 3
    assert np.all(a["gradient_relative_error"] < b["gradient_relative_error"] / 3)
    assert np.all(b["hessian_relative_error"] == 0) and np.all(a["hessian_relative_error"] > 0)
    assert np.all(a["hessian_relative_error"] < 0.5), a["hessian_relative_error"]
