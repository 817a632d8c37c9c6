"""FP64 CPU oracle (test infrastructure only).

ctypes wrapper of oracle/liboracle.so, the plain C++ restatement of the reference algorithms in oracle/oracle.cpp.
Only the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import this package,
and only as the checker or the timed CPU baseline; the product library never calls it.

Parity status of each function is stated in oracle/oracle.cpp's header.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

_libs = {}


def lib(native=False):
    name = "liboracle_native.so" if native else "liboracle.so"
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        L.orc_units.argtypes = [dp]
        L.orc_box.argtypes = [C.c_double] * 6 + [C.c_int, dp, dp, ip, dp, dp]
        L.orc_ff_pairs.argtypes = [C.c_int, dp, dp, ip, C.c_int, C.c_double, C.c_double, dp, dp, dp, ip]
        L.orc_ewald_auto.argtypes = [C.c_double, C.c_double, dp, dp, ip]
        L.orc_grid_dims.argtypes = [dp, C.c_double, ip]
        grid_common = [dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, ip, dp, C.c_int, dp, dp, dp,
                       ip, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_grid_build.argtypes = grid_common + [C.c_int, C.c_int, C.c_int, dp]
        L.orc_grid_build.restype = C.c_int
        L.orc_grid_points.argtypes = grid_common + [C.c_int, ip, C.c_int, dp]
        L.orc_grid_points.restype = C.c_int
        L.orc_energy_at_position.argtypes = [dp, C.c_int, C.c_int, C.c_int, dp, ip, dp, C.c_int, dp, dp, dp, ip,
                                             C.c_int, C.c_double, C.c_double, C.c_double, dp]
        L.orc_energy_at_position.restype = C.c_double
        L.orc_interpolate.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, dp, dp]
        L.orc_tricubic_matrix.argtypes = [dp]
        L.orc_grid_points_q.argtypes = grid_common + [C.c_int, ip, C.c_int, dp]
        L.orc_grid_points_q.restype = C.c_int
        L.orc_interpolate_hessian.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, dp, dp, dp]
        L.orc_interpolate_q.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, dp, dp, dp]
        L.orc_quintic_at_positions.argtypes = [dp, C.c_int, C.c_int, C.c_int, dp, ip, dp, C.c_int, dp, dp, dp, ip,
                                               C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]
        L.orc_grid_check.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, ip, dp, C.c_int, dp, dp,
                                     dp, ip, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp,
                                     C.c_double, C.c_double, C.c_int, dp, C.c_int, dp]
        L.orc_framework_energy.argtypes = [dp, C.c_int, C.c_int, dp, ip, dp, dp, dp, C.c_int, dp, dp, dp, ip,
                                           C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, dp, ip,
                                           dp, dp, dp, C.c_int, dp, C.c_int, ip, ip, C.POINTER(dp), dp, dp]
        L.orc_framework_energy.restype = C.c_int
        L.orc_intermolecular_energy.argtypes = [dp, C.c_int, C.c_int, dp, ip, dp, dp, dp, ip, C.c_int, dp, dp, dp,
                                                ip, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                                dp, ip, dp, dp, dp, ip, C.c_int, C.c_int, dp]
        L.orc_intermolecular_energy.restype = C.c_int
        L.orc_ewald_kvectors.argtypes = [dp, C.c_int, C.c_double, ip, C.c_longlong, C.c_double, dp]
        L.orc_ewald_kvectors.restype = C.c_int
        L.orc_structure_factor.argtypes = [dp, C.c_int, C.c_double, ip, C.c_longlong, C.c_double, C.c_int, dp, dp,
                                           dp]
        L.orc_structure_factor.restype = C.c_int
        L.orc_ewald_single_ion.argtypes = [dp, C.c_int, C.c_double, ip, C.c_longlong, C.c_double]
        L.orc_ewald_single_ion.restype = C.c_double
        L.orc_ewald_total.argtypes = [dp, C.c_int, C.c_double, ip, C.c_longlong, C.c_double, C.c_int, dp, dp,
                                      C.c_int, dp, dp, dp, ip, C.c_double, dp, dp]
        L.orc_ewald_total.restype = C.c_int
        L.orc_ewald_delta.argtypes = [dp, C.c_int, C.c_double, ip, C.c_longlong, C.c_double, dp, C.c_int, dp, dp,
                                      dp, ip, C.c_int, dp, dp, dp, ip, C.c_double, dp, dp]
        L.orc_ewald_delta.restype = C.c_int
        L.orc_select_trial.argtypes = [C.c_int, dp, C.c_double]
        L.orc_select_trial.restype = C.c_int
        _libs[name] = L
    return _libs[name]


def _d(a):
    return None if a is None else a.ctypes.data_as(dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(ip)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _n(a):
    return np.ascontiguousarray(a, dtype=np.int32)


INT_MAX_K = (1 << 62)
DBL_MAX = np.finfo(np.float64).max


def units():
    out = np.zeros(4)
    lib().orc_units(_d(out))
    return dict(energy_to_kelvin=out[0], kelvin_to_energy=out[1], coulomb=out[2], kB=out[3])


def box(a, b, c, alpha=90.0, beta=90.0, gamma=90.0, lengths_only=False):
    cell, inv, w = np.zeros(9), np.zeros(9), np.zeros(3)
    rect, vol = C.c_int(), C.c_double()
    lib().orc_box(a, b, c, alpha, beta, gamma, int(lengths_only), _d(cell), _d(inv), C.byref(rect), C.byref(vol),
                  _d(w))
    return dict(cell9=cell, cell=cell.reshape(3, 3).T.copy(), inv=inv.reshape(3, 3).T.copy(), rect=bool(rect.value),
                volume=vol.value, widths=w)


def ff_pairs(eps_K, sigma, is_framework, shifted=True, rc_framework=12.0, rc_molecule=12.0):
    n = len(eps_K)
    eps, sig, sh = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
    lj = np.zeros(n * n, np.int32)
    lib().orc_ff_pairs(n, _d(_f(eps_K)), _d(_f(sigma)), _i(_n(is_framework)), int(shifted), rc_framework,
                       rc_molecule, _d(eps), _d(sig), _d(sh), _i(lj))
    return dict(eps=eps, sigma=sig, shift=sh, lj=lj, n=n)


def pairs_from_spec(spec):
    return ff_pairs(spec.eps_K, spec.sigma, spec.is_framework, spec.shifted, spec.rc_framework, spec.rc_molecule)


def ewald_auto(precision, rc, widths):
    a = C.c_double()
    k = np.zeros(3, np.int32)
    lib().orc_ewald_auto(precision, rc, _d(_f(widths)), C.byref(a), _i(k))
    return a.value, k


def _fw_args(cell, rect, fw_pos, fw_type, fw_q, pairs):
    return (_d(_f(np.asarray(cell).T.reshape(9))), int(rect), len(fw_type), _d(_f(fw_pos)), _i(_n(fw_type)),
            _d(_f(fw_q)), pairs["n"], _d(pairs["eps"]), _d(pairs["sigma"]), _d(pairs["shift"]), _i(pairs["lj"]))


def grid_build(cell, rect, dims, grid_type, fw_pos, fw_type, fw_q, pairs, probe, rc_vdw, rc_supercell, alpha,
               overlap=1e6, k_range=None, nthreads=1, native=False):
    """Full-loop reference builder over z-planes k_range (default all).  Returns array (nz_sel, ny, nx, 8)."""
    nx, ny, nz = dims
    k0, k1 = (0, nz) if k_range is None else k_range
    out = np.zeros(8 * nx * ny * (k1 - k0))
    c9, r, na, p, t, q, nt, e, s, sh, lj = _fw_args(cell, rect, fw_pos, fw_type, fw_q, pairs)
    st = lib(native).orc_grid_build(c9, r, nx, ny, nz, grid_type, na, p, t, q, nt, e, s, sh, lj, probe, rc_vdw,
                                    rc_supercell, alpha, overlap, k0, k1, nthreads, _d(out))
    assert st == 0
    return out.reshape(k1 - k0, ny, nx, 8)


def grid_points(cell, rect, dims, grid_type, fw_pos, fw_type, fw_q, pairs, probe, rc_vdw, rc_supercell, alpha, ijk,
                overlap=1e6, nthreads=1, native=False):
    """Reference values at selected grid points ijk (m, 3).  Returns (m, 8)."""
    ijk = _n(ijk).reshape(-1, 3)
    out = np.zeros(8 * len(ijk))
    c9, r, na, p, t, q, nt, e, s, sh, lj = _fw_args(cell, rect, fw_pos, fw_type, fw_q, pairs)
    st = lib(native).orc_grid_points(c9, r, *dims, grid_type, na, p, t, q, nt, e, s, sh, lj, probe, rc_vdw,
                                     rc_supercell, alpha, overlap, len(ijk), _i(ijk), nthreads, _d(out))
    assert st == 0
    return out.reshape(-1, 8)


def grid_points_q(cell, rect, dims, grid_type, fw_pos, fw_type, fw_q, pairs, probe, rc_vdw, rc_supercell, alpha,
                  ijk, overlap=1e6, nthreads=1, native=False):
    """Triquintic reference values (27 per point, scaled and clamped as stored) at grid points ijk (m, 3)."""
    ijk = _n(ijk).reshape(-1, 3)
    out = np.zeros(27 * len(ijk))
    c9, r, na, p, t, q, nt, e, s, sh, lj = _fw_args(cell, rect, fw_pos, fw_type, fw_q, pairs)
    st = lib(native).orc_grid_points_q(c9, r, *dims, grid_type, na, p, t, q, nt, e, s, sh, lj, probe, rc_vdw,
                                       rc_supercell, alpha, overlap, len(ijk), _i(ijk), nthreads, _d(out))
    assert st == 0
    return out.reshape(-1, 27)


def quintic_at_positions(cell, rect, grid_type, fw_pos, fw_type, fw_q, pairs, probe, rc_vdw, rc_supercell, alpha,
                         pos):
    """The 27 unscaled fractional derivatives of the framework energy at Cartesian positions (m, 3)."""
    pos = _f(pos).reshape(-1, 3)
    out = np.zeros(27 * len(pos))
    c9, r, na, p, t, q, nt, e, s, sh, lj = _fw_args(cell, rect, fw_pos, fw_type, fw_q, pairs)
    lib().orc_quintic_at_positions(c9, r, grid_type, na, p, t, q, nt, e, s, sh, lj, probe, rc_vdw, rc_supercell,
                                   alpha, len(pos), _d(pos), _d(out))
    return out.reshape(-1, 27)


def interpolate_hessian(cell, rect, dims, grid, pos):
    """Tricubic interpolation of a (nz, ny, nx, 8) grid: energies (m,), gradients (m, 3), Hessians (m, 3, 3)."""
    pos = _f(pos).reshape(-1, 3)
    e = np.zeros(len(pos))
    g = np.zeros((len(pos), 3))
    h = np.zeros((len(pos), 3, 3))
    lib().orc_interpolate_hessian(_d(_f(np.asarray(cell).T.reshape(9))), int(rect), *dims, _d(_f(grid).reshape(-1)),
                                  len(pos), _d(pos), _d(e), _d(g), _d(h))
    return e, g, h


def interpolate_q(cell, rect, dims, grid, pos):
    """Triquintic interpolation of a (nz, ny, nx, 27) grid: energies (m,), gradients (m, 3), Hessians (m, 3, 3)."""
    pos = _f(pos).reshape(-1, 3)
    e = np.zeros(len(pos))
    g = np.zeros((len(pos), 3))
    h = np.zeros((len(pos), 3, 3))
    lib().orc_interpolate_q(_d(_f(np.asarray(cell).T.reshape(9))), int(rect), *dims, _d(_f(grid).reshape(-1)),
                            len(pos), _d(pos), _d(e), _d(g), _d(h))
    return e, g, h


def energy_at_position(cell, rect, grid_type, fw_pos, fw_type, fw_q, pairs, probe, rc_vdw, rc_supercell, alpha, p):
    c9, r, na, pp, t, q, nt, e, s, sh, lj = _fw_args(cell, rect, fw_pos, fw_type, fw_q, pairs)
    return lib().orc_energy_at_position(c9, r, grid_type, na, pp, t, q, nt, e, s, sh, lj, probe, rc_vdw,
                                        rc_supercell, alpha, _d(_f(p)))


def interpolate(cell, rect, dims, grid, pos, gradient=False):
    pos = _f(pos).reshape(-1, 3)
    e = np.zeros(len(pos))
    g = np.zeros((len(pos), 3)) if gradient else None
    lib().orc_interpolate(_d(_f(np.asarray(cell).T.reshape(9))), int(rect), *dims, _d(_f(grid).reshape(-1)),
                          len(pos), _d(pos), _d(e), _d(g))
    return (e, g) if gradient else e


def _check_stats(o):
    return dict(weight=o[0], energy_table=o[1], energy_full=o[2], energy_relative_error=o[3],
                gradient_table=o[4:7].copy(), gradient_full=o[7:10].copy(), gradient_relative_error=o[10:13].copy(),
                hessian_table=o[13:19].copy(), hessian_full=o[19:25].copy(), hessian_relative_error=o[25:31].copy())


def grid_check(cell, rect, dims, fw_pos, fw_type, fw_q, pairs, probe, rc_vdw, rc_supercell, alpha, lj_grid, frac,
               temperature, coulomb_grid=None, probe_charge=0.0, overlap=1e6):
    """The reference's grid test (system_interpolation_grids.cpp:240-840) at the given fractional points, for
    tricubic ((..., 8) arrays) or triquintic ((..., 27) arrays, with Hessians) grids: dict(n_points, lennard_jones,
    n_coulomb_points, coulomb) with Boltzmann-weighted averages and relative errors."""
    c9, r, na, p, t, q, nt, e, s, sh, lj = _fw_args(cell, rect, fw_pos, fw_type, fw_q, pairs)
    frac = _f(frac).reshape(-1, 3)
    g = _f(lj_grid).reshape(-1)
    values = g.size // (dims[0] * dims[1] * dims[2])  # 8 or 27 values per point
    assert values in (8, 27) and values * dims[0] * dims[1] * dims[2] == g.size
    cg = None if coulomb_grid is None else _f(coulomb_grid).reshape(-1)
    out = np.zeros(64)
    lib().orc_grid_check(c9, r, *dims, na, p, t, q, nt, e, s, sh, lj, probe, rc_vdw, rc_supercell, alpha, overlap,
                         _d(g), _d(cg) if cg is not None else None, probe_charge, temperature, len(frac), _d(frac),
                         int(values), _d(out))
    return dict(n_points=int(out[0]), lennard_jones=_check_stats(out[1:32]), n_coulomb_points=int(out[32]),
                coulomb=_check_stats(out[33:64]))


def tricubic_matrix():
    out = np.zeros(64 * 64)
    lib().orc_tricubic_matrix(_d(out))
    return out.reshape(64, 64)


def framework_energy(cell, rect, fw_pos, fw_type, fw_q, pairs, rc_vdw, rc_coulomb, alpha, use_charge, sites,
                     overlap=1e6, skip=-1, fw_scaling_vdw=None, fw_scaling_coul=None):
    """CBMC::computeFrameworkMoleculeEnergy, direct branch.  sites: dict(pos, type, charge[, scaling_vdw,
    scaling_coul]).  Returns (vdw, coulomb) or None on overlap."""
    nfw = len(fw_type)
    svdw = _f(np.ones(nfw) if fw_scaling_vdw is None else fw_scaling_vdw)
    scou = _f(np.ones(nfw) if fw_scaling_coul is None else fw_scaling_coul)
    ns = len(sites["type"])
    s_sv = _f(sites.get("scaling_vdw", np.ones(ns)))
    s_sc = _f(sites.get("scaling_coul", np.ones(ns)))
    out = np.zeros(2)
    c9 = _f(np.asarray(cell).T.reshape(9))
    st = lib().orc_framework_energy(_d(c9), int(rect), nfw, _d(_f(fw_pos)), _i(_n(fw_type)), _d(_f(fw_q)), _d(svdw),
                                    _d(scou), pairs["n"], _d(pairs["eps"]), _d(pairs["sigma"]), _d(pairs["shift"]),
                                    _i(pairs["lj"]), rc_vdw, rc_coulomb, alpha, int(use_charge), overlap, ns,
                                    _d(_f(sites["pos"])), _i(_n(sites["type"])), _d(_f(sites["charge"])), _d(s_sv),
                                    _d(s_sc), skip, None, 0, None, None, None, None, _d(out))
    return None if st else (out[0], out[1])


def intermolecular_energy(cell, rect, mol, pairs, rc_vdw, rc_coulomb, alpha, use_charge, sites, overlap=1e6,
                          skip=-1, skip_molecule=-1):
    """CBMC::computeInterMolecularEnergy.  mol, sites: dict(pos, type, charge, molecule[, scaling_*])."""
    nm = len(mol["type"])
    ns = len(sites["type"])
    out = np.zeros(2)
    c9 = _f(np.asarray(cell).T.reshape(9))
    st = lib().orc_intermolecular_energy(
        _d(c9), int(rect), nm, _d(_f(mol["pos"])), _i(_n(mol["type"])), _d(_f(mol["charge"])),
        _d(_f(mol.get("scaling_vdw", np.ones(nm)))), _d(_f(mol.get("scaling_coul", np.ones(nm)))),
        _i(_n(mol["molecule"])), pairs["n"], _d(pairs["eps"]), _d(pairs["sigma"]), _d(pairs["shift"]),
        _i(pairs["lj"]), rc_vdw, rc_coulomb, alpha, int(use_charge), overlap, ns, _d(_f(sites["pos"])),
        _i(_n(sites["type"])), _d(_f(sites["charge"])), _d(_f(sites.get("scaling_vdw", np.ones(ns)))),
        _d(_f(sites.get("scaling_coul", np.ones(ns)))), _i(_n(sites["molecule"])), skip, skip_molecule, _d(out))
    return None if st else (out[0], out[1])


def _kargs(kmax, int_cut2, cut2):
    return _i(_n(kmax)), C.c_longlong(INT_MAX_K if int_cut2 is None else int(int_cut2)), \
        (DBL_MAX if cut2 is None else float(cut2))


def ewald_kvectors(cell, rect, alpha, kmax, int_cut2=None, cut2=None):
    cap = 4 * (kmax[0] + 1) * (kmax[1] + 1) * (kmax[2] + 1) + 8
    temps = np.zeros(cap)
    k, ic, c2 = _kargs(kmax, int_cut2, cut2)
    nk = lib().orc_ewald_kvectors(_d(_f(np.asarray(cell).T.reshape(9))), int(rect), alpha, k, ic, c2, _d(temps))
    return temps[:nk]


def structure_factor(cell, rect, alpha, kmax, pos, q, int_cut2=None, cut2=None):
    cap = 4 * (kmax[0] + 1) * (kmax[1] + 1) * (kmax[2] + 1) + 8
    s = np.zeros(2 * cap)
    k, ic, c2 = _kargs(kmax, int_cut2, cut2)
    nk = lib().orc_structure_factor(_d(_f(np.asarray(cell).T.reshape(9))), int(rect), alpha, k, ic, c2, len(q),
                                    _d(_f(pos)), _d(_f(q)), _d(s))
    return s[:2 * nk].reshape(-1, 2)


def ewald_single_ion(cell, rect, alpha, kmax, int_cut2=None, cut2=None):
    k, ic, c2 = _kargs(kmax, int_cut2, cut2)
    return lib().orc_ewald_single_ion(_d(_f(np.asarray(cell).T.reshape(9))), int(rect), alpha, k, ic, c2)


def ewald_total(cell, rect, alpha, kmax, fw_pos, fw_q, ads_pos, ads_q, ads_scaling, ads_molecule,
                net_charge_framework=0.0, int_cut2=None, cut2=None):
    cap = 4 * (kmax[0] + 1) * (kmax[1] + 1) * (kmax[2] + 1) + 8
    out = np.zeros(3)
    stored = np.zeros(2 * cap)
    k, ic, c2 = _kargs(kmax, int_cut2, cut2)
    nk = lib().orc_ewald_total(_d(_f(np.asarray(cell).T.reshape(9))), int(rect), alpha, k, ic, c2, len(fw_q),
                               _d(_f(fw_pos)), _d(_f(fw_q)), len(ads_q), _d(_f(ads_pos)), _d(_f(ads_q)),
                               _d(_f(ads_scaling)), _i(_n(ads_molecule)), net_charge_framework, _d(out), _d(stored))
    return out, stored[:2 * nk].reshape(-1, 2)


def ewald_delta(cell, rect, alpha, kmax, stored, new, old, net_charge, int_cut2=None, cut2=None):
    """new/old: dict(pos, q, scaling, molecule).  Returns (fourier, self, exclusion), trial S_k."""
    nk = len(stored)
    trial = np.zeros(2 * nk)
    out = np.zeros(3)
    k, ic, c2 = _kargs(kmax, int_cut2, cut2)

    def arrs(x):
        n = len(x["q"])
        return (n, _d(_f(x["pos"]).reshape(-1)), _d(_f(x["q"])), _d(_f(x.get("scaling", np.ones(n)))),
                _i(_n(x.get("molecule", np.zeros(n)))))
    nn, npos, nq, ns, nm = arrs(new)
    no, opos, oq, os_, om = arrs(old)
    lib().orc_ewald_delta(_d(_f(np.asarray(cell).T.reshape(9))), int(rect), alpha, k, ic, c2,
                          _d(_f(stored).reshape(-1)), nn, npos, nq, ns, nm, no, opos, oq, os_, om, net_charge,
                          _d(out), _d(trial))
    return out, trial.reshape(-1, 2)


def select_trial(lbf, uniform):
    return lib().orc_select_trial(len(lbf), _d(_f(lbf)), uniform)


def philox4x32(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    up = C.POINTER(C.c_uint32)
    lib().orc_philox4x32(c.ctypes.data_as(up), k.ctypes.data_as(up), o.ctypes.data_as(up))
    return o


def _mc_lib():
    L = lib()
    if not getattr(L, "_mc_ready", False):
        vp = C.c_void_p
        L.orc_mc_create.restype = vp
        L.orc_mc_create.argtypes = [dp, C.c_int, C.c_int, dp, dp, dp, ip, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_int, C.c_double, C.c_int, dp, ip, dp, C.POINTER(dp), dp, dp, ip,
                                    ip, C.c_longlong, C.c_int, ip, dp, dp, C.POINTER(C.c_longlong), C.c_ulonglong]
        L.orc_mc_destroy.argtypes = [vp]
        L.orc_mc_set_proposal.argtypes = [vp, C.c_int, ip, dp]
        L.orc_mc_run.argtypes = [vp, C.c_int, C.c_int]
        L.orc_mc_state.argtypes = [vp, dp, dp]
        L.orc_mc_adsorbates.argtypes = [vp, dp, ip, ip]
        L.orc_mc_adsorbates.restype = C.c_int
        L.orc_widom.argtypes = [vp, C.c_int, C.c_longlong, C.c_int, dp, dp, dp]
        L.orc_mc_grow.restype = C.c_int
        L.orc_mc_grow.argtypes = [vp, C.c_int, C.c_longlong, dp, dp, dp]
        L.orc_mc_molecule_weights.restype = C.c_int
        L.orc_mc_molecule_weights.argtypes = [vp, C.c_int, C.c_int, dp, C.c_int, C.c_ulonglong, dp, dp]
        L._mc_ready = True
    return L


MAX_ATOMS = 8      # most atoms per Monte Carlo component; a system's adsorbate arrays hold max(n_atoms) per molecule
COMP_INT, COMP_DBL = 11, 41  # component record sizes of orc_mc_create


class MC:
    """One system of the oracle's Monte Carlo replay (same moves and Philox stream layout as the device engine).

    fw: workloads.Framework (direct sums) unless grids = (dict type -> (nz, ny, nx, 8) array, coulomb array, dims)
    comps: list of dicts (n_atoms, first_bead, position (n,3), type, charge, mol_fraction)
    params: dict(temperature, pressure_pa, k_first, k_orient, p_translation, p_rotation, max_translation,
                 max_rotation, min_steps_per_cycle, max_molecules, seed, production_cycles)."""

    def __init__(self, spec, fw, comps, params, kmax=(8, 8, 8), system=0, grids=None, int_cut2=0):
        L = _mc_lib()
        pairs = pairs_from_spec(spec)
        n = pairs["n"]
        self._keep = []
        cell9 = _f(np.asarray(fw.cell).T.reshape(9))
        gptrs = None
        cg = None
        gdims = None
        gcell9 = cell9
        if grids is not None:
            gmap, coul, dims = grids[:3]
            if len(grids) > 3:
                gcell9 = _f(np.asarray(grids[3]).T.reshape(9))
            arrs = [None] * n
            for t, g in gmap.items():
                arrs[t] = _f(g).reshape(-1)
            self._keep += [a for a in arrs if a is not None]
            gptrs = (dp * n)(*[(a.ctypes.data_as(dp) if a is not None else dp()) for a in arrs])
            if coul is not None:
                cg = _f(coul).reshape(-1)
                self._keep.append(cg)
            # the fourth entry is the scheme: values per grid point, 8 (tricubic) or 27 (triquintic)
            values = next(iter(gmap.values())).shape[-1] if gmap else (coul.shape[-1] if coul is not None else 8)
            gdims = _n(list(dims) + [values])
            self._keep.append(gdims)
        self.slots = max(int(d["n_atoms"]) for d in comps)
        ci = np.zeros(COMP_INT * len(comps), np.int32)
        cd = np.zeros(COMP_DBL * len(comps))
        for c, d in enumerate(comps):
            na = d["n_atoms"]
            i0, d0 = COMP_INT * c, COMP_DBL * c
            ci[i0] = na
            ci[i0 + 1] = d["first_bead"]
            ci[i0 + 2] = int(d.get("flexible", False))
            ci[i0 + 3:i0 + 3 + na] = d["type"]
            cd[d0:d0 + 3 * na] = np.asarray(d["position"], float).reshape(-1)
            cd[d0 + 24:d0 + 24 + na] = d["charge"]
            cd[d0 + 32] = d["mol_fraction"]
            if d.get("flexible", False):
                cd[d0 + 33] = d["bond_length"]
                cd[d0 + 34] = d["bend_theta0_deg"]
                cd[d0 + 35] = d["bend_k_K"]
                cd[d0 + 36:d0 + 40] = d["torsion_K"]
                cd[d0 + 40] = d.get("log_ideal_rosenbluth", 0.0)
        dpar = _f([params["temperature"], params["pressure_pa"], params["p_translation"], params["p_rotation"],
                   params["max_translation"], params["max_rotation"], params.get("p_reinsertion", 0.0)])
        ipar = (C.c_longlong * 6)(params["k_first"], params["k_orient"], params["min_steps_per_cycle"],
                                  params["max_molecules"], params["production_cycles"], system)
        self._keep += [ci, cd, dpar, cell9, gcell9, pairs]
        fpos, ft, fq = _f(fw.pos), _n(fw.type), _f(fw.charge)
        self._keep += [fpos, ft, fq]
        self.h = L.orc_mc_create(_d(cell9), int(fw.rect), n, _d(pairs["eps"]), _d(pairs["sigma"]), _d(pairs["shift"]),
                                 _i(pairs["lj"]), spec.rc_framework, spec.rc_molecule, spec.rc_coulomb, spec.alpha,
                                 int(spec.use_charge), spec.overlap, len(ft), _d(fpos), _i(ft), _d(fq), gptrs,
                                 _d(cg) if cg is not None else None, _d(gcell9),
                                 _i(gdims) if gdims is not None else None, _i(_n(kmax)), int(int_cut2), len(comps),
                                 _i(ci), _d(cd), _d(dpar), ipar, int(params["seed"]))

    def set_proposal(self, n3, prob, component=-1):
        p = _f(prob).reshape(-1)
        self._keep.append(p)
        _mc_lib().orc_mc_set_proposal(self.h, int(component), _i(_n(n3)), _d(p))

    def run(self, cycles, sample=False):
        _mc_lib().orc_mc_run(self.h, int(cycles), int(sample))

    def state(self):
        s = np.zeros(16)
        b = np.zeros(50)
        _mc_lib().orc_mc_state(self.h, _d(s), _d(b))
        return dict(n_molecules=int(s[0]), n_per_component=[int(s[1]), int(s[2])], energy=s[3], steps=int(s[4]),
                    cycles=int(s[5]), trials=s[6:11].astype(int), accepted=s[11:16].astype(int),
                    blocks=b.reshape(5, 10))

    def adsorbates(self, capacity=4096):
        pos = np.zeros((capacity, 3))
        t = np.zeros(capacity, np.int32)
        c = np.zeros(capacity, np.int32)
        n = _mc_lib().orc_mc_adsorbates(self.h, _d(pos), _i(t), _i(c))
        return pos[:n], t[:n], c[:n // self.slots]

    def grow(self, component, step, n_atoms):
        """One growth of a molecule at move 'step' (no acceptance): positions, log W, energy (None if it failed)."""
        pos = np.zeros((n_atoms, 3))
        lw, u = C.c_double(), C.c_double()
        ok = _mc_lib().orc_mc_grow(self.h, component, step, _d(pos), C.byref(lw), C.byref(u))
        return (pos, lw.value, u.value) if ok else None

    def widom(self, component, first, count):
        lw, u, ew = np.zeros(count), np.zeros(count), np.zeros(count)
        _mc_lib().orc_widom(self.h, component, first, count, _d(lw), _d(u), _d(ew))
        return lw, u, ew

    def molecule_weights(self, component, frac, n_orient, seed, mean_energy=False):
        """log of the orientation-averaged Boltzmann factor per point (and the Boltzmann-weighted mean energy over
        the orientations when mean_energy)."""
        frac = _f(frac).reshape(-1, 3)
        out = np.zeros(len(frac))
        mu = np.zeros(len(frac)) if mean_energy else None
        st = _mc_lib().orc_mc_molecule_weights(self.h, component, len(frac), _d(frac), n_orient, seed, _d(out),
                                               _d(mu) if mu is not None else None)
        assert st == 0, "every atom type of the component needs a grid"
        return (out, mu) if mean_energy else out

    def __del__(self):
        try:
            _mc_lib().orc_mc_destroy(self.h)
        except Exception:
            pass
