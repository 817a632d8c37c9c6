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
    out = np.zeros(8 * nx * ny * nz)
    c9, r, na, p, t, q, nt, e, s, sh, lj = _fw_args(cell, rect, fw_pos, fw_type, fw_q, pairs)
    st = lib(native).orc_grid_build(c9, r, nx, ny, nz, grid_type, na, p, t, q, nt, e, s, sh, lj, probe, rc_vdw,
                                    rc_supercell, alpha, overlap, k0, k1, nthreads, _d(out))
    assert st == 0
    return out.reshape(nz, ny, nx, 8)[k0:k1]


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
