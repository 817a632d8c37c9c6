"""ctypes binding of libraspa3_amd.so (the C ABI declared in include/raspa3_amd/rg.h).

The library must have been built (``__graft_entry__.build()`` or ``make -C raspa3_amd``).  Loading fails with an
error if it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libraspa3_amd.so")

_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p

RG_OK, RG_ERR_INVALID, RG_ERR_UNSUPPORTED, RG_ERR_DEVICE, RG_ERR_NOMEM = range(5)
RG_GRID_LENNARD_JONES, RG_GRID_EWALD_REAL = 0, 3
RG_SCHEME_TRICUBIC, RG_SCHEME_TRIQUINTIC = 8, 27
RG_GRID_FP64, RG_GRID_SPLIT_FP32 = 0, 1

# name -> (restype, argtypes); every symbol declared in include/raspa3_amd/rg.h
SIGNATURES = {
    "rg_version": (C.c_char_p, []),
    "rg_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "rg_ctx_destroy": (C.c_int, [vp]),
    "rg_last_error": (C.c_char_p, [vp]),
    "rg_sync": (C.c_int, [vp]),
    "rg_stream": (vp, [vp]),
    "rg_box": (C.c_int, [C.c_double] * 6 + [C.c_int, dp, dp, ip, dp, dp]),
    "rg_grid_dims": (C.c_int, [dp, C.c_double, ip]),
    "rg_forcefield_create": (C.c_int, [vp, C.c_int, dp, dp, ip, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_int, C.c_double, C.POINTER(vp)]),
    "rg_forcefield_destroy": (C.c_int, [vp]),
    "rg_forcefield_pairs": (C.c_int, [vp, dp, dp, dp, ip]),
    "rg_framework_create": (C.c_int, [vp, dp, C.c_int, C.c_int, dp, ip, dp, C.POINTER(vp)]),
    "rg_framework_destroy": (C.c_int, [vp]),
    "rg_grid_create": (C.c_int, [vp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    "rg_grid_create_scheme": (C.c_int, [vp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    "rg_grid_destroy": (C.c_int, [vp]),
    "rg_grid_device_data": (vp, [vp]),
    "rg_grid_build": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_double, vp]),
    "rg_grid_build_many": (C.c_int, [vp, vp, vp, C.c_int, ip, ip, dp, C.POINTER(vp)]),
    "rg_grid_build_many_async": (C.c_int, [vp, vp, vp, C.c_int, ip, ip, dp, C.POINTER(vp)]),
    "rg_grid_build_many_precision": (C.c_int, [vp, vp, vp, C.c_int, ip, ip, dp, C.POINTER(vp), C.c_int, C.c_double,
                                               C.c_int]),
    "rg_grid_download": (C.c_int, [vp, vp, C.c_size_t, C.c_size_t, dp]),
    "rg_grid_upload": (C.c_int, [vp, vp, C.c_size_t, C.c_size_t, dp]),
    "rg_grid_interpolate": (C.c_int, [vp, vp, C.c_int, dp, dp, dp]),
    "rg_grid_interpolate_hessian": (C.c_int, [vp, vp, C.c_int, dp, dp, dp, dp]),
    "rg_ewald_parameters": (C.c_int, [C.c_double, C.c_double, dp, dp, ip]),
    "rg_system_create": (C.c_int, [vp, vp, dp, C.c_int, vp, C.c_int, ip, C.POINTER(vp), vp, ip, C.c_longlong,
                                   C.c_int, C.POINTER(vp)]),
    "rg_system_destroy": (C.c_int, [vp]),
    "rg_system_info": (C.c_int, [vp, ip, dp, dp, ip]),
    "rg_system_set_adsorbates": (C.c_int, [vp, vp, C.c_int, dp, ip, dp, ip]),
    "rg_trial_energies": (C.c_int, [vp, vp, C.c_int, dp, ip, dp, ip, C.c_int, dp, ip]),
    "rg_trial_energies_noalloc": (C.c_int, [vp, vp, C.c_int, dp, ip, dp, ip, C.c_int, dp, ip]),
    "rg_rosenbluth_select": (C.c_int, [vp, C.c_int, C.c_int, dp, ip, C.c_double, dp, ip, dp]),
    "rg_ewald_delta": (C.c_int, [vp, vp, C.c_int, dp, dp, ip, C.c_int, dp, dp, ip, dp]),
    "rg_ewald_accept": (C.c_int, [vp, vp]),
    "rg_ewald_energy": (C.c_int, [vp, vp, dp]),
}


class RGError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RGError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C raspa3_amd)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def d(a):
    return a.ctypes.data_as(dp) if a is not None else None


def i(a):
    return a.ctypes.data_as(ip) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def check(status, ctx=None, what=""):
    if status != RG_OK:
        msg = lib().rg_last_error(ctx).decode() if ctx else ""
        raise RGError(f"{what}: status {status}: {msg}")


MAX_ATOMS = 8  # most atoms per Monte Carlo component; adsorbate arrays hold max(n_atoms) slots per molecule


class ComponentDesc(C.Structure):
    _fields_ = [("n_atoms", C.c_int), ("first_bead", C.c_int), ("position", C.c_double * 24),
                ("type", C.c_int * 8), ("charge", C.c_double * 8), ("mol_fraction", C.c_double),
                ("flexible", C.c_int), ("bond_length", C.c_double), ("bend_theta0_deg", C.c_double),
                ("bend_k_K", C.c_double), ("torsion_K", C.c_double * 4), ("log_ideal_rosenbluth", C.c_double)]


class McParams(C.Structure):
    _fields_ = [("temperature", C.c_double), ("pressure_pa", C.c_double), ("k_first", C.c_int),
                ("k_orient", C.c_int), ("p_translation", C.c_double), ("p_rotation", C.c_double),
                ("max_translation", C.c_double), ("max_rotation", C.c_double), ("min_steps_per_cycle", C.c_int),
                ("max_molecules", C.c_int), ("seed", C.c_ulonglong), ("production_cycles", C.c_longlong),
                ("p_reinsertion", C.c_double)]


class McState(C.Structure):
    _fields_ = [("n_molecules", C.c_int), ("n_per_component", C.c_int * 2), ("energy", C.c_double),
                ("steps", C.c_longlong), ("cycles_sampled", C.c_longlong), ("trials", C.c_longlong * 5),
                ("accepted", C.c_longlong * 5), ("block_count", C.c_double * 5), ("block_energy", C.c_double * 5),
                ("block_n", (C.c_double * 2) * 5), ("block_energy_n", (C.c_double * 2) * 5),
                ("block_nn", ((C.c_double * 2) * 2) * 5)]


class GridCheckStats(C.Structure):
    _fields_ = [("weight", C.c_double), ("energy_table", C.c_double), ("energy_full", C.c_double),
                ("energy_relative_error", C.c_double), ("gradient_table", C.c_double * 3),
                ("gradient_full", C.c_double * 3), ("gradient_relative_error", C.c_double * 3),
                ("hessian_table", C.c_double * 6), ("hessian_full", C.c_double * 6),
                ("hessian_relative_error", C.c_double * 6)]


class GridCheck(C.Structure):
    _fields_ = [("n_points", C.c_longlong), ("lennard_jones", GridCheckStats), ("n_coulomb_points", C.c_longlong),
                ("coulomb", GridCheckStats)]


SIGNATURES.update({
    "rg_grid_check": (C.c_int, [vp, vp, vp, vp, C.c_int, vp, C.c_double, C.c_double, C.c_int, dp, C.c_ulonglong,
                                C.POINTER(GridCheck)]),
    "rg_widom": (C.c_int, [vp, vp, C.POINTER(ComponentDesc), C.POINTER(McParams), C.c_longlong, C.c_int, dp, dp,
                           dp]),
    "rg_widom_device": (C.c_int, [vp, vp, C.POINTER(ComponentDesc), C.POINTER(McParams), C.c_longlong, C.c_int, vp, vp,
                           vp]),
    "rg_mc_create": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.c_int, C.POINTER(ComponentDesc), C.POINTER(McParams),
                               C.POINTER(vp)]),
    "rg_mc_destroy": (C.c_int, [vp]),
    "rg_mc_set_launch_shape": (C.c_int, [vp, vp, C.c_int]),
    "rg_mc_set_balance": (C.c_int, [vp, vp, C.c_int, C.c_double]),
    "rg_mc_set_proposal": (C.c_int, [vp, vp, C.c_int, ip, dp]),
    "rg_mc_set_component_proposal": (C.c_int, [vp, vp, C.c_int, C.c_int, ip, dp]),
    "rg_mc_run": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int]),
    "rg_mc_state": (C.c_int, [vp, vp, C.c_int, C.POINTER(McState)]),
    "rg_mc_adsorbates": (C.c_int, [vp, vp, C.c_int, C.c_int, dp, ip, ip, ip]),
    "rg_mc_total_energy": (C.c_int, [vp, vp, dp]),
    "rg_molecule_weights": (C.c_int, [vp, vp, C.POINTER(ComponentDesc), C.c_double, C.c_int, dp, C.c_int,
                                      C.c_ulonglong, dp]),
})

fp = C.POINTER(C.c_float)

SIGNATURES.update({
    "rg_mlp_num_params": (C.c_size_t, []),
    "rg_autotune_summary": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "rg_mlp_create": (C.c_int, [vp, C.c_ulonglong, C.POINTER(vp)]),
    "rg_mlp_destroy": (C.c_int, [vp]),
    "rg_mlp_get_params": (C.c_int, [vp, vp, fp]),
    "rg_mlp_set_params": (C.c_int, [vp, vp, fp]),
    "rg_mlp_forward": (C.c_int, [vp, vp, C.c_int, dp, dp]),
    "rg_mlp_gradients": (C.c_int, [vp, vp, C.c_int, dp, dp, C.c_double, fp, dp]),
    "rg_mlp_train": (C.c_int, [vp, vp, C.c_int, dp, dp, C.c_double, C.c_int, C.c_int, C.c_double, C.c_ulonglong,
                               dp]),
    "rg_mlp_proposal": (C.c_int, [vp, vp, ip, C.c_double, dp]),
})


def f32p(a):
    return a.ctypes.data_as(fp) if a is not None else None


class WidomStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rosenbluth_weight", "rosenbluth_weight_ci", "mu_excess", "mu_excess_ci",
                                          "mu_ideal_gas", "mu_ideal_gas_ci", "mu_total", "mu_total_ci", "fugacity",
                                          "fugacity_ci")]


SIGNATURES.update({
    "rg_block_error": (C.c_int, [C.c_int, dp, C.c_double, dp]),
    "rg_block_average": (C.c_int, [C.c_int, dp, dp, dp, dp]),
    "rg_enthalpy_of_adsorption": (C.c_int, [C.c_int, C.c_int, dp, dp, dp, dp, dp, C.c_double, dp, dp]),
    "rg_widom_statistics": (C.c_int, [C.c_int, dp, dp, dp, C.c_double, C.POINTER(WidomStats)]),
})


SIGNATURES.update({
    "rg_grid_key": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.POINTER(C.c_ulonglong)]),
    "rg_grid_save": (C.c_int, [vp, vp, C.c_ulonglong, C.c_char_p]),
    "rg_grid_load": (C.c_int, [vp, vp, C.c_ulonglong, C.c_char_p]),
    "rg_grid_write_cube": (C.c_int, [vp, vp, C.c_char_p]),
})
