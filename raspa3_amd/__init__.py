"""raspa3_amd: RASPA3's framework–adsorbate energy path on MI355X (gfx950).

Python host layer over the C ABI (include/raspa3_amd/rg.h), used by the tests and the benchmark.  All compute runs
in libraspa3_amd.so on the GPU; this module only moves host arrays in and out.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import RGError, lib

__all__ = ["Context", "ForceField", "Framework", "Grid", "box", "grid_dims", "RGError", "lib"]


def box(a, b, c, alpha=90.0, beta=90.0, gamma=90.0, lengths_only=False):
    """SimulationBox (src/raspakit/simulationbox.cpp:15-60): returns dict(cell, inv, rect, volume, widths)."""
    cell = np.zeros(9)
    inv = np.zeros(9)
    rect = C.c_int()
    vol = C.c_double()
    w = np.zeros(3)
    L.check(lib().rg_box(a, b, c, alpha, beta, gamma, int(lengths_only), L.d(cell), L.d(inv), C.byref(rect),
                         C.byref(vol), L.d(w)), None, "rg_box")
    return dict(cell=cell.reshape(3, 3).T.copy(), inv=inv.reshape(3, 3).T.copy(), rect=bool(rect.value),
                volume=vol.value, widths=w)


def grid_dims(cell, spacing):
    n = np.zeros(3, np.int32)
    c9 = L.f64(np.asarray(cell).T.reshape(9))
    L.check(lib().rg_grid_dims(L.d(c9), spacing, L.i(n)), None, "rg_grid_dims")
    return [int(x) for x in n]


class Context:
    """rg_ctx: one HIP device and stream."""

    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib().rg_ctx_create(device, C.byref(h))
        if st != L.RG_OK:
            raise RGError(f"rg_ctx_create failed with status {st} (no usable HIP device?)")
        self.h = h

    def close(self):
        if self.h:
            lib().rg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        L.check(lib().rg_sync(self.h), self.h, "rg_sync")

    @property
    def stream(self):
        return lib().rg_stream(self.h)

    def error(self):
        return lib().rg_last_error(self.h).decode()


class ForceField:
    """rg_forcefield built from a workloads.ForceFieldSpec (or any object with the same fields)."""

    def __init__(self, ctx, spec):
        self.ctx = ctx
        self.spec = spec
        h = C.c_void_p()
        eps, sig, isf = L.f64(spec.eps_K), L.f64(spec.sigma), L.i32(spec.is_framework)
        L.check(lib().rg_forcefield_create(ctx.h if ctx else None, len(eps), L.d(eps), L.d(sig), L.i(isf),
                                           int(spec.shifted), spec.rc_framework, spec.rc_molecule, spec.rc_coulomb,
                                           spec.alpha, int(spec.use_charge), spec.overlap, C.byref(h)),
                ctx.h if ctx else None, "rg_forcefield_create")
        self.h = h
        self.n = len(eps)

    def pairs(self):
        n = self.n
        eps, sig, sh = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
        lj = np.zeros(n * n, np.int32)
        L.check(lib().rg_forcefield_pairs(self.h, L.d(eps), L.d(sig), L.d(sh), L.i(lj)), None, "pairs")
        return eps.reshape(n, n), sig.reshape(n, n), sh.reshape(n, n), lj.reshape(n, n)

    def __del__(self):
        try:
            if self.h:
                lib().rg_forcefield_destroy(self.h)
        except Exception:
            pass


class Framework:
    """rg_framework from a workloads.Framework."""

    def __init__(self, ctx, fw):
        self.ctx = ctx
        self.fw = fw
        h = C.c_void_p()
        c9 = L.f64(fw.cell.T.reshape(9))
        pos, typ, q = L.f64(fw.pos), L.i32(fw.type), L.f64(fw.charge)
        L.check(lib().rg_framework_create(ctx.h, L.d(c9), int(fw.rect), len(typ), L.d(pos), L.i(typ), L.d(q),
                                          C.byref(h)), ctx.h, "rg_framework_create")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().rg_framework_destroy(self.h)
        except Exception:
            pass


class Grid:
    """rg_grid: a device-resident tricubic grid (8 values per point, layout_left)."""

    def __init__(self, ctx, cell, rect, n):
        self.ctx = ctx
        self.cell = np.asarray(cell, float)
        self.rect = bool(rect)
        self.n = tuple(int(x) for x in n)
        h = C.c_void_p()
        c9 = L.f64(self.cell.T.reshape(9))
        L.check(lib().rg_grid_create(ctx.h, L.d(c9), int(rect), *self.n, C.byref(h)), ctx.h, "rg_grid_create")
        self.h = h

    @property
    def count(self):
        return 8 * self.n[0] * self.n[1] * self.n[2]

    def download(self, offset=0, count=None):
        count = self.count - offset if count is None else count
        out = np.empty(count)
        L.check(lib().rg_grid_download(self.ctx.h, self.h, offset, count, L.d(out)), self.ctx.h, "download")
        return out

    def upload(self, data, offset=0):
        data = L.f64(data).reshape(-1)
        L.check(lib().rg_grid_upload(self.ctx.h, self.h, offset, data.size, L.d(data)), self.ctx.h, "upload")

    def interpolate(self, pos, gradient=False):
        pos = L.f64(pos).reshape(-1, 3)
        e = np.empty(len(pos))
        g = np.empty((len(pos), 3)) if gradient else None
        L.check(lib().rg_grid_interpolate(self.ctx.h, self.h, len(pos), L.d(pos), L.d(e), L.d(g)), self.ctx.h,
                "rg_grid_interpolate")
        return (e, g) if gradient else e

    def __del__(self):
        try:
            if self.h:
                lib().rg_grid_destroy(self.h)
        except Exception:
            pass


def build_grids(ctx, ff, fw, specs, sync=True):
    """Builds several equal-dimension grids in one fused pass.  specs: list of (Grid, type, probe, cutoff)."""
    n = len(specs)
    types = L.i32([s[1] for s in specs])
    probes = L.i32([s[2] for s in specs])
    cuts = L.f64([s[3] for s in specs])
    arr = (C.c_void_p * n)(*[s[0].h.value for s in specs])
    fn = lib().rg_grid_build_many if sync else lib().rg_grid_build_many_async
    L.check(fn(ctx.h, ff.h, fw.h, n, L.i(types), L.i(probes), L.d(cuts), arr), ctx.h, "rg_grid_build_many")
