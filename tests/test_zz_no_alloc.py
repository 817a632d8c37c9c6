"""rg_trial_energies_noalloc (device staging in the library's static array, no allocation) against
rg_trial_energies (per-call allocation) through the C ABI: the same bits on two calls in a row, the second with fewer
sites, and the size limit rejected on the host."""
import ctypes as C

import numpy as np
import pytest

import raspa3_amd as ra
from raspa3_amd import _lib as L
from raspa3_amd import workloads as w

pytestmark = pytest.mark.gpu

MAX_SITES = 16384  # RG_TRIAL_NOALLOC_MAX_SITES in include/raspa3_amd/rg.h


def call(name, ctx, sysm, pos, t, q, m, skip):
    n = len(t)
    e = np.full((n, 4), np.nan)
    ov = np.full(n, -7, np.int32)
    st = getattr(ra.lib(), name)(ctx.h, sysm.h, n, L.d(pos), L.i(t), L.d(q), L.i(m), skip, L.d(e), L.i(ov))
    return st, e, ov


def sites(rng, spec, cell, n):
    frac = rng.random((n, 3))
    pos = L.f64(frac @ cell.T)
    t = L.i32(rng.choice([spec.index("C_co2"), spec.index("O_co2")], n))
    q = L.f64(np.where(t == spec.index("C_co2"), 0.6512, -0.3256))
    m = L.i32(rng.integers(0, 6, n))
    return pos, t, q, m


def test_noalloc_trial_energies_equal_allocating_entry_point(gpu_ctx):
    spec = w.forcefield_zeolite()
    fw, unit = w.framework_itq29(spec, 3)
    ff = ra.ForceField(gpu_ctx, spec)
    f = ra.Framework(gpu_ctx, fw)
    sysm = ra.System(gpu_ctx, ff, fw.cell, True, framework=f, kmax=(0, 0, 0), capacity=64)
    rng = np.random.default_rng(2024)
    # adsorbates of five CO2-like molecules, so that the inter-molecular sums and the molecule exclusions take part
    apos, at, aq, am = sites(rng, spec, fw.cell, 15)
    sysm.set_adsorbates(apos, at, aq, np.repeat(np.arange(5), 3).astype(np.int32))
    for n, skip in ((3000, 2), (777, -1)):
        pos, t, q, m = sites(rng, spec, fw.cell, n)
        st0, e0, ov0 = call("rg_trial_energies", gpu_ctx, sysm, pos, t, q, m, skip)
        st1, e1, ov1 = call("rg_trial_energies_noalloc", gpu_ctx, sysm, pos, t, q, m, skip)
        assert st0 == L.RG_OK and st1 == L.RG_OK, gpu_ctx.error()
        assert np.array_equal(ov0, ov1)
        assert np.array_equal(e0.view(np.uint64), e1.view(np.uint64))
        assert ov0.any() and not ov0.all() and np.isfinite(e0[~ov0.astype(bool)]).all()
    # beyond the static array: rejected before any device work, and the entry point stays usable afterwards
    pos, t, q, m = sites(rng, spec, fw.cell, MAX_SITES + 1)
    st, _, _ = call("rg_trial_energies_noalloc", gpu_ctx, sysm, pos, t, q, m, -1)
    assert st == L.RG_ERR_INVALID and "exceed" in gpu_ctx.error()
    st, e, ov = call("rg_trial_energies_noalloc", gpu_ctx, sysm, pos[:30], t[:10], q[:10], m[:10], -1)
    st0, e0, ov0 = call("rg_trial_energies", gpu_ctx, sysm, pos[:30], t[:10], q[:10], m[:10], -1)
    assert st == L.RG_OK and np.array_equal(e.view(np.uint64), e0.view(np.uint64)) and np.array_equal(ov, ov0)
    sysm.close()
