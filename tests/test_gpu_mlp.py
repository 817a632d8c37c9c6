"""Learned proposal network on the GPU (BF16 MFMA kernels through the C ABI) against the oracle's BF16-emulating
restatement (oracle/mlp.py).

Tolerances: the device accumulates the BF16 products in FP32 in another order than the oracle's exact sums, and
uses the hardware exponential in SiLU, so a few activations round to the neighbouring BF16 value (2^-8 relative).
Forward values agree to 5e-3 absolute; each gradient block to 2 % in L2 norm; the loss to 1e-4 relative.  The
FP64 network is further away (it is not the device's arithmetic), which the tests also check, so that the
comparison discriminates.
"""
import numpy as np
import pytest

import raspa3_amd as ra
from oracle import mlp as om

pytestmark = pytest.mark.gpu


def problem(n, seed):
    rng = np.random.default_rng(seed)
    w = om.init_params(seed).astype(np.float64)
    P = om.unpack(w)
    for b in ("b1", "b2", "b3", "b4"):
        P[b][:] = rng.normal(scale=0.1, size=P[b].shape)
    w = om.pack(P).astype(np.float32)
    frac = rng.random((n, 3))
    energy = rng.normal(scale=500.0, size=n)
    return w, frac, energy, 1.0 / 300.0


def test_mlp_init_matches_oracle(gpu_ctx):
    print("case: mlp init", flush=True)
    net = ra.Mlp(gpu_ctx, seed=12345)
    assert net.n_params == om.NPARAM
    assert np.array_equal(net.params(), om.init_params(12345))
    net.close()


def test_mlp_forward_matches_oracle(gpu_ctx):
    print("case: mlp forward", flush=True)
    w, frac, _, _ = problem(300, 1)  # 300 points: a ragged last row tile
    net = ra.Mlp(gpu_ctx)
    net.set_params(w)
    assert np.array_equal(net.params(), w)
    f = net.forward(frac)
    fe, _ = om.forward(w, frac, emulate=True)
    f64, _ = om.forward(w, frac)
    err = np.abs(f - fe).max()
    assert err < 5e-3, err
    assert np.abs(f64 - fe).max() > err
    net.close()


def test_mlp_gradients_match_oracle(gpu_ctx):
    print("case: mlp gradients", flush=True)
    for n, seed in ((256, 2), (1024, 3)):
        w, frac, energy, beta = problem(n, seed)
        net = ra.Mlp(gpu_ctx)
        net.set_params(w)
        g, loss = net.gradients(frac, energy, beta)
        le, ge = om.loss_and_grad(w, frac, energy, beta, emulate=True)
        _, g64 = om.loss_and_grad(w, frac, energy, beta)
        assert abs(loss - le) <= 1e-4 * abs(le), (loss, le)
        for name, (o, s) in om.OFFSETS.items():
            if name == "b4":  # sum(q - p) = 0 up to rounding
                assert abs(g[o]) < 1e-5
                continue
            c = int(np.prod(s))
            a, b = g[o:o + c], ge[o:o + c]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            rel64 = np.linalg.norm(g64[o:o + c] - b) / np.linalg.norm(b)
            assert rel < 0.02, (n, name, rel)
            assert rel < rel64 or rel < 2e-3, (n, name, rel, rel64)
        net.close()


def test_mlp_gradients_with_infinite_energies(gpu_ctx):
    """Rows of infinite energy (overlaps) weigh nothing: a whole workgroup's 16 rows and one single row at +inf give
    finite gradients and loss equal to the oracle's."""
    print("case: mlp gradients, infinite energies", flush=True)
    w, frac, energy, beta = problem(256, 5)
    energy = energy.copy()
    energy[32:48] = np.inf
    energy[100] = np.inf
    net = ra.Mlp(gpu_ctx)
    net.set_params(w)
    g, loss = net.gradients(frac, energy, beta)
    le, ge = om.loss_and_grad(w, frac, energy, beta, emulate=True)
    assert np.isfinite(loss) and np.all(np.isfinite(g))
    assert abs(loss - le) <= 1e-4 * abs(le), (loss, le)
    for name, (o, s) in om.OFFSETS.items():
        if name == "b4":
            continue
        c = int(np.prod(s))
        a, b = g[o:o + c], ge[o:o + c]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 0.02, name
    net.close()


def test_mlp_proposal_matches_oracle(gpu_ctx):
    print("case: mlp proposal", flush=True)
    w, _, _, _ = problem(64, 4)
    net = ra.Mlp(gpu_ctx)
    net.set_params(w)
    n3 = (10, 9, 7)
    p = net.proposal(n3, 0.05)
    pe = om.proposal(w, n3, 0.05, emulate=True)
    assert abs(p.sum() - 1.0) < 1e-12 and p.min() >= 0.05 / p.size * (1 - 1e-12)
    assert np.abs(p / pe - 1.0).max() < 1e-2
    net.close()


def test_mlp_training_learns_boltzmann_density(gpu_ctx):
    """Adam on the device fits f to -beta U + const for a smooth periodic energy landscape."""
    print("case: mlp training", flush=True)
    rng = np.random.default_rng(8)
    beta = 1.0 / 300.0

    def energy(s):
        c = np.cos(2 * np.pi * s)
        return -300.0 * (1.5 * c[:, 0] + c[:, 1] * c[:, 2] + 0.5 * np.cos(4 * np.pi * s[:, 0]) * c[:, 2])

    train = rng.random((1 << 17, 3))
    net = ra.Mlp(gpu_ctx, seed=3)
    losses = net.train(train, energy(train), beta, batch=4096, epochs=12, lr=1e-3, seed=5)
    assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]
    test = rng.random((4096, 3))
    f = net.forward(test)
    target = -beta * energy(test)
    r = np.corrcoef(f, target)[0, 1]
    assert r > 0.98, (r, losses)
    # the minibatch loss approaches its minimum, the entropy of the Boltzmann weights of a batch
    a = target[:4096]
    p = np.exp(a - a.max())
    p /= p.sum()
    assert losses[-1] < -(p * np.log(p)).sum() + 0.05, losses
    net.close()
