"""Pins the oracle's proposal network (oracle/mlp.py).

* SplitMix64 against its published first outputs for seed 0.
* BF16 rounding against torch's float32 -> bfloat16 cast, exact ties included.
* FP64 forward, loss and every gradient against torch autograd in FP64 (an independent implementation of the same
  network), and against central finite differences of the loss.  A dropped SiLU' factor, a transposed weight or a
  sign error in q - p fails all three.
* The BF16-emulating mode stays close to FP64 (it is the device's arithmetic, not a different network).
"""
import numpy as np
import pytest
import torch

from oracle import mlp

RNG = np.random.default_rng(11)


def test_splitmix64_known_outputs():
    assert [int(x) for x in mlp.splitmix64(0, 3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_bf16_rounding_matches_torch():
    x = RNG.normal(size=20000).astype(np.float32) * np.float32(10.0) ** RNG.integers(-8, 8, 20000)
    # exact ties: low 16 bits 0x8000 with even and odd bit 16
    u = (RNG.integers(0, 1 << 6, 64, dtype=np.uint32) << np.uint32(17)) | np.uint32(0x3F800000)
    ties = np.concatenate([u | np.uint32(0x8000), u | np.uint32(0x18000)]).view(np.float32)
    assert np.isfinite(ties).all()
    x = np.concatenate([x, ties])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(x), ref)


def test_init_params_he_uniform():
    w = mlp.init_params(7)
    P = mlp.unpack(w)
    for n, fan in (("W1", mlp.NIN), ("W2", 256), ("W3", 256), ("w4", 256)):
        a = np.sqrt(6.0 / fan)
        sd, c = a / np.sqrt(3.0), P[n].size
        assert np.abs(P[n]).max() <= a and abs(P[n].mean()) < 4 * sd / np.sqrt(c)
        assert abs(P[n].std() / sd - 1.0) < 2.0 / np.sqrt(c)
    for n in ("b1", "b2", "b3", "b4"):
        assert not P[n].any()
    assert np.array_equal(w, mlp.init_params(7)) and not np.array_equal(w, mlp.init_params(8))


def random_problem(n=64, seed=3):
    rng = np.random.default_rng(seed)
    w = mlp.init_params(seed).astype(np.float64)
    P = mlp.unpack(w)
    for b in ("b1", "b2", "b3", "b4"):
        P[b][:] = rng.normal(scale=0.1, size=P[b].shape)
    w = mlp.pack(P)
    frac = rng.random((n, 3))
    energy = rng.normal(scale=400.0, size=n)
    return w, frac, energy, 1.0 / 300.0


def torch_loss_and_grad(w, frac, energy, beta):
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in mlp.unpack(w).items()}
    s = torch.tensor(frac, dtype=torch.float64)
    m = torch.arange(1, mlp.NFREQ + 1, dtype=torch.float64)
    ph = 2 * torch.pi * s[:, :, None] * m
    x = torch.stack([torch.sin(ph), torch.cos(ph)], dim=-1).reshape(len(frac), -1)
    h = x
    for l in (1, 2, 3):
        h = torch.nn.functional.silu(h @ P[f"W{l}"] + P[f"b{l}"])
    f = h @ P["w4"] + P["b4"]
    p = torch.softmax(torch.tensor(-beta * energy, dtype=torch.float64), 0)
    loss = -(p * torch.log_softmax(f, 0)).sum()
    loss.backward()
    return loss.item(), mlp.pack({k: v.grad.numpy() for k, v in P.items()}), f.detach().numpy()


def test_fp64_matches_torch_autograd():
    w, frac, energy, beta = random_problem()
    loss, g = mlp.loss_and_grad(w, frac, energy, beta)
    f, _ = mlp.forward(w, frac)
    tl, tg, tf = torch_loss_and_grad(w, frac, energy, beta)
    assert np.allclose(f, tf, rtol=1e-12, atol=1e-12)
    assert abs(loss - tl) <= 1e-12 * abs(tl)
    for n, (o, s) in mlp.OFFSETS.items():
        c = int(np.prod(s))
        a, b = g[o:o + c], tg[o:o + c]
        assert np.linalg.norm(a - b) <= 1e-11 * np.linalg.norm(b) + 1e-14, n  # d loss / d b4 = sum(q - p) = 0


def test_fp64_gradient_finite_differences():
    w, frac, energy, beta = random_problem(seed=5)
    _, g = mlp.loss_and_grad(w, frac, energy, beta)
    rng = np.random.default_rng(9)
    h = 1e-6
    for n, (o, s) in mlp.OFFSETS.items():
        c = int(np.prod(s))
        idx = o + rng.choice(c, size=min(c, 6), replace=False)
        for k in idx:
            wp, wm = w.copy(), w.copy()
            wp[k] += h
            wm[k] -= h
            fd = (mlp.loss_and_grad(wp, frac, energy, beta)[0] - mlp.loss_and_grad(wm, frac, energy, beta)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(g[k])) + 1e-9, (n, k, fd, g[k])


def test_loss_minimum_is_boltzmann():
    """At f = -beta U + c the minibatch gradient vanishes and the loss equals the entropy of p."""
    w, frac, energy, beta = random_problem(seed=6)
    P = mlp.unpack(w)
    f, _ = mlp.forward(w, frac)
    # choose energies that the current network reproduces exactly
    energy = -(f - 3.0) / beta
    loss, g = mlp.loss_and_grad(w, frac, energy, beta)
    p = np.exp(f - f.max())
    p /= p.sum()
    assert abs(loss + (p * np.log(p)).sum()) < 1e-10
    assert np.abs(g).max() < 1e-12
    assert P  # unpack views are valid


def test_emulated_bf16_close_to_fp64():
    w, frac, energy, beta = random_problem(seed=4)
    f64, _ = mlp.forward(w, frac)
    fem, _ = mlp.forward(w, frac, emulate=True)
    assert np.abs(fem - f64).max() < 0.05 * np.abs(f64).max()
    _, g64 = mlp.loss_and_grad(w, frac, energy, beta)
    _, gem = mlp.loss_and_grad(w, frac, energy, beta, emulate=True)
    cos = g64 @ gem / (np.linalg.norm(g64) * np.linalg.norm(gem))
    assert cos > 0.99


def test_proposal_normalised_with_floor():
    w = mlp.init_params(2)
    n3 = (6, 5, 4)
    p = mlp.proposal(w, n3, 0.1)
    assert p.shape == (120,) and abs(p.sum() - 1.0) < 1e-12 and p.min() >= 0.1 / 120 - 1e-15
