"""Learned first-bead proposal network, restated in numpy.  TEST INFRASTRUCTURE: the checker for the device's
BF16-MFMA training and inference kernels; nothing in the product imports it.

The reference has no learned proposal; the network hooks into the first-bead trials that the reference draws
uniformly (src/raspakit/cbmc/multiple_first_bead.cpp:26-27 grow, :56-75 retrace).  Definition:

    x(s)  = [sin 2 pi m s_a, cos 2 pi m s_a], input q = 64 a + 2 (m - 1) + c, a = 0..2, m = 1..32, c = 0 sin / 1 cos
    h1 = SiLU(x W1 + b1), h2 = SiLU(h1 W2 + b2), h3 = SiLU(h2 W3 + b3), f = h3 . w4 + b4
    p = softmax(-beta U), q = softmax(f) over a minibatch, loss = -sum_i p_i log q_i,  dloss/df_i = q_i - p_i

Two modes:
* ``emulate=False``: everything in FP64.  This mode is pinned by (1) torch autograd in FP64 on the CPU, an
  independent implementation of the same network and loss, and (2) central finite differences of the loss.
* ``emulate=True``: the device's arithmetic: features and weight matrices rounded to BF16 (round to nearest even,
  via FP32), products accumulated exactly (the device accumulates in FP32), pre-activations in FP32, activations
  and back-propagated pre-activation gradients rounded to BF16, biases and the output weight bias in FP32.
  The rounding helper is pinned against torch's float32 -> bfloat16 cast.

Parameter layout (one FP32 array): W1 (192 x 256), b1, W2 (256 x 256), b2, W3 (256 x 256), b3, w4 (256), b4;
W[in][out] row-major.  Initialisation: He-uniform from a splitmix64 stream, zero biases.
"""
from __future__ import annotations

import numpy as np

NFREQ = 32
NIN = 6 * NFREQ
WIDTH = 256
SHAPES = [("W1", (NIN, WIDTH)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)), ("W3", (WIDTH, WIDTH)),
          ("b3", (WIDTH,)), ("w4", (WIDTH,)), ("b4", (1,))]
OFFSETS = {}
_o = 0
for _n, _s in SHAPES:
    OFFSETS[_n] = (_o, _s)
    _o += int(np.prod(_s))
NPARAM = _o


def unpack(w):
    return {n: np.asarray(w[o:o + int(np.prod(s))]).reshape(s) for n, (o, s) in OFFSETS.items()}


def pack(d):
    out = np.zeros(NPARAM)
    for n, (o, s) in OFFSETS.items():
        out[o:o + int(np.prod(s))] = np.asarray(d[n]).reshape(-1)
    return out


def splitmix64(seed, count):
    """Steele, Lea & Flood's SplitMix64 outputs (uint64) of a generator seeded with 'seed'."""
    s = np.uint64(seed)
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = s + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _uniform(seed, count):
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def init_params(seed):
    """He-uniform weights a (2u - 1), a = sqrt(6 / fan_in), from one splitmix64 stream over W1, W2, W3, w4."""
    sizes = [("W1", NIN), ("W2", WIDTH), ("W3", WIDTH), ("w4", WIDTH)]
    total = sum(int(np.prod(OFFSETS[n][1])) for n, _ in sizes)
    u = _uniform(seed, total)
    w = np.zeros(NPARAM, np.float32)
    k = 0
    for n, fan in sizes:
        o, s = OFFSETS[n]
        c = int(np.prod(s))
        w[o:o + c] = (np.sqrt(6.0 / fan) * (2.0 * u[k:k + c] - 1.0)).astype(np.float32)
        k += c
    return w


def bf16(x):
    """Round to the nearest BF16 value (ties to even) through FP32; returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def features(frac, emulate=False):
    frac = np.asarray(frac, float).reshape(-1, 3)
    m = np.arange(1, NFREQ + 1)
    ph = 2.0 * np.pi * frac[:, :, None] * m[None, None, :]          # (n, 3, F)
    x = np.stack([np.sin(ph), np.cos(ph)], axis=-1).reshape(len(frac), NIN)
    return bf16(x) if emulate else x


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(w, frac, emulate=False):
    P = unpack(np.asarray(w, np.float64))
    rw = bf16 if emulate else (lambda a: a)
    ra = f32 if emulate else (lambda a: a)
    x = features(frac, emulate)
    cache = {"X": x}
    h = x
    for l in (1, 2, 3):
        z = ra(h @ rw(P[f"W{l}"]) + P[f"b{l}"])
        h = rw(z * _sig(z))
        cache[f"Z{l}"], cache[f"H{l}"] = z, h
    f = ra(h @ rw(P["w4"]) + P["b4"][0])
    cache["f"] = f
    return f, cache


def loss_and_grad(w, frac, energy, beta, emulate=False):
    """Minibatch loss -sum_i p_i log softmax(f)_i and its gradient with respect to every parameter."""
    P = unpack(np.asarray(w, np.float64))
    rw = bf16 if emulate else (lambda a: a)
    ra = f32 if emulate else (lambda a: a)
    f, c = forward(w, frac, emulate)
    a = -beta * np.asarray(energy, float)
    p = np.exp(a - a.max())
    p /= p.sum()
    lse = f.max() + np.log(np.exp(f - f.max()).sum())
    q = np.exp(f - lse)
    loss = -(p * (f - lse)).sum()
    df = ra(q - p)
    G = {"w4": c["H3"].T @ df, "b4": np.array([df.sum()])}
    dsilu = lambda z: _sig(z) * (1.0 + z * (1.0 - _sig(z)))
    dz = rw(df[:, None] * rw(P["w4"])[None, :] * dsilu(c["Z3"]))
    for l in (3, 2, 1):
        hin = c[f"H{l - 1}"] if l > 1 else c["X"]
        G[f"W{l}"] = hin.T @ dz
        G[f"b{l}"] = dz.sum(axis=0)
        if l > 1:
            dz = rw((dz @ rw(P[f"W{l}"]).T) * dsilu(c[f"Z{l - 1}"]))
    return loss, pack(G)


def proposal(w, n3, floor, emulate=False):
    """Voxel probabilities (1 - floor) softmax(f(centres)) + floor / nvox, x fastest."""
    nx, ny, nz = n3
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    frac = np.stack([(x + 0.5) / nx, (y + 0.5) / ny, (z + 0.5) / nz], axis=-1).reshape(-1, 3)
    f, _ = forward(w, frac, emulate)
    e = np.exp(f - f.max())
    return (1.0 - floor) * e / e.sum() + floor / len(f)
