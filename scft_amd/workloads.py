"""Synthetic SCFT workloads of the benchmark configurations (host side, FP64).

Every instance is a homopolymer melt against a hard wall at x = 0 with a
zero-flux end at x = L (Dirichlet-Neumann), on a linear-FEM mesh, with the
EXP wall target phi_t(x) = tanh^2(2 tau x / (tau^2 - x^2)) for x <= tau and 1
beyond (DEALII_SCFT/src/scft.cc:188-215, no mirror image for one wall).

Noise is generated once on the host from a splitmix64 stream with Box-Muller
pairs, so the CPU oracle and the device path see identical inputs.
"""
from __future__ import annotations

import numpy as np

TAU = 0.530252230020752
MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_uniform(seed: int, count: int) -> np.ndarray:
    """`count` uniform doubles in (0, 1) from splitmix64 started at `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        s = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * np.uint64(0x9E3779B97F4A7C15)
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) / 9007199254740992.0


def gaussian(seed: int, count: int, sigma: float) -> np.ndarray:
    m = count + (count & 1)
    u = splitmix64_uniform(seed, m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return sigma * z[:count]


def exp_target(x: np.ndarray, tau: float = TAU) -> np.ndarray:
    """phi_t on the nodes, one wall at x = 0 (same arithmetic as the reference)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.ones_like(x)
    inside = x <= tau
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        u = 4.0 * tau * x[inside] / (tau * tau - x[inside] * x[inside])
        e = np.exp(u)
        v = (e - 1.0) ** 2 / (e + 1.0) ** 2
    v[np.isnan(v)] = 1.0
    out[inside] = v
    return out


def uniform_mesh(m: int, L: float) -> np.ndarray:
    return np.arange(m + 1) * L / m


def graded_mesh(L: float, h0: float, ell: float, m_min: int = 63, m_max: int = 4095):
    """h(x) = h0 (1 + x/ell): x_i = ell (exp(i g) - 1) with g = ln(1 + L/ell)/m."""
    m = int(np.clip(round(ell * np.log1p(L / ell) / h0), m_min, m_max))
    g = np.log1p(L / ell) / m
    x = ell * np.expm1(g * np.arange(m + 1))
    x[0], x[-1] = 0.0, L
    return x


def w0_field(x: np.ndarray, seed: int, sigma: float = 1e-3) -> np.ndarray:
    """Initial field 0.1 exp(-x/0.2) + N(0, sigma^2) on every node; w(0) = 0."""
    w = 0.1 * np.exp(-np.asarray(x) / 0.2) + gaussian(seed, len(x), sigma)
    w[0] = 0.0
    return w


def config1():
    """m = 32, N = 2048, L = 3.72, w0 = 0 (single instance)."""
    x = uniform_mesh(32, 3.72)
    return dict(xs=[x], Ns=[2048], pts=[exp_target(x)], w0=[np.zeros(33)], name="config1")


def config2(seed: int = 20261202):
    """m = 1024, N = 2048, L = 3.72, w0 = 0.1 exp(-x/0.2) + noise (single instance)."""
    x = uniform_mesh(1024, 3.72)
    return dict(xs=[x], Ns=[2048], pts=[exp_target(x)], w0=[w0_field(x, seed)],
                name="config2")


def config3(n_inst: int = 4096, seed: int = 20261202, Ns=(512, 1024, 2048, 4096)):
    """Batched plain method: m log-uniform in [32, 4096], N uniform over Ns,
    L uniform in [2, 8], w0 as config 2 with seed base + k."""
    rng = np.random.default_rng(seed)
    ms = np.rint(np.exp(rng.uniform(np.log(32), np.log(4096), n_inst))).astype(int)
    Nk = rng.choice(np.asarray(Ns), n_inst)
    Ls = rng.uniform(2.0, 8.0, n_inst)
    xs = [uniform_mesh(int(m), float(L)) for m, L in zip(ms, Ls)]
    return dict(xs=xs, Ns=[int(v) for v in Nk], pts=[exp_target(x) for x in xs],
                w0=[w0_field(x, seed + 1 + k) for k, x in enumerate(xs)], Ls=Ls, ms=ms,
                name=f"config3_{n_inst}")


def config5(n_inst: int = 16384, seed: int = 20261205, Ns=(512, 1024, 2048, 4096, 8192)):
    """Production mix: graded meshes h(x) = h0 (1 + x/ell), h0 log-uniform in
    [1e-4, 1e-2], ell in [0.05, 0.5], 64-4096 nodes, N over powers of two."""
    rng = np.random.default_rng(seed)
    h0 = np.exp(rng.uniform(np.log(1e-4), np.log(1e-2), n_inst))
    ell = rng.uniform(0.05, 0.5, n_inst)
    Ls = rng.uniform(2.0, 8.0, n_inst)
    Nk = rng.choice(np.asarray(Ns), n_inst)
    xs = [graded_mesh(float(L), float(a), float(b)) for L, a, b in zip(Ls, h0, ell)]
    return dict(xs=xs, Ns=[int(v) for v in Nk], pts=[exp_target(x) for x in xs],
                w0=[w0_field(x, seed + 1 + k) for k, x in enumerate(xs)], Ls=Ls,
                name=f"config5_{n_inst}")
