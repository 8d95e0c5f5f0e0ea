"""Shared builders for the parity tests: meshes, targets and fields of the
benchmark configurations, generated identically for the oracle and the
device path (seeded, FP64, on the host)."""
from __future__ import annotations

import json
import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixtures():
    with open(os.path.join(GOLDEN, "reference_fixtures.json")) as f:
        return json.load(f)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """splitmix64 stream -> uniform doubles in (0, 1)."""
    out = np.empty(count)
    s = seed & 0xFFFFFFFFFFFFFFFF
    mask = 0xFFFFFFFFFFFFFFFF
    for i in range(count):
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        out[i] = ((z >> 11) + 0.5) / 9007199254740992.0
    return out


def gaussian_noise(seed: int, count: int, sigma: float) -> np.ndarray:
    """Box-Muller normals from the splitmix64 stream."""
    m = count + (count & 1)
    u = splitmix64(seed, m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return sigma * z[:count]


def w0_field(x, seed, sigma=1e-3):
    """Initial field 0.1 exp(-x/0.2) + N(0, sigma) on every node."""
    return 0.1 * np.exp(-np.asarray(x) / 0.2) + gaussian_noise(seed, len(x), sigma)


def instance(m, L, N, bc=O.BC_DN, graded=None, seed=None, w=None):
    """Mesh, target (on all nodes) and field (on all nodes) of one instance."""
    M = m + 1
    if graded is not None:
        x = O.mesh_graded(M, L, graded)
    else:
        x = O.mesh_uniform(M, L)
    pt = O.target_position(x, O.TAU, bc == O.BC_DD)
    if w is None:
        w = w0_field(x, seed if seed is not None else 1234)
    return dict(x=x, pt=pt, w=np.asarray(w, dtype=np.float64), N=N, bc=bc, M=M)


def unknowns(a, bc):
    return a[1:-1] if bc == O.BC_DD else a[1:]


def workloads():
    """scft_amd/workloads.py loaded on its own (pure numpy), so CPU tests do
    not need the device library that `import scft_amd` loads."""
    import importlib.util
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "scft_amd", "workloads.py")
    spec = importlib.util.spec_from_file_location("scft_workloads", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod
