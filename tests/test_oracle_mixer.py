"""Mixer extensions in the oracle (oracle_solvers.c:orc_anderson): the
Laplacian preconditioner of the mixing residual and step backtracking.

Neither exists in the reference (adm.c mixes the plain residual); both are
options that default to off.  They are pinned here by closed forms (the
stencil), by an exact identity (the halving), and by convergence checked
with the oracle's own evaluator.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from tests.helpers import workloads

WL = workloads()


def _stencil(x, bc):
    M = len(x)
    n = M - 1 if bc == O.BC_DN else M - 2
    st = np.zeros(3 * n)
    O.lib().orc_precond_stencil(np.ascontiguousarray(x, dtype=np.float64), M, bc, st)
    return st.reshape(n, 3)


def _apply(st, f):
    fl = np.concatenate([[0.0], f[:-1]])
    fr = np.concatenate([f[1:], [0.0]])
    return st[:, 1] * f - st[:, 0] * fl - st[:, 2] * fr


@pytest.mark.parametrize("bc", [O.BC_DN, O.BC_DD])
@pytest.mark.parametrize("graded", [False, True])
def test_precond_stencil_closed_forms(bc, graded):
    x = WL.graded_mesh(3.0, 2e-3, 0.2) if graded else WL.uniform_mesh(97, 3.0)
    st = _stencil(x, bc)
    n = st.shape[0]
    xu = x[1:1 + n]
    # constants are in the kernel (natural ends everywhere)
    assert np.abs(_apply(st, np.ones(n))).max() <= 1e-9 * np.abs(st).max()
    # interior rows: -f'' of x^2 is exactly -2 on any mesh with lumped mass
    lap = _apply(st, xu * xu)
    assert np.abs(lap[1:-1] + 2.0).max() <= 1e-9 * max(1.0, (xu * xu).max() * np.abs(st).max())
    # M_L K-form is symmetric: m_u r_u = m_{u+1} l_{u+1} = 1 / h
    h = np.diff(xu)
    m = np.empty(n)
    m[0], m[-1] = 0.5 * h[0], 0.5 * h[-1]
    m[1:-1] = 0.5 * (h[:-1] + h[1:])
    assert np.allclose(m[:-1] * st[:-1, 2], 1.0 / h, rtol=1e-12)
    assert np.allclose(m[1:] * st[1:, 0], 1.0 / h, rtol=1e-12)
    assert st[0, 0] == 0.0 and st[-1, 2] == 0.0


def _opts(**kw):
    base = dict(history=10, beta=1.0, tol=1e-6, max_iter=200, step_cap=0.0,
                transform=O.MIX_LOG, floor=1e-8, precond=0.5)
    base.update(kw)
    return O.anderson_opts(**base)


def test_preconditioned_log_mixing_converges_in_few_evaluations():
    """Config-3 instances from m = 32 to m = 4000: the preconditioned mixer
    needs at most a few dozen evaluations where plain mixing needs hundreds
    to thousands; every solution is checked by the evaluator."""
    cfg = WL.config3(n_inst=4096)
    ms = np.array([len(x) - 1 for x in cfg["xs"]])
    pick = []
    for lo, hi in [(30, 60), (200, 400), (1000, 1500), (3000, 4100)]:
        pick.append(next(k for k in range(4096) if lo <= ms[k] < hi and cfg["Ns"][k] == 512))
    out = O.batch_solve([ms[k] + 1 for k in pick], [cfg["Ns"][k] for k in pick],
                        [cfg["xs"][k] for k in pick], [cfg["pts"][k] for k in pick],
                        [cfg["w0"][k] for k in pick], _opts(), bc=O.BC_DN, threads=4)
    assert (out["status"] == 0).all(), out
    assert out["iters"].max() <= 40, out["iters"]
    for i, k in enumerate(pick):
        r = O.evaluate(cfg["xs"][k], out["w"][i], cfg["pts"][k], cfg["Ns"][k], bc=O.BC_DN)
        assert r["rnorm"] <= 1e-6


def test_plain_mixing_unchanged_when_extensions_off():
    """precond = 0 and backtrack = 0 give the reference's plain update."""
    cfg = WL.config3(n_inst=3, seed=7, Ns=(256,))
    k = 0
    a = O.solve_anderson(cfg["xs"][k], cfg["pts"][k], 256, cfg["w0"][k],
                         O.anderson_opts(max_iter=6), bc=O.BC_DN)
    b = O.solve_anderson(cfg["xs"][k], cfg["pts"][k], 256, cfg["w0"][k],
                         O.anderson_opts(max_iter=6, precond=0.0, backtrack=0.0), bc=O.BC_DN)
    assert np.array_equal(a["w"], b["w"])


BT_CASE = dict(n_inst=24, seed=11, Ns=(256,))
BT_INST = 5  # m = 2891, N = 256: ||r|| jumps 0.156 -> 15.8 at evaluation 4


def _bt_run(max_iter, backtrack):
    cfg = WL.config3(**BT_CASE)
    k = BT_INST
    return O.solve_anderson(cfg["xs"][k], cfg["pts"][k], 256, cfg["w0"][k],
                            _opts(max_iter=max_iter, backtrack=backtrack), bc=O.BC_DN,
                            trace_len=8)


def test_backtracking_halves_toward_last_accepted_iterate():
    """Evaluation 4 (at W_3) grows ||r|| 100-fold, past backtrack = 10: the
    mixer returns W_2 + (W_3 - W_2) / 2 with W_2 the last accepted iterate."""
    tr = _bt_run(4, 0.0)["trace"]
    assert tr[3] > 10.0 * tr[2] and tr[2] < tr[1]
    W2, W3 = _bt_run(2, 10.0)["w"], _bt_run(3, 10.0)["w"]
    W4 = _bt_run(4, 10.0)
    assert W4["restarts"] == 1
    assert np.array_equal(W4["w"], W2 + 0.5 * (W3 - W2))
    # without the safeguard the same evaluation is an ordinary update
    assert not np.array_equal(_bt_run(4, 0.0)["w"], W4["w"])
