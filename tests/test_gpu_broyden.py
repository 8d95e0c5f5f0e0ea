"""Broyden with the batched finite-difference Jacobian (scft_broyden_solve)
against the oracle's restatement of broydn.c (itself checked against the
reference's compiled broydn in tests/test_oracle.py)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from tests.helpers import fixtures

pytestmark = pytest.mark.gpu


def _S():
    import scft_amd
    return scft_amd


def _oracle_broydn(x, pt, N, w_full, tolf, bc):
    ctx = O.EvalCtx(x, pt, N, bc=bc)
    n = len(w_full) - 2
    O.global_callbacks(ctx, n, trace_len=0)
    cb = C.cast(O.lib().orc_global_residual_ctx, C.c_void_p)
    xm = w_full[1:-1].copy() if bc == O.BC_DD else w_full[1:].copy()
    n = len(xm)
    jc = C.c_int(0)
    qt, r, d = np.zeros(n * n), np.zeros(n * n), np.zeros(n)
    st = O.BroydnStats()
    chk = O.lib().orc_broydn(xm, n, cb, None, tolf, C.byref(jc), qt, r, d, C.byref(st))
    ctx.close()
    return chk, xm, st


def test_reference_driver_problem_matches_oracle():
    """1D_FEM.c's main: DD, m = 32, N = 2048, TOLF 1e-8 (1D_FEM.c:349), from
    the .res field; both solvers converge to the same field."""
    S = _S()
    m, N = 32, 2048
    x = O.mesh_uniform(m + 1, O.L_REF)
    pt = O.target_position(x, O.TAU, True)
    eta = np.array(fixtures()["res_m32"]["eta"])
    out = S.broyden_solve(x, pt[1:-1], N, eta[1:-1], tolf=1e-8, bc=S.BC_DD)
    chk, w_ref, st = _oracle_broydn(x, pt, N, eta, 1e-8, O.BC_DD)
    print(f"device: rc {out['rc']} its {out['iterations']} evals {out['evaluations']} "
          f"sweeps {out['sweeps']}; oracle: check {chk} its {st.iterations} evals {st.evaluations}")
    assert out["rc"] == 0 and chk == 0
    assert out["final_err"] < 1e-8
    assert np.abs(out["w"] - w_ref).max() <= 1e-6 * np.abs(w_ref).max()
    # one device launch per Jacobian instead of n evaluations in sequence
    assert out["sweeps"] < out["evaluations"] - (m - 1) + 2
    w = np.concatenate([[0.0], out["w"], [0.0]])
    assert O.evaluate(x, w, pt, N, bc=O.BC_DD)["rnorm"] < 1e-8


def test_dn_config1_from_zero_field():
    S = _S()
    from scft_amd import workloads as WL
    cfg = WL.config1()
    x, N, pt = cfg["xs"][0], cfg["Ns"][0], cfg["pts"][0]
    out = S.broyden_solve(x, pt[1:], N, np.zeros(len(x) - 1), tolf=1e-8, bc=S.BC_DN)
    chk, w_ref, st = _oracle_broydn(x, pt, N, np.zeros(len(x)), 1e-8, O.BC_DN)
    print(f"device its {out['iterations']} evals {out['evaluations']}; oracle its {st.iterations}")
    assert out["rc"] == 0 and chk == 0
    assert np.abs(out["w"] - w_ref).max() <= 1e-6 * max(1.0, np.abs(w_ref).max())


def test_batched_jacobian_on_a_larger_instance():
    """m = 256 (n = 256 Jacobian columns per restart), config-2 style start:
    converges, and every Jacobian is one device launch."""
    S = _S()
    from scft_amd import workloads as WL
    x = WL.uniform_mesh(256, 3.72)
    pt = WL.exp_target(x)
    w0 = WL.w0_field(x, 20261202)
    out = S.broyden_solve(x, pt[1:], 1024, w0[1:], tolf=1e-8, bc=S.BC_DN)
    n_jac = (out["evaluations"] - out["sweeps"]) // 255  # each Jacobian: n evals, 1 sweep
    print(f"its {out['iterations']} evals {out['evaluations']} sweeps {out['sweeps']} "
          f"jacobians {n_jac} err {out['final_err']:.2e}")
    assert out["rc"] == 0 and out["final_err"] < 1e-8
    assert n_jac >= 1 and out["evaluations"] - out["sweeps"] == 255 * n_jac
    w = np.concatenate([[0.0], out["w"]])
    assert O.evaluate(x, w, pt, 1024, bc=O.BC_DN)["rnorm"] < 1e-8


def test_consistent_mass_walls_follow_the_field():
    """The consistent w-mass (deal.II, scft.cc:641-651) reads the field at the
    walls; the deal.II evaluator extrapolates them from the interior by a
    natural spline before every assembly (scft.cc:454-480), and so does
    scft_broyden_solve for every field it evaluates, Jacobian columns
    included.  Oracle: the same Broyden iteration (host-callback entry)
    driving the CPU oracle evaluator with the walls extrapolated by scipy's
    natural spline, from the deal.II fixture's field (DD, 33 nodes)."""
    from scipy.interpolate import CubicSpline
    S = _S()
    fx = fixtures()["dealii_N33"]
    x, w_full = np.array(fx["x"]), np.array(fx["w"])
    N = 2048
    pt = O.target_position(x, O.TAU, True)

    def full(wu):
        cs = CubicSpline(x[1:-1], wu, bc_type="natural")
        return np.concatenate([[float(cs(x[0]))], wu, [float(cs(x[-1]))]])

    def F(n, w, f, ctx):
        wu = np.ctypeslib.as_array(w, shape=(n,)).copy()
        r = O.evaluate(x, full(wu), pt, N, bc=O.BC_DD, cmat=O.C_CONSISTENT)["r"][1:-1]
        for i in range(n):
            f[i] = r[i]
        return 0

    cb = S.BROYDEN_FN(F)
    ref = S.broyden_solve_fn(cb, w_full[1:-1], tolf=1e-8)
    out = S.broyden_solve(x, pt[1:-1], N, w_full[1:-1], tolf=1e-8, bc=S.BC_DD,
                          cmat=S.C_CONSISTENT)
    print(f"device its {out['iterations']} evals {out['evaluations']}; oracle its "
          f"{ref['iterations']} evals {ref['evaluations']}", flush=True)
    assert out["rc"] == 0 and ref["rc"] == 0
    assert np.abs(out["w"] - ref["w"]).max() <= 1e-6 * np.abs(ref["w"]).max()
    r = O.evaluate(x, full(out["w"]), pt, N, bc=O.BC_DD, cmat=O.C_CONSISTENT)["rnorm"]
    assert r < 1.01e-8
    # with frozen zero walls the same start converges elsewhere: the walls matter
    r0 = O.evaluate(x, np.concatenate([[0.0], out["w"], [0.0]]), pt, N, bc=O.BC_DD,
                    cmat=O.C_CONSISTENT)["rnorm"]
    assert r0 > 1e-7
