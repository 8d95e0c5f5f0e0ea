"""The Broyden iteration of libscft_hip.so (scft_amd/csrc/broyden.cpp,
written from Dennis & Schnabel's secant method with the constants of
broydn.c) against the oracle's restatement of the reference's broydn
(oracle_solvers.c:orc_broydn, itself checked against the compiled reference
in tests/test_oracle.py), both driving the same CPU oracle evaluator through
the host-callback entry scft_broyden_solve_fn.  No device work: the callback
is the oracle.

Band: the two implementations take the same Newton / secant steps while the
line search accepts full steps, and may part where it backtracks or stalls
(the published search tests sufficient decrease before the step floor, the
reference after).  Required: the same root to 1e-6 relative, iterations within
3 of the reference's, evaluations between 0.5x and 1.25x of its count.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from tests.helpers import fixtures, workloads

WL = workloads()


def _S():
    import scft_amd
    return scft_amd


def _unknowns(w_full, bc):
    return (w_full[1:-1] if bc == O.BC_DD else w_full[1:]).copy()


def _oracle_broydn(x, pt, N, w_full, tolf, bc):
    ctx = O.EvalCtx(x, pt, N, bc=bc)
    xm = _unknowns(w_full, bc)
    n = len(xm)
    O.global_callbacks(ctx, n, trace_len=0)
    cb = C.cast(O.lib().orc_global_residual_ctx, C.c_void_p)
    jc = C.c_int(0)
    qt, r, d = np.zeros(n * n), np.zeros(n * n), np.zeros(n)
    st = O.BroydnStats()
    chk = O.lib().orc_broydn(xm, n, cb, None, tolf, C.byref(jc), qt, r, d, C.byref(st))
    ctx.close()
    return chk, xm, st


def _library_broyden(x, pt, N, w_full, tolf, bc):
    S = _S()
    ctx = O.EvalCtx(x, pt, N, bc=bc)
    xm = _unknowns(w_full, bc)
    O.global_callbacks(ctx, len(xm), trace_len=0)
    out = S.broyden_solve_fn(O.lib().orc_global_residual_ctx, xm, tolf=tolf)
    ctx.close()
    return out


def _cases():
    eta = np.array(fixtures()["res_m32"]["eta"])
    x32 = O.mesh_uniform(33, O.L_REF)
    c1 = WL.config1()
    x256 = WL.uniform_mesh(256, 3.72)
    return [
        # 1D_FEM.c's own driver problem (1D_FEM.c:346-356): DD, m = 32, TOLF 1e-8
        ("driver_m32_DD", x32, O.target_position(x32, O.TAU, True), 2048, eta, O.BC_DD),
        ("config1_DN_zero", c1["xs"][0], c1["pts"][0], 2048, np.zeros(33), O.BC_DN),
        ("m256_DN_w0", x256, WL.exp_target(x256), 1024, WL.w0_field(x256, 20261202), O.BC_DN),
    ]


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_broyden_matches_reference_iteration(case):
    name, x, pt, N, w_full, bc = case
    chk, w_ref, st = _oracle_broydn(x, pt, N, w_full, 1e-8, bc)
    out = _library_broyden(x, pt, N, w_full, 1e-8, bc)
    print(f"{name}: library rc {out['rc']} its {out['iterations']} evals {out['evaluations']}; "
          f"reference iteration check {chk} its {st.iterations} evals {st.evaluations}")
    assert chk == 0 and out["rc"] == 0 and out["check"] == 0
    assert out["final_err"] < 1e-8
    assert np.abs(out["w"] - w_ref).max() <= 1e-6 * max(1.0, np.abs(w_ref).max())
    assert abs(out["iterations"] - st.iterations) <= 3
    assert 0.5 * st.evaluations <= out["evaluations"] <= 1.25 * st.evaluations
    # every solution solves the evaluator
    wf = w_full.copy()
    if bc == O.BC_DD:
        wf[1:-1] = out["w"]
    else:
        wf[1:] = out["w"]
    assert O.evaluate(x, wf, pt, N, bc=bc)["rnorm"] < 1e-8


def _solve_fn(F, w0, tolf=1e-12):
    S = _S()
    calls = [0]

    def cb(n, w, f, ctx):
        calls[0] += 1
        wv = np.ctypeslib.as_array(w, shape=(n,)).copy()
        fv = F(wv)
        if not np.all(np.isfinite(fv)):
            return 1
        for i in range(n):
            f[i] = fv[i]
        return 0

    fn = S.BROYDEN_FN(cb)
    out = S.broyden_solve_fn(fn, np.asarray(w0, dtype=np.float64), tolf=tolf)
    out["calls"] = calls[0]
    return out


def test_linear_system_is_one_newton_step():
    """F = A w - b: the difference Jacobian is exact up to rounding, so the
    first full step lands on the root: 1 + n + 1 evaluations."""
    n = 7
    rng = np.random.default_rng(4)
    A = np.eye(n) * 3.0 + rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    out = _solve_fn(lambda w: A @ w - b, np.zeros(n), tolf=1e-9)
    assert out["rc"] == 0
    np.testing.assert_allclose(out["w"], np.linalg.solve(A, b), rtol=0, atol=1e-8)
    assert out["evaluations"] == out["calls"] == n + 2
    assert out["iterations"] == 1


def test_secant_updates_reach_a_nonlinear_root():
    """Broyden's test problem (Dennis & Schnabel, sec. 8.2): a mildly
    nonlinear system whose root is known; after the first Jacobian only
    secant updates (no further n-evaluation Jacobians) are needed."""
    def F(w):
        return np.array([w[0] ** 2 + w[1] ** 2 - 2.0, np.exp(w[0] - 1.0) + w[1] ** 3 - 2.0])
    out = _solve_fn(F, [2.0, 0.5], tolf=1e-12)
    assert out["rc"] == 0
    np.testing.assert_allclose(out["w"], [1.0, 1.0], rtol=0, atol=1e-10)
    assert out["evaluations"] - 2 < 2 * out["iterations"] + 2  # one Jacobian only


def test_failed_evaluation_is_reported():
    out = _solve_fn(lambda w: np.array([np.nan, 1.0]), [0.0, 0.0])
    assert out["rc"] == 3  # SCFT_E_NAN
