"""CPU tests of the oracle (the checker), pinned against the reference's own
data files, its compiled Numerical-Recipes routines (oracle/_ref, built from
/root/reference; the committed fixtures carry their outputs where the tree is
absent) and closed forms that a plausible slip (dropped term, wrong sign or
index, transposed operand) would violate."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from tests.helpers import fixtures

FX = fixtures()
REF = O.ref.available()


# ---------------------------------------------------------------- target (a1)
@pytest.mark.parametrize("key,m,tol", [("res_m32", 32, 1e-10), ("res_m1024", 1024, 3e-12)])
def test_target_matches_res_phi(key, m, tol):
    """.res `phi` = phi_t on x_i = i L/m (Exp_m*_n2048_IE.res columns 2)."""
    phi = np.array(FX[key]["phi"])
    x = np.arange(m + 1) * O.L_REF / m
    pt = O.target_position(x, O.TAU, True)
    assert np.abs(pt - phi).max() < tol


def test_target_nan_guard_and_mirror():
    x = O.mesh_uniform(33, O.L_REF)
    pt = O.target_position(x, O.TAU, True)
    assert pt[0] == 0.0 and pt[-1] == 0.0
    assert np.array_equal(pt, pt[::-1])
    assert np.all(pt[(x > O.TAU) & (x < O.L_REF - O.TAU)] == 1.0)
    # at x = tau the profile's exponent overflows; the reference maps NaN -> 1
    assert O.exp_profile(O.TAU, O.TAU) == 1.0
    # count-based variant (1D_FEM.c:312): ceil(33*0.5302/3.72374) = 5 wall nodes
    pc = O.target_count(x, O.TAU_SHORT, O.L_SHORT, nan_guard=False)
    assert np.sum(pc < 1.0) == 10


def test_f0bar_known_answer():
    """romint(phi_t on 2^16+1 points, L/2^16)/L = 0.892581217773656
    (DEALII_SCFT/src/scft.cc:368, testFiBar.cc:50)."""
    n = 2 ** 16
    x = np.arange(n + 1) * O.L_REF / n
    pt = O.target_position(x, O.TAU, True)
    val = O.romint(pt, n, O.L_REF / n) / O.L_REF
    assert abs(val - FX["constants"]["f0bar"]) < 1e-14


# ---------------------------------------------------------------- Romberg (a6)
def test_romint_matches_reference_vectors():
    for v in FX["ref_nr"]["romint"]:
        f = np.array(v["f"])
        got = O.romint(f, v["m"], 1.0 / v["m"])
        assert abs(got - v["value"]) <= 1e-15 * max(1.0, abs(v["value"])), v["m"]


@pytest.mark.skipif(not REF, reason="oracle/_ref not built (no /root/reference)")
def test_romint_matches_live_reference():
    rng = np.random.default_rng(1)
    for m in (16, 64, 512, 4096, 8192):
        f = rng.standard_normal(m + 1)
        a = O.romint(f, m, 1.0 / m)
        b = O.ref.dealii.romint(f, m, 1.0 / m)
        assert abs(a - b) <= 2e-16 * max(1.0, np.abs(f).sum() / m)


@pytest.mark.parametrize("m", [16, 32, 256, 2048, 8192])
def test_romberg_weights_reproduce_romint(m):
    rng = np.random.default_rng(m)
    c = O.romberg_weights(m)
    assert np.allclose(c, c[::-1], rtol=0, atol=1e-18)  # symmetric: pairing fold is exact
    assert abs(c.sum() - 1.0) < 1e-14
    for _ in range(3):
        f = rng.standard_normal(m + 1)
        assert abs(c @ f - O.romint(f, m, 1.0 / m)) < 1e-14


def test_romberg_polynomial_exactness():
    """Five trapezoid levels extrapolated along h^2: exact for degree <= 9."""
    m = 64
    s = np.linspace(0.0, 1.0, m + 1)
    for p in range(10):
        assert abs(O.romint(s ** p, m, 1.0 / m) - 1.0 / (p + 1)) < 1e-13, p
    assert abs(O.romint(s ** 10, m, 1.0 / m) - 1.0 / 11) > 1e-15


def test_romint_rejects_short_arrays():
    assert np.isnan(O.romint(np.ones(9), 8, 1.0 / 8))  # romint.c:29-33 exits there


# ---------------------------------------------------------------- assembly (a3)
def test_uniform_stencils_closed_form():
    M, L = 11, 2.5
    h = L / (M - 1)
    x = O.mesh_uniform(M, L)
    w = np.linspace(-1.0, 2.0, M)
    A, B, Cm = O.assemble(x, w, O.C_ROWSCALED)
    for i in range(1, M - 1):  # 1D_FEM.c:95-98
        assert np.allclose(A[i], [h / 6, 2 * h / 3, h / 6], rtol=1e-15)
        assert np.allclose(B[i], [-1 / h, 2 / h, -1 / h], rtol=1e-15)
        assert np.allclose(Cm[i], w[i] * A[i], rtol=1e-15)  # row-scaled (1D_FEM.c:104-105)
    # natural boundary rows (simple_FEM_1D_transient.m:42-47,58-63)
    assert np.allclose(A[0], [0, h / 3, h / 6]) and np.allclose(A[-1], [h / 6, h / 3, 0])
    assert np.allclose(B[0], [0, 1 / h, -1 / h]) and np.allclose(B[-1], [-1 / h, 1 / h, 0])


def test_general_mesh_invariants():
    rng = np.random.default_rng(2)
    x = np.concatenate([[0.0], np.cumsum(rng.uniform(0.05, 0.3, 20))])
    M = len(x)
    hs = np.diff(x)
    w = rng.uniform(-2, 3, M)
    A, B, Cm = O.assemble(x, w, O.C_CONSISTENT)
    # stiffness annihilates constants; mass rows sum to the nodal patch length / 2
    assert np.abs(B.sum(axis=1)).max() < 1e-12
    patch = np.concatenate([[hs[0]], hs[:-1] + hs[1:], [hs[-1]]]) / 2
    assert np.allclose(A.sum(axis=1), patch, rtol=1e-14)
    # consistent w-mass == exact integral of w_h phi_i phi_j (3-point Gauss is exact)
    g, gw = np.polynomial.legendre.leggauss(3)
    Cx = np.zeros((M, 3))
    for e in range(M - 1):
        xa, xb = x[e], x[e + 1]
        for t, wt in zip(g, gw):
            s = 0.5 * (t + 1)
            pa, pb = 1 - s, s
            wq = w[e] * pa + w[e + 1] * pb
            jw = 0.5 * (xb - xa) * wt
            Cx[e, 1] += wq * pa * pa * jw
            Cx[e, 2] += wq * pa * pb * jw
            Cx[e + 1, 0] += wq * pb * pa * jw
            Cx[e + 1, 1] += wq * pb * pb * jw
    assert np.allclose(Cm, Cx, rtol=1e-13, atol=1e-15)
    # for constant w the consistent and row-scaled forms agree: C = c A
    Ac, _, Cc = O.assemble(x, np.full(M, 1.7), O.C_CONSISTENT)
    assert np.allclose(Cc, 1.7 * Ac, rtol=1e-14)


def test_system_boundary_rows():
    x = O.mesh_uniform(9, 1.0)
    w = np.ones(9)
    D, R = O.system(x, w, 1 / 64, bc=O.BC_DD)
    assert np.array_equal(D[0], [0, 1, 0]) and np.array_equal(D[-1], [0, 1, 0])
    assert not R[0].any() and not R[-1].any()
    D2, R2 = O.system(x, w, 1 / 64, bc=O.BC_DN)
    assert np.array_equal(D2[0], [0, 1, 0]) and D2[-1][1] != 1.0 and R2[-1][1] > 0


# ---------------------------------------------------------------- march (a4-a5)
@pytest.mark.parametrize("c,M,N", [(0.0, 33, 2048), (0.7, 33, 2048), (-2.0, 65, 512),
                                   (5.0, 129, 256)])
def test_sine_mode_exact_solution(c, M, N):
    """Constant w, uniform mesh, DD: q^j has a closed form in sine modes."""
    L = 3.72
    x = O.mesh_uniform(M, L)
    q = O.march(x, np.full(M, c), N, bc=O.BC_DD)
    for j in (1, 7, N // 2, N):
        ex = O.sine_mode_exact(M, L, c, N, j)
        assert np.abs(q[j] - ex).max() < 1e-12 * max(1.0, np.abs(ex).max()), j


def test_march_matches_dense_solve():
    rng = np.random.default_rng(4)
    for bc in (O.BC_DD, O.BC_DN):
        x = np.concatenate([[0.0], np.cumsum(rng.uniform(0.05, 0.2, 15))])
        M = len(x)
        w = rng.uniform(-3, 5, M)
        N = 32
        D3, R3 = O.system(x, w, 1 / N, bc=bc)
        Dm = np.diag(D3[:, 1]) + np.diag(D3[1:, 0], -1) + np.diag(D3[:-1, 2], 1)
        Rm = np.diag(R3[:, 1]) + np.diag(R3[1:, 0], -1) + np.diag(R3[:-1, 2], 1)
        q = O.march(x, w, N, bc=bc)
        qd = q[0].copy()
        for j in range(1, N + 1):
            qd = np.linalg.solve(Dm, Rm @ qd)
            assert np.abs(q[j] - qd).max() < 1e-13


def test_march_symmetry_and_half_domain():
    """Symmetric w under DD gives q(x) = q(L-x); DN on [0, L/2] reproduces it."""
    m, L, N = 64, 4.0, 256
    x = O.mesh_uniform(m + 1, L)
    w = np.cos(2 * np.pi * x / L)
    w = 0.5 * (w + w[::-1])
    q = O.march(x, w, N, bc=O.BC_DD)
    assert np.abs(q - q[:, ::-1]).max() < 1e-13
    qh = O.march(x[: m // 2 + 1], w[: m // 2 + 1], N, bc=O.BC_DN)
    assert np.abs(qh - q[:, : m // 2 + 1]).max() < 1e-13


def test_march_positivity_and_extended_precision():
    x = O.mesh_uniform(257, 6.0)
    w = 2 * np.sin(x)
    q = O.march(x, w, 512, bc=O.BC_DN)
    assert (q[1:, 1:] > 0).all()
    qld = O.march_ld(x, w, 512, bc=O.BC_DN)
    assert np.abs(q - qld).max() < 1e-12


# ---------------------------------------------------------------- evaluator
def test_phi_romint_equals_phi_weights():
    x = O.mesh_uniform(65, 3.0)
    w = 0.3 * np.exp(-x)
    q = O.march(x, w, 256, bc=O.BC_DN)
    a = O.phi_romint(q)
    b = O.phi_weights(q, O.romberg_weights(256))
    assert np.abs(a - b).max() < 1e-14


def test_mode_m_golden_field():
    """solution_matlab_N=33 solves the MATLAB-convention IE residual to ~2.3e-7
    (offset trapezoid over 2049 steps, count-based target, phi_t - phi)."""
    f = FX["matlab_N33"]
    x, w = np.array(f["x"]), np.array(f["w"])
    pt = O.target_count(x, O.TAU_SHORT, O.L_SHORT, nan_guard=True)
    out = O.eval_matlab(x, w, pt, 2049, 1.0 / 2048)
    res = np.abs(out).max()
    assert 1e-7 < res < 3e-7
    # the same field is not a Mode R solution (pairing/quadrature/target differ)
    ref = O.evaluate(x, w, O.target_position(x, O.TAU, True), 2048, bc=O.BC_DD)
    assert ref["rnorm"] > 1e-4


def test_res_fields_are_warm_starts_only():
    """Wang's solver fields are solutions of another discretisation."""
    for key, lo, hi in (("res_m32", 1e-2, 0.2), ("res_m1024", 1e-5, 1e-3)):
        eta = np.array(FX[key]["eta"])
        m = len(eta) - 1
        x = np.arange(m + 1) * O.L_REF / m
        pt = O.target_position(x, O.TAU, True)
        w = eta.copy()
        w[0] = w[-1] = 0.0
        r = O.evaluate(x, w, pt, 2048, bc=O.BC_DD)["rnorm"]
        assert lo < r < hi, (key, r)


# ---------------------------------------------------------------- mixers (a8-a11)
def test_gaussj_matches_reference_vector():
    g = FX["ref_nr"]["gaussj"]
    A, b = np.array(g["A"]), np.array(g["b"])
    rc, inv, x = O.gaussj(A, b.reshape(-1, 1), mode=1)
    assert rc == 0
    assert np.allclose(x.ravel(), g["x"], rtol=1e-14, atol=1e-15)
    assert np.allclose(inv, g["inv"], rtol=1e-13, atol=1e-15)
    assert np.allclose(A @ x.ravel(), b)
    rc, _, _ = O.gaussj(np.zeros((2, 2)), np.ones((2, 1)), mode=1)
    assert rc == 1  # DEALII_SCFT/src/gaussj.c:46-50


def test_adm_linear_toy():
    """1D_FEM.c:372-379: x1 = x1/2 - 2, x2 = x2/2 + 3 -> (-4, 6); the reference
    adm returns f(x) of the converged iterate (fixture from oracle/_ref)."""
    def f(n, x, xn, ctx):
        xn[0] = x[0] * 0.5 - 2.0
        xn[1] = x[1] * 0.5 + 3.0
    cb = O.FIXED_POINT_FN(f)
    x = np.array([1.0, 2.0])
    xnew = np.zeros(2)
    its, err = C.c_int(), C.c_double()
    chk = O.lib().orc_adm(x, 2, C.cast(cb, C.c_void_p), None, 1e-10, 100, xnew, C.byref(its),
                          C.byref(err))
    assert chk == 0
    assert np.allclose(xnew, FX["ref_nr"]["adm_toy"]["xnew"], rtol=0, atol=1e-13)
    assert np.allclose(xnew, [-4.0, 6.0], atol=1e-12)


@pytest.mark.skipif(not REF, reason="oracle/_ref not built (no /root/reference)")
def test_adm_trajectory_matches_reference_on_scft():
    """The reference's compiled adm and the restatement drive the same oracle
    evaluator (m = 32, DD, fixed-point x + 0.1 (phi - phi_t)); the inputs of
    the first 25 evaluations agree."""
    m, N = 32, 2048
    x = O.mesh_uniform(m + 1, O.L_REF)
    pt = O.target_position(x, O.TAU, True)
    eta = np.array(FX["res_m32"]["eta"])
    K = 25
    ctx = O.EvalCtx(x, pt, N, bc=O.BC_DD, beta=0.1)
    fp, _, trace_ref = O.global_callbacks(ctx, m - 1, trace_len=K + 2, limit=K)
    x0 = (C.c_double * (m - 1))(*eta[1:-1])
    chk = C.c_int(1)
    O.ref.top.adm(x0, m - 1, C.byref(chk), fp, 0)
    t_ref = trace_ref.reshape(-1, m - 1)[:K].copy()
    _, _, trace_me = O.global_callbacks(ctx, m - 1, trace_len=K + 2, limit=K)
    xm = eta[1:-1].copy()
    xn = np.zeros(m - 1)
    its, err = C.c_int(), C.c_double()
    O.lib().orc_adm(xm, m - 1, C.cast(O.lib().orc_global_fixed_point_ctx, C.c_void_p), None,
                    1e-10, K, xn, C.byref(its), C.byref(err))
    t_me = trace_me.reshape(-1, m - 1)[:K]
    scale = np.abs(t_ref).max(axis=1, keepdims=True)
    rel = np.abs(t_me - t_ref) / scale
    assert rel.max() < 1e-9, rel.max(axis=1)
    ctx.close()


@pytest.mark.skipif(not REF, reason="oracle/_ref not built (no /root/reference)")
def test_broydn_matches_reference_on_scft():
    """1D_FEM.c's main: broydn from the .res eta on m = 32 (DD, TOLF 1e-8).
    The restatement and the reference's compiled broydn make the same first
    n+1 evaluations (initial + forward-difference Jacobian) and converge to
    the same field."""
    m, N = 32, 2048
    n = m - 1
    x = O.mesh_uniform(m + 1, O.L_REF)
    pt = O.target_position(x, O.TAU, True)
    eta = np.array(FX["res_m32"]["eta"])
    ctx = O.EvalCtx(x, pt, N, bc=O.BC_DD)
    L = O.ref.dealii
    L.nrref_broydn_setup(n, 1e-8)
    _, vf, tr_ref = O.global_callbacks(ctx, n, trace_len=400)
    xr = (C.c_double * (n + 1))(0.0, *eta[1:-1])
    chk = C.c_int(1)
    L.broydn.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p]
    L.broydn(xr, n, C.byref(chk), vf)
    calls_ref = O.lib().orc_global_calls()
    t_ref = tr_ref.reshape(-1, n)[: n + 1].copy()
    x_ref = np.array(xr[1:])
    # restatement through the same traced callback
    _, _, tr_me = O.global_callbacks(ctx, n, trace_len=400)

    cb = C.cast(O.lib().orc_global_residual_ctx, C.c_void_p)
    xm = eta[1:-1].copy()
    jc = C.c_int(0)
    qt, r, d = np.zeros(n * n), np.zeros(n * n), np.zeros(n)
    st = O.BroydnStats()
    chk2 = O.lib().orc_broydn(xm, n, cb, None, 1e-8, C.byref(jc), qt, r, d, C.byref(st))
    t_me = tr_me.reshape(-1, n)[: n + 1]
    assert np.abs(t_me - t_ref).max() <= 1e-13 * np.abs(t_ref).max()
    assert chk.value == 0 and chk2 == 0
    assert abs(st.evaluations - calls_ref) <= 2
    assert np.abs(xm - x_ref).max() < 1e-6 * np.abs(x_ref).max()
    ctx.close()


def test_safeguarded_anderson_converges_config1():
    """History 10, mixing 0.1, step cap 0.5 (documented safeguard), w0 = 0."""
    m, N, L = 32, 2048, 3.72
    x = O.mesh_uniform(m + 1, L)
    pt = O.target_position(x, O.TAU, False)
    r = O.solve_anderson(x, pt, N, np.zeros(m + 1), O.anderson_opts(step_cap=0.5),
                         bc=O.BC_DN)
    assert r["rc"] == 0 and r["final_err"] <= 1e-6
    e = O.evaluate(x, r["w"], pt, N, bc=O.BC_DN)
    assert e["rnorm"] <= 1e-6


# ---------------------------------------------------------------- MLP oracle
def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(5)
    dims = [5, 7, 6, 3]
    P = O.mlp_param_count(dims)
    assert P == 5 * 7 + 7 + 7 * 6 + 6 + 6 * 3 + 3
    params = rng.standard_normal(P) * 0.5
    X = rng.standard_normal((4, 5))
    T = rng.standard_normal((4, 3))
    loss, g = O.mlp_loss_grad(dims, params, X, T)
    eps = 1e-6
    for k in rng.choice(P, 25, replace=False):
        pp, pm = params.copy(), params.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (O.mlp_loss_grad(dims, pp, X, T)[0] - O.mlp_loss_grad(dims, pm, X, T)[0]) / (2 * eps)
        assert abs(fd - g[k]) < 1e-8 * max(1.0, abs(fd)), k
    Y = O.mlp_forward(dims, params, X)
    assert abs(np.mean((Y - T) ** 2) - loss) < 1e-15


def test_adam_step_closed_form():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    m, v = np.zeros(2), np.zeros(2)
    O.adam_step(p, g, m, v, 1, lr=0.1)
    # first bias-corrected step is lr * sign(g) (up to eps)
    assert np.allclose(p, [0.9, -1.9], atol=1e-7)
