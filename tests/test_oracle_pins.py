"""Pins of the oracle parts that no reference fixture covers.

* The graded mesh of the production mix (oracle/scft_oracle.c:
  orc_graded_elements, orc_mesh_graded; scft_amd/workloads.py:graded_mesh).
  The reference has no graded generator; its general-mesh stencils
  (Matlab_files/simple_FEM_1D_transient.m:35-63) take any node vector.  The
  specification h(x) = h0 (1 + x/ell) is pinned by what it fixes exactly:
  h_i / (1 + x_i/ell) is one constant, the end points are exact, the element
  count is round(ell ln(1 + L/ell) / h0) clamped to [63, 4095], and the nodes
  solve the recurrence x_{i+1} = x_i + h(x_i) written out independently.
* The residual-form Anderson mixer (oracle_solvers.c:orc_anderson), which
  the device mixer follows, and its ground-state preconditioner (no
  reference counterpart), pinned by the properties of -(1/d^2)(d^2 f')':
  constants in its null space, self-adjoint and nonnegative in the product
  weighted by the lumped masses and d^2, and Anderson's finite termination
  with it as a fixed preconditioner.  On an affine residual r(w) = b - A w with the
  history at least n and no regularisation, Anderson mixing is GMRES in
  disguise (Walker & Ni, SIAM J. Numer. Anal. 49 (2011) 1715-1735, Thm 2.2):
  the iterate after n+1 evaluations is the exact fixed point, so the
  (n+2)-th evaluation reads a residual at rounding level.  A wrong sign of
  the least-squares coefficients, a dropped history column or a transposed
  Gram system loses that finite termination.  History 0 is plain mixing
  w + beta r, and the reference's linear toy (1D_FEM.c:372-379, fixed point
  (-4, 6)) is reached in at most 4 evaluations.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from tests.helpers import workloads

WL = workloads()


# ------------------------------------------------------------------ graded mesh
def _draw(rng):
    L = rng.uniform(2.0, 8.0)
    h0 = math.exp(rng.uniform(math.log(1e-4), math.log(1e-2)))
    ell = rng.uniform(0.05, 0.5)
    return L, h0, ell


def test_graded_mesh_closed_form():
    rng = np.random.default_rng(20261205)
    for _ in range(300):
        L, h0, ell = _draw(rng)
        m = O.graded_elements(L, h0, ell, 63, 4095)
        x = O.mesh_graded(m + 1, L, ell)
        assert x[0] == 0.0 and x[-1] == L and len(x) == m + 1
        h = np.diff(x)
        assert (h > 0).all()
        g = math.log1p(L / ell) / m
        c = ell * math.expm1(g)
        # h_i / (1 + x_i/ell) is the same constant c = ell (e^g - 1) everywhere
        ratio = h / (1.0 + x[:-1] / ell)
        assert np.abs(ratio / c - 1.0).max() <= 1e-9, (L, h0, ell)
        # the recurrence x_{i+1} = x_i + c (1 + x_i/ell), summed independently
        xr = np.empty(m + 1)
        xr[0] = 0.0
        for i in range(m):
            xr[i + 1] = xr[i] + c * (1.0 + xr[i] / ell)
        assert np.abs(xr - x).max() <= 1e-12 * L
        # element count: round(ell ln(1+L/ell)/h0) before clamping, so the first
        # element is h0 up to the rounding of the count (|dm| <= 1/2)
        m_raw = ell * math.log1p(L / ell) / h0
        assert m == min(max(int(round(m_raw)), 63), 4095)
        if 63 < m < 4095:
            assert abs(c / h0 - 1.0) <= 0.5 / m + g, (c, h0)


def test_graded_mesh_clamps():
    # huge h0: too few elements -> 63; tiny h0: too many -> 4095
    assert O.graded_elements(2.0, 0.5, 0.5, 63, 4095) == 63
    assert O.graded_elements(8.0, 1e-6, 0.5, 63, 4095) == 4095
    x = WL.graded_mesh(8.0, 1e-6, 0.5)
    assert len(x) == 4096 and x[-1] == 8.0


def test_graded_mesh_workloads_match_oracle():
    """The workload generator (numpy) and the oracle (C libm) build the same
    nodes to a few ulp (numpy's and libm's expm1 round differently; every
    parity test hands one array to both sides, so this never enters a bar)."""
    rng = np.random.default_rng(7)
    for _ in range(300):
        L, h0, ell = _draw(rng)
        xw = WL.graded_mesh(L, h0, ell)
        m = O.graded_elements(L, h0, ell, 63, 4095)
        assert len(xw) == m + 1
        xo = O.mesh_graded(m + 1, L, ell)
        assert xw[0] == xo[0] == 0.0 and xw[-1] == xo[-1] == L
        assert (np.abs(xw - xo) <= 16.0 * np.spacing(np.maximum(np.abs(xo), 1e-300))).all()


# ------------------------------------------------------------------ Anderson
def _affine_solve(A, b, w0, history, beta, max_iter, transform=O.MIX_PLAIN):
    n = len(w0)
    calls = []

    def F(nn, w, r, ctx):
        wv = np.ctypeslib.as_array(w, shape=(nn,))
        rv = b - A @ wv
        for i in range(nn):
            r[i] = rv[i]
        calls.append(float(np.abs(rv).max()))
        return 0

    cb = O.RESIDUAL_FN(F)
    opts = O.anderson_opts(history=history, beta=beta, tol=0.0, max_iter=max_iter,
                           step_cap=0.0, regularization=0.0, transform=transform)
    st = O.AndersonStats(0, 0, 0.0, None, 0)
    w = np.ascontiguousarray(w0, dtype=np.float64).copy()
    O.lib().orc_anderson(w, n, C.cast(cb, C.c_void_p), None, C.byref(opts), None, C.byref(st))
    return w, calls


@pytest.mark.parametrize("n,beta", [(2, 1.0), (5, 1.0), (9, 1.0), (16, 1.0),
                                    (2, 0.3), (5, 0.3), (9, 0.3), (12, 0.3)])
def test_anderson_finite_termination_on_affine_map(n, beta):
    """GMRES reaches x* in n steps, so evaluation n+2 (index n+1) reads a zero
    residual.  With beta < 1 the differences of the residuals are shorter and
    the normal equations of the Gram system (condition squared) leave that
    step at ~1e-2 on these draws; the next update lands, so one evaluation
    more is allowed there."""
    rng = np.random.default_rng(100 + n)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(np.geomspace(0.1, 3.0, n)) @ Q.T + 0.2 * rng.standard_normal((n, n)) / n
    b = rng.standard_normal(n)
    wstar = np.linalg.solve(A, b)
    last = n + 1 if beta == 1.0 else n + 2
    w, calls = _affine_solve(A, b, np.zeros(n), history=16, beta=beta, max_iter=last + 1)
    scale = np.abs(b).max()
    assert calls[last] <= 1e-9 * scale, calls
    assert np.abs(w - wstar).max() <= 1e-9 * np.abs(wstar).max()
    # halfway the residual is far from rounding level (no early accident)
    assert calls[n // 2] > 1e-6 * scale


def test_anderson_truncated_history_does_not_terminate():
    """Finite termination needs the whole history: with history 2 on n = 9
    the (n+2)-th residual is still far from zero (the bound above is sharp
    enough to catch a dropped column)."""
    n = 9
    rng = np.random.default_rng(109)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(np.linspace(0.5, 3.0, n)) @ Q.T
    b = rng.standard_normal(n)
    _, calls = _affine_solve(A, b, np.zeros(n), history=2, beta=1.0, max_iter=n + 2)
    assert calls[n + 1] > 1e-6 * np.abs(b).max()


def test_anderson_history_zero_is_plain_mixing():
    n = 6
    rng = np.random.default_rng(3)
    A = np.eye(n) + 0.1 * rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    w = np.zeros(n)
    for _ in range(5):
        w = w + 0.4 * (b - A @ w)
    wm, _ = _affine_solve(A, b, np.zeros(n), history=0, beta=0.4, max_iter=5)
    np.testing.assert_allclose(wm, w, rtol=0, atol=1e-14)


def test_anderson_reference_linear_toy():
    """1D_FEM.c:372-379: x -> (x1/2 - 2, x2/2 + 3); in residual form
    r = g(x) - x = b - A x with A = I/2, b = (-2, 3).  Fixed point (-4, 6)."""
    A = 0.5 * np.eye(2)
    b = np.array([-2.0, 3.0])
    w, calls = _affine_solve(A, b, np.array([1.0, 2.0]), history=10, beta=0.1, max_iter=4)
    assert calls[-1] <= 1e-13
    np.testing.assert_allclose(w, [-4.0, 6.0], rtol=0, atol=1e-12)


# ------------------------------------------------------------ ground-state form
def _gs_operator(x, phi_t_unk, w0, a=0.5, c=0.0, floor=1e-8, shift=0.0):
    """P of the ground-state form, column by column: one plain-mixing step
    (history 0, beta 1, plain transform) of orc_anderson on the constant
    residual e_j returns w0 + P e_j."""
    n = len(w0)
    gs = np.zeros(3 * n)
    O.lib().orc_gs_stencil(np.ascontiguousarray(x), len(x), O.BC_DN, gs)
    P = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0

        def F(nn, w, r, ctx, e=e):
            for i in range(nn):
                r[i] = e[i]
            return 0

        cb = O.RESIDUAL_FN(F)
        opts = O.anderson_opts(history=0, beta=1.0, tol=0.0, max_iter=1, step_cap=0.0,
                               regularization=0.0, transform=O.MIX_PLAIN, floor=floor,
                               gs_precond=a, gs_field_weight=c, gs_shift=shift)
        opts.gstencil = gs.ctypes.data
        w = np.array(w0, dtype=np.float64)
        stt = O.AndersonStats(0, 0, 0.0, None, 0)
        O.lib().orc_anderson(w, n, C.cast(cb, C.c_void_p), None, C.byref(opts),
                             np.ascontiguousarray(phi_t_unk).ctypes.data, C.byref(stt))
        P[:, j] = w - w0
    return P


def _graded_case(n=40):
    x = WL.graded_mesh(3.0, 2e-3, 0.2)[: n + 1].copy()  # DN: unknowns at nodes 1..n
    pt = O.target_position(x, O.TAU, False)[1:]
    return x, pt


def test_ground_state_form_maps_constants_to_constants():
    """g = f + a [-(1/d^2)(d^2 f')'] with c = 0: a constant residual is mixed
    unchanged, whatever d (here phi_t next to a wall, where d spans two
    decades and the wall node's d is zero)."""
    x, pt = _graded_case()
    P = _gs_operator(x, pt, np.zeros(len(pt)))
    np.testing.assert_allclose(P @ np.ones(len(pt)), np.ones(len(pt)), rtol=0, atol=1e-9)


def test_ground_state_form_is_self_adjoint_in_the_weighted_product():
    """-(1/d^2)(d^2 f')' is self-adjoint and nonnegative in the product
    <f, g> = sum_u m_u d_u^2 f_u g_u (m_u the lumped mass (h_l + h_r)/2 of
    node u + 1, the wall spacing included, h_r = 0 at the natural end):
    W (P - I) is symmetric positive semidefinite.  A swapped ratio
    d_u/d_{u-1}, a wrong mass or a sign error breaks it."""
    x, pt = _graded_case()
    n = len(pt)
    P = _gs_operator(x, pt, np.zeros(n), a=0.5)
    hl = np.diff(x)[:n]
    hr = np.concatenate([np.diff(x)[1:n], [0.0]])
    m = 0.5 * (hl + hr)
    d = np.sqrt(np.maximum(pt, 1e-8))
    A = np.diag(m * d * d) @ (P - np.eye(n))
    scale = np.abs(A).max()
    assert np.abs(A - A.T).max() <= 1e-12 * scale
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev.min() >= -1e-12 * scale
    # the coupling of neighbours u, u+1 is -a d_u d_{u+1} / h: nonzero, local
    assert np.abs(np.triu(A, 2)).max() == 0.0
    np.testing.assert_allclose(np.diag(A, 1), -0.5 * d[:-1] * d[1:] / np.diff(x)[1:n],
                               rtol=1e-12)


def test_ground_state_form_field_term():
    """The field term adds a c max(w, 0) on the diagonal, w the current
    iterate."""
    x, pt = _graded_case(24)
    rng = np.random.default_rng(8)
    w0 = rng.uniform(-3.0, 5.0, len(pt))
    P0 = _gs_operator(x, pt, w0, a=0.5, c=0.0)
    P1 = _gs_operator(x, pt, w0, a=0.5, c=1.5)
    # (differences of entries up to ~1e5: rounding at 1e-15 of that scale)
    np.testing.assert_allclose(P1 - P0, np.diag(0.5 * 1.5 * np.maximum(w0, 0.0)), rtol=0,
                               atol=1e-15 * np.abs(P0).max())


@pytest.mark.parametrize("n", [6, 12])
def test_ground_state_form_keeps_finite_termination(n):
    """With c = 0 the form is a fixed linear operator P, so Anderson on the
    preconditioned affine residual P (b - A w) is still GMRES and terminates
    by evaluation n+2 (plain transform: the test drives orc_anderson with an
    affine residual, phi_t only weights d)."""
    rng = np.random.default_rng(200 + n)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(np.geomspace(0.2, 2.0, n)) @ Q.T
    b = rng.standard_normal(n)
    x = np.linspace(0.0, 1.0, n + 1)
    gs = np.zeros(3 * n)
    O.lib().orc_gs_stencil(np.ascontiguousarray(x), n + 1, O.BC_DN, gs)
    phi_t = np.geomspace(1e-4, 1.0, n)  # d spans two decades, as next to a wall
    calls = []

    def F(nn, w, r, ctx):
        wv = np.ctypeslib.as_array(w, shape=(nn,))
        rv = b - A @ wv
        for i in range(nn):
            r[i] = rv[i]
        calls.append(float(np.abs(rv).max()))
        return 0

    cb = O.RESIDUAL_FN(F)
    opts = O.anderson_opts(history=16, beta=1.0, tol=0.0, max_iter=n + 2, step_cap=0.0,
                           regularization=0.0, transform=O.MIX_PLAIN, precond=0.0,
                           gs_precond=0.05, gs_field_weight=0.0)
    opts.gstencil = gs.ctypes.data
    w = np.zeros(n)
    stt = O.AndersonStats(0, 0, 0.0, None, 0)
    O.lib().orc_anderson(w, n, C.cast(cb, C.c_void_p), None, C.byref(opts),
                         np.ascontiguousarray(phi_t).ctypes.data, C.byref(stt))
    assert calls[n + 1] <= 1e-8 * np.abs(b).max(), calls
    np.testing.assert_allclose(w, np.linalg.solve(A, b), rtol=0, atol=1e-7)


def test_rational_form_is_the_resolvent_combination():
    """The rational form (1 - a) I + a G + a g (g I + G0)^-1 (G0: G without
    the field term) against the same combination built from G and G0
    themselves (read off the linear form with a = 1) and a dense inverse:
    pins the Thomas solve of g I + G0 (its forward and backward sweeps).  Constants are preserved and
    diag(m d^2)(P - I) is symmetric positive semidefinite, as for any
    rational function of G with a nonnegative symbol x (x + g - 1)/(g + x)."""
    x, pt = _graded_case(30)
    n = len(pt)
    rng = np.random.default_rng(12)
    w0 = rng.uniform(-2.0, 40.0, n)
    G0 = _gs_operator(x, pt, w0, a=1.0, c=0.0) - np.eye(n)
    for c in (0.0, 0.25):
        G = _gs_operator(x, pt, w0, a=1.0, c=c) - np.eye(n)
        for a, g in ((0.5, 3.0), (0.4, 2.0)):
            P = _gs_operator(x, pt, w0, a=a, c=c, shift=g)
            ref = (1 - a) * np.eye(n) + a * G + a * g * np.linalg.inv(g * np.eye(n) + G0)
            scale = np.abs(ref).max()
            assert np.abs(P - ref).max() <= 1e-10 * scale, (c, a, g, np.abs(P - ref).max())
    # on a coarse mesh G is O(1), so the resolvent part itself is checked:
    # P - (1 - a) I - a G = a g (g I + G)^-1 to rounding
    xc = np.linspace(0.0, 6.0, 25)
    ptc = O.target_position(xc, O.TAU, False)[1:]
    wc = rng.uniform(-1.0, 3.0, len(ptc))
    nc = len(ptc)
    Gc = _gs_operator(xc, ptc, wc, a=1.0, c=0.25) - np.eye(nc)
    G0c = _gs_operator(xc, ptc, wc, a=1.0, c=0.0) - np.eye(nc)
    Pc = _gs_operator(xc, ptc, wc, a=0.5, c=0.25, shift=3.0)
    res = Pc - 0.5 * np.eye(nc) - 0.5 * Gc
    ref = 1.5 * np.linalg.inv(3.0 * np.eye(nc) + G0c)
    assert np.abs(ref).max() > 0.1 and np.abs(Gc).max() < 1e3
    assert np.abs(res - ref).max() <= 1e-12, np.abs(res - ref).max()
    P = _gs_operator(x, pt, np.zeros(n), a=0.5, shift=3.0)
    np.testing.assert_allclose(P @ np.ones(n), np.ones(n), rtol=0, atol=1e-9)
    hl = np.diff(x)[:n]
    hr = np.concatenate([np.diff(x)[1:n], [0.0]])
    d = np.sqrt(np.maximum(pt, 1e-8))
    A = np.diag(0.5 * (hl + hr) * d * d) @ (P - np.eye(n))
    sc = np.abs(A).max()
    assert np.abs(A - A.T).max() <= 1e-10 * sc
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() >= -1e-10 * sc


def test_rational_form_threshold_switch():
    """gs_shift_below: the linear form until ||r|| first falls below the
    threshold, then the rational one with the history cleared.  A threshold
    below every residual gives the linear form's iterates bit for bit, one
    above the first residual the rational form's (the switch at the first
    evaluation clears an empty history), and one in between neither."""
    n = 12
    rng = np.random.default_rng(77)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(np.geomspace(0.3, 2.0, n)) @ Q.T
    b = rng.standard_normal(n)
    x = np.linspace(0.0, 1.0, n + 1)
    gs = np.zeros(3 * n)
    O.lib().orc_gs_stencil(np.ascontiguousarray(x), n + 1, O.BC_DN, gs)
    phi_t = np.geomspace(1e-3, 1.0, n)

    def run(**kw):
        calls = []

        def F(nn, w, r, ctx):
            wv = np.ctypeslib.as_array(w, shape=(nn,))
            rv = b - A @ wv
            for i in range(nn):
                r[i] = rv[i]
            calls.append(float(np.abs(rv).max()))
            return 0

        cb = O.RESIDUAL_FN(F)
        opts = O.anderson_opts(history=4, beta=1.0, tol=0.0, max_iter=6, step_cap=0.0,
                               regularization=0.0, transform=O.MIX_PLAIN, gs_precond=0.02,
                               gs_field_weight=0.0, **kw)
        opts.gstencil = gs.ctypes.data
        w = np.zeros(n)
        stt = O.AndersonStats(0, 0, 0.0, None, 0)
        O.lib().orc_anderson(w, n, C.cast(cb, C.c_void_p), None, C.byref(opts),
                             np.ascontiguousarray(phi_t).ctypes.data, C.byref(stt))
        return w, calls

    lin, calls = run()
    rat, _ = run(gs_shift=3.0)
    assert not np.array_equal(lin, rat)
    low, _ = run(gs_shift=3.0, gs_shift_below=1e-12)
    high, _ = run(gs_shift=3.0, gs_shift_below=10.0 * max(calls))
    assert np.array_equal(low, lin) and np.array_equal(high, rat)
    # a threshold the first residual is above and a later one below
    assert min(calls[1:]) < 0.5 * calls[0]
    mid, _ = run(gs_shift=3.0, gs_shift_below=np.sqrt(calls[0] * min(calls[1:])))
    assert not np.array_equal(mid, lin) and not np.array_equal(mid, rat)
