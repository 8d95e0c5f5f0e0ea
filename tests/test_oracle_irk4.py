"""Pins of the oracle's two-stage Gauss-Legendre (IRK4) march.

orc_march_irk4 restates the deal.II path's stepper (DEALII_SCFT/src/scft.cc:
671-695, drivescft.cc:130-146; Matlab_files/IRK4.m:96-120) by solving the
2n x 2n stage system as a block-tridiagonal system.  It is checked against:

  * the closed form for constant w on a uniform mesh: sine mode k is
    multiplied per step by the Gauss method's stability factor, the (2,2)
    Pade approximant (1 - y/2 + y^2/12) / (1 + y/2 + y^2/12), y = dt (lam_k + c);
  * fourth-order convergence to the exact semi-discrete propagator
    expm(-s A^{-1} D') (scipy), for a non-constant field, both boundary
    conditions and both w-mass variants;
  * a dense solve of the same stage system (numpy), one step on a tiny mesh;
  * the reference's own fixture DEALII_SCFT/inputFiles/N=33_for_read.txt: a
    field the deal.II IRK solver converged to ||r|| = 1.42e-9 (its header).
    With the consistent w-mass, the file's wall values and N = 2048 steps
    (drivescft.cc:130) the oracle's residual of that field is 1.7e-9, against
    1.1e-4 for implicit Euler, and it is smallest at N = 2048;
  * the MATLAB refinement flow's field solution_yita_1D_N=141: plausibility
    only (IRK4 residual below implicit Euler's).
"""
import os
import re

import numpy as np
import pytest
from scipy.linalg import expm

import oracle as O

FILES = os.path.join(os.path.dirname(__file__), "golden", "files")


def _read_rows(name):
    rows = [l for l in open(os.path.join(FILES, name)) if re.match(r"^\s*\d+\s*,", l)]
    d = np.array([[float(v) for v in r.split(",")] for r in rows])
    return d[:, 1], d[:, 2]


def _full(T, M):
    F = np.zeros((M, M))
    for i in range(M):
        if i > 0:
            F[i, i - 1] = T[3 * i]
        F[i, i] = T[3 * i + 1]
        if i < M - 1:
            F[i, i + 1] = T[3 * i + 2]
    return F


def _matrices(x, w, cmat):
    M = len(x)
    A3, B3, C3 = np.zeros(3 * M), np.zeros(3 * M), np.zeros(3 * M)
    O.lib().orc_assemble(M, np.ascontiguousarray(x), np.ascontiguousarray(w), cmat, A3, B3, C3)
    return _full(A3, M), _full(B3, M) + _full(C3, M)


@pytest.mark.parametrize("M,L,c,N", [(33, 3.7, 0.7, 16), (33, 3.7, 0.7, 256), (65, 2.0, -1.5, 64),
                                     (129, 6.0, 3.0, 2048)])
def test_sine_modes_pade(M, L, c, N):
    x = np.linspace(0.0, L, M)
    q = O.march_irk4(x, np.full(M, c), N)
    n = M - 1
    h, dt = L / n, 1.0 / N
    i = np.arange(1, n)
    for j in (1, N // 2, N):
        ex = np.zeros(M)
        for k in range(1, n):
            t = k * np.pi / n
            a = 2.0 / n * np.sin(t * i).sum()
            lam = 6.0 / h**2 * (1.0 - np.cos(t)) / (2.0 + np.cos(t))
            y = dt * (lam + c)
            g = (1.0 - y / 2 + y * y / 12) / (1.0 + y / 2 + y * y / 12)
            ex[1:n] += a * g**j * np.sin(t * i)
        assert np.abs(q[j] - ex).max() <= 1e-12 * max(1.0, np.abs(ex).max()), j


@pytest.mark.parametrize("bc,cmat", [(O.BC_DD, O.C_ROWSCALED), (O.BC_DN, O.C_CONSISTENT),
                                     (O.BC_DN, O.C_ROWSCALED)])
def test_fourth_order(bc, cmat):
    x = np.array([0.0, 0.2, 0.5, 0.9, 1.3, 1.7, 2.0])
    w = np.sin(3 * x) + 0.5
    M = len(x)
    A, D = _matrices(x, w, cmat)
    hi = M - 1 if bc == O.BC_DD else M
    sl = slice(1, hi)
    exact = expm(-np.linalg.solve(A[sl, sl], D[sl, sl])) @ np.ones(hi - 1)
    errs = [np.abs(O.march_irk4(x, w, N, bc=bc, cmat=cmat)[N][1:hi] - exact).max()
            for N in (16, 32, 64)]
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 15.0 < r1 < 17.5 and 15.0 < r2 < 17.0, errs


def test_dense_stage_system_one_step():
    x = np.array([0.0, 0.3, 0.7, 1.0, 1.6, 2.0])
    w = np.array([0.0, 2.0, -1.0, 0.5, 3.0, 0.0])
    M, N = len(x), 8
    dt = 1.0 / N
    A, D = _matrices(x, w, O.C_ROWSCALED)
    sl = slice(1, M - 1)
    A, D = A[sl, sl], D[sl, sl]
    a12, a21 = 0.25 - np.sqrt(3) / 6, 0.25 + np.sqrt(3) / 6
    big = np.block([[A + dt / 4 * D, a12 * dt * D], [a21 * dt * D, A + dt / 4 * D]])
    q0 = np.ones(M - 2)
    k = np.linalg.solve(big, np.concatenate([-D @ q0, -D @ q0]))
    n = M - 2
    q1 = q0 + dt / 2 * (k[:n] + k[n:])
    np.testing.assert_allclose(O.march_irk4(x, w, N)[1][1:-1], q1, rtol=1e-13, atol=1e-14)


def test_dealii_fixture_residual():
    x, w = _read_rows("N=33_for_read.txt")
    pt = O.target_position(x, mirror=True)
    r = {N: O.evaluate(x, w, pt, N, cmat=O.C_CONSISTENT, stepper=O.STEP_IRK4)["rnorm"]
         for N in (1024, 2048, 4096)}
    ie = O.evaluate(x, w, pt, 2048, cmat=O.C_CONSISTENT)["rnorm"]
    assert r[2048] < 3e-9, r        # the reference printed ERROR = 1.42e-9
    assert r[1024] > 1e-8 and r[4096] > r[2048], r
    assert ie > 1e-5, ie


def test_matlab_refinement_field_plausible():
    x, w = _read_rows("solution_yita_1D_N=141.txt")
    w = w.copy()
    w[0] = w[-1] = 0.0
    pt = O.target_position(x, mirror=True)
    for N in (512, 2048):
        irk = O.evaluate(x, w, pt, N, stepper=O.STEP_IRK4)["rnorm"]
        ie = O.evaluate(x, w, pt, N)["rnorm"]
        assert irk < ie and irk < 6e-4, (N, irk, ie)
