"""Device parity of the two-stage Gauss-Legendre (IRK4) march against the oracle.

The device solves each step as one complex tridiagonal system (the stage
system in the eigenbasis of the Butcher matrix); the oracle solves the
reference's real 2n x 2n block system as written (orc_march_irk4), so these
tests also check the decoupling.  Bars (FP64):
  * propagator, per stored step j:  ||q_dev,j - q_cpu,j||_inf <= 1e-10 ||q_cpu,j||_inf
  * density and residual:           |phi_dev - phi_cpu| <= 1e-10
plus the reference's deal.II fixture (a field its IRK solver converged to
||r|| = 1.42e-9) through the batch and through the drop-in callback.
"""
import os
import re

import numpy as np
import pytest

import oracle as O
from tests.helpers import instance, unknowns, w0_field

pytestmark = pytest.mark.gpu

STEP_RTOL = 1e-10
PHI_ATOL = 1e-10
FILES = os.path.join(os.path.dirname(__file__), "golden", "files")


def _scft():
    import scft_amd
    return scft_amd


def _cases():
    rng = np.random.default_rng(17)
    c = [("m32_DD", 32, 2048, O.L_REF, O.BC_DD, O.C_ROWSCALED, None, None),
         ("m32_DN", 32, 256, 3.72, O.BC_DN, O.C_ROWSCALED, None, None),
         ("m1024_DN", 1024, 512, 3.72, O.BC_DN, O.C_ROWSCALED, None, None)]
    for m in (2, 3, 63, 64, 65, 129, 255, 257, 511, 513, 1023, 1025):
        bc = O.BC_DD if m % 2 else O.BC_DN
        c.append((f"edge_m{m}", m, 128, 2.0 + m % 7, bc, O.C_ROWSCALED, None, None))
    c.append(("max_m2048_DN", 2048, 256, 8.0, O.BC_DN, O.C_ROWSCALED, None, None))
    c.append(("max_m2049_DD", 2049, 128, 8.0, O.BC_DD, O.C_ROWSCALED, None, None))
    # the 4096-unknown bucket (the spilling (8, 8) variant)
    c.append(("m2049_DN", 2049, 64, 8.0, O.BC_DN, O.C_ROWSCALED, None, None))
    c.append(("m3001_DD", 3001, 128, 8.0, O.BC_DD, O.C_ROWSCALED, None, None))
    c.append(("max_m4096_DN", 4096, 64, 8.0, O.BC_DN, O.C_ROWSCALED, None, None))
    c.append(("max_m4097_DD", 4097, 64, 8.0, O.BC_DD, O.C_CONSISTENT, None, None))
    c.append(("consistent_m100_DD", 100, 256, 3.0, O.BC_DD, O.C_CONSISTENT, None, None))
    c.append(("consistent_m700_DN", 700, 512, 4.0, O.BC_DN, O.C_CONSISTENT, None, None))
    c.append(("graded_DN", None, 1024, 5.0, O.BC_DN, O.C_ROWSCALED, (1e-3, 0.2), None))
    c.append(("rough_w_m300", 300, 1024, 3.5, O.BC_DN, O.C_ROWSCALED, None,
              rng.uniform(-5.0, 20.0, 301)))
    c.append(("steps_4096", 200, 4096, 3.0, O.BC_DN, O.C_ROWSCALED, None, None))
    c.append(("steps_16", 40, 16, 3.0, O.BC_DD, O.C_ROWSCALED, None, None))
    return c


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_irk4_march_matches_oracle(case):
    S = _scft()
    name, m, N, L, bc, cmat, graded, w = case
    if graded is not None:
        h0, ell = graded
        m = O.graded_elements(L, h0, ell, 63, 2047)
        inst = instance(m, L, N, bc=bc, graded=ell, seed=11)
    else:
        inst = instance(m, L, N, bc=bc, seed=3, w=w)
    x, pt = inst["x"], inst["pt"]
    wf = inst["w"].copy()
    wf[0] = 0.0
    if bc == O.BC_DD:
        wf[-1] = 0.0
    print(f"case {name}: M={len(x)} N={N} bc={bc} cmat={cmat}", flush=True)
    ref = O.evaluate(x, wf, pt, N, bc=bc, cmat=cmat, want_q=True, stepper=O.STEP_IRK4)
    assert ref["rc"] == 0
    b = S.Batch([x], [N], [unknowns(pt, bc)], bc=bc, cmat=cmat, stepper=S.STEP_IRK4)
    out = b.eval([unknowns(wf, bc)])
    assert out["status"][0] == 0, out["status"]
    slabs = b.debug_slabs(0)
    qc = ref["q"][1:N // 2]
    qc = qc[:, 1:-1] if bc == O.BC_DD else qc[:, 1:]
    err = np.abs(slabs - qc).max(axis=1) / np.abs(qc).max(axis=1)
    assert err.max() <= STEP_RTOL, f"step {err.argmax() + 1}: {err.max():.3e}"
    np.testing.assert_allclose(out["phi"][0], unknowns(ref["phi"], bc), rtol=0, atol=PHI_ATOL)
    np.testing.assert_allclose(out["r"][0], unknowns(ref["r"], bc), rtol=0, atol=PHI_ATOL)
    assert abs(out["rnorm"][0] - ref["rnorm"]) <= PHI_ATOL
    b.close()


def test_irk4_mixed_batch_matches_oracle():
    """Every IRK4 bucket (64 .. 4096 unknowns) populated in one sweep."""
    S = _scft()
    rng = np.random.default_rng(4242)
    ms = [32, 64, 65, 127, 128, 200, 256, 300, 511, 512, 700, 1023, 1024, 1500, 2047, 2048,
          2049, 3500, 4096]
    Ns = [int(rng.choice([64, 128, 256])) for _ in ms]
    Ls = rng.uniform(2.0, 8.0, len(ms))
    insts = [instance(int(m), float(L), N, bc=O.BC_DN, seed=300 + i)
             for i, (m, L, N) in enumerate(zip(ms, Ls, Ns))]
    for it in insts:
        it["w"][0] = 0.0
    b = S.Batch([it["x"] for it in insts], Ns, [it["pt"][1:] for it in insts], bc=O.BC_DN,
                stepper=S.STEP_IRK4)
    out = b.eval([it["w"][1:] for it in insts])
    assert (out["status"] == 0).all()
    for k, it in enumerate(insts):
        ref = O.evaluate(it["x"], it["w"], it["pt"], Ns[k], bc=O.BC_DN, stepper=O.STEP_IRK4)
        np.testing.assert_allclose(out["phi"][k], ref["phi"][1:], rtol=0, atol=PHI_ATOL,
                                   err_msg=f"instance {k} m={ms[k]} N={Ns[k]}")
    out2 = b.eval([it["w"][1:] for it in insts])
    for k in range(len(insts)):
        assert np.array_equal(out["phi"][k], out2["phi"][k])
    b.close()


def _dealii_field():
    rows = [l for l in open(os.path.join(FILES, "N=33_for_read.txt"))
            if re.match(r"^\s*\d+\s*,", l)]
    d = np.array([[float(v) for v in r.split(",")] for r in rows])
    return d[:, 1], d[:, 2]


def test_irk4_dealii_fixture_batch_and_dropin():
    """DEALII_SCFT/inputFiles/N=33_for_read.txt: consistent w-mass, wall values
    from the file, N = 2048 (drivescft.cc:130).  The reference printed
    ||r|| = 1.42e-9 for this field; the oracle gives 1.7e-9."""
    S = _scft()
    x, w = _dealii_field()
    pt = O.target_position(x, mirror=True)
    ref = O.evaluate(x, w, pt, 2048, bc=O.BC_DD, cmat=O.C_CONSISTENT, stepper=O.STEP_IRK4)
    assert ref["rnorm"] < 3e-9
    b = S.Batch([x], [2048], [pt[1:-1]], bc=S.BC_DD, cmat=S.C_CONSISTENT, stepper=S.STEP_IRK4)
    b.set_walls([w[0]], [w[-1]])
    out = b.eval([w[1:-1]])
    np.testing.assert_allclose(out["r"][0], ref["r"][1:-1], rtol=0, atol=1e-12)
    assert out["rnorm"][0] < 3e-9
    b.close()
    # the deal.II callback form (drivescft.cc:218-243: out = phi_t - phi), with
    # the walls extrapolated by the natural spline as scft.cc:454-480 does
    S.init(N=2048, L=O.L_REF, bc=S.BC_DD, cmat=S.C_CONSISTENT, stepper=S.STEP_IRK4)
    n = len(x) - 2
    o = np.zeros(n)
    S.lib.scft_adm_chen_f(n, np.ascontiguousarray(w[1:-1]), o)
    np.testing.assert_allclose(o, -ref["r"][1:-1], rtol=0, atol=1e-12)
    assert np.abs(o).max() < 3e-9
    S.finalize()


def test_irk4_rejects_more_than_4096_unknowns():
    S = _scft()
    x = O.mesh_uniform(4098, 4.0)
    with pytest.raises(S.ScftError):
        S.Batch([x], [64], [np.ones(4097)], bc=S.BC_DN, stepper=S.STEP_IRK4)
    b = S.Batch([x], [64], [np.ones(4096)], bc=S.BC_DD, stepper=S.STEP_IRK4)  # n = 4096
    b.close()


def test_irk4_batched_solve_converges():
    """Preconditioned Anderson over IRK4 evaluations; the returned fields are
    re-evaluated by the oracle's IRK4 march.  Meshes and step counts are in the
    regime the reference runs IRK4 in (33 nodes, N = 2048): the Gauss method is
    not L-stable (its factor tends to 1 for stiff modes), so dt * 12/h^2 stays
    below ~25 here; far beyond that the wall layer oscillates and no mixer
    converges (implicit Euler is the stepper for fine meshes)."""
    S = _scft()
    rng = np.random.default_rng(99)
    ms = [32, 48, 64, 96, 128]
    Ls = rng.uniform(2.0, 6.0, len(ms))
    Ns = [1024, 2048, 1024, 2048, 2048]
    xs = [O.mesh_uniform(m + 1, L) for m, L in zip(ms, Ls)]
    pts = [O.target_position(x, O.TAU, False) for x in xs]
    w0 = [w0_field(x, 700 + k) for k, x in enumerate(xs)]
    b = S.Batch(xs, Ns, [p[1:] for p in pts], bc=S.BC_DN, stepper=S.STEP_IRK4)
    opts = S.anderson_opts(history=10, beta=1.0, tol=1e-6, max_iter=300, step_cap=0.0,
                           transform=S.MIX_LOG, precond=0.5, backtrack=1e30, check_every=4)
    out = b.solve([v[1:] for v in w0], opts)
    assert (out["status"] & S.ST_CONVERGED).all(), out["status"]
    for k in range(len(ms)):
        wf = np.concatenate([[0.0], out["w"][k]])
        r = O.evaluate(xs[k], wf, pts[k], Ns[k], bc=O.BC_DN, stepper=O.STEP_IRK4)["rnorm"]
        assert r <= 1.01e-6, (k, r)
    b.close()
