"""The reference's own outer loop driving the drop-in: the compiled
reference broydn (DEALII_SCFT/src/broydn.c with its NR helpers, built from
the reference's sources into oracle/_ref) calls libscft_hip.so's
simple_FEM_1D_transient exactly as 1D_FEM.c:346-356 calls the PETSc
evaluator -- m = 32, DD, N = 2048, TOLF 1e-8, from the .res field -- and
must reach the field and the evaluation count that the same compiled broydn
reaches on the CPU oracle's evaluator (and the oracle's restatement,
orc_broydn)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from tests.helpers import fixtures

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.ref.available(),
                                 reason="oracle/_ref not built (needs the reference sources)")]


def _reference_broydn(vecfunc, n, x0, tolf):
    L = O.ref.dealii
    L.nrref_broydn_setup(n, tolf)
    L.broydn.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p]
    xr = (C.c_double * (n + 1))(0.0, *x0)
    chk = C.c_int(1)
    L.broydn(xr, n, C.byref(chk), vecfunc)
    return chk.value, np.array(xr[1:])


def test_reference_broydn_drives_the_dropin():
    import scft_amd as S
    m, N, tolf = 32, 2048, 1e-8
    n = m - 1
    x = O.mesh_uniform(m + 1, O.L_REF)
    pt = O.target_position(x, O.TAU, True)
    eta = np.array(fixtures()["res_m32"]["eta"])

    # the compiled reference broydn on the device evaluator (the drop-in)
    S.init(N=N, L=O.L_REF, bc=S.BC_DD)
    before = S.local_iteration()
    chk_dev, w_dev = _reference_broydn(C.cast(S.lib.simple_FEM_1D_transient, C.c_void_p), n,
                                       eta[1:-1], tolf)
    calls_dev = S.local_iteration() - before
    assert S.funcerr() == 0
    S.finalize()

    # the same compiled broydn on the CPU oracle's evaluator
    ctx = O.EvalCtx(x, pt, N, bc=O.BC_DD)
    _, vf, _ = O.global_callbacks(ctx, n, trace_len=0)
    chk_cpu, w_cpu = _reference_broydn(vf, n, eta[1:-1], tolf)
    calls_cpu = O.lib().orc_global_calls()
    ctx.close()
    print(f"reference broydn: drop-in check {chk_dev}, {calls_dev} evaluations; oracle "
          f"evaluator check {chk_cpu}, {calls_cpu} evaluations; max |w_dev - w_cpu| "
          f"{np.abs(w_dev - w_cpu).max():.2e}", flush=True)
    assert chk_dev == 0 and chk_cpu == 0
    assert abs(calls_dev - calls_cpu) <= 2
    assert np.abs(w_dev - w_cpu).max() <= 1e-6 * np.abs(w_cpu).max()
    # and the field solves the oracle's evaluator to the reference's tolerance
    wf = np.concatenate([[0.0], w_dev, [0.0]])
    assert O.evaluate(x, wf, pt, N, bc=O.BC_DD)["rnorm"] < tolf
