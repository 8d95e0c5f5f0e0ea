"""The learned warm start against the method without it: a network trained
(on the device, by the library's kernels) on fields the batched solver
converged cuts the Anderson evaluations of held-out instances, and both
starts converge to the same solution."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_warm_start_reduces_iterations_on_held_out_instances():
    import scft_amd as S
    from scft_amd import warmstart as WS, workloads as WL
    opts = S.anderson_opts(history=10, beta=1.0, tol=1e-6, max_iter=300, step_cap=0.0,
                           transform=S.MIX_LOG, precond=0.5, backtrack=1e30, check_every=4)
    train = WL.config3(n_inst=2048, seed=4242, Ns=(512, 1024))
    out = WS.solve(S, train, opts)
    conv = (out["status"] & S.ST_CONVERGED) != 0
    assert conv.all()
    X, Y = WS.dataset(S, train, out["w_nodes"], conv)
    ws = WS.WarmStart(S, seed=5)
    hist = ws.fit(X, Y, epochs=100, batch=256, lr=1e-3, seed=6)
    print(f"training loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    assert hist[-1] < 0.5 * hist[0]

    test = WL.config3(n_inst=512, seed=99, Ns=(512, 1024))  # disjoint seeds
    cold = WS.solve(S, test, opts)
    wi = ws.predict(test)
    wopts = S.anderson_opts(history=10, beta=1.0, tol=1e-6, max_iter=300, step_cap=0.0,
                            transform=S.MIX_LOG, precond=0.5, backtrack=1e30, check_every=4,
                            fallback_factor=10.0)
    warm = WS.solve(S, test, wopts, w_init=wi, w_cold=test["w0"])
    ic, iw = cold["iters"].mean(), warm["iters"].mean()
    print(f"held-out mean evaluations: cold {ic:.2f}, warm {iw:.2f}; "
          f"fallbacks {int(((warm['status'] & S.ST_FALLBACK) != 0).sum())}")
    assert cold["converged"] == 512 and warm["converged"] == 512
    assert iw <= 0.9 * ic
    # same solutions: the oracle's evaluator on a sample of both results
    for k in range(0, 512, 64):
        x, N, pt = test["xs"][k], test["Ns"][k], test["pts"][k]
        ea = O.evaluate(x, cold["w_nodes"][k], pt, N, bc=O.BC_DN)
        eb = O.evaluate(x, warm["w_nodes"][k], pt, N, bc=O.BC_DN)
        assert ea["rnorm"] <= 1e-6 + 1e-10 and eb["rnorm"] <= 1e-6 + 1e-10
        assert np.abs(ea["phi"] - eb["phi"]).max() <= 3e-6
    ws.close()
