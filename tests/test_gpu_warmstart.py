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


def test_device_warm_plan_matches_host_mapping():
    """scft_warm_plan_predict (network forward on resident inputs, then the
    mapping kernel writing the batch's field) against the host path
    (network forward, host spline resampling, node samples): the same field
    to FP32 rounding of the network's outputs, on graded meshes whose first
    unknowns come from the wall samples and on coarse ones without them."""
    import scft_amd as S
    from scft_amd import warmstart as WS, workloads as WL
    c5, c3 = WL.config5(n_inst=256, seed=31), WL.config3(n_inst=256, seed=32)
    cfg = {k: list(c5[k]) + list(c3[k]) for k in ("xs", "Ns", "pts", "w0")}
    prep = WS.Prepared(S, cfg)
    ws = WS.WarmStart(S, seed=3)
    rng = np.random.default_rng(4)
    Yfake = rng.standard_normal((64, WS.DIMS[-1])).astype(np.float32) * 3.0 + 1.0
    ws.mlp.set_normalisation(*WS.normalisation(prep.F, Yfake))
    host = ws.predict_unknowns(prep)
    b = S.Batch(cfg["xs"], cfg["Ns"], [p[1:] for p in cfg["pts"]], bc=S.BC_DN)
    b.upload(np.zeros(b.total))
    plan = ws.device_plan(b, prep)
    plan.predict()
    dev = np.concatenate(b.download()["w"])
    plan.close()
    b.close()
    ws.close()
    assert any(k > 0 for _, _, k in prep.spikes) and len(prep.spikes) < len(cfg["xs"])
    err = np.abs(dev - host) / np.maximum(1.0, np.abs(host))
    print(f"device vs host prediction: max relative difference {err.max():.2e} over "
          f"{len(host)} unknowns", flush=True)
    assert err.max() <= 2e-6
