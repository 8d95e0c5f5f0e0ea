"""The learned warm start against the method without it: a network trained
(on the device, by the library's kernels) on fields the batched solver
converged cuts the Anderson evaluations of held-out instances, and both
starts converge to the same solution."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _train_fields(S, WS, WL, opts, n_fields, seed0):
    Xs, Ys, left, part = [], [], n_fields, 0
    while left > 0:
        n = min(8192, left)
        cfg = WL.config3(n_inst=n, seed=seed0 + 101 * part)
        out = WS.solve(S, cfg, opts)
        conv = (out["status"] & S.ST_CONVERGED) != 0
        assert conv.all()
        X, Y = WS.dataset(S, cfg, out["w_nodes"], conv)
        Xs.append(X)
        Ys.append(Y)
        left -= n
        part += 1
    return np.concatenate(Xs), np.concatenate(Ys)


def test_warm_start_on_held_out_instances_as_specified():
    """The surrogate evaluation as the workload specifies it: 20 000 fields
    the batched solver converged (configuration-3 distribution), Adam at rate
    1e-3, batch 256, 50 epochs; then 2048 instances from disjoint seeds solved
    from w0 and from the prediction (device plan, outliers back to w0), with
    the evaluation counts of both, each with the ground-state preconditioner
    (from w0 the linear form until ||r|| < 0.5, then the rational one; from
    the predictions the rational form from the first update, history 8 and
    the plain form after 24 evaluations).  Both solves converge every instance to the same
    solution."""
    import scft_amd as S
    from scft_amd import warmstart as WS, workloads as WL
    kw = dict(beta=1.0, tol=1e-6, max_iter=300, step_cap=0.0, transform=S.MIX_LOG,
              precond=0.5, backtrack=1e30, check_every=4, floor=1e-8, field_floor=0.5,
              gs_field_weight=0.25, gs_precond=0.5, gs_shift=3.0)
    opts = S.anderson_opts(history=16, gs_shift_below=0.5, **kw)
    X, Y = _train_fields(S, WS, WL, opts, 20000, 4242)
    ws = WS.WarmStart(S, seed=5)
    hist = ws.fit(X, Y, epochs=50, batch=256, lr=1e-3, seed=6)
    print(f"{len(X)} training fields, loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    assert hist[-1] < 0.2 * hist[0]

    test = WL.config3(n_inst=2048, seed=99)  # seeds disjoint from the training draws
    w0 = [w[1:] for w in test["w0"]]
    b = S.Batch(test["xs"], test["Ns"], [p[1:] for p in test["pts"]], bc=S.BC_DN)
    b.upload(w0)
    b.solve_resident(opts)
    cold = b.download()
    w0f = b.join(w0)
    b.upload(w0f, w0f)
    prep = WS.Prepared(S, test)
    plan = ws.device_plan(b, prep)
    plan.predict()
    wopts = S.anderson_opts(history=8, fallback_factor=10.0, gs_switch=24, **kw)
    b.solve_resident(wopts)
    warm = b.download()
    plan.close()
    b.close()
    ws.close()
    ic, iw = cold["iters"].mean(), warm["iters"].mean()
    print(f"held-out mean evaluations: cold {ic:.2f} (max {cold['iters'].max()}), warm "
          f"{iw:.2f} (max {warm['iters'].max()}); fallbacks "
          f"{int(((warm['status'] & S.ST_FALLBACK) != 0).sum())}", flush=True)
    assert cold["converged"] == 2048 and warm["converged"] == 2048
    assert iw <= 0.6 * ic
    # same solutions: the oracle's evaluator on a sample of both results
    for k in range(0, 2048, 256):
        x, N, pt = test["xs"][k], test["Ns"][k], test["pts"][k]
        wa = np.concatenate([[0.0], cold["w"][k]])
        wb = np.concatenate([[0.0], warm["w"][k]])
        ea = O.evaluate(x, wa, pt, N, bc=O.BC_DN)
        eb = O.evaluate(x, wb, pt, N, bc=O.BC_DN)
        assert ea["rnorm"] <= 1e-6 + 1e-10 and eb["rnorm"] <= 1e-6 + 1e-10
        assert np.abs(ea["phi"] - eb["phi"]).max() <= 3e-6


@pytest.mark.parametrize("cuts", [(), (128, 1024)], ids=["one_network", "three_ranges"])
def test_device_warm_plan_matches_host_mapping(cuts):
    """scft_warm_plan_predict (network forward on resident inputs, then the
    mapping kernel writing the batch's field) against the host path
    (network forward, host spline resampling, node samples): the same field
    to FP32 rounding of the network's outputs, on graded meshes whose first
    unknowns come from the wall samples and on coarse ones without them;
    with one network per mesh-size range, one plan per range writes its
    instances (scft_warm_plan_create_sel) and leaves the others alone."""
    import scft_amd as S
    from scft_amd import warmstart as WS, workloads as WL
    c5, c3 = WL.config5(n_inst=256, seed=31), WL.config3(n_inst=256, seed=32)
    cfg = {k: list(c5[k]) + list(c3[k]) for k in ("xs", "Ns", "pts", "w0")}
    prep = WS.Prepared(S, cfg)
    ws = WS.WarmStart(S, seed=3, cuts=cuts)
    rng = np.random.default_rng(4)
    for net in ws.nets:
        Yfake = rng.standard_normal((64, WS.DIMS[-1])).astype(np.float32) * 3.0 + 1.0
        net.set_normalisation(*WS.normalisation(prep.F, Yfake))
    host = ws.predict_unknowns(prep)
    b = S.Batch(cfg["xs"], cfg["Ns"], [p[1:] for p in cfg["pts"]], bc=S.BC_DN)
    g = ws.group(prep.m)
    assert len(np.unique(g)) == len(cuts) + 1
    plan = ws.device_plan(b, prep)
    assert len(plan.plans) == len(cuts) + 1
    for k, p in enumerate(plan.plans):  # each range's plan alone: its instances only
        b.upload(np.full(b.total, 7.0))
        p.predict()
        part = b.download()["w"]
        for i in range(len(cfg["xs"])):
            if g[i] != k:
                assert np.all(part[i] == 7.0), (k, i)
    b.upload(np.zeros(b.total))
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


def test_saved_network_predicts_the_same(tmp_path):
    """WarmStart.save / load: parameters, normalisation and shape in an .npz
    (no pickled objects); the reloaded network predicts bit for bit the
    same fields, and a file of another shape is refused."""
    import scft_amd as S
    from scft_amd import warmstart as WS, workloads as WL
    cfg = WL.config5(n_inst=64, seed=41)
    prep = WS.Prepared(S, cfg)
    rng = np.random.default_rng(9)
    X = rng.standard_normal((512, WS.DIMS[0])).astype(np.float32)
    Y = rng.standard_normal((512, WS.DIMS[-1])).astype(np.float32)
    a = WS.WarmStart(S, seed=3)
    a.fit(X, Y, epochs=2, batch=128, lr=1e-3, seed=4)
    path = str(tmp_path / "net.npz")
    a.save(path, training_fields=512, epochs=2)
    b = WS.WarmStart(S, seed=99)  # different initial weights
    rec = b.load(path)
    assert rec == {"training_fields": 512, "epochs": 2}
    assert np.array_equal(a.predict_unknowns(prep), b.predict_unknowns(prep))
    z = dict(np.load(path))
    z["dims"] = np.asarray([WS.DIMS[0], 128, WS.DIMS[-1]])
    bad = str(tmp_path / "bad.npz")
    np.savez(bad, **z)
    with pytest.raises(ValueError):
        b.load(bad)
    # one network per mesh-size range: cuts and every network round trip
    Xr = X.copy()
    Xr[:, 2] = np.log(rng.choice([64, 200, 3000], size=len(X))).astype(np.float32)
    c = WS.WarmStart(S, seed=3, cuts=(128, 1024))
    c.fit(Xr, Y, epochs=2, batch=128, lr=1e-3, seed=4)
    c.save(path, training_fields=512)
    d = WS.WarmStart(S, seed=99, cuts=(128, 1024))
    assert d.load(path) == {"training_fields": 512}
    assert np.array_equal(c.predict_unknowns(prep), d.predict_unknowns(prep))
    with pytest.raises(ValueError):
        a.load(path)  # a one-network object refuses a file of three
    for w in (a, b, c, d):
        w.close()
    a.close()
    b.close()
